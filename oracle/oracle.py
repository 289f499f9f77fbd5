"""ORACLE -- test infrastructure only (imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, never by the product).

Python front end of the CPU restatement in oracle/alt3_oracle.c plus the harmonic part
(dynamical matrix, Hermitian eigensolve, finite-difference velocities) restated with numpy:

  eigen()   get_frequencies (rest.f90:2230-2405) -> finitedif_vel (:2587-2631)
            -> get_freq (:4319-4445) -> set_dynamical_matrix (:445-556), zheev('V','U')
            (zhegv.f:213, here numpy.linalg.eigh(UPLO='U')), sort (extratools.f90:154-175);
            subst_eivecs (rest.f90:3262-3293).
            Parity: eigenvalues/velocities against the reference dumps; eigenvectors
            through gauge-invariant projectors (degenerate subspaces are not unique).

All arrays use the reference's column-major layouts (numpy order='F') with 0-based indices.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "liboracle.so")
_lib = None


def build(force: bool = False) -> str:
    src = os.path.join(_HERE, "alt3_oracle.c")
    if force or not os.path.exists(LIB_PATH) or os.path.getmtime(LIB_PATH) < os.path.getmtime(src):
        subprocess.check_call(["gcc", "-O2", "-fopenmp", "-ffp-contract=off", "-shared", "-fPIC",
                               src, "-lm", "-o", LIB_PATH])
    return LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        _lib = C.CDLL(LIB_PATH)
        _lib.or_const33.restype = C.c_double
        _lib.or_cnst.restype = C.c_double
        _lib.or_nbe.restype = C.c_double
        _lib.or_nbe.argtypes = [C.c_double, C.c_double, C.c_int]
        _lib.or_temp_cm.restype = C.c_double
        _lib.or_temp_cm.argtypes = [C.c_double]
        _lib.or_delta_g.restype = C.c_double
        _lib.or_delta_g.argtypes = [C.c_double, C.c_double]
    return _lib


def _p(a):
    return a.ctypes.data_as(C.c_void_p)


def _f64(a):
    return np.require(a, dtype=np.float64, requirements=["F_CONTIGUOUS", "ALIGNED"])


def _i32(a):
    return np.require(a, dtype=np.int32, requirements=["F_CONTIGUOUS", "ALIGNED"])


CNST = float(np.float32(521.1098918))   # modules.f90:18, a REAL*4 literal
DQ_FD = float(np.float32(0.001))        # rest.f90:2600, a REAL*4 literal


def const33() -> float:
    return lib().or_const33()


# ---------------------------------------------------------------------------- indexing
def reciprocal(r: np.ndarray):
    g = np.zeros((3, 3)); vol = C.c_double()
    # r: rows r1, r2, r3 (C order, as lat_fc.dat lists them)
    lib().or_reciprocal_2pi(_p(np.ascontiguousarray(r, dtype=np.float64)), _p(g), C.byref(vol))
    return g, vol.value


def lattice_ops(latparams, primlatt):
    opk = np.zeros((48, 3, 3))   # [op][col][row] == Fortran op_kmatrix(row,col,op)
    n = lib().or_lattice_ops(_p(_f64(np.asarray(latparams, float))), _p(_f64(primlatt)), _p(opk))
    if n <= 0:
        raise RuntimeError(f"lattice_ops failed {n}")
    return opk[:n]


def make_kp_reg(nc, g):
    nk = int(np.prod(nc))
    kp = np.zeros((nk, 3))
    lib().or_make_kp_reg(_p(_i32(np.asarray(nc))), _p(np.ascontiguousarray(g)), _p(kp))
    return kp


def make_kp_reg_tet(nc, g):
    nk = int(np.prod(nc))
    kp = np.zeros((nk, 3)); tet = np.zeros((6 * nk, 4), dtype=np.int32)
    rc = lib().or_make_kp_reg_tet(_p(_i32(np.asarray(nc))), _p(np.ascontiguousarray(g)), _p(kp), _p(tet))
    if rc != 0:
        raise RuntimeError("make_kp_reg_tet: corner is not a mesh point (odd mesh?)")
    return kp, tet


def get_k_info(q, nc, g):
    nk = C.c_int(); ijk = np.zeros(3, dtype=np.int32)
    ins = lib().or_get_k_info(_p(np.ascontiguousarray(q, dtype=np.float64)), _p(_i32(np.asarray(nc))),
                              _p(np.ascontiguousarray(g)), C.byref(nk), _p(ijk))
    return nk.value, ijk, ins


def get_negatives(kp, nc, g):
    nk = len(kp); m = np.zeros(nk, dtype=np.int32)
    npos = lib().or_get_negatives(nk, _p(np.ascontiguousarray(kp)), _p(_i32(np.asarray(nc))),
                                  _p(np.ascontiguousarray(g)), _p(m))
    return npos, m


def get_weights(kp, primlatt, opk):
    nk = len(kp)
    kibz = np.zeros((nk, 3)); wibz = np.zeros(nk); mapibz = np.zeros(nk, np.int32); mapinv = np.zeros(nk, np.int32)
    n = lib().or_get_weights(nk, _p(np.ascontiguousarray(kp)), _p(_f64(primlatt)), len(opk),
                             _p(np.ascontiguousarray(opk)), _p(kibz), _p(wibz), _p(mapibz), _p(mapinv))
    return n, kibz[:n].copy(), wibz[:n].copy(), mapibz, mapinv[:n].copy()


def kstars(q, kp, primlatt, opk):
    st = np.zeros((48, 3)); ops = np.zeros(48, np.int32); n0 = np.zeros(48, np.int32)
    n = lib().or_kstarsinfbz1(_p(np.ascontiguousarray(q, dtype=np.float64)), len(kp), _p(np.ascontiguousarray(kp)),
                              _p(_f64(primlatt)), len(opk), _p(np.ascontiguousarray(opk)), _p(st), _p(ops), _p(n0))
    return st[:n].copy(), ops[:n].copy(), n0[:n].copy()


# ---------------------------------------------------------------------------- eigen
def dynmat_batch(qs, masses, atompos, iatomcell0, at2, xyz2, fc2):
    """set_dynamical_matrix (rest.f90:445-556) for many q at once; returns (nq,nd,nd)
    indexed [q, row, col]."""
    qs = np.atleast_2d(qs)
    nat0 = len(masses); nd = 3 * nat0
    i0 = at2[:, 0]; j = at2[:, 1]; al = xyz2[:, 0]; be = xyz2[:, 1]
    home = i0 <= nat0
    i0, j, al, be, v = i0[home], j[home], al[home], be[home], fc2[home]
    j0 = iatomcell0[j - 1]
    rr = atompos[j - 1] - atompos[j0 - 1]
    amp = v / np.sqrt(masses[i0 - 1] * masses[j0 - 1])
    ph = np.exp(1j * (qs @ rr.T))                        # (nq, nt)
    D = np.zeros((len(qs), nd, nd), dtype=np.complex128)
    rows = al - 1 + 3 * (i0 - 1); cols = be - 1 + 3 * (j0 - 1)
    for t in range(len(v)):
        D[:, rows[t], cols[t]] += amp[t] * ph[:, t]
    allm = np.abs(D).sum(axis=(1, 2)) / (nd * nd)
    for t in range(nd):
        dt = D[:, t, t]
        big = np.abs(dt.imag) > 9e-4 * np.abs(dt.real)
        D[big, t, t] = dt[big].real
        for jj in range(t + 1, nd - 1):
            a = D[:, t, jj].copy(); b = D[:, jj, t].copy()
            mi = (a.imag - b.imag) / 2; mj = (a.real + b.real) / 2
            D[:, t, jj] = mj + 1j * mi
            D[:, jj, t] = mj - 1j * mi
    si = np.abs(D.imag).sum(axis=(1, 2)) / (nd * nd); sr = np.abs(D.real).sum(axis=(1, 2)) / (nd * nd)
    real = si < sr * 1e-10
    D[real] = (D[real] + np.conj(D[real])) / 2
    return D


def _ref_sort(r):
    """sort (extratools.f90:154-175): selection swaps on an index map."""
    n = len(r); m = list(range(n))
    for i in range(n):
        for j in range(i + 1, n):
            if r[m[i]] > r[m[j]]:
                m[i], m[j] = m[j], m[i]
    return m


def born_params(inp):
    """What nonanal (rest.f90:3517-3578) uses: epsil, the charges, the masses and the constants
    it derives from the lattice (rho from g1,g2,g3 with the REAL*4 exponent 0.33, the cell
    volume of r1,r2,r3, eps_scale); None when born_flag is 0."""
    if not getattr(inp, "born_flag", 0):
        return None
    g, _ = reciprocal(inp.r123)
    gg = (np.linalg.norm(g[0]) * np.linalg.norm(g[1]) * np.linalg.norm(g[2])) ** float(np.float32(0.33))
    rho = gg / 4.0
    r1, r2, r3 = inp.r123
    om0 = abs(np.dot(r3, np.cross(r1, r2)))
    eps_scale = 8.8541878176e-12 / 1e10 / 1.60217653e-19
    return dict(epsil=np.asarray(inp.epsil, float), zeu=np.asarray(inp.zeu, float), masses=np.asarray(inp.masses, float),
                rho2=rho * rho, om0=om0, eps_scale=eps_scale)


def nonanal_batch(qs, D, b):
    """nonanal (rest.f90:3517-3578) on a batch of dynamical matrices D[q, row, col], after
    get_freq's shift of q = 0 to q + 1e-10 (rest.f90:4342-4346): D += zag(i) zbg(j) /
    (qeq eps_scale) / om0 / sqrt(ma mb) exp(-q^2/rho^2), zag(i) = sum_k q_k charge_a(k,i)."""
    qs = np.array(np.atleast_2d(qs), dtype=float)
    zero = np.all(qs == 0.0, axis=1)
    qs[zero] += 1.0e-10
    E = b["epsil"]
    qeq = (qs[:, 0] * (E[0, 0] * qs[:, 0] + E[0, 1] * qs[:, 1] + E[0, 2] * qs[:, 2])
           + qs[:, 1] * (E[1, 0] * qs[:, 0] + E[1, 1] * qs[:, 1] + E[1, 2] * qs[:, 2])
           + qs[:, 2] * (E[2, 0] * qs[:, 0] + E[2, 1] * qs[:, 1] + E[2, 2] * qs[:, 2]))
    q2 = qs[:, 0] * qs[:, 0] + qs[:, 1] * qs[:, 1] + qs[:, 2] * qs[:, 2]
    nat0 = len(b["masses"])
    zg = np.einsum("qk,aki->qai", qs, b["zeu"])          # (nq, atom, i)
    for na in range(nat0):
        for nb_ in range(nat0):
            ms = np.sqrt(b["masses"][na] * b["masses"][nb_])
            for i in range(3):
                for j in range(3):
                    t = zg[:, na, i] * zg[:, nb_, j] / (qeq * b["eps_scale"]) / b["om0"] / ms * np.exp(-q2 / b["rho2"])
                    D[:, i + 3 * na, j + 3 * nb_] += t
    return D


def get_freq_batch(qs, model, born=None):
    """get_freq (rest.f90:4319-4445): |eigenvalues| ascending and eigenvectors (columns);
    born (born_params) adds nonanal's term."""
    D = dynmat_batch(qs, *model)
    if born is not None:
        D = nonanal_batch(qs, D, born)
    w, v = np.linalg.eigh(D, UPLO="U")
    w = np.abs(w)  # eigh returns ascending; the reference sort is then the identity map
    re = np.abs(v.real).sum(axis=1)               # (nq, nb) per column
    v = np.where((re < 1e-3)[:, None, :], 1j * v, v)
    return w, v


def eigen(kp_mesh, model, born=None):
    """get_frequencies with finite-difference velocities.  Returns eigenval (nb,nk),
    eigenvec (nb,nb,nk) [component, branch, k], veloc (3,nb,nk), all order='F'."""
    kp_mesh = np.asarray(kp_mesh)
    nk = len(kp_mesh)
    w0, v0 = get_freq_batch(kp_mesh, model, born)
    nb = w0.shape[1]
    vg = np.zeros((nk, nb, 3))
    for a in range(3):
        qp = kp_mesh.copy(); qp[:, a] += DQ_FD
        qm = kp_mesh.copy(); qm[:, a] -= DQ_FD
        wp, _ = get_freq_batch(qp, model, born)
        wm, _ = get_freq_batch(qm, model, born)
        vg[:, :, a] = (wp - wm) / 2 / DQ_FD / 2 / np.sqrt(w0) * CNST * 1e-8 * 2 * np.pi
    eival = np.zeros((nb, nk), order="F"); eivec = np.zeros((nb, nb, nk), dtype=np.complex128, order="F")
    for i in range(nk):
        mp = _ref_sort(w0[i])
        eival[:, i] = w0[i][mp]
        eivec[:, :, i] = v0[i][:, mp]
    veloc = np.asfortranarray(vg.transpose(2, 1, 0))
    return eival, eivec, veloc


def subst_eivecs(eivec, kp_mesh, nc, g):
    """subst_eivecs (rest.f90:3262-3293): e(k_pos) <- conj(e(k_neg))."""
    npos, mp = get_negatives(kp_mesh, nc, g)
    nk = len(kp_mesh)
    for n in range(npos, nk):
        n3 = mp[n]
        mk3, _, _ = get_k_info(-kp_mesh[n3], nc, g)
        eivec[:, :, mk3] = np.conj(eivec[:, :, n3])
    return eivec


# ---------------------------------------------------------------------------- hot path
def v3sq(ibz_lo, ibz_hi, mapinv, kp, nc, g, eigenval, eigenvec, masses, atompos, iatomcell0, at3, xyz3, fc3):
    nk = len(kp); nb = eigenval.shape[0]; ksub = ibz_hi - ibz_lo + 1
    out = np.zeros((ksub, nk, nb, nb, nb), order="F")
    ev = np.asfortranarray(eigenvec, dtype=np.complex128)
    rc = lib().or_v3sq(ibz_lo, ibz_hi, _p(_i32(mapinv)), nk, _p(np.ascontiguousarray(kp)), _p(_i32(np.asarray(nc))),
                       _p(np.ascontiguousarray(g)), nb, _p(_f64(eigenval)), _p(ev), _p(_f64(masses)),
                       _p(np.ascontiguousarray(atompos)), _p(_i32(iatomcell0)), len(fc3),
                       _p(np.ascontiguousarray(at3, dtype=np.int32)), _p(np.ascontiguousarray(xyz3, dtype=np.int32)),
                       _p(_f64(fc3)), _p(out))
    if rc:
        raise RuntimeError(f"or_v3sq error {rc}")
    return out


def freq_from_eigenval(eigenval):
    return np.asfortranarray(np.sqrt(eigenval.T) * CNST)


def fgr(ibz_lo, ibz_hi, mapinv, kp, nc, g, freq, tet, v33):
    nk = len(kp); nb = freq.shape[1]; ksub = ibz_hi - ibz_lo + 1
    P1 = np.zeros((ksub, nk, nb, nb, nb), order="F"); P2 = np.zeros_like(P1)
    lib().or_fgr(ibz_lo, ibz_hi, _p(_i32(mapinv)), nk, _p(np.ascontiguousarray(kp)), _p(_i32(np.asarray(nc))),
                 _p(np.ascontiguousarray(g)), nb, _p(_f64(freq)), _p(np.ascontiguousarray(tet, dtype=np.int32)),
                 _p(_f64(v33)), _p(P1), _p(P2))
    return P1, P2


def fgr_gauss(ibz_lo, ibz_hi, mapinv, kp, nc, g, freq, sigma, v33):
    nk = len(kp); nb = freq.shape[1]; ksub = ibz_hi - ibz_lo + 1
    P1 = np.zeros((ksub, nk, nb, nb, nb), order="F"); P2 = np.zeros_like(P1)
    lib().or_fgr_gauss(ibz_lo, ibz_hi, _p(_i32(mapinv)), nk, _p(np.ascontiguousarray(kp)), _p(_i32(np.asarray(nc))),
                       _p(np.ascontiguousarray(g)), nb, _p(_f64(freq)), C.c_double(sigma), _p(_f64(v33)), _p(P1), _p(P2))
    return P1, P2


def delta_g(x, w):
    return lib().or_delta_g(x, w)


def weight_tet(w4, e):
    c = np.zeros(4)
    lib().or_weight_tet1(_p(np.ascontiguousarray(w4, dtype=np.float64)), C.c_double(e), _p(c))
    return c


def distribution(eigenval, tempk, classical=0):
    nb, nk = eigenval.shape
    freq = np.zeros((nk, nb), order="F"); dist = np.zeros((nk, nb), order="F")
    lib().or_distribution(nk, nb, _p(_f64(eigenval)), C.c_double(tempk), classical, _p(freq), _p(dist))
    return freq, dist


def tau(ibz_lo, ibz_hi, mapinv, kp, nc, g, dist, P1sm, P2sm, want_p=True):
    nk, nb = dist.shape; ksub = ibz_hi - ibz_lo + 1
    P1 = np.zeros_like(P1sm) if want_p else None
    P2 = np.zeros_like(P2sm) if want_p else None
    Q = np.zeros((ksub, nb), order="F"); QN = np.zeros_like(Q); QU = np.zeros_like(Q)
    lib().or_tau(ibz_lo, ibz_hi, _p(_i32(mapinv)), nk, _p(np.ascontiguousarray(kp)), _p(_i32(np.asarray(nc))),
                 _p(np.ascontiguousarray(g)), nb, _p(_f64(dist)), _p(_f64(P1sm)), _p(_f64(P2sm)),
                 _p(P1) if want_p else None, _p(P2) if want_p else None, _p(Q), _p(QN), _p(QU))
    return P1, P2, Q, QN, QU


def qandtau(mapinv, dist, QN, QU):
    nibz, nb = QN.shape; nk = dist.shape[0]
    tN = np.zeros((nibz, nb), order="F"); tU = np.zeros_like(tN); tt = np.zeros_like(tN)
    lib().or_qandtau(nibz, nb, nk, _p(_i32(mapinv)), _p(_f64(dist)), _p(_f64(QN)), _p(_f64(QU)), _p(tN), _p(tU), _p(tt))
    return tN, tU, tt


def frta(mapinv, freq, dist, veloc, Q, tempk):
    nibz, nb = Q.shape; nk = freq.shape[0]
    F = np.zeros((nibz, nb, 3), order="F")
    lib().or_frta(nibz, nb, nk, _p(_i32(mapinv)), _p(_f64(freq)), _p(_f64(dist)), _p(_f64(veloc)), _p(_f64(Q)),
                  C.c_double(tempk), _p(F))
    return F


def mode_kappa(kibz, mapinv, kp, nc, g, primlatt, opk, eigenval, veloc, FFF, tempk, volume_r, classical=0):
    nibz = len(kibz); nb, nk = eigenval.shape
    kap = np.zeros((nb, 3, 3), order="F"); te = np.zeros((nibz, nb), order="F")
    lib().or_mode_kappa(nibz, nb, nk, _p(np.ascontiguousarray(kibz)), _p(_i32(mapinv)), _p(np.ascontiguousarray(kp)),
                        _p(_i32(np.asarray(nc))), _p(np.ascontiguousarray(g)), _p(_f64(primlatt)), len(opk),
                        _p(np.ascontiguousarray(opk)), _p(_f64(eigenval)), _p(_f64(veloc)), _p(_f64(FFF)),
                        C.c_double(tempk), classical, C.c_double(volume_r), _p(kap), _p(te))
    return kap, te


def collision(ibz_lo, ibz_hi, kibz, mapinv, kp, nc, g, primlatt, opk, P1, P2, Piso=None):
    """CollisionMatrix_mpi, or CollisionMatrix_iso when Piso (nibz, nb) is given."""
    nibz = len(kibz); nk = len(kp); nb = P1.shape[2]; ksub = ibz_hi - ibz_lo + 1
    CM = np.zeros((ksub * nb * 3, nibz * nb * 3), order="F")
    pi = None if Piso is None else _f64(Piso)
    lib().or_collision(ibz_lo, ibz_hi, nibz, _p(np.ascontiguousarray(kibz)), _p(_i32(mapinv)), nk,
                       _p(np.ascontiguousarray(kp)), _p(_i32(np.asarray(nc))), _p(np.ascontiguousarray(g)),
                       _p(_f64(primlatt)), len(opk), _p(np.ascontiguousarray(opk)), nb, _p(_f64(P1)), _p(_f64(P2)),
                       None if pi is None else _p(pi), _p(CM))
    return CM


def piso(ibz_lo, ibz_hi, mapinv, freq, tet, eigenvec, giso):
    """P_iso_split: Piso (ksub, nb) from the mesh eigenvectors (nb, nb, nk), the tetrahedra
    (4, 6 nk, 0-based) and the per-atom g factors."""
    nk, nb = freq.shape; ksub = ibz_hi - ibz_lo + 1
    out = np.zeros((ksub, nb), order="F")
    ev = np.require(eigenvec, dtype=np.complex128, requirements=["F_CONTIGUOUS", "ALIGNED"])
    gi = np.ascontiguousarray(giso, dtype=np.float64)
    lib().or_piso(ibz_lo, ibz_hi, _p(_i32(mapinv)), nk, nb, len(gi), _p(_f64(freq)),
                  _p(np.ascontiguousarray(tet, dtype=np.int32)), _p(ev), _p(gi), _p(out))
    return out


def update(ibz_lo, ibz_hi, kibz, mapibz, kp, nc, g, primlatt, opk, P1, P2, FFF, FRTA, Q, Piso=None):
    """update_mpi: F2 (ksub, nb, 3) = F_RTA + A(FFF)/Q for the irreducible points ibz_lo..hi;
    with Piso (nibz, nb) update_iso."""
    nibz = len(kibz); nk = len(kp); nb = P1.shape[2]; ksub = ibz_hi - ibz_lo + 1
    F2 = np.zeros((ksub, nb, 3), order="F")
    pi = None if Piso is None else _f64(Piso)
    rc = lib().or_update(ibz_lo, ibz_hi, nibz, _p(np.ascontiguousarray(kibz)), _p(_i32(mapibz)), nk,
                         _p(np.ascontiguousarray(kp)), _p(_i32(np.asarray(nc))), _p(np.ascontiguousarray(g)),
                         _p(_f64(primlatt)), len(opk), _p(np.ascontiguousarray(opk)), nb, _p(_f64(P1)), _p(_f64(P2)),
                         _p(_f64(FFF)), _p(_f64(FRTA)), _p(_f64(Q)), None if pi is None else _p(pi), _p(F2))
    if rc != 0:
        raise ValueError("update: a mesh point q1 +- q2 is not a member of its star")
    return F2


def bte_iter(kibz, mapinv, mapibz, kp, nc, g, primlatt, opk, eigenval, veloc, P1, P2, FRTA, Q, tempk, volume_r,
             max_iter=150, update_mix=float(np.float32(0.4)), tol=float(np.float32(1e-4)), classical=0, Piso=None):
    """The iterative BTE of job 3 (kappa.f90:1373-1487) on the whole irreducible zone:
    F <- F + update_mix (update(F) - F), kappa by mode_kappa after every update, stop once
    sqrt(sum (kappa_new - kappa_old)^2 / (9 nb)) < tol.  Returns (F, kappa history (nb,3,3,n),
    tauinv_eff(nibz,nb) after each iteration (n,)).
    The defaults are read_params' (rest.f90:36,44): max_iter = 150 and update_mix = 0.4, a
    REAL*4 literal stored in a REAL*8 (0.4000000059604645); the stop test normk < 0.0001
    (kappa.f90:1476) compares with a REAL*4 literal too.  With Piso the update is update_iso
    (kappa.f90:1401-1407)."""
    nibz = len(kibz); nb = eigenval.shape[0]
    F = np.array(FRTA, order="F", copy=True)
    kap, _ = mode_kappa(kibz, mapinv, kp, nc, g, primlatt, opk, eigenval, veloc, F, tempk, volume_r, classical)
    hist, tes = [], []
    for _ in range(max_iter):
        old = kap
        F2 = update(0, nibz - 1, kibz, mapibz, kp, nc, g, primlatt, opk, P1, P2, F, FRTA, Q, Piso)
        F = np.asfortranarray(F + update_mix * (F2 - F))
        kap, te = mode_kappa(kibz, mapinv, kp, nc, g, primlatt, opk, eigenval, veloc, F, tempk, volume_r, classical)
        hist.append(kap)
        tes.append(te[-1, -1])
        if np.sqrt(np.sum((kap - old) ** 2 / (nb * 9))) < tol:
            break
    return F, np.stack(hist, axis=-1), np.array(tes)


def self_w3(qs, model, kp, g, eigenval, eigenvec, fc3_terms, tempk, eta, v3thr=float(np.float32(1e-7)), classical=0,
            kind=0):
    """function_self_w3 at every branch of each q (jobs 2/6/7, kappa.f90:900-925): get_freq at q
    and at -q-k2 for every mesh k2 (LAPACK here), omega = omega(q, la), T in K.  Returns
    omega0 (nb, nq), nself and uself (complex (nb, nq)).  fc3_terms = (masses, atompos,
    iatomcell0, at3, xyz3, fc3).  v3_threshold is 1E-7, a REAL*4 literal (rest.f90:47).
    kind = 1: function_self_w4 (rest.f90:4126-4240), the Gaussian imaginary part."""
    qs = np.atleast_2d(qs)
    masses, atompos, iatomcell0, at3, xyz3, fc3 = fc3_terms
    nb = eigenval.shape[0]; nk = len(kp)
    temp = float(lib().or_temp_cm(C.c_double(tempk)))
    wq, vq = get_freq_batch(qs, model)
    om0 = np.zeros((nb, len(qs)), order="F")
    ns = np.zeros((nb, len(qs)), dtype=np.complex128, order="F"); us = np.zeros_like(ns)
    ev = np.asfortranarray(eigenvec, dtype=np.complex128)
    for iq, q in enumerate(qs):
        w3, v3 = get_freq_batch(-q[None, :] - kp, model)          # (nk, nb), (nk, nb, nb) [k, comp, branch]
        e3v = np.asfortranarray(w3.T)
        e3c = np.asfortranarray(np.transpose(v3, (1, 2, 0)))      # [comp, branch, k]
        for la in range(nb):
            omq = np.sqrt(abs(wq[iq, la])) * CNST
            om0[la, iq] = omq
            evq = np.ascontiguousarray(vq[iq, :, la])
            n2 = np.zeros(2); u2 = np.zeros(2)
            lib().or_self_w3(_p(np.ascontiguousarray(q, dtype=np.float64)), nb, C.c_double(omq), _p(evq), C.c_double(omq),
                             C.c_double(temp), classical, C.c_double(eta), C.c_double(v3thr), nk,
                             _p(np.ascontiguousarray(kp)), _p(np.ascontiguousarray(g)), _p(_f64(eigenval)), _p(ev),
                             _p(e3v), _p(e3c), _p(_f64(masses)), _p(np.ascontiguousarray(atompos)), _p(_i32(iatomcell0)),
                             len(fc3), _p(np.ascontiguousarray(at3, dtype=np.int32)),
                             _p(np.ascontiguousarray(xyz3, dtype=np.int32)), _p(_f64(fc3)), kind, _p(n2), _p(u2))
            ns[la, iq] = n2[0] + 1j * n2[1]
            us[la, iq] = u2[0] + 1j * u2[1]
    return om0, ns, us


def rhs(mapinv, freq, dist, veloc, tempk):
    nk, nb = freq.shape; nibz = len(mapinv)
    R = np.zeros(nibz * nb * 3)
    lib().or_rhs(nibz, nb, nk, _p(_i32(mapinv)), _p(_f64(freq)), _p(_f64(dist)), _p(_f64(veloc)), C.c_double(tempk), _p(R))
    return R


def cg(A, b, x0, max_iter=100, ndig=2):
    x = np.array(x0, dtype=np.float64, copy=True)
    fm = C.c_double()
    it = lib().or_cg(len(b), _p(_f64(A)), _p(np.ascontiguousarray(b, dtype=np.float64)), _p(x), max_iter, ndig, C.byref(fm))
    return x, it, fm.value


def mode_kappa_d(kibz, mapinv, kp, nc, g, primlatt, opk, eigenval, veloc, x, tempk, volume_r, taueff_in, classical=0):
    nibz = len(kibz); nb, nk = eigenval.shape
    kap = np.zeros((nb, 3, 3), order="F"); te = np.asfortranarray(np.array(taueff_in, dtype=np.float64, copy=True))
    lib().or_mode_kappa_d(nibz, nb, nk, _p(np.ascontiguousarray(kibz)), _p(_i32(mapinv)), _p(np.ascontiguousarray(kp)),
                          _p(_i32(np.asarray(nc))), _p(np.ascontiguousarray(g)), _p(_f64(primlatt)), len(opk),
                          _p(np.ascontiguousarray(opk)), _p(_f64(eigenval)), _p(_f64(veloc)),
                          _p(np.ascontiguousarray(x, dtype=np.float64)), C.c_double(tempk), classical,
                          C.c_double(volume_r), _p(kap), _p(te))
    return kap, te
