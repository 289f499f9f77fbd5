"""ctypes binding of libalt3.so (include/alt3.h) and a thin Context wrapper.

Array conventions follow the C ABI: every array argument/result is in the reference's
Fortran column-major layout (numpy order='F' with the Fortran shape) with 0-based index
values.  The library must be built (thermacond.build); there is no CPU fallback: loading
fails loudly when libalt3.so is missing.
"""
from __future__ import annotations

import ctypes as C
import os
from typing import Optional, Sequence

import numpy as np

from .refio import InputSet

PKG_DIR = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.path.join(PKG_DIR, "libalt3.so")

V3_FP64_REF, V3_FP64, V3_MIXED = 0, 1, 2
DELTA_TETRA, DELTA_GAUSS = 0, 1
MMA_F16X3, MMA_F32 = 0, 1
THZ_TO_CM = 33.35640951981521  # 1 THz in cm^-1

ERRORS = {-1: "ALT3_E_ARG", -2: "ALT3_E_HIP", -3: "ALT3_E_MESH", -4: "ALT3_E_STATE", -5: "ALT3_E_DEN0",
          -6: "ALT3_E_NOMEM", -7: "ALT3_E_NBE"}


class Alt3Error(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"{ERRORS.get(code, code)}: {msg}")
        self.code = code


class System(C.Structure):
    _fields_ = [
        ("natoms0", C.c_int), ("mass", C.c_void_p), ("r123", C.c_void_p), ("latparams", C.c_void_p),
        ("primlatt", C.c_void_p), ("natoms", C.c_int), ("atompos", C.c_void_p), ("iatomcell0", C.c_void_p),
        ("nt2", C.c_int), ("fc2_atoms", C.c_void_p), ("fc2_xyz", C.c_void_p), ("fc2", C.c_void_p),
        ("nt3", C.c_int), ("fc3_atoms", C.c_void_p), ("fc3_xyz", C.c_void_p), ("fc3", C.c_void_p),
        ("nc", C.c_int * 3), ("classical", C.c_int), ("host_only", C.c_int),
        ("born_flag", C.c_int), ("epsil", C.c_void_p), ("zeu", C.c_void_p),
    ]


EXPORTS = ["alt3_create", "alt3_destroy", "alt3_last_error", "alt3_device_bytes", "alt3_sizes", "alt3_get_mesh",
           "alt3_kstars", "alt3_eig", "alt3_set_eig", "alt3_v3sq_ibz", "alt3_fgr", "alt3_tau", "alt3_kappa_rta",
           "alt3_collision", "alt3_cg", "alt3_kappa_bte", "alt3_solve", "alt3_last_kernel_ms", "alt3_fgr_gauss",
           "alt3_set_delta", "alt3_footprint", "alt3_update", "alt3_kappa_iter", "alt3_self_w3", "alt3_self_w4", "alt3_set_fgr",
           "alt3_set_iso", "alt3_isotope_g", "alt3_device_free", "alt3_cg_static", "alt3_kernel_ms",
           "alt3_v3_batch", "alt3_get_v3sq", "alt3_set_v3_mma"]

# read_params (rest.f90:36,44): max_iter = 150, update_mix = 0.4; the job-3 stop normk < 0.0001
# (kappa.f90:1476).  Both constants are REAL*4 literals.
ITER_MAX = 150
UPDATE_MIX = float(np.float32(0.4))
ITER_TOL = float(np.float32(1e-4))
# read_params (rest.f90:47): v3_threshold = 1E-7, a REAL*4 literal
V3_THRESHOLD = float(np.float32(1e-7))

_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise OSError(f"{LIB_PATH} is missing: build it with thermacond.build (hipcc, gfx950)")
        L = C.CDLL(LIB_PATH)
        L.alt3_last_error.restype = C.c_char_p
        for name in EXPORTS:
            if name != "alt3_last_error":
                getattr(L, name).restype = C.c_int
        _lib = L
    return _lib


def isotope_g(element: str) -> float:
    """cal_g (sub.f90:3766-4693): the natural-isotope mass variance of an element (0 if the
    table has no data for it)."""
    g = C.c_double()
    _check(lib().alt3_isotope_g(element.encode(), C.byref(g)))
    return g.value


def device_free_bytes() -> int:
    """Free memory of the current device (alt3_device_free: hipMemGetInfo)."""
    f = C.c_longlong()
    _check(lib().alt3_device_free(C.byref(f), None))
    return int(f.value)


def v3_batch(contexts: Sequence["Context"], precision: int = V3_FP64, tile: int = 0) -> int:
    """|V3|^2 over the whole IBZ of every context in one shared persistent launch
    (alt3_v3_batch); each context then holds what v3sq(0, nibz - 1, precision) leaves and
    solve() reuses it.  Returns the q2 tile the launch used (0: one launch per context)."""
    hs = (C.c_void_p * len(contexts))(*[c.h for c in contexts])
    used = C.c_int()
    _check(lib().alt3_v3_batch(hs, len(contexts), precision, tile, C.byref(used)))
    for c in contexts:
        c._last_ksub = c.nibz
    return used.value


def giso_of(inp: InputSet):
    """cal_giso (sub.f90:4695-4725): g of every atom of the primitive cell from the element
    names and atom types of structure.params."""
    return np.array([isotope_g(inp.names[t - 1]) for t in inp.types])


def _check(rc: int):
    if rc != 0:
        raise Alt3Error(rc, lib().alt3_last_error().decode())


def _p(a: Optional[np.ndarray]):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


class Context:
    """A device-resident crystal + force constants + mesh (alt3_ctx)."""

    def __init__(self, inp: InputSet, nc: Optional[Sequence[int]] = None, classical: Optional[int] = None,
                 host_only: bool = False):
        self.inp = inp
        nc = tuple(inp.params.nc if nc is None else nc)
        self._keep = []

        def arr(a, dt):
            a = np.ascontiguousarray(a, dtype=dt)
            self._keep.append(a)
            return a.ctypes.data_as(C.c_void_p)

        s = System()
        s.natoms0 = inp.natoms0
        s.mass = arr(inp.masses, np.float64)
        s.r123 = arr(inp.r123, np.float64)
        # structure.params line 1 with a,b,c times the scale, as read_input_fit holds it (rest.f90:89-92)
        lp = np.array(inp.latparams, dtype=np.float64)
        lp[:3] *= inp.lattice_parameter
        s.latparams = arr(lp, np.float64)
        s.primlatt = arr(np.asfortranarray(inp.primlatt).reshape(-1, order="F"), np.float64)
        s.natoms = len(inp.iatomcell0)
        s.atompos = arr(inp.atompos, np.float64)
        s.iatomcell0 = arr(inp.iatomcell0, np.int32)
        s.nt2 = len(inp.fc2)
        s.fc2_atoms = arr(inp.fc2_atoms, np.int32)
        s.fc2_xyz = arr(inp.fc2_xyz, np.int32)
        s.fc2 = arr(inp.fc2, np.float64)
        s.nt3 = len(inp.fc3)
        s.fc3_atoms = arr(inp.fc3_atoms, np.int32)
        s.fc3_xyz = arr(inp.fc3_xyz, np.int32)
        s.fc3 = arr(inp.fc3, np.float64)
        s.nc = (C.c_int * 3)(*nc)
        s.classical = int(inp.params.classical if classical is None else classical)
        s.host_only = 1 if host_only else 0
        # dielectric.params (read_born): the non-analytic term when born_flag != 0
        s.born_flag = int(inp.born_flag)
        if inp.born_flag:
            s.epsil = arr(inp.epsil, np.float64)
            s.zeu = arr(inp.zeu, np.float64)
        self._sys = s
        h = C.c_void_p()
        _check(lib().alt3_create(C.byref(s), C.byref(h)))
        self.h = h
        nk, nibz, nb, nops = C.c_int(), C.c_int(), C.c_int(), C.c_int()
        _check(lib().alt3_sizes(h, C.byref(nk), C.byref(nibz), C.byref(nb), C.byref(nops)))
        self.nk, self.nibz, self.nb, self.nops = nk.value, nibz.value, nb.value, nops.value
        self.nc = nc

    def close(self):
        if getattr(self, "h", None):
            lib().alt3_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    # ---------------------------------------------------------------- tables
    def mesh(self) -> dict:
        nk, ni, no = self.nk, self.nibz, self.nops
        d = dict(kpc_eig=np.zeros((nk, 3)), kpc=np.zeros((nk, 3)), mapibz=np.zeros(nk, np.int32),
                 mapinv=np.zeros(ni, np.int32), kibz=np.zeros((ni, 3)), wibz=np.zeros(ni),
                 tet=np.zeros((6 * nk, 4), np.int32), opk=np.zeros((no, 3, 3)), mappos=np.zeros(nk, np.int32))
        npos = C.c_int()
        _check(lib().alt3_get_mesh(self.h, *[_p(d[k]) for k in ("kpc_eig", "kpc", "mapibz", "mapinv", "kibz", "wibz",
                                                                 "tet", "opk", "mappos")], C.byref(npos)))
        d["npos"] = npos.value
        return d

    def kstars(self, i: int):
        n = C.c_int(); ops = np.zeros(48, np.int32); n0 = np.zeros(48, np.int32)
        _check(lib().alt3_kstars(self.h, i, C.byref(n), _p(ops), _p(n0)))
        return ops[:n.value].copy(), n0[:n.value].copy()

    # ---------------------------------------------------------------- path
    def eig(self, want: bool = True):
        nk, nb = self.nk, self.nb
        if not want:
            _check(lib().alt3_eig(self.h, None, None, None))
            return None
        ev = np.zeros((nb, nk), order="F"); vec = np.zeros((nb, nb, nk), dtype=np.complex128, order="F")
        vel = np.zeros((3, nb, nk), order="F")
        _check(lib().alt3_eig(self.h, _p(ev), _p(vec), _p(vel)))
        return ev, vec, vel

    def set_eig(self, eigenval, eigenvec, veloc):
        ev = np.asfortranarray(eigenval, dtype=np.float64)
        vec = np.asfortranarray(eigenvec, dtype=np.complex128)
        vel = np.asfortranarray(veloc, dtype=np.float64)
        _check(lib().alt3_set_eig(self.h, _p(ev), _p(vec), _p(vel)))

    def v3sq(self, lo: int, hi: int, precision: int = V3_FP64, want: bool = True):
        out = np.zeros((hi - lo + 1, self.nk, self.nb, self.nb, self.nb), order="F") if want else None
        _check(lib().alt3_v3sq_ibz(self.h, lo, hi, precision, _p(out)))
        self._last_ksub = hi - lo + 1
        return out

    def get_v3sq(self):
        """The |V3|^2 the context holds (alt3_get_v3sq), shape (ksub, nk, nb, nb, nb)."""
        out = np.zeros((self._ksub(), self.nk, self.nb, self.nb, self.nb), order="F")
        _check(lib().alt3_get_v3sq(self.h, _p(out)))
        return out

    def fgr(self, want: bool = True):
        if not want:
            _check(lib().alt3_fgr(self.h, None, None))
            return None
        ksub = self._ksub()
        P1 = np.zeros((ksub, self.nk, self.nb, self.nb, self.nb), order="F"); P2 = np.zeros_like(P1)
        _check(lib().alt3_fgr(self.h, _p(P1), _p(P2)))
        return P1, P2

    def set_fgr(self, lo: int, P1sm, P2sm):
        """Load P1sm / P2sm (ksub, nk, nb, nb, nb) of irreducible points lo.. (FP64)."""
        P1 = np.asfortranarray(P1sm, dtype=np.float64); P2 = np.asfortranarray(P2sm, dtype=np.float64)
        _check(lib().alt3_set_fgr(self.h, lo, lo + P1.shape[0] - 1, _p(P1), _p(P2)))
        self._last_ksub = P1.shape[0]

    def set_iso(self, giso=None, want: bool = True):
        """Isotope scattering (iso = 1): giso per atom of the primitive cell, None to switch it
        off.  Returns Piso (nibz, nb) of P_iso_split when want."""
        if giso is None:
            _check(lib().alt3_set_iso(self.h, None, None))
            return None
        g = np.ascontiguousarray(giso, dtype=np.float64)
        if g.shape != (self.inp.natoms0,):
            raise ValueError(f"giso needs {self.inp.natoms0} values")
        out = np.zeros((self.nibz, self.nb), order="F") if want else None
        _check(lib().alt3_set_iso(self.h, _p(g), _p(out)))
        return out

    def _ksub(self):
        return getattr(self, "_last_ksub", self.nibz)

    def fgr_gauss(self, sigma: float, want: bool = True):
        if not want:
            _check(lib().alt3_fgr_gauss(self.h, C.c_double(sigma), None, None))
            return None
        ksub = self._ksub()
        P1 = np.zeros((ksub, self.nk, self.nb, self.nb, self.nb), order="F"); P2 = np.zeros_like(P1)
        _check(lib().alt3_fgr_gauss(self.h, C.c_double(sigma), _p(P1), _p(P2)))
        return P1, P2

    def set_delta(self, kind: int, sigma: float = 0.0):
        _check(lib().alt3_set_delta(self.h, kind, C.c_double(sigma)))

    def set_v3_mma(self, kind: int):
        """Matrix-core arithmetic of the mixed V3 on 24-branch cells: MMA_F16X3 (split FP16,
        the default) or MMA_F32."""
        _check(lib().alt3_set_v3_mma(self.h, kind))

    def tau(self, tempk: float):
        ksub = self._ksub()
        outs = [np.zeros((ksub, self.nb), order="F") for _ in range(6)]
        _check(lib().alt3_tau(self.h, C.c_double(tempk), *[_p(o) for o in outs]))
        return dict(zip(("Q", "QN", "QU", "tauN", "tauU", "tautot"), outs))

    def kappa_rta(self, tempk: float):
        k = np.zeros((self.nb, 3, 3), order="F"); F = np.zeros((self.nibz, self.nb, 3), order="F")
        te = np.zeros((self.nibz, self.nb), order="F")
        _check(lib().alt3_kappa_rta(self.h, C.c_double(tempk), _p(k), _p(F), _p(te)))
        return k, F, te

    def update(self, tempk: float, FFF, FRTA):
        """update_mpi: F2 (ksub, nb, 3) = F_RTA + A(FFF)/Q for the irreducible points of the context."""
        FFF = np.asfortranarray(FFF, dtype=np.float64); FRTA = np.asfortranarray(FRTA, dtype=np.float64)
        F2 = np.zeros((self._ksub(), self.nb, 3), order="F")
        _check(lib().alt3_update(self.h, C.c_double(tempk), _p(FFF), _p(FRTA), _p(F2)))
        return F2

    def kappa_iter(self, tempk: float, max_iter: int = ITER_MAX, update_mix: float = UPDATE_MIX,
                   tol: float = ITER_TOL):
        """The iterative BTE of job 3: final kappa (nb,3,3), kappa after every iteration
        (nb,3,3,iters), the surviving tauinv_eff(nibz,nb) per iteration, final F (nibz,nb,3)
        and the iteration count."""
        k = np.zeros((self.nb, 3, 3), order="F"); F = np.zeros((self.nibz, self.nb, 3), order="F")
        hist = np.zeros(max_iter * self.nb * 9); te = np.zeros(max_iter); it = C.c_int()
        _check(lib().alt3_kappa_iter(self.h, C.c_double(tempk), max_iter, C.c_double(update_mix), C.c_double(tol),
                                     _p(k), _p(hist), _p(te), _p(F), C.byref(it)))
        n = it.value
        return dict(kappa=k, hist=hist[:n * self.nb * 9].reshape((self.nb, 3, 3, n), order="F"), taueff=te[:n],
                    F=F, iters=n)

    def self_w3(self, q, tempk: float, eta: float, v3_threshold: float = V3_THRESHOLD):
        """Three-phonon self-energy at points q (nq,3) Cartesian: omega0 (nb,nq), nself and
        uself (complex (nb,nq)); the linewidth of jobs 2/6/7 is 2 Im(nself + uself)."""
        q = np.ascontiguousarray(np.atleast_2d(q), dtype=np.float64)
        nq = len(q)
        om = np.zeros((self.nb, nq), order="F")
        ns = np.zeros((self.nb, nq), dtype=np.complex128, order="F"); us = np.zeros_like(ns)
        _check(lib().alt3_self_w3(self.h, nq, _p(q), C.c_double(tempk), C.c_double(eta), C.c_double(v3_threshold),
                                  _p(om), _p(ns), _p(us)))
        return om, ns, us

    def self_w4(self, q, tempk: float, eta: float, v3_threshold: float = V3_THRESHOLD):
        """The Gaussian form (function_self_w4): imaginary parts only, as self_w3 otherwise."""
        q = np.ascontiguousarray(np.atleast_2d(q), dtype=np.float64)
        nq = len(q)
        om = np.zeros((self.nb, nq), order="F")
        ns = np.zeros((self.nb, nq), dtype=np.complex128, order="F"); us = np.zeros_like(ns)
        _check(lib().alt3_self_w4(self.h, nq, _p(q), C.c_double(tempk), C.c_double(eta), C.c_double(v3_threshold),
                                  _p(om), _p(ns), _p(us)))
        return om, ns, us

    def collision(self, tempk: float, want: bool = True):
        n = self.nibz * self.nb * 3
        CM = np.zeros((n, n), order="F") if want else None
        R = np.zeros(n)
        _check(lib().alt3_collision(self.h, C.c_double(tempk), _p(CM), _p(R)))
        return CM, R

    def cg(self, x0, A=None, b=None, max_iter: int = 100, ndig: int = 2):
        x = np.array(x0, dtype=np.float64, copy=True)
        n = len(x)
        fm = C.c_double(); it = C.c_int()
        Af = None if A is None else np.asfortranarray(A, dtype=np.float64)
        bf = None if b is None else np.ascontiguousarray(b, dtype=np.float64)
        _check(lib().alt3_cg(self.h, n, _p(Af), _p(bf), _p(x), max_iter, ndig, C.byref(fm), C.byref(it)))
        return x, it.value, fm.value

    def cg_static(self, x0, A=None, b=None, max_iter: int = 100, ndig: int = 2):
        """alt3_cg_static: alt3_cg without device allocations (static scratch, n <= 2505 with A)."""
        x = np.array(x0, dtype=np.float64, copy=True)
        n = len(x)
        Af = None if A is None else np.asfortranarray(A, dtype=np.float64)
        bf = None if b is None else np.ascontiguousarray(b, dtype=np.float64)
        it = C.c_int(); fm = C.c_double()
        _check(lib().alt3_cg_static(self.h, n, _p(Af), _p(bf), _p(x), max_iter, ndig, C.byref(fm), C.byref(it)))
        return x, it.value, fm.value

    def kappa_bte(self, tempk: float, x, taueff=None):
        k = np.zeros((self.nb, 3, 3), order="F")
        te = None if taueff is None else np.asfortranarray(np.array(taueff, dtype=np.float64, copy=True))
        _check(lib().alt3_kappa_bte(self.h, C.c_double(tempk), _p(np.ascontiguousarray(x, dtype=np.float64)), _p(k), _p(te)))
        return k, te

    def solve(self, temps: Sequence[float], precision: int = V3_FP64):
        nt = len(temps)
        kr = np.zeros((nt, self.nb, 3, 3)); kb = np.zeros_like(kr)
        its = np.zeros(nt, np.int32); ms = np.zeros(8)
        t = np.ascontiguousarray(temps, dtype=np.float64)
        kr_f = np.zeros(nt * self.nb * 9); kb_f = np.zeros_like(kr_f)
        _check(lib().alt3_solve(self.h, nt, _p(t), precision, _p(kr_f), _p(kb_f), _p(its), _p(ms)))
        self._last_ksub = self.nibz     # the solve leaves the whole zone's arrays in the context
        for i in range(nt):
            kr[i] = kr_f[i * self.nb * 9:(i + 1) * self.nb * 9].reshape((self.nb, 3, 3), order="F")
            kb[i] = kb_f[i * self.nb * 9:(i + 1) * self.nb * 9].reshape((self.nb, 3, 3), order="F")
        return dict(kappa_rta=kr, kappa_bte=kb, cg_iters=its, stage_ms=ms)

    def last_ms(self) -> float:
        ms = C.c_double()
        _check(lib().alt3_last_kernel_ms(self.h, C.byref(ms)))
        return ms.value

    def kernel_ms(self):
        """HIP-event times (ms) of the last |V3|^2 contraction kernel and the last FGR kernel."""
        v, f = C.c_double(), C.c_double()
        _check(lib().alt3_kernel_ms(self.h, C.byref(v), C.byref(f)))
        return v.value, f.value

    def footprint(self, ntemps: int, precision: int = V3_FP64) -> int:
        b = C.c_longlong()
        _check(lib().alt3_footprint(self.h, ntemps, precision, C.byref(b)))
        return b.value

    def device_bytes(self) -> int:
        b = C.c_longlong()
        _check(lib().alt3_device_bytes(self.h, C.byref(b)))
        return b.value
