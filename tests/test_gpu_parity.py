"""GPU parity: the HIP path (through the C ABI) against the reference THERMACOND build
(golden fixtures written by tests/golden/make_golden.py) and the oracle.

Tolerances (relative to the largest magnitude of the compared array unless stated):
  eigenvalues  1e-12 vs the oracle (LAPACK); 1e-6 vs the reference build, whose vendored
               ZHEEV stops at ~1e-8 residuals
  velocities   1e-6 vs the oracle (Gamma excluded); 1e-4 vs the reference
  |V3|^2       1e-10 with the reference eigenvectors loaded (FP64_REF); 1e-8 for the batched
               FP64 path, whose phase e^{-i q1.Rk} e^{-i q2.Rk} replaces the reference's
               e^{i q3.Rk} at the folded q3 = -q1-q2+G: the two differ by e^{i G.Rk}, which
               is 1 only to ~1e-10 because lat_fc.dat stores positions to 10 decimals
  P1sm/P2sm, Q, tau, F_RTA, kappa_RTA, CM, RHS   1e-8 with the reference eigen data
  CG solution, kappa_BTE   1e-6 with the reference collision matrix
  kappa (own eigenvectors, whole path)   0.5 %
"""
import numpy as np
import pytest

import oracle as O
from thermacond import capi

pytestmark = pytest.mark.gpu


def rel(a, b):
    a = np.asarray(a); b = np.asarray(b)
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


def check(a, b, tol, what):
    r = rel(a, b)
    assert r < tol, f"{what}: max rel diff {r:.3e} >= {tol:.1e}"


@pytest.fixture(scope="module")
def ctx4(si_inputs):
    print("case: SW-Si 4x4x4 context", flush=True)
    c = capi.Context(si_inputs[4])
    yield c
    c.close()


@pytest.fixture(scope="module")
def ctx8(si_inputs):
    print("case: SW-Si 8x8x8 context", flush=True)
    c = capi.Context(si_inputs[8])
    yield c
    c.close()


def _load_ref_eig(ctx, R):
    ctx.set_eig(R["eigenval"], R["eigenvec"], R["veloc"])


def test_eig_vs_oracle_and_reference(ctx8, ref8, si_inputs):
    inp = si_inputs[8]
    ev, vec, vel = ctx8.eig()
    m = ctx8.mesh()
    model = (inp.masses, inp.atompos, inp.iatomcell0, inp.fc2_atoms, inp.fc2_xyz, inp.fc2)
    oev, ovec, ovel = O.eigen(m["kpc_eig"], model)
    check(ev, oev, 1e-12, "eigenval vs oracle")
    check(ev, ref8["eigenval"], 1e-6, "eigenval vs reference")
    gam = np.all(np.abs(m["kpc_eig"]) < 1e-12, axis=1)
    check(vel[:, :, ~gam], ovel[:, :, ~gam], 1e-6, "veloc vs oracle")
    check(vel[:, :, ~gam], ref8["veloc"][:, :, ~gam], 1e-4, "veloc vs reference")
    # eigenvectors: orthonormal, and the projector of every degenerate group equals the oracle's
    worst = 0.0
    for k in range(ctx8.nk):
        V = vec[:, :, k]
        assert np.abs(V.conj().T @ V - np.eye(ctx8.nb)).max() < 1e-12
        w = ev[:, k]; l = 0
        while l < len(w):
            u = l + 1
            while u < len(w) and abs(w[u] - w[l]) < 1e-8 * max(abs(w).max(), 1e-30):
                u += 1
            P1 = V[:, l:u] @ V[:, l:u].conj().T
            P2 = ovec[:, l:u, k] @ ovec[:, l:u, k].conj().T
            worst = max(worst, np.abs(P1 - P2).max())
            l = u
    assert worst < 1e-8


def test_eig_time_reversal(ctx4):
    ev, vec, vel = ctx4.eig()
    m = ctx4.mesh()
    n = ctx4.nc
    for k in range(ctx4.nk):
        c2 = k % n[2]; c1 = (k // n[2]) % n[1]; c0 = k // (n[1] * n[2])
        mk = ((n[2] - c2) % n[2]) + n[2] * (((n[1] - c1) % n[1]) + n[1] * ((n[0] - c0) % n[0]))
        assert np.allclose(vec[:, :, mk], np.conj(vec[:, :, k])) or np.allclose(vec[:, :, k], np.conj(vec[:, :, mk]))
        assert np.allclose(ev[:, mk], ev[:, k], rtol=1e-10, atol=1e-14)


@pytest.mark.parametrize("prec", [capi.V3_FP64_REF, capi.V3_FP64])
def test_v3_vs_reference_4(ctx4, ref4, prec):
    print(f"case: V3 4^3 prec={prec}", flush=True)
    _load_ref_eig(ctx4, ref4)
    v = ctx4.v3sq(0, ctx4.nibz - 1, prec)
    r = ref4["v33s8"]
    check(v, r, 1e-10 if prec == capi.V3_FP64_REF else 1e-8, "v33s8")
    big = r > 1e-6 * r.max()
    worst = float(np.max(np.abs(v[big] - r[big]) / r[big]))
    assert worst < 1e-6, f"elementwise |V3|^2 rel {worst:.3e}"


def test_v3_vs_reference_8_subset(ctx8, ref8):
    print("case: V3 8^3 subset", flush=True)
    _load_ref_eig(ctx8, ref8)
    for n1 in ref8["ibz_sel"]:
        v = ctx8.v3sq(int(n1), int(n1), capi.V3_FP64)
        r = ref8["v33s8"][list(ref8["ibz_sel"]).index(n1)]
        check(v[0], r, 1e-8, f"v33s8 ibz {n1}")


def test_fgr_tau_kappa_cm_vs_reference_4(ctx4, ref4):
    print("case: FGR/tau/kappa/CM 4^3", flush=True)
    _load_ref_eig(ctx4, ref4)
    ctx4.v3sq(0, ctx4.nibz - 1, capi.V3_FP64)
    P1, P2 = ctx4.fgr()
    check(P1, ref4["P1sm"], 1e-8, "P1sm")
    check(P2, ref4["P2sm"], 1e-8, "P2sm")
    for T in ref4["temps"]:
        t = f"T{T}_"
        out = ctx4.tau(float(T))
        for k in ("Q", "QN", "QU", "tauN", "tauU", "tautot"):
            mine = out[k].copy()
            if k in ("tauN", "tauU"):
                # the dump follows write_kappa, which stores 1e10 for zero rates (sub.f90:1880-1886)
                mine[mine == 0] = 1.0e10
            check(mine, ref4[t + k], 1e-8, t + k)
        kap, F, te = ctx4.kappa_rta(float(T))
        check(F, ref4[t + "FRTA"], 1e-8, t + "FRTA")
        check(kap, ref4[t + "kappaRTA"], 1e-8, t + "kappaRTA")
        assert np.allclose(te, ref4[t + "taueff"], rtol=1e-8, atol=0)
        CM, RHS = ctx4.collision(float(T))
        check(CM, ref4[t + "CM"], 1e-8, t + "CM")
        check(RHS, ref4[t + "RHS"], 1e-8, t + "RHS")
        x0 = np.array([F[i, l, a] for i in range(ctx4.nibz) for l in range(ctx4.nb) for a in range(3)])
        x, it, fm = ctx4.cg(x0)
        check(x, ref4[t + "x"], 1e-6, t + "x")
        kb, te2 = ctx4.kappa_bte(float(T), x, te)
        check(kb, ref4[t + "kappaBTE"], 1e-6, t + "kappaBTE")


def test_cg_external_matrix(ctx4, ref4):
    """alt3_cg on a host matrix equals the oracle's cg_quadratic on the same system."""
    T = ref4["temps"][1]
    A = ref4[f"T{T}_CM"]; b = ref4[f"T{T}_RHS"]
    x0 = np.linspace(-1, 1, len(b)) * 1e-9
    xg, itg, fmg = ctx4.cg(x0, A, b)
    xo, ito, fmo = O.cg(A, b, x0)
    assert itg == ito
    check(xg, xo, 1e-9, "cg x")


def test_solve_own_eigenvectors_vs_reference(si_inputs, ref4, ref8):
    """Whole path with the GPU's own eigen solve against the reference.

    Inside degenerate branches eigenvectors are defined only up to a unitary rotation, and
    the reference's per-branch |V3|^2, tau and star-rotated velocities depend on that
    choice; the loose CG stop ((dx/x)^2 < 1e-4) amplifies it.  Feeding the reference
    routines LAPACK eigenvectors instead of its own moves kappa_RTA by up to 0.49 % (4^3)
    / 0.29 % (8^3) and kappa_BTE by up to 1.19 % / 0.88 % (oracle measurement), so the
    bound here is 1 % (RTA) and 2 % (BTE).  Exact parity is tested above with the
    reference eigenvectors loaded."""
    for n, R in ((4, ref4), (8, ref8)):
        print(f"case: solve {n}^3", flush=True)
        with capi.Context(si_inputs[n]) as c:
            out = c.solve([float(t) for t in R["temps"]], capi.V3_FP64)
            for i, T in enumerate(R["temps"]):
                kr = np.trace(out["kappa_rta"][i].sum(0)) / 3
                kb = np.trace(out["kappa_bte"][i].sum(0)) / 3
                rr = np.trace(R[f"T{T}_kappaRTA"].sum(0)) / 3
                rb = np.trace(R[f"T{T}_kappaBTE"].sum(0)) / 3
                assert abs(kr / rr - 1) < 1e-2, (n, T, kr, rr)
                assert abs(kb / rb - 1) < 2e-2, (n, T, kb, rb)


def test_mixed_v3_vs_fp64(si_inputs):
    """ALT3_V3_MIXED (6 branches: FP32 phase sum and contractions on the vector ALUs, FP64
    scale) against the FP64 path on the same eigenvectors:
    |V3|^2 within 1e-5 of the largest element, kappa within 0.5 %."""
    print("case: mixed vs fp64 8^3", flush=True)
    with capi.Context(si_inputs[8]) as c:
        c.eig(want=False)
        v64 = c.v3sq(0, c.nibz - 1, capi.V3_FP64)
        v32 = c.v3sq(0, c.nibz - 1, capi.V3_MIXED)
        check(v32, v64, 1e-5, "mixed |V3|^2")
        temps = [100.0, 300.0, 800.0]
        o64 = c.solve(temps, capi.V3_FP64)
        o32 = c.solve(temps, capi.V3_MIXED)
        for i in range(3):
            for k in ("kappa_rta", "kappa_bte"):
                a = np.trace(o32[k][i].sum(0)) / 3
                b = np.trace(o64[k][i].sum(0)) / 3
                assert abs(a / b - 1) < 5e-3, (k, temps[i], a, b)


def test_solve_matches_reference_8(si_inputs, ref8):
    """alt3_solve (one tau pass over all temperatures) against the reference at 8^3 with the
    reference eigen data loaded: kappa_RTA tensors 1e-8, kappa_BTE 1e-5 (CG stop 1e-4)."""
    print("case: solve 8^3 vs reference, all temperatures", flush=True)
    with capi.Context(si_inputs[8]) as c:
        c.set_eig(ref8["eigenval"], ref8["eigenvec"], ref8["veloc"])
        temps = [float(t) for t in ref8["temps"]]
        out = c.solve(temps, capi.V3_FP64)
        for i, T in enumerate(ref8["temps"]):
            check(out["kappa_rta"][i], ref8[f"T{T}_kappaRTA"], 1e-8, f"T{T} kappaRTA")
            check(out["kappa_bte"][i], ref8[f"T{T}_kappaBTE"], 1e-5, f"T{T} kappaBTE")


def test_solve_temperature_groups(si_inputs, ref4):
    """Five temperatures (a tau pass of four and one of one) give the same kappa as five
    single-temperature solves, and the stagewise entry points agree with both."""
    print("case: solve 4^3 temperature groups", flush=True)
    temps = [100.0, 200.0, 300.0, 500.0, 800.0]
    with capi.Context(si_inputs[4]) as c:
        c.set_eig(ref4["eigenval"], ref4["eigenvec"], ref4["veloc"])
        many = c.solve(temps, capi.V3_FP64)
        for i, T in enumerate(temps):
            one = c.solve([T], capi.V3_FP64)
            check(many["kappa_rta"][i], one["kappa_rta"][0], 1e-13, f"T{T} kappa_rta")
            check(many["kappa_bte"][i], one["kappa_bte"][0], 1e-12, f"T{T} kappa_bte")
            assert many["cg_iters"][i] == one["cg_iters"][0]
        c.tau(300.0)
        kap, F, te = c.kappa_rta(300.0)
        check(kap, many["kappa_rta"][2], 1e-13, "stagewise kappa_rta")
        with pytest.raises(capi.Alt3Error):
            c.kappa_rta(200.0)   # not held by the last tau pass


def test_fgr_gauss_vs_oracle_4(ctx4, ref4, si_inputs):
    """alt3_fgr_gauss against the oracle's Gaussian form on the same |V3|^2 (FP64 both, 1e-10)."""
    print("case: Gaussian FGR 4^3", flush=True)
    _load_ref_eig(ctx4, ref4)
    v = ctx4.v3sq(0, ctx4.nibz - 1, capi.V3_FP64)
    sigma = 0.1 * capi.THZ_TO_CM
    P1, P2 = ctx4.fgr_gauss(sigma)
    inp = si_inputs[4]
    nc = np.array(inp.params.nc)
    g, _ = O.reciprocal(inp.r123)
    kp, _ = O.make_kp_reg_tet(nc, g)
    mapinv = (ref4["mapinv"] - 1).astype(np.int32)
    freq = O.freq_from_eigenval(ref4["eigenval"])
    O1, O2 = O.fgr_gauss(0, ctx4.nibz - 1, mapinv, kp, nc, g, freq, sigma, v)
    check(P1, O1, 1e-10, "gauss P1sm")
    check(P2, O2, 1e-10, "gauss P2sm")


def test_gauss_rta_kappa_8(si_inputs, ref8):
    """Gaussian-smeared RTA kappa on SW-Si 8^3 (sigma = 0.1 THz, T = 300 K) with the reference
    eigen data: alt3_solve with ALT3_DELTA_GAUSS equals the oracle chain (tau, F_RTA,
    mode_kappa) on the device's P to 1e-8, and lies within 25 % of the tetrahedron result."""
    print("case: Gaussian RTA kappa 8^3", flush=True)
    inp = si_inputs[8]
    sigma, T = 0.1 * capi.THZ_TO_CM, 300.0
    with capi.Context(inp) as c:
        c.set_eig(ref8["eigenval"], ref8["eigenvec"], ref8["veloc"])
        c.set_delta(capi.DELTA_GAUSS, sigma)
        out = c.solve([T], capi.V3_FP64)
        c.v3sq(0, c.nibz - 1, capi.V3_FP64, want=False)
        P1, P2 = c.fgr_gauss(sigma)
    nc = np.array(inp.params.nc)
    g, vol = O.reciprocal(inp.r123)
    opk = O.lattice_ops(inp.latparams, inp.primlatt)
    kp, _ = O.make_kp_reg_tet(nc, g)
    mapinv = (ref8["mapinv"] - 1).astype(np.int32)
    kibz = ref8["kibz"].T.copy()
    E, V = ref8["eigenval"], ref8["veloc"]
    freq, dist = O.distribution(E, T)
    _, _, Q, QN, QU = O.tau(0, len(mapinv) - 1, mapinv, kp, nc, g, dist, P1, P2, want_p=False)
    F = O.frta(mapinv, freq, dist, V, Q, T)
    kap, _ = O.mode_kappa(kibz, mapinv, kp, nc, g, inp.primlatt, opk, E, V, F, T, vol)
    check(out["kappa_rta"][0], kap, 1e-8, "Gaussian kappa_RTA vs oracle chain")
    kg = np.trace(out["kappa_rta"][0].sum(0)) / 3
    kt = np.trace(ref8["T300_kappaRTA"].sum(0)) / 3
    assert 0.75 < kg / kt < 1.25, (kg, kt)


def test_mixed_fgr_decay_mirror_vs_fp64(si_inputs):
    """The mixed path builds the decay rows of (l2 > l3) from those of (l3, l2) at q1 - q2.
    Against the FP64 kernel, which computes every row itself: P1sm and P2sm to 1e-4 of the
    largest element, and the sums over q2 of every (q1, l1, l2, l3) row of P2sm, the quantity
    the scattering rates integrate, to 1e-4 of the largest sum.  (Point by point a row can
    differ more: on a mirrored row the corners of a tetrahedron come in another order, and
    where two have the same REAL*4 energy weight_tet gives the tie's weights to the other
    corner; and the FP32 weights of a nearly flat tetrahedron lose digits.  Neither moves the
    row sums.)"""
    print("case: mixed FGR mirror vs fp64 8^3", flush=True)
    with capi.Context(si_inputs[8]) as c:
        c.eig(want=False)
        c.v3sq(0, c.nibz - 1, capi.V3_FP64, want=False)
        P1d, P2d = c.fgr()
        c.v3sq(0, c.nibz - 1, capi.V3_MIXED, want=False)
        P1m, P2m = c.fgr()
    check(P1m, P1d, 1e-4, "mixed P1sm")
    check(P2m, P2d, 1e-4, "mixed P2sm")
    sm, sd = P2m.sum(axis=1), P2d.sum(axis=1)            # (q1, l1, l2, l3)
    check(sm, sd, 1e-4, "P2sm row sums")
    nb = P2d.shape[2]
    for l2 in range(nb):
        for l3 in range(l2):                            # the mirrored rows on their own
            r = float(np.max(np.abs(sm[:, :, l2, l3] - sd[:, :, l2, l3])) / np.max(np.abs(sd)))
            assert r < 1e-4, (l2, l3, r)


def test_set_fgr_from_split_files(si_inputs, ref4, tmp_path):
    """P1sm / P2sm written to v33sq_delta.NNN.dat blocks (the reference's split format) and
    loaded back with alt3_set_fgr give the same scattering rates and kappa_RTA as the FGR
    kernel's own P (bit for bit: the same FP64 values)."""
    print("case: alt3_set_fgr from split files 4^3", flush=True)
    from thermacond import splitio
    T = 300.0
    with capi.Context(si_inputs[4]) as c:
        _load_ref_eig(c, ref4)
        c.v3sq(0, c.nibz - 1, capi.V3_FP64, want=False)
        P1, P2 = c.fgr()
        q0 = c.tau(T)["Q"]
        k0, _, _ = c.kappa_rta(T)
        cut = c.nibz // 2
        for idx, (lo, hi) in enumerate(((0, cut), (cut, c.nibz)), 1):
            splitio.write_v33sq_delta(str(tmp_path / splitio.block_name("v33sq_delta", idx)), P1[lo:hi], P2[lo:hi])
        parts = [splitio.read_v33sq_delta(str(tmp_path / splitio.block_name("v33sq_delta", i)), n, c.nk, c.nb)
                 for i, n in ((1, cut), (2, c.nibz - cut))]
        c.set_fgr(0, np.concatenate([parts[0][1], parts[1][1]]), np.concatenate([parts[0][2], parts[1][2]]))
        q1 = c.tau(T)["Q"]
        k1, _, _ = c.kappa_rta(T)
    assert np.array_equal(q1, q0)
    assert np.array_equal(k1, k0)
