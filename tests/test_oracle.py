"""The oracle (oracle/alt3_oracle.c + oracle/oracle.py) pinned against the reference
THERMACOND build (golden fixtures) and against closed forms.  CPU only."""
import numpy as np
import pytest

import oracle as O


def rel(a, b):
    return float(np.max(np.abs(np.asarray(a) - np.asarray(b))) / max(np.max(np.abs(b)), 1e-300))


@pytest.fixture(scope="module")
def geo4(si_inputs):
    inp = si_inputs[4]
    nc = np.array(inp.params.nc)
    g, vol = O.reciprocal(inp.r123)
    opk = O.lattice_ops(inp.latparams, inp.primlatt)
    kpm = O.make_kp_reg(nc, g)
    nibz, kibz, wibz, mapibz, mapinv = O.get_weights(kpm, inp.primlatt, opk)
    kp, tet = O.make_kp_reg_tet(nc, g)
    return dict(inp=inp, nc=nc, g=g, vol=vol, opk=opk, kpm=kpm, nibz=nibz, kibz=kibz, wibz=wibz,
                mapibz=mapibz, mapinv=mapinv, kp=kp, tet=tet)


# ------------------------------------------------------------------ indexing
@pytest.mark.parametrize("n", [4, 8])
def test_indexing_bit_exact(si_inputs, ref4, ref8, n):
    R = ref4 if n == 4 else ref8
    inp = si_inputs[n]
    nc = np.array(inp.params.nc)
    g, _ = O.reciprocal(inp.r123)
    opk = O.lattice_ops(inp.latparams, inp.primlatt)
    assert np.array_equal(opk, R["op_kmatrix"].transpose(2, 1, 0))
    kpm = O.make_kp_reg(nc, g)
    assert np.abs(kpm - R["kpc_mesh"].T).max() < 1e-14
    npos, mp = O.get_negatives(kpm, nc, g)
    assert npos == int(R["npos"]) and np.array_equal(mp + 1, R["mappos"])
    nibz, kibz, wibz, mapibz, mapinv = O.get_weights(kpm, inp.primlatt, opk)
    assert nibz == int(R["nibz"])
    assert np.array_equal(mapibz + 1, R["mapibz"]) and np.array_equal(mapinv + 1, R["mapinv"])
    assert np.array_equal(wibz, R["wibz"])
    kp, tet = O.make_kp_reg_tet(nc, g)
    assert np.array_equal(tet + 1, R["tet"].T)
    for i in range(nibz):
        st, ops, n0 = O.kstars(kibz[i], kp, inp.primlatt, opk)
        m = R["star_n"][i]
        assert len(ops) == m and np.array_equal(ops + 1, R["star_op"][:m, i]) and np.array_equal(n0 + 1, R["star_n0"][:m, i])


def test_get_k_info_cases(geo4):
    nc, g, kp = geo4["nc"], geo4["g"], geo4["kp"]
    for k in range(len(kp)):
        nk, ijk, ins = O.get_k_info(kp[k], nc, g)
        assert nk == k and ins == 1
    # a half-step vector is not a mesh point (get_k_info inside=2)
    _, _, ins = O.get_k_info(kp[1] + 0.5 * (kp[1] - kp[0]), nc, g)
    assert ins == 2
    # Gamma sits at i = j = k = n/2 + 1 on even meshes (sub.f90:7351-7356)
    nk, ijk, ins = O.get_k_info(np.zeros(3), nc, g)
    assert list(ijk) == [nc[0] // 2 + 1] * 3


# ------------------------------------------------------------------ eigen
def test_eigen_vs_reference(geo4, ref4):
    inp = geo4["inp"]
    model = (inp.masses, inp.atompos, inp.iatomcell0, inp.fc2_atoms, inp.fc2_xyz, inp.fc2)
    ev, vec, vel = O.eigen(geo4["kpm"], model)
    # the reference's vendored ZHEEV (as built with amdflang) stops at ~1e-8 residuals
    assert rel(ev, ref4["eigenval"]) < 1e-6
    # acoustic sum rule: three zero modes at Gamma
    gam = int(np.where(np.all(np.abs(geo4["kpm"]) < 1e-12, axis=1))[0][0])
    assert np.all(ev[:3, gam] < 1e-10 * ev.max())
    # eigenvectors solve H e = lambda e for the Hermitian matrix zheev('U') sees: the upper
    # triangle of D (set_dynamical_matrix leaves the last column unsymmetrised)
    D = O.dynmat_batch(geo4["kpm"][5][None], *model)[0]
    H = np.triu(D) + np.triu(D, 1).conj().T
    H[np.diag_indices(6)] = D.diagonal().real
    assert np.abs(H @ vec[:, :, 5] - vec[:, :, 5] * ev[:, 5]).max() < 1e-12 * ev.max()


# ------------------------------------------------------------------ V3
def test_v3_vs_reference(geo4, ref4):
    G, inp = geo4, geo4["inp"]
    v = O.v3sq(0, G["nibz"] - 1, G["mapinv"], G["kp"], G["nc"], G["g"], ref4["eigenval"], ref4["eigenvec"],
               inp.masses, inp.atompos, inp.iatomcell0, inp.fc3_atoms, inp.fc3_xyz, inp.fc3)
    assert rel(v, ref4["v33s8"]) < 1e-13


def test_v3_acoustic_sum_rule_and_time_reversal(geo4, ref4):
    """Numerator of V3 vanishes for a rigid translation at q3 = Gamma (FC3 acoustic sum
    rule), and |V3(-q1,-q2)|^2 = |V3(q1,q2)|^2 with e(-q) = e(q)*."""
    G, inp = geo4, geo4["inp"]
    nk, nb = len(G["kp"]), 6
    ones = np.ones((nb, nk), order="F")
    gam = int(np.where(np.all(np.abs(G["kp"]) < 1e-12, axis=1))[0][0])
    vec = np.array(ref4["eigenvec"], order="F", copy=True)
    vec[:, :3, gam] = 0
    for a in range(3):  # pure translations along x, y, z (equal masses)
        vec[a::3, a, gam] = 1 / np.sqrt(2)
    # q1 = k, q2 = -k  ->  q3 = Gamma
    k = 5
    mk, _, _ = O.get_k_info(-G["kp"][k], G["nc"], G["g"])
    v = O.v3sq(0, 0, np.array([k], np.int32), G["kp"], G["nc"], G["g"], ones, vec, inp.masses, inp.atompos,
               inp.iatomcell0, inp.fc3_atoms, inp.fc3_xyz, inp.fc3)
    typical = np.abs(v[0, :, :, :, 3:]).max()
    # exact up to the O(h^2) asymmetry of the finite-difference FC3 (|V|^2 ~ (1e-6)^2)
    assert np.abs(v[0, mk, :, :, :3]).max() < 1e-10 * typical
    # time reversal
    a = O.v3sq(0, 0, np.array([k], np.int32), G["kp"], G["nc"], G["g"], ref4["eigenval"], ref4["eigenvec"],
               inp.masses, inp.atompos, inp.iatomcell0, inp.fc3_atoms, inp.fc3_xyz, inp.fc3)
    b = O.v3sq(0, 0, np.array([mk], np.int32), G["kp"], G["nc"], G["g"], ref4["eigenval"], ref4["eigenvec"],
               inp.masses, inp.atompos, inp.iatomcell0, inp.fc3_atoms, inp.fc3_xyz, inp.fc3)
    neg = np.array([O.get_k_info(-G["kp"][j], G["nc"], G["g"])[0] for j in range(nk)])
    assert rel(b[0][neg], a[0]) < 1e-10


# ------------------------------------------------------------------ tetrahedra
def test_weight_tet_closed_form():
    """sum_i c_i(e) equals the linear-tetrahedron density of states g(e) and integrates to 1."""
    rng = np.random.default_rng(3)
    for _ in range(20):
        w = rng.uniform(10, 100, 4).astype(np.float32).astype(np.float64)
        e1, e2, e3, e4 = np.sort(w)
        es = np.linspace(e1, e4, 4001)
        s = np.array([O.weight_tet(w, e).sum() for e in es])
        g = np.where(es <= e2, 3 * (es - e1) ** 2 / ((e2 - e1) * (e3 - e1) * (e4 - e1)),
                     np.where(es <= e3, 3 / ((e3 - e1) * (e4 - e1)) * (e2 - e1 + 2 * (es - e2) - (e3 - e1 + e4 - e2) * (es - e2) ** 2 / ((e3 - e2) * (e4 - e2))),
                              3 * (e4 - es) ** 2 / ((e4 - e1) * (e4 - e2) * (e4 - e3))))
        assert np.max(np.abs(s - g)) < 1e-4 * g.max()
        assert abs(np.trapezoid(s, es) - 1) < 1e-5
        assert np.all(O.weight_tet(w, e1 - 1) == 0) and np.all(O.weight_tet(w, e4 + 1) == 0)


def test_fgr_vs_reference(geo4, ref4):
    G = geo4
    freq = O.freq_from_eigenval(ref4["eigenval"])
    P1, P2 = O.fgr(0, G["nibz"] - 1, G["mapinv"], G["kp"], G["nc"], G["g"], freq, G["tet"], ref4["v33s8"])
    assert np.array_equal(P1, ref4["P1sm"]) and np.array_equal(P2, ref4["P2sm"])


def test_fgr_vs_reference_8_subset(si_inputs, ref8):
    inp = si_inputs[8]
    nc = np.array(inp.params.nc)
    g, _ = O.reciprocal(inp.r123)
    kp, tet = O.make_kp_reg_tet(nc, g)
    mapinv = (ref8["mapinv"] - 1).astype(np.int32)
    freq = O.freq_from_eigenval(ref8["eigenval"])
    for j, n1 in enumerate(ref8["ibz_sel"][:2]):
        P1, P2 = O.fgr(int(n1), int(n1), mapinv, kp, nc, g, freq, tet, ref8["v33s8"][j:j + 1])
        assert np.array_equal(P1[0], ref8["P1sm"][j]) and np.array_equal(P2[0], ref8["P2sm"][j])


def test_delta_g_closed_form():
    """delta_g (extratools.f90:239-246) is a normalised Gaussian of variance w^2."""
    for w in (0.5, 3.3356, 10.0):
        x = np.linspace(-12 * w, 12 * w, 200001)
        y = np.array([O.delta_g(float(v), w) for v in x[::50]])
        xs = x[::50]
        assert abs(np.trapezoid(y, xs) - 1) < 1e-8
        assert abs(np.trapezoid(y * xs * xs, xs) / w ** 2 - 1) < 1e-6
        assert abs(O.delta_g(0.0, w) - 1 / (np.sqrt(2 * np.pi) * w)) < 1e-15 / w


def test_fgr_gauss_vs_tetrahedra(si_inputs, ref8):
    """With |V3|^2 = 1 the summed P1sm/P2sm are two-phonon densities of states at omega(q1,l1)
    (times (2 pi)^2); the Gaussian form (sigma = 5 cm^-1) and the tetrahedron form must agree
    to within the mesh error (20 % at 8^3 for general q1; a wrong normalisation -- a missing
    1/nk, 2pi or 1/sqrt(2pi) -- is off by factors of 2 or more)."""
    inp = si_inputs[8]
    nc = np.array(inp.params.nc)
    g, _ = O.reciprocal(inp.r123)
    kp, tet = O.make_kp_reg_tet(nc, g)
    mapinv = (ref8["mapinv"] - 1).astype(np.int32)
    freq = O.freq_from_eigenval(ref8["eigenval"])
    nk, nb = freq.shape
    v33 = np.ones((1, nk, nb, nb, nb), order="F")
    for s in (1, 3):
        T1, T2 = O.fgr(s, s, mapinv, kp, nc, g, freq, tet, v33)
        G1, G2 = O.fgr_gauss(s, s, mapinv, kp, nc, g, freq, 5.0, v33)
        for t, gg in ((T1, G1), (T2, G2)):
            r = gg.sum() / t.sum()
            assert 0.8 < r < 1.2, (s, r)
        # acoustic branches l2 < 3 at q2 = Gamma carry no weight (sub.f90:7349-7356)
        gam = int(np.argmin(np.abs(kp).sum(axis=1)))
        assert np.all(G1[0, gam, :, :3, :] == 0) and np.all(G2[0, gam, :, :3, :] == 0)


# ------------------------------------------------------------------ Q, tau, kappa, CM, CG
def test_tau_kappa_cm_cg_vs_reference(geo4, ref4):
    G = geo4
    E, V = ref4["eigenval"], ref4["veloc"]
    for T in ref4["temps"]:
        t = f"T{T}_"
        freq, dist = O.distribution(E, float(T))
        assert rel(dist, ref4[t + "dist"]) == 0
        P1, P2, Q, QN, QU = O.tau(0, G["nibz"] - 1, G["mapinv"], G["kp"], G["nc"], G["g"], dist, ref4["P1sm"], ref4["P2sm"])
        assert np.array_equal(Q, ref4[t + "Q"]) and np.array_equal(QN, ref4[t + "QN"]) and np.array_equal(QU, ref4[t + "QU"])
        tN, tU, tt = O.qandtau(G["mapinv"], dist, QN, QU)
        assert np.array_equal(tt, ref4[t + "tautot"])
        F = O.frta(G["mapinv"], freq, dist, V, Q, float(T))
        assert rel(F, ref4[t + "FRTA"]) < 1e-15
        kap, te = O.mode_kappa(G["kibz"], G["mapinv"], G["kp"], G["nc"], G["g"], G["inp"].primlatt, G["opk"], E, V, F,
                               float(T), G["vol"])
        assert rel(kap, ref4[t + "kappaRTA"]) < 1e-14
        CM = O.collision(0, G["nibz"] - 1, G["kibz"], G["mapinv"], G["kp"], G["nc"], G["g"], G["inp"].primlatt, G["opk"], P1, P2)
        assert np.array_equal(CM, ref4[t + "CM"])
        R = O.rhs(G["mapinv"], freq, dist, V, float(T))
        assert rel(R, ref4[t + "RHS"]) < 1e-15
        x0 = np.array([F[i, l, a] for i in range(G["nibz"]) for l in range(6) for a in range(3)])
        x, it, fm = O.cg(CM, R, x0)
        assert rel(x, ref4[t + "x"]) < 1e-12
        kb, _ = O.mode_kappa_d(G["kibz"], G["mapinv"], G["kp"], G["nc"], G["g"], G["inp"].primlatt, G["opk"], E, V, x,
                               float(T), G["vol"], te)
        assert rel(kb, ref4[t + "kappaBTE"]) < 1e-12


def test_collision_rows_vs_reference_8(si_inputs, ref8):
    inp = si_inputs[8]
    nc = np.array(inp.params.nc)
    g, _ = O.reciprocal(inp.r123)
    opk = O.lattice_ops(inp.latparams, inp.primlatt)
    kp, _ = O.make_kp_reg_tet(nc, g)
    mapinv = (ref8["mapinv"] - 1).astype(np.int32)
    kibz = ref8["kibz"].T.copy()
    T = 300
    _, dist = O.distribution(ref8["eigenval"], float(T))
    rows = ref8[f"T{T}_CM_rows"]
    for j, n1 in enumerate(ref8["ibz_sel"][:2]):
        P1, P2, Q, _, _ = O.tau(int(n1), int(n1), mapinv, kp, nc, g, dist, ref8["P1sm"][j:j + 1], ref8["P2sm"][j:j + 1])
        assert np.array_equal(Q[0], ref8[f"T{T}_Q"][n1])
        CM = O.collision(int(n1), int(n1), kibz, mapinv, kp, nc, g, inp.primlatt, opk, P1, P2)
        sel = np.where((rows >= 18 * n1) & (rows < 18 * (n1 + 1)))[0]
        assert np.array_equal(CM, ref8[f"T{T}_CM"][sel])


def test_cg_against_direct_solve():
    rng = np.random.default_rng(0)
    n = 60
    M = rng.standard_normal((n, n))
    A = np.asfortranarray(M @ M.T + n * np.eye(n))
    b = rng.standard_normal(n)
    x, it, fm = O.cg(A, b, np.zeros(n), max_iter=200, ndig=6)
    xs = np.linalg.solve(A, b)
    assert rel(x, xs) < 1e-6
    assert abs(fm - (0.5 * xs @ A @ xs - xs @ b)) < 1e-8 * abs(fm)


def test_nbe_limits():
    t = O.lib().or_temp_cm(300.0)
    assert abs(O.lib().or_nbe(1e-6 * t, t, 0) - 1e6) / 1e6 < 1e-5
    assert O.lib().or_nbe(100 * t, t, 0) == 0.0
    assert abs(O.lib().or_nbe(2 * t, t, 0) - 1 / (np.exp(2) - 1)) < 1e-15
    assert abs(O.lib().or_nbe(3 * t, t, 1) - 1 / 3) < 1e-15


# ------------------------------------------------------------------ iterative BTE (job 3)
@pytest.fixture(scope="module")
def ref4_iter():
    from conftest import load_fixture
    return load_fixture("si4_iter")


def test_update_and_bte_iter_vs_reference(geo4, ref4, ref4_iter):
    """update_mpi and the job-3 loop (kappa.f90:1373-1487) against the reference build: the
    first update F2, kappa after every iteration and the iteration count, per temperature."""
    G = geo4
    E, V = ref4["eigenval"], ref4["veloc"]
    for T in ref4["temps"]:
        t = f"T{T}_"
        freq, dist = O.distribution(E, float(T))
        P1, P2, Q, QN, QU = O.tau(0, G["nibz"] - 1, G["mapinv"], G["kp"], G["nc"], G["g"], dist, ref4["P1sm"], ref4["P2sm"])
        F = O.frta(G["mapinv"], freq, dist, V, Q, float(T))
        F2 = O.update(0, G["nibz"] - 1, G["kibz"], G["mapibz"], G["kp"], G["nc"], G["g"], G["inp"].primlatt, G["opk"],
                      P1, P2, F, F, Q)
        assert rel(F2, ref4_iter[t + "F2first"]) < 1e-13
        Fi, hist, tes = O.bte_iter(G["kibz"], G["mapinv"], G["mapibz"], G["kp"], G["nc"], G["g"], G["inp"].primlatt,
                                   G["opk"], E, V, P1, P2, F, Q, float(T), G["vol"])
        ref_hist = ref4_iter[t + "kappaIter"]
        assert hist.shape == ref_hist.shape, (T, hist.shape, ref_hist.shape)
        assert rel(hist, ref_hist) < 1e-11
        assert rel(tes, ref4_iter[t + "taueffIter"]) < 1e-11
        assert rel(Fi, ref4_iter[t + "Fiter"]) < 1e-11


def test_update_fixed_point_is_the_bte_solution(geo4, ref4):
    """A planted check on the update's algebra: with F_RTA = F - A(F)/Q for an arbitrary F,
    F is a fixed point of the update, and the map is affine in F (update(F1) - update(F0) is
    linear in F1 - F0).  A sign or index slip in either process breaks the second property
    only if it breaks linearity, so both are checked with random F."""
    G = geo4
    T = 300.0
    freq, dist = O.distribution(ref4["eigenval"], T)
    P1, P2, Q, _, _ = O.tau(0, G["nibz"] - 1, G["mapinv"], G["kp"], G["nc"], G["g"], dist, ref4["P1sm"], ref4["P2sm"])
    rng = np.random.default_rng(3)
    args = (0, G["nibz"] - 1, G["kibz"], G["mapibz"], G["kp"], G["nc"], G["g"], G["inp"].primlatt, G["opk"], P1, P2)
    Z = np.zeros((G["nibz"], 6, 3), order="F")
    Fa = np.asfortranarray(rng.standard_normal((G["nibz"], 6, 3)))
    Fb = np.asfortranarray(rng.standard_normal((G["nibz"], 6, 3)))
    Aa = O.update(*args, Fa, Z, Q)
    Ab = O.update(*args, Fb, Z, Q)
    Aab = O.update(*args, np.asfortranarray(2 * Fa - 3 * Fb), Z, Q)
    assert rel(Aab, 2 * Aa - 3 * Ab) < 1e-12
    # fixed point: F_RTA := F - A(F)/Q  =>  update(F) = F
    Frta = np.asfortranarray(Fa - Aa)
    assert rel(O.update(*args, Fa, Frta, Q), Fa) < 1e-12


# ------------------------------------------------------------------ self-energy (jobs 2/6/7)
def _groups(w, tol=1e-3):
    """Branches with frequencies within tol of the largest: per-branch values inside such a
    group depend on the eigenvector gauge, their sum does not."""
    out, l = [], 0
    while l < len(w):
        u = l + 1
        while u < len(w) and abs(w[u] - w[l]) < tol * w.max():
            u += 1
        out.append((l, u))
        l = u
    return out


def test_self_w3_vs_reference(geo4, ref4):
    """function_self_w3 with matrix_elt_full and get_freq against the reference build at points
    off the mesh (tests/golden/ref_si4_self.npz): omega(q) to 1e-7 (ZHEEV vs LAPACK), the
    normal + umklapp self-energy summed over (near-)degenerate branches to 2e-5, Im > 0."""
    from conftest import load_fixture
    S = load_fixture("si4_self")
    G = geo4
    inp = G["inp"]
    model = (inp.masses, inp.atompos, inp.iatomcell0, inp.fc2_atoms, inp.fc2_xyz, inp.fc2)
    fc3 = (inp.masses, inp.atompos, inp.iatomcell0, inp.fc3_atoms, inp.fc3_xyz, inp.fc3)
    sel = [0, 1, 4, 6]
    om, ns, us = O.self_w3(S["q"][sel], model, G["kp"], G["g"], ref4["eigenval"], ref4["eigenvec"], fc3,
                           float(S["tempk"]), float(S["etaz"]))
    assert rel(om, S["omega"][:, sel]) < 1e-7
    tot, rtot = ns + us, (S["nself"] + S["uself"])[:, sel]
    assert np.all(rtot.imag > 0) and np.all(tot.imag > 0)
    for i in range(len(sel)):
        for a, b in _groups(S["omega"][:, sel[i]]):
            y = rtot[a:b, i].sum()
            assert abs(tot[a:b, i].sum() - y) / abs(y) < 2e-5, (sel[i], a, b)


def test_self_w4_vs_reference(geo4, ref4):
    """function_self_w4 (the Gaussian imaginary part, rest.f90:4126-4240) against the reference
    build at the same points: real part zero, the normal + umklapp sum over (near-)degenerate
    branches to 2e-5; and on a single branch against the Lorentzian w3 in the limit of a narrow
    width both estimate the same delta function (pinned by the reference values themselves)."""
    from conftest import load_fixture
    S = load_fixture("si4_self")
    G = geo4
    inp = G["inp"]
    model = (inp.masses, inp.atompos, inp.iatomcell0, inp.fc2_atoms, inp.fc2_xyz, inp.fc2)
    fc3 = (inp.masses, inp.atompos, inp.iatomcell0, inp.fc3_atoms, inp.fc3_xyz, inp.fc3)
    sel = [0, 1, 4, 6]
    om, ns, us = O.self_w3(S["q"][sel], model, G["kp"], G["g"], ref4["eigenval"], ref4["eigenvec"], fc3,
                           float(S["tempk"]), float(S["etaz"]), kind=1)
    tot, rtot = ns + us, (S["nself4"] + S["uself4"])[:, sel]
    assert not np.any(tot.real) and not np.any(rtot.real)
    for i in range(len(sel)):
        for a, b in _groups(S["omega"][:, sel[i]]):
            y = rtot[a:b, i].sum()
            if abs(y) == 0:
                assert abs(tot[a:b, i].sum()) < 1e-12
                continue
            assert abs(tot[a:b, i].sum() - y) / abs(y) < 2e-5, (sel[i], a, b)
