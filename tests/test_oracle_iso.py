"""Isotope scattering (iso = 1) on the CPU: the oracle's restatement of P_iso_split,
calculate_tauRTA_iso, CollisionMatrix_iso and update_iso, and the library's element table
(alt3_isotope_g, host code), pinned against the reference build (tests/golden/ref_si4_iso.npz
and ref_zb4_iso.npz, written by make_golden.py iso from the reference routines).

Tolerances (relative to the largest magnitude of the compared array):
  g per element            bit-exact vs cal_g, all 83 elements of the table
  Piso                     1e-12 (the reference sums nk identical copies, then divides by nk)
  Q, QN, QU, tau, F_RTA, kappa_RTA, CM, RHS, update F2, the job-3 kappa history and
  final F, CG x, kappa_BTE 1e-12, iteration count equal
"""
import numpy as np
import pytest

import oracle as O
from conftest import load_fixture


def rel(a, b):
    a = np.asarray(a); b = np.asarray(b)
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


@pytest.fixture(scope="module")
def iso4():
    return load_fixture("si4_iso")


@pytest.fixture(scope="module")
def geo4(si_inputs):
    inp = si_inputs[4]
    nc = np.array(inp.params.nc)
    g, vol = O.reciprocal(inp.r123)
    opk = O.lattice_ops(inp.latparams, inp.primlatt)
    kpm = O.make_kp_reg(nc, g)
    nibz, kibz, _, mapibz, mapinv = O.get_weights(kpm, inp.primlatt, opk)
    kp, tet = O.make_kp_reg_tet(nc, g)
    return dict(inp=inp, nc=nc, g=g, vol=vol, opk=opk, nibz=nibz, kibz=kibz, mapibz=mapibz, mapinv=mapinv,
                kp=kp, tet=tet)


def test_isotope_g_vs_reference(iso4):
    """alt3_isotope_g (the library's host table) equals cal_g bit for bit for every element,
    a one-letter name padded with a blank, and a name without data (g = 0)."""
    from thermacond import capi
    names = [str(n) for n in iso4["calg_names"]]
    mine = np.array([capi.isotope_g(n) for n in names])
    assert np.array_equal(mine, iso4["calg"])
    assert capi.isotope_g("B") == capi.isotope_g("B ") and capi.isotope_g("Xx") == 0.0
    # a monoisotopic element has no mass variance; natural Si is the textbook 2.01e-4
    assert abs(capi.isotope_g("As")) < 1e-30 and abs(capi.isotope_g("Si") - 2.01e-4) < 1e-6


def test_giso_of_inputs(si_inputs, cells, tmp_path_factory, iso4):
    """cal_giso: g per atom from the element names and atom types of structure.params."""
    from conftest import inputs_for
    from thermacond import capi
    assert np.array_equal(capi.giso_of(si_inputs[4]), iso4["giso"])
    zb = load_fixture("zb4_iso")
    assert np.array_equal(capi.giso_of(inputs_for(cells["zb4"], 4, tmp_path_factory, "zb4g")), zb["giso"])


def test_piso_vs_reference(geo4, ref4, iso4):
    G = geo4
    freq, _ = O.distribution(ref4["eigenval"], 300.0)
    P = O.piso(0, G["nibz"] - 1, G["mapinv"], freq, G["tet"], ref4["eigenvec"], iso4["giso"])
    assert rel(P, iso4["Piso"]) < 1e-12
    # linear in g, and zero without isotopes
    P2 = O.piso(0, G["nibz"] - 1, G["mapinv"], freq, G["tet"], ref4["eigenvec"], 2 * iso4["giso"])
    assert rel(P2, 2 * P) < 1e-15
    assert not np.any(O.piso(0, G["nibz"] - 1, G["mapinv"], freq, G["tet"], ref4["eigenvec"], 0 * iso4["giso"]))


def test_piso_per_atom_vs_reference(cells, tmp_path_factory):
    """The Ga/As cell: g differs per atom (As has one isotope), nb = 12."""
    from conftest import inputs_for
    zb = load_fixture("zb4_iso")
    inp = inputs_for(cells["zb4"], 4, tmp_path_factory, "zb4p")
    nc = np.array(inp.params.nc)
    g, _ = O.reciprocal(inp.r123)
    kp, tet = O.make_kp_reg_tet(nc, g)
    freq, _ = O.distribution(zb["eigenval"], 300.0)
    mapinv = (zb["mapinv"] - 1).astype(np.int32)
    P = O.piso(0, len(mapinv) - 1, mapinv, freq, tet, zb["eigenvec"], zb["giso"])
    assert rel(P, zb["Piso"]) < 1e-12


def test_iso_path_vs_reference(geo4, ref4, iso4):
    """calculate_tauRTA_iso (Q + Piso in Q, QN and QU), Qandtau, F_RTA, kappa_RTA,
    CollisionMatrix_iso, RHS, CG and kappa_BTE, update_iso and the job-3 loop with it."""
    G = geo4
    E, V = ref4["eigenval"], ref4["veloc"]
    T = 300.0
    t = "T300_"
    freq, dist = O.distribution(E, T)
    Pi = iso4["Piso"]
    P1, P2, Q, QN, QU = O.tau(0, G["nibz"] - 1, G["mapinv"], G["kp"], G["nc"], G["g"], dist, ref4["P1sm"], ref4["P2sm"])
    Q, QN, QU = Q + Pi, QN + Pi, QU + Pi
    for k, v in (("Q", Q), ("QN", QN), ("QU", QU)):
        assert rel(v, iso4[t + k]) < 1e-12, k
    tN, tU, tt = O.qandtau(G["mapinv"], dist, QN, QU)
    assert rel(tt, iso4[t + "tautot"]) < 1e-12
    F = O.frta(G["mapinv"], freq, dist, V, Q, T)
    assert rel(F, iso4[t + "FRTA"]) < 1e-12
    args = (G["kibz"], G["mapinv"], G["kp"], G["nc"], G["g"], G["inp"].primlatt, G["opk"])
    kap, te = O.mode_kappa(*args, E, V, F, T, G["vol"])
    assert rel(kap, iso4[t + "kappaRTA"]) < 1e-12
    CM = O.collision(0, G["nibz"] - 1, G["kibz"], G["mapinv"], G["kp"], G["nc"], G["g"], G["inp"].primlatt, G["opk"],
                     P1, P2, Pi)
    assert rel(CM, iso4[t + "CM"]) < 1e-12
    # with Piso = 0 the older iso form (P1 at q1+q0, P2 in full) and CollisionMatrix_mpi (P1 at
    # q0-q1, P2 split between q0 and q1-q0) sum to the same star totals on this crystal
    CM0 = O.collision(0, G["nibz"] - 1, G["kibz"], G["mapinv"], G["kp"], G["nc"], G["g"], G["inp"].primlatt, G["opk"],
                      P1, P2, 0 * Pi)
    assert rel(CM0, ref4[t + "CM"]) < 1e-9
    R = O.rhs(G["mapinv"], freq, dist, V, T)
    assert rel(R, iso4[t + "RHS"]) < 1e-12
    x0 = np.array([F[i, l, a] for i in range(G["nibz"]) for l in range(6) for a in range(3)])
    x, _, _ = O.cg(CM, R, x0)
    assert rel(x, iso4[t + "x"]) < 1e-12
    kb, _ = O.mode_kappa_d(*args, E, V, x, T, G["vol"], te)
    assert rel(kb, iso4[t + "kappaBTE"]) < 1e-12
    uargs = (0, G["nibz"] - 1, G["kibz"], G["mapibz"], G["kp"], G["nc"], G["g"], G["inp"].primlatt, G["opk"], P1, P2)
    F2 = O.update(*uargs, F, F, Q, Pi)
    assert rel(F2, iso4[t + "F2first"]) < 1e-12
    Fi, hist, tes = O.bte_iter(G["kibz"], G["mapinv"], G["mapibz"], G["kp"], G["nc"], G["g"], G["inp"].primlatt,
                               G["opk"], E, V, P1, P2, F, Q, T, G["vol"], Piso=Pi)
    assert hist.shape == iso4[t + "kappaIter"].shape
    assert rel(hist, iso4[t + "kappaIter"]) < 1e-12 and rel(Fi, iso4[t + "Fiter"]) < 1e-12
    assert rel(tes, iso4[t + "taueffIter"]) < 1e-12


def test_update_iso_term_is_one_column(geo4, ref4, iso4):
    """update_iso = update_mpi + Piso(n1,l1) (op F(nibz, nb))(xyz) / Q: the difference of the
    two updates depends only on the last branch of the last irreducible point."""
    G = geo4
    T = 300.0
    _, dist = O.distribution(ref4["eigenval"], T)
    P1, P2, Q, _, _ = O.tau(0, G["nibz"] - 1, G["mapinv"], G["kp"], G["nc"], G["g"], dist, ref4["P1sm"], ref4["P2sm"])
    Pi = iso4["Piso"]
    uargs = (0, G["nibz"] - 1, G["kibz"], G["mapibz"], G["kp"], G["nc"], G["g"], G["inp"].primlatt, G["opk"], P1, P2)
    rng = np.random.default_rng(5)
    F = np.asfortranarray(rng.standard_normal((G["nibz"], 6, 3)))
    Z = np.zeros_like(F)
    d = O.update(*uargs, F, Z, Q, Pi) - O.update(*uargs, F, Z, Q)
    F[:-1] = rng.standard_normal(F[:-1].shape)
    F[-1, :-1] = rng.standard_normal(F[-1, :-1].shape)
    d2 = O.update(*uargs, F, Z, Q, Pi) - O.update(*uargs, F, Z, Q)
    assert rel(d2, d) < 1e-9 and np.max(np.abs(d)) > 0
