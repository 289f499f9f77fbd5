"""The batched SW-Si sweep and the shared persistent |V3|^2 launch (alt3_v3_batch).

The sweep: 64 instances, seeds 0..63, lattice constant ~ U[5.4, 5.9] A and mesh from
{8,10,12,14}^3 per seed (workloads.sweep_draw), temperatures 100, 200, 300, 500 and 800 K,
mixed V3, RTA and CG kappa.  Force constants from tests/golden/sweep_fc.npz (make_sweep.py).

CPU: the pack holds seeds 0..63 with the draws of sweep_draw, all four meshes occur, and
instance 0 regenerated by models.sw_silicon equals the pack bit for bit.
GPU: every instance of the packed sweep (all 64 resident, |V3|^2 of all of them in one
persistent launch, autotuned tile) gives kappa_RTA, kappa_BTE and CG iteration counts equal
to its lone solve (the tile is computed by the same code; the FP32 FGR adds tetrahedron
weights into double W with atomics whose order varies from run to run, a last-bit change of
W flips the float rounding of a P element now and then, and two lone solves of one instance
differ by up to 4e-11 in kappa_RTA and 3e-10 in kappa_BTE: relative 1e-9 and 1e-8); two 8^3 instances through the oracle chain (eigen 1e-10,
|V3|^2 of two q1 1e-8, P 1e-8, kappa_RTA 1e-8, kappa_BTE 1e-5) and their packed mixed kappa
within 1e-5 of FP64; |V3|^2 after a batched launch equal to the lone path's bit for bit
(and P1/P2 to 1e-12, the FGR's atomics order) for every tile candidate, FP64 and mixed, 6
branches (SW-Si) and 12 (the 4-atom cell); cells of 24 branches run one launch per context
(tile 0) with the lone results.
"""
import os

import numpy as np
import pytest

import oracle as O
from conftest import GOLDEN, inputs_for
from thermacond import capi, models, workloads

FC_FILE = os.path.join(GOLDEN, "sweep_fc.npz")
TEMPS = list(workloads.SWEEP_TEMPS)


def rel(a, b):
    a = np.asarray(a); b = np.asarray(b)
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


@pytest.fixture(scope="module")
def pack():
    with np.load(FC_FILE, allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


@pytest.fixture(scope="module")
def sweep(pack):
    return workloads.sweep_instances(64, 0, FC_FILE)


def test_pack_holds_the_draws(pack):
    assert list(pack["seeds"]) == list(range(64))
    meshes = set()
    for s in range(64):
        a, n = workloads.sweep_draw(s)
        assert tuple(pack[f"s{s}_a_n"]) == (a, n)
        assert 5.4 <= a <= 5.9 and n in workloads.SWEEP_MESHES
        meshes.add(n)
    assert meshes == set(workloads.SWEEP_MESHES)


def test_pack_equals_generator(pack):
    a, _ = workloads.sweep_draw(0)
    fc = models.sw_silicon(a, batched=True)
    got = workloads.sweep_fc_unpack(pack, 0)
    for k in ("atompos", "iatomcell0", "iatomcell", "fc2", "fc3"):
        assert np.array_equal(getattr(got, k), getattr(fc, k)), k
    assert got.crystal.lattice_parameter == fc.crystal.lattice_parameter
    assert np.array_equal(got.crystal.primlatt, fc.crystal.primlatt)


@pytest.mark.gpu
def test_sweep_packed_equals_lone(sweep):
    print("case: SW-Si sweep, 64 instances packed vs lone", flush=True)
    inps = [x[3] for x in sweep]
    r = workloads.solve_packed(inps, TEMPS, capi.V3_MIXED, log=print)
    assert r["tile"] in (16, 32, 64)
    for (s, a, n, inp), o in zip(sweep, r["results"]):
        with capi.Context(inp) as c:
            lone = c.solve(TEMPS, capi.V3_MIXED)
        for k, tol in (("kappa_rta", 1e-9), ("kappa_bte", 1e-8)):
            assert rel(o[k], lone[k]) <= tol, (s, a, n, k, rel(o[k], lone[k]))
        assert np.array_equal(o["cg_iters"], lone["cg_iters"]), (s, o["cg_iters"], lone["cg_iters"])


def _geo(inp):
    nc = np.array(inp.params.nc)
    g, vol = O.reciprocal(inp.r123)
    opk = O.lattice_ops(inp.latparams, inp.primlatt)
    kpm = O.make_kp_reg(nc, g)
    nibz, kibz, wibz, mapibz, mapinv = O.get_weights(kpm, inp.primlatt, opk)
    kp, tet = O.make_kp_reg_tet(nc, g)
    model = (inp.masses, inp.atompos, inp.iatomcell0, inp.fc2_atoms, inp.fc2_xyz, inp.fc2)
    return dict(nc=nc, g=g, vol=vol, opk=opk, kpm=kpm, nibz=nibz, kibz=kibz, mapinv=mapinv, kp=kp, tet=tet, model=model)


@pytest.mark.gpu
def test_sweep_8cubed_oracle_chain(sweep):
    eight = [x for x in sweep if x[2] == 8][:2]
    assert len(eight) == 2
    T = 300.0
    for s, a, n, inp in eight:
        print(f"case: sweep instance {s} (a = {a}, 8^3) through the oracle chain", flush=True)
        G = _geo(inp)
        with capi.Context(inp) as c:
            ev, vec, vel = c.eig()
            oev, _, _ = O.eigen(G["kpm"], G["model"])
            assert rel(ev, oev) < 1e-10
            v2 = c.v3sq(0, 1, capi.V3_FP64)
            o2 = O.v3sq(0, 1, G["mapinv"], G["kp"], G["nc"], G["g"], ev, vec, inp.masses, inp.atompos,
                        inp.iatomcell0, inp.fc3_atoms, inp.fc3_xyz, inp.fc3)
            assert rel(v2, o2) < 1e-8
            v = c.v3sq(0, c.nibz - 1, capi.V3_FP64)
            P1, P2 = c.fgr()
            c.tau(T)
            kap, F, te = c.kappa_rta(T)
            c.collision(T, want=False)
            x0 = np.array([F[i, l, k] for i in range(c.nibz) for l in range(c.nb) for k in range(3)])
            x, it, _ = c.cg(x0)
            kb, _ = c.kappa_bte(T, x, te)
        freq, dist = O.distribution(ev, T)
        q1, q2 = O.fgr(0, G["nibz"] - 1, G["mapinv"], G["kp"], G["nc"], G["g"], freq, G["tet"], v)
        assert rel(P1, q1) < 1e-8 and rel(P2, q2) < 1e-8
        p1, p2, Q, QN, QU = O.tau(0, G["nibz"] - 1, G["mapinv"], G["kp"], G["nc"], G["g"], dist, q1, q2)
        Fo = O.frta(G["mapinv"], freq, dist, vel, Q, T)
        ko, teo = O.mode_kappa(G["kibz"], G["mapinv"], G["kp"], G["nc"], G["g"], inp.primlatt, G["opk"], ev, vel,
                               Fo, T, G["vol"])
        assert rel(kap, ko) < 1e-8
        CM = O.collision(0, G["nibz"] - 1, G["kibz"], G["mapinv"], G["kp"], G["nc"], G["g"], inp.primlatt, G["opk"],
                         p1, p2)
        R = O.rhs(G["mapinv"], freq, dist, vel, T)
        xo, ito, _ = O.cg(CM, R, np.array([Fo[i, l, k] for i in range(G["nibz"]) for l in range(6) for k in range(3)]))
        assert it == ito
        kbo, _ = O.mode_kappa_d(G["kibz"], G["mapinv"], G["kp"], G["nc"], G["g"], inp.primlatt, G["opk"], ev, vel, xo,
                                T, G["vol"], teo)
        assert rel(kb, kbo) < 1e-5
    # the two instances packed, mixed V3: kappa within 1e-5 of the FP64 solve
    pk = workloads.solve_packed([x[3] for x in eight], [T], capi.V3_MIXED)
    for (s, a, n, inp), o in zip(eight, pk["results"]):
        with capi.Context(inp) as c:
            f = c.solve([T], capi.V3_FP64)
        for k in ("kappa_rta", "kappa_bte"):
            m, d = np.trace(o[k][0].sum(0)) / 3, np.trace(f[k][0].sum(0)) / 3
            assert abs(m / d - 1) < 1e-5, (s, k, m, d)


def _after_batch(inps, prec, tile):
    """(tile used, [(|V3|^2, P1, P2)] per context) after one batched launch."""
    ctxs = [capi.Context(inp) for inp in inps]
    try:
        for c in ctxs:
            c.eig(want=False)
        used = capi.v3_batch(ctxs, prec, tile)
        return used, [(c.get_v3sq(),) + c.fgr() for c in ctxs]
    finally:
        for c in ctxs:
            c.close()


def _lone(inp, prec):
    with capi.Context(inp) as c:
        c.eig(want=False)
        v = c.v3sq(0, c.nibz - 1, prec)
        return (v,) + c.fgr()


def _same(got, lone, what):
    """|V3|^2 bit for bit; P1/P2 to 1e-12 (the FGR accumulates tetrahedron weights with LDS
    atomics, whose order varies from run to run at the 1e-16 level)."""
    for (v, p1, p2), (w, q1, q2) in zip(got, lone):
        assert np.array_equal(v, w), what
        assert rel(p1, q1) <= 1e-12 and rel(p2, q2) <= 1e-12, (what, rel(p1, q1), rel(p2, q2))


@pytest.mark.gpu
@pytest.mark.parametrize("prec", [capi.V3_FP64, capi.V3_MIXED])
def test_batch_tiles_bit_identical_6_branches(sweep, prec):
    print(f"case: batched V3, 6 branches, precision {prec}, every tile", flush=True)
    small = [x[3] for x in sweep if x[2] in (8, 10)][:3]
    lone = [_lone(inp, prec) for inp in small]
    for tile in (16, 32, 64):
        used, got = _after_batch(small, prec, tile)
        assert used == tile
        _same(got, lone, tile)


@pytest.mark.gpu
@pytest.mark.parametrize("prec", [capi.V3_FP64, capi.V3_MIXED])
def test_batch_tiles_bit_identical_12_branches(cells, tmp_path_factory, prec):
    print(f"case: batched V3, 12 branches, precision {prec}, every tile", flush=True)
    inps = [inputs_for(cells["zb4"], n, tmp_path_factory, f"zb4b{prec}") for n in (4, 6)]
    lone = [_lone(inp, prec) for inp in inps]
    ran = 0
    for tile in (0, 8, 16, 32):
        try:
            used, got = _after_batch(inps, prec, tile)
        except capi.Alt3Error as e:   # FP64 tiles of 32 points exceed the LDS of a workgroup
            assert tile == 32 and prec == capi.V3_FP64 and "does not fit in LDS" in str(e), (tile, str(e))
            continue
        assert used in (8, 16, 32) and (tile == 0 or used == tile)
        _same(got, lone, tile)
        ran += 1
    assert ran >= 3
    # contexts of different branch counts do not share a launch
    with capi.Context(inps[0]) as c12, capi.Context(cells_si4(tmp_path_factory)) as c6:
        c12.eig(want=False)
        c6.eig(want=False)
        with pytest.raises(capi.Alt3Error, match="same branch count"):
            capi.v3_batch([c12, c6], prec)


def cells_si4(tmp_path_factory):
    return inputs_for(models.sw_silicon(), 4, tmp_path_factory, "si4b")


@pytest.mark.gpu
def test_batch_24_branches_per_context(cells, tmp_path_factory):
    print("case: batched V3 request on 24 branches (one launch per context)", flush=True)
    inp = inputs_for(cells["rs8"], 4, tmp_path_factory, "rs8b")
    lone = _lone(inp, capi.V3_MIXED)
    used, got = _after_batch([inp], capi.V3_MIXED, 0)
    assert used == 0
    _same(got, [lone], "24 branches")


def test_batch_argument_checks(si_inputs):
    """Argument checks run before any device work: host-only contexts and an empty list are
    refused."""
    with capi.Context(si_inputs[4], host_only=True) as a:
        with pytest.raises(capi.Alt3Error):
            capi.v3_batch([a], capi.V3_MIXED)
    with pytest.raises(capi.Alt3Error):
        capi.v3_batch([], capi.V3_MIXED)
