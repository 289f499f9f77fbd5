"""Meshes beyond 16^3.  Up to nk = 8192 (20^3) the tetrahedron FGR holds the corner energies and
weight sums of the whole mesh in LDS (12 bytes per mesh point, 16-bit tetrahedron codes);
beyond, k_fgr_big keeps them in a per-workgroup slice of global scratch.

GPU, SW-Si at 20^3 (nk = 8000, LDS kernel), 22^3 and 30^3 (nk = 10648 and 27000, global
scratch): |V3|^2 of the first irreducible q1 against the oracle (1e-8), P1sm/P2sm of the first
two q1 against the oracle's calculate_FGR_mpi2 on the device's eigen data and |V3|^2 (1e-8);
the whole path in FP64 and mixed precision at 20^3 and 30^3 (kappa within 1e-5 of each other).
"""
import numpy as np
import pytest

import oracle as O
from conftest import inputs_for
from thermacond import capi


def rel(a, b):
    return float(np.max(np.abs(np.asarray(a) - np.asarray(b))) / max(np.max(np.abs(b)), 1e-300))


@pytest.fixture(scope="module")
def si20(si_inputs, tmp_path_factory):
    return inputs_for(si_inputs["fc"], 20, tmp_path_factory, "si20")


@pytest.fixture(scope="module")
def si30(si_inputs, tmp_path_factory):
    return inputs_for(si_inputs["fc"], 30, tmp_path_factory, "si30")


def _fgr_vs_oracle(inp, nk):
    nc = np.array(inp.params.nc)
    g, _ = O.reciprocal(inp.r123)
    kp, tet = O.make_kp_reg_tet(nc, g)
    with capi.Context(inp) as c:
        assert c.nk == nk
        ev, vec, vel = c.eig()
        mapinv = c.mesh()["mapinv"]
        v = c.v3sq(0, 1, capi.V3_FP64)
        o = O.v3sq(0, 0, mapinv, kp, nc, g, ev, vec, inp.masses, inp.atompos, inp.iatomcell0,
                   inp.fc3_atoms, inp.fc3_xyz, inp.fc3)
        assert rel(v[:1], o) < 1e-8
        P1, P2 = c.fgr()
        fgr_ms = c.kernel_ms()[1]
    freq, _ = O.distribution(ev, 300.0)
    q1, q2 = O.fgr(0, 1, mapinv, kp, nc, g, freq, tet, v)
    assert rel(P1, q1) < 1e-8 and rel(P2, q2) < 1e-8
    print(f"  FGR kernel {fgr_ms:.2f} ms for two q1", flush=True)


@pytest.mark.gpu
def test_fgr_20cubed_vs_oracle(si20):
    print("case: SW-Si 20^3 FGR vs oracle (LDS kernel)", flush=True)
    _fgr_vs_oracle(si20, 8000)


@pytest.mark.gpu
def test_fgr_22cubed_vs_oracle(si_inputs, tmp_path_factory):
    print("case: SW-Si 22^3 FGR vs oracle (global-scratch kernel)", flush=True)
    _fgr_vs_oracle(inputs_for(si_inputs["fc"], 22, tmp_path_factory, "si22"), 10648)


@pytest.mark.gpu
def test_fgr_30cubed_vs_oracle(si30):
    print("case: SW-Si 30^3 FGR vs oracle (global-scratch kernel)", flush=True)
    _fgr_vs_oracle(si30, 27000)


def _mixed_vs_fp64(inp):
    with capi.Context(inp) as c:
        o64 = c.solve([300.0], capi.V3_FP64)
        om = c.solve([300.0], capi.V3_MIXED)
        print(f"  stage ms FP64 {np.round(o64['stage_ms'], 2)}, mixed {np.round(om['stage_ms'], 2)}", flush=True)
    for k in ("kappa_rta", "kappa_bte"):
        a, b = np.trace(om[k][0].sum(0)) / 3, np.trace(o64[k][0].sum(0)) / 3
        assert abs(a / b - 1) < 1e-5, (k, a, b)


@pytest.mark.gpu
def test_solve_20cubed_mixed_vs_fp64(si20):
    print("case: SW-Si 20^3 solve, FP64 and mixed", flush=True)
    _mixed_vs_fp64(si20)


@pytest.mark.gpu
def test_solve_30cubed_mixed_vs_fp64(si30):
    print("case: SW-Si 30^3 solve, FP64 and mixed", flush=True)
    _mixed_vs_fp64(si30)
