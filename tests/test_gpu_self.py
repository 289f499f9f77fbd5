"""GPU: the three-phonon self-energy of jobs 2/6/7 (alt3_self_w3, and the Gaussian form
alt3_self_w4) against the reference build
(tests/golden/ref_si{4,8}_self.npz: function_self_w3 with get_freq at points off the mesh,
T = 300 K, etaz of latdyn.params) with the reference's mesh eigen data loaded.

Tolerances: omega(q) 1e-7 (Jacobi vs ZHEEV); nself + uself summed over branches whose
frequencies agree to 1e-3 (per-branch values inside such a group depend on the eigenvector
gauge, the group sum does not) 2e-5, the spread the oracle shows against the reference with
LAPACK eigenvectors at the same points; normal and umklapp parts separately 2e-5 of the total.
"""
import numpy as np
import pytest

from conftest import load_fixture
from thermacond import capi

pytestmark = pytest.mark.gpu


def _groups(w, tol=1e-3):
    out, l = [], 0
    while l < len(w):
        u = l + 1
        while u < len(w) and abs(w[u] - w[l]) < tol * w.max():
            u += 1
        out.append((l, u))
        l = u
    return out


@pytest.mark.parametrize("n", [4, 8])
def test_self_w3_vs_reference(si_inputs, n):
    print(f"case: alt3_self_w3 SW-Si {n}^3 vs reference", flush=True)
    R = load_fixture(f"si{n}")
    S = load_fixture(f"si{n}_self")
    with capi.Context(si_inputs[n]) as c:
        c.set_eig(R["eigenval"], R["eigenvec"], R["veloc"])
        om, ns, us = c.self_w3(S["q"], float(S["tempk"]), float(S["etaz"]))
    assert np.max(np.abs(om - S["omega"])) / np.max(S["omega"]) < 1e-7
    tot, rtot = ns + us, S["nself"] + S["uself"]
    assert np.all(tot.imag > 0)
    for i in range(len(S["q"])):
        for a, b in _groups(S["omega"][:, i]):
            y = rtot[a:b, i].sum()
            assert abs(tot[a:b, i].sum() - y) / abs(y) < 2e-5, (i, a, b, tot[a:b, i].sum(), y)
            yn = S["nself"][a:b, i].sum()
            assert abs(ns[a:b, i].sum() - yn) / abs(y) < 2e-5, (i, a, b, "normal part")


@pytest.mark.parametrize("n", [4, 8])
def test_self_w4_vs_reference(si_inputs, n):
    """function_self_w4 (Gaussian deltas, imaginary part only) against the reference build:
    real parts exactly zero, group sums of nself + uself and of nself to 2e-5 of the total."""
    print(f"case: alt3_self_w4 SW-Si {n}^3 vs reference", flush=True)
    R = load_fixture(f"si{n}")
    S = load_fixture(f"si{n}_self")
    with capi.Context(si_inputs[n]) as c:
        c.set_eig(R["eigenval"], R["eigenvec"], R["veloc"])
        om, ns, us = c.self_w4(S["q"], float(S["tempk"]), float(S["etaz"]))
        with pytest.raises(capi.Alt3Error):
            c.self_w4(S["q"][:1], float(S["tempk"]), 0.0)
    assert np.max(np.abs(om - S["omega"])) / np.max(S["omega"]) < 1e-7
    assert not np.any(ns.real) and not np.any(us.real)
    tot, rtot = ns + us, S["nself4"] + S["uself4"]
    scale = np.max(np.abs(rtot))
    assert scale > 0
    for i in range(len(S["q"])):
        for a, b in _groups(S["omega"][:, i]):
            y = rtot[a:b, i].sum()
            d = max(abs(y), 1e-3 * scale)
            assert abs(tot[a:b, i].sum() - y) / d < 2e-5, (i, a, b, tot[a:b, i].sum(), y)
            yn = S["nself4"][a:b, i].sum()
            assert abs(ns[a:b, i].sum() - yn) / d < 2e-5, (i, a, b, "normal part")


def test_self_w3_batches_agree(si_inputs):
    """One call over many points equals one call per point (the batch split of the kernel)."""
    print("case: alt3_self_w3 batches 4^3", flush=True)
    S = load_fixture("si4_self")
    rng = np.random.default_rng(5)
    q = np.concatenate([S["q"], S["q"][:, :] * 0.93 + rng.uniform(-0.05, 0.05, S["q"].shape)] * 10)
    with capi.Context(si_inputs[4]) as c:
        c.eig(want=False)
        om, ns, us = c.self_w3(q, 300.0, float(S["etaz"]))
        for i in (0, 9, 77, len(q) - 1):
            o1, n1, u1 = c.self_w3(q[i], 300.0, float(S["etaz"]))
            assert np.allclose(o1[:, 0], om[:, i], rtol=1e-14, atol=0)
            assert np.allclose(n1[:, 0], ns[:, i], rtol=1e-12, atol=0)
            assert np.allclose(u1[:, 0], us[:, i], rtol=1e-12, atol=0)


def test_job2_band_path_files(si_inputs, tmp_path):
    """run_job2: kpbs.params -> path (make_kp_bs) -> lw-selfenergy.dat / lw-selfenergy2.dat with
    the linewidth 2 Im(nself + uself) of alt3_self_w3 at every (point, branch)."""
    print("case: job 2 driver, SW-Si 4^3", flush=True)
    from thermacond import refio, selfenergy
    refio.write_inputs(str(tmp_path), si_inputs["fc"], refio.RunParams(nc=(4, 4, 4)))
    (tmp_path / "kpbs.params").write_text("1\n5\n2\nA 0.11 0.13 0.17\nB 0.41 0.07 0.33\nC 0.2 0.3 0.45\n")
    r = selfenergy.run_job2(str(tmp_path))
    rows = np.loadtxt(tmp_path / "lw-selfenergy.dat")
    nb = r["omega"].shape[0]
    assert rows.shape == (10 * nb, 12)
    lw = 2 * np.imag(r["nself"] + r["uself"])
    assert np.allclose(rows[:, 9], lw.T.ravel(), rtol=1e-12, atol=0)
    assert np.allclose(rows[:, 7], r["omega"].T.ravel(), rtol=1e-12, atol=0)
    rows2 = np.loadtxt(tmp_path / "lw-selfenergy2.dat")
    assert rows2.shape == (10, 2 + 2 * nb)
    with capi.Context(si_inputs[4]) as c:
        c.eig(want=False)
        om, ns, us = c.self_w3(r["kp"][3], 300.0, 30.0)
    assert np.allclose(2 * np.imag(ns + us)[:, 0], lw[:, 3], rtol=1e-12, atol=0)
