"""alt3_cg_static, the entry point that allocates no device memory (static scratch), against
alt3_cg through the C ABI: the same kernels in the same order, so x, the iteration count and
funcmin are identical (bit for bit).  Two calls in a row with external matrices, the second
smaller (the scratch keeps the first call's contents), then the context's own collision
matrix (A = NULL), then a size beyond the scratch, refused on the host."""
import numpy as np
import pytest

from thermacond import capi

pytestmark = pytest.mark.gpu


def _spd(n, seed):
    rng = np.random.default_rng(seed)
    m = rng.standard_normal((n, n))
    return m @ m.T / n + np.eye(n), rng.standard_normal(n), rng.standard_normal(n)


def test_static_cg_equals_cg(si_inputs):
    print("case: alt3_cg_static vs alt3_cg, n = 600 then 200, then the collision matrix", flush=True)
    with capi.Context(si_inputs[4]) as c:
        for n, seed in ((600, 1), (200, 2)):
            A, b, x0 = _spd(n, seed)
            xs, its, fs = c.cg_static(x0, A, b)
            xr, itr, fr = c.cg(x0, A, b)
            assert its == itr and its > 1
            assert np.array_equal(xs, xr), (n, float(np.abs(xs - xr).max()))
            assert fs == fr
            assert np.abs(A @ xs - b).max() < 1e-2 * np.abs(b).max()
        c.eig(want=False)
        c.v3sq(0, c.nibz - 1, capi.V3_FP64, want=False)
        c.fgr(want=False)
        c.tau(300.0)
        _, F, _ = c.kappa_rta(300.0)
        c.collision(300.0, want=False)
        x0 = np.array([F[i, l, a] for i in range(c.nibz) for l in range(c.nb) for a in range(3)])
        xs, its, fs = c.cg_static(x0)
        xr, itr, fr = c.cg(x0)
        assert its == itr and np.array_equal(xs, xr) and fs == fr
        big = np.zeros(2506)
        with pytest.raises(capi.Alt3Error, match="static scratch"):
            c.cg_static(big, np.eye(2506), big)
