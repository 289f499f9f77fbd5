"""The two matrix-core forms of the mixed-precision V3 on the 24-branch rock-salt-type cell:
split FP16 (ALT3_MMA_F16X3, k_v3_chain16: every FP32 operand as hi + lo halves, three FP16
products per product) and FP32 (ALT3_MMA_F32, k_v3_chain).

GPU, on the same eigenvectors:
* |V3|^2 of both forms against the FP64 path: at most 1e-6 of the largest element anywhere but
  the triples with an acoustic mode at Gamma (round-off of the sum rule in every precision),
  and a relative error below 1e-4 on every element above 1e-4 of the largest;
* kappa_RTA and kappa_BTE within 1e-6 of the FP64 path (100 and 300 K);
* 4^3 (one workgroup of 64 points) and 6^3 (216 points: a workgroup with a partial chunk);
* the range handling of the FP16 form: with the cubic force constants scaled by 2^20 or
  2^-30 (G' would overflow or underflow FP16 unscaled) |V3|^2 is the unscaled result times
  2^40 or 2^-60, bit for bit, since the power-of-two scales of G' absorb the factor exactly
  (at 2^-60 but for elements that leave FP32's normal range in the final product);
* the mixed path's P1sm/P2sm, whose zero rows the FGR does not store (the P row map), copied
  out equal to the FP64 path's (1e-5 of the largest element, zero rows exactly zero).
"""
import dataclasses

import numpy as np
import pytest

from conftest import inputs_for
from thermacond import capi


def _gamma_acoustic(c, hi):
    """Mask (ksub, nk, nb, nb, nb) of the triples with an acoustic mode at Gamma."""
    nc = c.nc
    k = np.arange(c.nk)
    c2, c1, c0 = k % nc[2], (k // nc[2]) % nc[1], k // (nc[1] * nc[2])
    neg = ((nc[2] - c2) % nc[2]) + nc[2] * (((nc[1] - c1) % nc[1]) + nc[1] * ((nc[0] - c0) % nc[0]))
    gam = (nc[2] // 2) + nc[2] * ((nc[1] // 2) + nc[1] * (nc[0] // 2))
    mapinv = c.mesh()["mapinv"]
    ac = np.arange(c.nb) < 3
    m = np.zeros((hi + 1, c.nk, c.nb, c.nb, c.nb), bool)
    for s in range(hi + 1):
        k1 = mapinv[s]
        if k1 == gam:
            m[s] |= ac[None, :, None, None]
        m[s] |= (k == gam)[:, None, None, None] & ac[None, None, :, None]
        m[s] |= (k == neg[k1])[:, None, None, None] & ac[None, None, None, :]
    return m


def _v3_errors(v, ref, mask):
    mx = np.abs(ref).max()
    d = np.abs(v - ref)
    big = np.abs(ref) > 1e-4 * mx
    return float(d[~mask].max() / mx), float((d[big & ~mask] / np.abs(ref[big & ~mask])).max())


def _kiso(k):
    return np.trace(k.sum(0)) / 3


@pytest.mark.gpu
@pytest.mark.parametrize("n", [4, 6])
def test_mma_forms_vs_fp64(cells, tmp_path_factory, n):
    print(f"case: rock-salt {n}^3 mixed V3, F16X3 and F32 matrix cores vs FP64", flush=True)
    inp = inputs_for(cells["rs8"], n, tmp_path_factory, f"rs8mma{n}")
    temps = [100.0, 300.0]
    with capi.Context(inp) as c:
        c.eig(want=False)
        hi = min(3, c.nibz - 1)
        vf = c.v3sq(0, hi, capi.V3_FP64)
        mask = _gamma_acoustic(c, hi)
        res = {}
        for kind in (capi.MMA_F32, capi.MMA_F16X3):
            c.set_v3_mma(kind)
            v = c.v3sq(0, hi, capi.V3_MIXED)
            assert np.isfinite(v).all()
            res[kind] = _v3_errors(v, vf, mask)
            print(f"  kind {kind}: |V3|^2 max abs/largest {res[kind][0]:.2e}, max rel (>1e-4) {res[kind][1]:.2e}")
            assert res[kind][0] < 1e-6 and res[kind][1] < 1e-4, res[kind]
        o64 = c.solve(temps, capi.V3_FP64)
        for kind in (capi.MMA_F32, capi.MMA_F16X3):
            c.set_v3_mma(kind)
            om = c.solve(temps, capi.V3_MIXED)
            for i in range(len(temps)):
                for key in ("kappa_rta", "kappa_bte"):
                    a, b = _kiso(om[key][i]), _kiso(o64[key][i])
                    print(f"  kind {kind} {key} {temps[i]} K: rel {abs(a / b - 1):.2e}")
                    assert abs(a / b - 1) < 1e-6, (kind, key, temps[i], a, b)


@pytest.mark.gpu
def test_f16x3_scale_invariance(cells, tmp_path_factory):
    print("case: rock-salt 4^3 F16X3 with FC3 scaled by 2^20 and 2^-30", flush=True)
    inp = inputs_for(cells["rs8"], 4, tmp_path_factory, "rs8scale")
    out = {}
    for e in (0, 20, -30):
        sc = dataclasses.replace(inp, fc3=inp.fc3 * 2.0 ** e)
        with capi.Context(sc) as c:
            c.eig(want=False)
            c.set_v3_mma(capi.MMA_F16X3)
            out[e] = c.v3sq(0, 1, capi.V3_MIXED)
    for e in (20, -30):
        ref = out[0] * 2.0 ** (2 * e)
        mx = np.abs(ref).max()
        # bit for bit but where the FP32 result itself falls below FP32's normal range (2^-60
        # takes elements under 1e-22 of the largest there)
        big = np.abs(ref) > 1e-20 * mx
        assert np.array_equal(out[e][big], ref[big]), e
        assert np.abs(out[e] - ref).max() <= 1e-25 * mx, e


@pytest.mark.gpu
def test_mixed_p_rows_vs_fp64(cells, tmp_path_factory):
    """The FP32 FGR leaves the rows of branches l1 outside an item's energy range unstored (the
    P row map); copied out they must read as zeros, and every row must match the FP64 path's."""
    print("case: rock-salt 4^3 P1sm/P2sm mixed (row map) vs FP64", flush=True)
    inp = inputs_for(cells["rs8"], 4, tmp_path_factory, "rs8prow")
    with capi.Context(inp) as c:
        c.eig(want=False)
        c.v3sq(0, c.nibz - 1, capi.V3_FP64, want=False)
        P1d, P2d = c.fgr()
        c.v3sq(0, c.nibz - 1, capi.V3_MIXED, want=False)
        P1m, P2m = c.fgr()
    for a, b, name in ((P1m, P1d, "P1sm"), (P2m, P2d, "P2sm")):
        assert np.isfinite(a).all(), name
        mx = np.abs(b).max()
        assert np.abs(a - b).max() < 1e-5 * mx, (name, np.abs(a - b).max() / mx)
        # rows (s, l1, l2, l3) that are zero on the FP64 path are zero on the mixed path
        zd = np.all(b == 0, axis=1)
        assert np.all(a.transpose(0, 2, 3, 4, 1)[zd] == 0), name
        print(f"  {name}: max diff {np.abs(a - b).max() / mx:.2e} of the largest, {zd.mean():.0%} zero rows")
