"""Host-side logic (CPU only): the C ABI exports, the library's mesh/symmetry tables
(host-only context) against the reference's integer dumps, the force-constant generator
and the THERMACOND file formats."""
import ctypes as C
import os
import re

import numpy as np
import pytest

from thermacond import capi, models, refio

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_library_exports_every_declared_symbol():
    with open(os.path.join(REPO, "include", "alt3.h")) as f:
        hdr = f.read()
    declared = set(re.findall(r"^\s*(?:int|const char \*)\s*(alt3_\w+)\s*\(", hdr, re.M))
    assert declared, "no declarations parsed"
    assert declared == set(capi.EXPORTS)
    lib = C.CDLL(capi.LIB_PATH)
    for name in declared:
        assert hasattr(lib, name), name


@pytest.mark.parametrize("n", [4, 8])
def test_host_tables_bit_exact(si_inputs, ref4, ref8, n):
    R = ref4 if n == 4 else ref8
    with capi.Context(si_inputs[n], host_only=True) as c:
        assert c.nibz == int(R["nibz"]) and c.nops == 48 and c.nk == n ** 3
        m = c.mesh()
        assert np.array_equal(m["mapibz"] + 1, R["mapibz"])
        assert np.array_equal(m["mapinv"] + 1, R["mapinv"])
        assert np.array_equal(m["wibz"], R["wibz"])
        assert np.array_equal(m["tet"] + 1, R["tet"].T)
        assert np.array_equal(m["opk"], R["op_kmatrix"].transpose(2, 1, 0))
        assert np.array_equal(m["mappos"] + 1, R["mappos"]) and m["npos"] == int(R["npos"])
        assert np.abs(m["kpc_eig"] - R["kpc_mesh"].T).max() < 1e-14
        assert np.abs(m["kpc"] - R["kpc"].T).max() < 1e-14
        assert np.abs(m["kibz"] - R["kibz"].T).max() < 1e-14
        for i in range(c.nibz):
            ops, n0 = c.kstars(i)
            k = R["star_n"][i]
            assert np.array_equal(ops + 1, R["star_op"][:k, i]) and np.array_equal(n0 + 1, R["star_n0"][:k, i])


def test_host_only_context_has_no_device_state(si_inputs):
    with capi.Context(si_inputs[4], host_only=True) as c:
        with pytest.raises(capi.Alt3Error) as e:
            c.eig()
        assert e.value.code == -4


def test_odd_mesh_rejected(si_inputs):
    with pytest.raises(capi.Alt3Error) as e:
        capi.Context(si_inputs[4], nc=(5, 5, 5), host_only=True)
    assert e.value.code == -3


def test_generator_reproduces_fixture_force_constants(si_inputs, ref4):
    fc = si_inputs["fc"]
    assert np.array_equal(fc.fc2, ref4["fc2"]) and np.array_equal(fc.fc3, ref4["fc3"])
    assert np.array_equal(fc.atompos, ref4["atompos"])
    # acoustic sum rules of the finite-difference tensors
    for i in (1, 2):
        for al in (1, 2, 3):
            for be in (1, 2, 3):
                s = fc.fc2[(fc.fc2[:, 0] == i) & (fc.fc2[:, 1] == al) & (fc.fc2[:, 3] == be), 4].sum()
                assert abs(s) < 1e-9
    # FC3: the raw finite-difference tensor sums to zero over the force index exactly; the
    # (j,k) symmetrisation mixes in the displaced-atom index, whose sum vanishes only to the
    # O(h^2) finite-difference error (h = 0.01 A)
    f3 = fc.fc3
    for key in {(int(r[0]), int(r[1]), int(r[2]), int(r[3]), int(r[5])) for r in f3[:400]}:
        i, al, j, be, ga = key
        m = (f3[:, 0] == i) & (f3[:, 1] == al) & (f3[:, 2] == j) & (f3[:, 3] == be) & (f3[:, 5] == ga)
        assert abs(f3[m, 6].sum()) < 1e-5 * np.abs(f3[:, 6]).max()


def test_input_files_round_trip(tmp_path, si_inputs):
    fc = si_inputs["fc"]
    rp = refio.RunParams(nc=(6, 6, 6), tmin=123.0, tmax=456.0, classical=1)
    refio.write_inputs(str(tmp_path), fc, rp)
    inp = refio.read_inputs(str(tmp_path))
    assert tuple(inp.params.nc) == (6, 6, 6) and inp.params.tmin == 123.0 and inp.params.classical == 1
    assert np.array_equal(inp.fc3, fc.fc3[:, 6]) and np.array_equal(inp.fc2_atoms, fc.fc2[:, [0, 2]].astype(int))
    assert np.allclose(inp.atompos, fc.atompos, atol=1e-10)
    assert np.allclose(inp.r123, fc.crystal.prim_cart, atol=1e-10)


def test_fc_file_count_mismatch_raises(tmp_path, si_inputs):
    refio.write_inputs(str(tmp_path), si_inputs["fc"], refio.RunParams(nc=(4, 4, 4)))
    p = tmp_path / "fc3.dat"
    lines = p.read_text().splitlines()
    p.write_text("\n".join(lines[:-1]) + "\n")
    with pytest.raises(ValueError, match="READ_FC3"):
        refio.read_inputs(str(tmp_path))


def test_make_kp_bs_vs_reference_ge_example():
    """make_kp_bs (kpoints.f90:37-173) on the reference's Ge example: the 360 path points equal
    those of its bs_freqs.dat (written to 3 decimals).  That file's path coordinate is |q| per
    segment, while the source divides by |g01| (kpoints.f90:133): the two differ by exactly that
    factor, which is checked too."""
    from thermacond import selfenergy as se
    with np.load(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "ge_bandpath.npz"),
                 allow_pickle=False) as z:
        d = {k: z[k] for k in z.files}
    kp, dk, names = se.make_kp_bs(d["latparams"], d["primlatt"], str(d["kpbs"]))
    assert kp.shape == d["kp"].shape
    assert np.abs(kp - d["kp"]).max() < 6e-4
    g01 = np.linalg.norm(se.reciprocal_2pi(se.conventional_cell(d["latparams"]) @ d["primlatt"])[0])
    assert np.abs(dk * g01 - d["dk"]).max() < 2e-3
    assert names[0] == "L" and names[-1] == "G"


def test_set_iso_needs_device_state(si_inputs):
    """alt3_set_iso on a host-only context is refused like every device stage; a wrong
    number of g values is refused by the binding before any call."""
    with capi.Context(si_inputs[4], host_only=True) as c:
        with pytest.raises(capi.Alt3Error) as e:
            c.set_iso(capi.giso_of(si_inputs[4]))
        assert e.value.code == -4
        with pytest.raises(ValueError):
            c.set_iso([2e-4])
