"""The host C++ of libalt3 under AddressSanitizer and UndefinedBehaviorSanitizer.

host_tables.cpp (mesh, irreducible zone, stars, tetrahedra, negatives map, FC3 tables),
capi.cpp (argument checks, copies, footprint arithmetic) and isotopes.cpp are compiled with
-fsanitize=address,undefined for the host only (-Xarch_host; the device code of the other
sources is the regular build's) and linked with tests/sanitize/host_check.cpp, which drives the
host-only entry points (alt3_create with host_only = 1, alt3_sizes, alt3_get_mesh, alt3_kstars
for every irreducible point, alt3_footprint, bad arguments) on the cells of the workload set,
including a polar cell (Born charges); an odd mesh is refused with ALT3_E_MESH.  CPU only: any
sanitizer report fails the run, and the tables must equal those of the regular library.
"""
import os
import struct
import subprocess
from concurrent.futures import ThreadPoolExecutor

import numpy as np
import pytest

from conftest import REPO, inputs_for
from thermacond import build, capi, models

HIPCC = build.HIPCC
SAN = ["-Xarch_host", "-fsanitize=address", "-Xarch_host", "-fsanitize=undefined",
       "-Xarch_host", "-fno-sanitize-recover=all", "-Xarch_host", "-fno-omit-frame-pointer", "-g"]
HOST_SOURCES = ["host_tables.cpp", "capi.cpp", "isotopes.cpp"]


@pytest.fixture(scope="module")
def harness(tmp_path_factory):
    if not os.path.exists(HIPCC):
        pytest.skip("hipcc not found")
    d = str(tmp_path_factory.mktemp("asan"))
    devobjs = [os.path.join(build.OBJ_DIR, os.path.splitext(s)[0] + ".o") for s in build.SOURCES if s.endswith(".hip")]
    if not all(os.path.exists(o) for o in devobjs):
        build.build(verbose=False)
    base = [HIPCC] + [f for f in build.FLAGS if f != "-Wall"] + SAN + ["-x", "hip", "-c"]

    def cc(src):
        obj = os.path.join(d, os.path.basename(src) + ".o")
        r = subprocess.run(base + [src, "-o", obj], capture_output=True, text=True)
        assert r.returncode == 0, r.stderr[-3000:]
        return obj

    srcs = [os.path.join(build.CSRC, s) for s in HOST_SOURCES] + [os.path.join(REPO, "tests", "sanitize", "host_check.cpp")]
    with ThreadPoolExecutor(4) as ex:
        objs = list(ex.map(cc, srcs))
    exe = os.path.join(d, "host_check")
    r = subprocess.run([HIPCC, "--offload-arch=gfx950"] + SAN + ["-o", exe] + objs + devobjs,
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-3000:]
    return exe


def _write_input(path, inp):
    born = int(getattr(inp, "born_flag", 0))
    with open(path, "wb") as f:
        nc = inp.params.nc
        f.write(struct.pack("<9i", len(inp.masses), len(inp.iatomcell0), len(inp.fc2), len(inp.fc3),
                            nc[0], nc[1], nc[2], int(inp.params.classical), born))
        for a, t in ((inp.masses, "<f8"), (inp.r123, "<f8"), (inp.latparams, "<f8"),
                     (np.asfortranarray(inp.primlatt).reshape(-1, order="F"), "<f8"), (inp.atompos, "<f8"),
                     (inp.iatomcell0, "<i4"), (inp.fc2_atoms, "<i4"), (inp.fc2_xyz, "<i4"), (inp.fc2, "<f8"),
                     (inp.fc3_atoms, "<i4"), (inp.fc3_xyz, "<i4"), (inp.fc3, "<f8")):
            f.write(np.ascontiguousarray(a, dtype=t).tobytes())
        if born:
            f.write(np.ascontiguousarray(inp.epsil, dtype="<f8").tobytes())
            f.write(np.ascontiguousarray(inp.zeu, dtype="<f8").tobytes())


def _read_output(path):
    b = open(path, "rb").read()
    off = 0

    def take(dt, n):
        nonlocal off
        a = np.frombuffer(b, dtype=dt, count=n, offset=off)
        off += a.nbytes
        return a
    nk, nibz, nb, nops, npos = take("<i4", 5)
    out = dict(nk=nk, nibz=nibz, nb=nb, nops=nops, npos=npos, kpc_eig=take("<f8", 3 * nk), kpc=take("<f8", 3 * nk),
               mapibz=take("<i4", nk), mapinv=take("<i4", nibz), kibz=take("<f8", 3 * nibz), wibz=take("<f8", nibz),
               tet=take("<i4", 24 * nk), opk=take("<f8", 9 * nops), mappos=take("<i4", nk))
    stars = []
    for _ in range(nibz):
        n = int(take("<i4", 1)[0])
        stars.append((take("<i4", n), take("<i4", n)))
    out["stars"] = stars
    out["footprint"] = take("<i8", 6)
    assert off == len(b)
    return out


def _cases(si_inputs, cells, tmp_path_factory):
    from make_golden import polar_zb4
    from thermacond import refio
    fc, born = polar_zb4()
    d = str(tmp_path_factory.mktemp("asan_born"))
    refio.write_inputs(d, fc, refio.RunParams(nc=(4, 4, 4)), born=born)
    return {"sw-si 4^3": si_inputs[4], "sw-si 8^3": si_inputs[8],
            "zb4 6^3": inputs_for(cells["zb4"], 6, tmp_path_factory, "asan_zb4"),
            "rs8 4^3": inputs_for(cells["rs8"], 4, tmp_path_factory, "asan_rs8"),
            "rs16 2^3": inputs_for(cells["rs16"], 2, tmp_path_factory, "asan_rs16"),
            "zb4 polar 4^3": refio.read_inputs(d)}


def test_host_code_is_clean_under_asan_ubsan(harness, si_inputs, cells, tmp_path_factory):
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=0:exitcode=90",
               UBSAN_OPTIONS="print_stacktrace=1:halt_on_error=1")
    d = str(tmp_path_factory.mktemp("asan_io"))
    for name, inp in _cases(si_inputs, cells, tmp_path_factory).items():
        fin, fout = os.path.join(d, "in.bin"), os.path.join(d, "out.bin")
        _write_input(fin, inp)
        r = subprocess.run([harness, fin, fout], capture_output=True, text=True, env=env, timeout=600)
        assert r.returncode == 0, f"{name}: rc {r.returncode}\n{r.stderr[-4000:]}"
        assert "ERROR: AddressSanitizer" not in r.stderr and "runtime error" not in r.stderr, r.stderr[-4000:]
        got = _read_output(fout)
        with capi.Context(inp, host_only=True) as h:
            m = h.mesh()
            assert (got["nk"], got["nibz"], got["nb"], got["nops"]) == (h.nk, h.nibz, h.nb, h.nops), name
            assert np.array_equal(got["mapinv"], m["mapinv"]) and np.array_equal(got["mapibz"], m["mapibz"]), name
            assert np.array_equal(got["kpc"].reshape(-1, 3), m["kpc"]), name
            assert np.array_equal(got["tet"].reshape(-1, 4), m["tet"]), name
            for i in range(h.nibz):
                ops, n0 = h.kstars(i)
                assert np.array_equal(got["stars"][i][0], ops) and np.array_equal(got["stars"][i][1], n0), (name, i)
            fp = [h.footprint(nt, p) for nt in (1, 3, 7) for p in (capi.V3_FP64, capi.V3_MIXED)]
            assert list(got["footprint"]) == fp, name
    # an odd mesh: refused in alt3_create, nothing else runs
    fin, fout = os.path.join(d, "in.bin"), os.path.join(d, "out.bin")
    _write_input(fin, inputs_for(si_inputs["fc"], 5, tmp_path_factory, "asan_si5"))
    r = subprocess.run([harness, fin, fout], capture_output=True, text=True, env=env, timeout=600)
    assert r.returncode == 3 and "mesh must be even" in r.stderr, (r.returncode, r.stderr[-2000:])
