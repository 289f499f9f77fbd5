"""The split-job files this build writes, read back by the reference's own readers.

oracle/_ref/ref_readers (built from the reference sources by oracle/Makefile.ref) runs
read_Collision_Matrix (sub.f90:5303-5363, the Cs_sycollisionm.NNN.dat blocks under v3path),
read_CM_1D_mpi, read_RHS and read_DS (sub.f90:5366-5469, the job-5 inputs) and dumps what they
read.

CPU: files written by thermacond.splitio from seeded values (including values that need all 17
digits, signed zeros and subnormals) come back bit for bit.  read_Collision_Matrix opens unit
1004 for every block without closing it (its close is commented out); with this image's flang
the second open of the connected unit keeps the first file, so the reader is run on the
one-block layout (nibz <= ksub_size), whose records are those of every block.
GPU: the files the Fortran driver's job 4 writes (CM-1D.dat, RHS_IBZ.dat, DS.dat, 17 digits)
come back from the reference's readers bit for bit as written (the same values splitio parses),
and equal to the binding's alt3_collision matrix and right-hand side and alt3_kappa_rta's F_RTA
on the same inputs and temperature to 1e-13 (a separate run: the FGR's LDS atomics order its
FP64 sums differently from run to run at the 1e-16 level).
"""
import os
import subprocess

import numpy as np
import pytest

from conftest import REPO
from thermacond import build, capi, splitio

READERS = os.path.join(REPO, "oracle", "_ref", "ref_readers")


def _need():
    if not os.path.exists(READERS):
        pytest.skip("oracle/_ref/ref_readers is built only where the reference sources are")


def _run(wd, nibz, ndn, ksub, v3path="./"):
    with open(os.path.join(wd, "readers.txt"), "w") as f:
        f.write(f"{nibz} {ndn} {ksub}\n{v3path}\n")
    r = subprocess.run([READERS], cwd=wd, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0 and "ref_readers: done" in r.stdout, r.stdout + r.stderr
    n = nibz * ndn * 3
    out = {}
    for name, shape in (("cs", (n, n)), ("cm1d", (n, n)), ("rhs", (n,)), ("ds", (n,))):
        p = os.path.join(wd, name + ".bin")
        if os.path.exists(p):
            out[name] = np.fromfile(p, dtype="<f8").reshape(shape, order="F")
    return out


def _awkward(rng, shape):
    v = rng.standard_normal(shape) * 10.0 ** rng.integers(-30, 30, shape)
    flat = v.reshape(-1)
    flat[:4] = [0.0, -0.0, 5e-324, -2.2250738585072014e-308]
    flat[4] = np.nextafter(1.0, 2.0)
    return v


@pytest.mark.parametrize("nibz,ksub", [(5, 5), (11, 11)])
def test_collision_blocks_read_by_the_reference(tmp_path, nibz, ksub):
    _need()
    ndn = 6
    n = nibz * ndn * 3
    rng = np.random.default_rng(nibz * 100 + ksub)
    cm = _awkward(rng, (n, n))
    nfiles = nibz // ksub + nibz % ksub
    rows = ksub * ndn * 3
    os.makedirs(tmp_path / "v3", exist_ok=True)
    for b in range(nfiles):
        lo = b * rows
        hi = n if b == nfiles - 1 else lo + rows
        splitio.write_cs_sycollisionm(str(tmp_path / "v3" / splitio.block_name("Cs_sycollisionm", b + 1)), cm[lo:hi])
    got = _run(str(tmp_path), nibz, ndn, ksub, "v3/")["cs"]
    assert np.array_equal(got.view(np.uint64), cm.view(np.uint64))


def test_cg_system_read_by_the_reference(tmp_path):
    _need()
    nibz, ndn = 7, 6
    n = nibz * ndn * 3
    rng = np.random.default_rng(7)
    cm, rhs, frta = _awkward(rng, (n, n)), _awkward(rng, (n,)), _awkward(rng, (nibz, ndn, 3))
    splitio.write_cg_system(str(tmp_path), cm, rhs, frta)
    got = _run(str(tmp_path), nibz, ndn, 1)
    assert np.array_equal(got["cm1d"].view(np.uint64), cm.view(np.uint64))
    assert np.array_equal(got["rhs"].view(np.uint64), rhs.view(np.uint64))
    assert np.array_equal(got["ds"].view(np.uint64), frta.reshape(-1).view(np.uint64))


@pytest.mark.gpu
def test_driver_job4_files_read_by_the_reference(si_inputs, tmp_path):
    _need()
    print("case: driver job 4 files through the reference readers, SW-Si 4^3", flush=True)
    from thermacond import refio
    T = 300.0
    refio.write_inputs(str(tmp_path), si_inputs["fc"], refio.RunParams(nc=(4, 4, 4), tmin=T, tmax=T, job=4))
    r = subprocess.run([build.FORTRAN_EXE, "fp64"], cwd=tmp_path, capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    with capi.Context(si_inputs[4]) as c:
        c.eig(want=False)
        c.v3sq(0, c.nibz - 1, capi.V3_FP64, want=False)
        c.fgr(want=False)
        c.tau(T)
        _, F, _ = c.kappa_rta(T)
        CM, RHS = c.collision(T)
        nibz, nb = c.nibz, c.nb
    got = _run(str(tmp_path), nibz, nb, 1)
    n = nibz * nb * 3
    fcm, frhs, fds = splitio.read_cg_system(str(tmp_path), n)
    for k, v in (("cm1d", fcm), ("rhs", frhs), ("ds", fds)):
        assert np.array_equal(got[k].view(np.uint64), np.ascontiguousarray(v).view(np.uint64)), k
    for k, v in (("cm1d", CM), ("rhs", RHS), ("ds", np.asarray(F).reshape(-1))):
        d = np.abs(got[k] - v).max() / np.abs(v).max()
        assert d < 1e-13, (k, d)
