"""The reference's split formats (thermacond/splitio.py): round trips, and the unformatted
layouts against files written and read by Fortran itself (the image's amdflang, with the
record statements of sub.f90:7161-7177 and :3371-3377, and the list-directed reads of
read_CM_1D_mpi / read_RHS / read_DS).  CPU only."""
import os
import shutil
import subprocess

import numpy as np
import pytest

from thermacond import splitio

FLANG = shutil.which("amdflang") or "/opt/rocm/bin/amdflang"

PROG = """
program t
  implicit none
  integer, parameter :: ks=2, nk=3, nd=2
  real(8) :: p1(ks,nk,nd,nd,nd), p2(ks,nk,nd,nd,nd), cs(4,5), r1, r2, x
  integer :: n1,n2,l1,l2,l3,nv,i,j,bad
  do n1=1,ks; do n2=1,nk; do l1=1,nd; do l2=1,nd; do l3=1,nd
    p1(n1,n2,l1,l2,l3) = n1 + 10*n2 + 100*l1 + 1000*l2 + 10000*l3 + 0.25d0
    p2(n1,n2,l1,l2,l3) = -p1(n1,n2,l1,l2,l3) / 3
  enddo; enddo; enddo; enddo; enddo
  do i=1,4; do j=1,5; cs(i,j) = i + j/7d0; enddo; enddo
  open(11,file='v33sq_delta.fortran.dat',form='unformatted',status='replace')
  write(11) ks*nk*nd**3
  do n1=1,ks; do l1=1,nd; do n2=1,nk; do l2=1,nd; do l3=1,nd
    write(11) p1(n1,n2,l1,l2,l3), p2(n1,n2,l1,l2,l3)
  enddo; enddo; enddo; enddo; enddo
  close(11)
  open(12,file='Cs.fortran.dat',form='unformatted',status='replace')
  do i=1,4; do j=1,5; write(12) cs(i,j); enddo; enddo
  close(12)
  ! read the Python-written files back as the reference does
  bad = 0
  open(13,file='v33sq_delta.python.dat',form='unformatted',status='old')
  read(13) nv
  if (nv /= ks*nk*nd**3) bad = bad + 1
  do n1=1,ks; do l1=1,nd; do n2=1,nk; do l2=1,nd; do l3=1,nd
    read(13) r1, r2
    if (r1 /= p1(n1,n2,l1,l2,l3) .or. r2 /= p2(n1,n2,l1,l2,l3)) bad = bad + 1
  enddo; enddo; enddo; enddo; enddo
  close(13)
  open(14,file='CM-1D.dat',status='old')
  do i=1,4; do j=1,5; read(14,*) x; if (x /= cs(i,j)) bad = bad + 1; enddo; enddo
  close(14)
  write(*,*) 'BAD', bad
end program t
"""


def _expected():
    ks, nk, nd = 2, 3, 2
    n1, n2, l1, l2, l3 = np.meshgrid(*(np.arange(1, k + 1) for k in (ks, nk, nd, nd, nd)), indexing="ij")
    p1 = n1 + 10 * n2 + 100 * l1 + 1000 * l2 + 10000 * l3 + 0.25
    cs = np.add.outer(np.arange(1, 5), np.arange(1, 6) / 7.0)
    return p1.astype(np.float64), -p1 / 3, cs


def test_round_trips(tmp_path):
    rng = np.random.default_rng(0)
    P1 = np.asfortranarray(rng.standard_normal((3, 8, 2, 2, 2))); P2 = np.asfortranarray(rng.standard_normal(P1.shape))
    f = str(tmp_path / splitio.block_name("v33sq_delta", 1))
    assert f.endswith("v33sq_delta.001.dat")
    splitio.write_v33sq_delta(f, P1, P2)
    nv, Q1, Q2 = splitio.read_v33sq_delta(f, 3, 8, 2)
    assert nv == P1.size and np.array_equal(Q1, P1) and np.array_equal(Q2, P2)
    CM = rng.standard_normal((12, 12)); R = rng.standard_normal(12); F = rng.standard_normal((2, 2, 3))
    splitio.write_cg_system(str(tmp_path), CM, R, F)
    C2, R2, D2 = splitio.read_cg_system(str(tmp_path), 12)
    assert np.array_equal(C2, CM) and np.array_equal(R2, R)
    assert np.array_equal(D2, np.array([F[i, l, a] for i in range(2) for l in range(2) for a in range(3)]))
    cs = str(tmp_path / "Cs_sycollisionm.002.dat")
    splitio.write_cs_sycollisionm(cs, CM[:6])
    assert np.array_equal(splitio.read_cs_sycollisionm(cs, 6, 12), CM[:6])


def test_layouts_against_fortran(tmp_path):
    if not os.path.exists(FLANG):
        pytest.skip("no Fortran compiler in this image")
    p1, p2, cs = _expected()
    splitio.write_v33sq_delta(str(tmp_path / "v33sq_delta.python.dat"), p1, p2)
    splitio.write_column(str(tmp_path / "CM-1D.dat"), cs.reshape(-1))
    (tmp_path / "t.f90").write_text(PROG)
    subprocess.run([FLANG, "-O0", "t.f90", "-o", "t"], cwd=tmp_path, check=True, capture_output=True)
    out = subprocess.run(["./t"], cwd=tmp_path, check=True, capture_output=True, text=True).stdout
    assert out.split()[-1] == "0", out                       # Fortran read the Python files
    nv, Q1, Q2 = splitio.read_v33sq_delta(str(tmp_path / "v33sq_delta.fortran.dat"), 2, 3, 2)
    assert nv == 48 and np.array_equal(Q1, p1) and np.array_equal(Q2, p2)   # and the reverse
    assert np.array_equal(splitio.read_cs_sycollisionm(str(tmp_path / "Cs.fortran.dat"), 4, 5), cs)
    assert (tmp_path / "v33sq_delta.fortran.dat").read_bytes() == (tmp_path / "v33sq_delta.python.dat").read_bytes()
