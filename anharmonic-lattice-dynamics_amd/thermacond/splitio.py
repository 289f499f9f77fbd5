"""The reference's on-disk split formats, for runs that pass P or the collision system between
jobs or processes (the job 4 -> job 5 route of kappa.f90 and the split v33sq files).

  v33sq_delta.NNN.dat   P1sm, P2sm of one irreducible block: Fortran unformatted sequential,
                        one record with the element count, then one record (P1sm, P2sm) per
                        (n1, l1, n2, l2, l3), l3 fastest (writer sub.f90:7161-7177, reader
                        :7598-7613; NNN = the block index, i3.3)
  Cs_sycollisionm.NNN.dat  collision-matrix rows of one block: unformatted, one record per
                        value, row by row (writer sub.f90:3371-3377, reader :5304-5363)
  CM-1D.dat             the whole collision matrix, one value per line, row by row
                        (kappa.f90:2577-2587, reader read_CM_1D_mpi sub.f90:5401-5434)
  RHS_IBZ.dat           the right-hand side, one value per line (sub.f90:2681-2685, :5367-5399)
  DS.dat                the CG starting point F_RTA flattened as (i, l, alpha), alpha fastest
                        (kappa.f90:2614-2625, reader read_DS sub.f90:5439-5470)

Unformatted records are written with 4-byte little-endian length markers, the gfortran and
flang default on this platform.  Text values are written with 17 significant digits and read
with Fortran 'D' exponents accepted.
"""
from __future__ import annotations

import os
from typing import Tuple

import numpy as np


def block_name(stem: str, index: int) -> str:
    """'v33sq_delta', 3 -> 'v33sq_delta.003.dat' (format (a,i3.3,a))."""
    return f"{stem}.{index:03d}.dat"


# ---------------------------------------------------------------------------- v33sq_delta
def write_v33sq_delta(path: str, P1sm: np.ndarray, P2sm: np.ndarray, nv3: int = None) -> None:
    """P1sm, P2sm (ksub, nk, nb, nb, nb) as the reference indexes them (n1, n2, l1, l2, l3)."""
    P1sm = np.asarray(P1sm, dtype=np.float64); P2sm = np.asarray(P2sm, dtype=np.float64)
    ksub, nk, nb = P1sm.shape[0], P1sm.shape[1], P1sm.shape[2]
    if nv3 is None:
        nv3 = ksub * nk * nb ** 3
    # record order n1, l1, n2, l2, l3 with l3 fastest
    a = np.transpose(P1sm, (0, 2, 1, 3, 4)).reshape(-1)
    b = np.transpose(P2sm, (0, 2, 1, 3, 4)).reshape(-1)
    rec = np.empty(a.size, dtype=[("m0", "<i4"), ("p1", "<f8"), ("p2", "<f8"), ("m1", "<i4")])
    rec["m0"] = 16; rec["m1"] = 16; rec["p1"] = a; rec["p2"] = b
    with open(path, "wb") as f:
        np.array([4, nv3, 4], dtype="<i4").tofile(f)
        rec.tofile(f)


def read_v33sq_delta(path: str, ksub: int, nk: int, nb: int) -> Tuple[int, np.ndarray, np.ndarray]:
    """(nv3, P1sm, P2sm) with P in the reference's (ksub, nk, nb, nb, nb) layout, order='F'."""
    with open(path, "rb") as f:
        head = np.fromfile(f, dtype="<i4", count=3)
        if head[0] != 4 or head[2] != 4:
            raise ValueError(f"{path}: not a Fortran unformatted file with 4-byte markers")
        n = ksub * nb * nk * nb * nb
        rec = np.fromfile(f, dtype=[("m0", "<i4"), ("p1", "<f8"), ("p2", "<f8"), ("m1", "<i4")], count=n)
    if rec.size != n or np.any(rec["m0"] != 16) or np.any(rec["m1"] != 16):
        raise ValueError(f"{path}: expected {n} records of two doubles")
    shape = (ksub, nb, nk, nb, nb)
    P1 = np.asfortranarray(np.transpose(rec["p1"].reshape(shape), (0, 2, 1, 3, 4)))
    P2 = np.asfortranarray(np.transpose(rec["p2"].reshape(shape), (0, 2, 1, 3, 4)))
    return int(head[1]), P1, P2


# ---------------------------------------------------------------------------- Cs_sycollisionm
def write_cs_sycollisionm(path: str, rows: np.ndarray) -> None:
    """Collision-matrix rows (nrow, ncol): one unformatted record per value, row by row."""
    v = np.ascontiguousarray(rows, dtype=np.float64).reshape(-1)
    rec = np.empty(v.size, dtype=[("m0", "<i4"), ("v", "<f8"), ("m1", "<i4")])
    rec["m0"] = 8; rec["m1"] = 8; rec["v"] = v
    rec.tofile(path)


def read_cs_sycollisionm(path: str, nrow: int, ncol: int) -> np.ndarray:
    rec = np.fromfile(path, dtype=[("m0", "<i4"), ("v", "<f8"), ("m1", "<i4")])
    if rec.size != nrow * ncol or np.any(rec["m0"] != 8) or np.any(rec["m1"] != 8):
        raise ValueError(f"{path}: expected {nrow * ncol} records of one double")
    return rec["v"].reshape(nrow, ncol)


# ---------------------------------------------------------------------------- text files
def write_column(path: str, values) -> None:
    """One value per line (list-directed output of the reference)."""
    v = np.asarray(values, dtype=np.float64).reshape(-1)
    with open(path, "w") as f:
        f.write("".join(f"   {x:.17G}\n" for x in v))


def read_column(path: str, n: int = None) -> np.ndarray:
    vals = []
    with open(path) as f:
        for line in f:
            t = line.strip()
            if t:
                vals.append(float(t.replace("D", "E").replace("d", "e")))
    v = np.array(vals)
    if n is not None and v.size != n:
        raise ValueError(f"{path}: expected {n} values, found {v.size}")
    return v


def write_cg_system(workdir: str, CM: np.ndarray, RHS: np.ndarray, FRTA: np.ndarray) -> None:
    """CM-1D.dat (CM(i,j) row by row), RHS_IBZ.dat and DS.dat (F_RTA (nibz,nb,3) flattened
    with alpha fastest) as job 4 leaves them for job 5."""
    write_column(os.path.join(workdir, "CM-1D.dat"), np.asarray(CM).reshape(-1, order="C"))
    write_column(os.path.join(workdir, "RHS_IBZ.dat"), RHS)
    write_column(os.path.join(workdir, "DS.dat"), np.asarray(FRTA).reshape(-1, order="C"))


def read_cg_system(workdir: str, n: int):
    """(CM (n, n), RHS (n,), DS (n,)) as job 5 reads them (read_CM_1D_mpi, read_RHS, read_DS)."""
    CM = read_column(os.path.join(workdir, "CM-1D.dat"), n * n).reshape(n, n)
    RHS = read_column(os.path.join(workdir, "RHS_IBZ.dat"), n)
    DS = read_column(os.path.join(workdir, "DS.dat"), n)
    return CM, RHS, DS
