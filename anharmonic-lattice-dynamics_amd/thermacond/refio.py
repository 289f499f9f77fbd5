"""THERMACOND input files: writers for structure.params, latdyn.params, dielectric.params,
kpbs.params, lat_fc.dat, fc2.dat, fc3.dat and readers for the same files.

Layouts follow the reference readers (positional, list-directed):
  structure.params  rest.f90:87-127 (8 skipped records after natoms0, rest.f90:105-112)
  latdyn.params     rest.f90:13-51
  dielectric.params rest.f90:3586-3597
  lat_fc.dat        rest.f90:166-289 (writer FOCEX/ios.f90:2278-2333)
  fc2.dat, fc3.dat  rest.f90:337-402 (writer FOCEX/ios.f90:1183-1195)
"""
from __future__ import annotations

import dataclasses
import os
from typing import Dict, List, Optional, Sequence, Tuple

import numpy as np

from .models import Crystal, ForceConstants


@dataclasses.dataclass
class RunParams:
    """latdyn.params records (rest.f90:13-51)."""
    nc: Sequence[int] = (8, 8, 8)
    dos_bs_only: bool = False
    shft: Sequence[float] = (0.0, 0.0, 0.0)
    normald: Sequence[float] = (0.0, 0.0, 1.0)
    wmesh: int = 300
    wmax: float = 700.0
    width: float = 10.0
    etaz: float = 30.0
    verbose: bool = False
    tmin: float = 300.0
    tmax: float = 300.0
    ntemp: int = 1
    job: int = 1
    ksub_size: int = 2
    v3path: str = "./"
    iso: int = 0
    classical: int = 0
    calc_cross: int = 0
    q_cross: Sequence[float] = (0.0, 0.0, 0.0)
    lmicron: float = 100.0


def _fl(x: float) -> str:
    return repr(float(x))


@dataclasses.dataclass
class Born:
    """dielectric.params (read_born, rest.f90:3580-3600): born_flag, the dielectric tensor
    epsil(i,j) and the effective charges charge(i,j) of every atom of the primitive cell."""
    born_flag: int = 0
    epsil: np.ndarray = dataclasses.field(default_factory=lambda: np.eye(3))
    zeu: Optional[np.ndarray] = None            # (natoms0, 3, 3)


def write_inputs(dirname: str, fcs: ForceConstants, rp: RunParams, nsmax: int = 15,
                 born: Optional[Born] = None) -> None:
    os.makedirs(dirname, exist_ok=True)
    c: Crystal = fcs.crystal
    with open(os.path.join(dirname, "structure.params"), "w") as f:
        f.write(" ".join(_fl(v) for v in c.latparams) + "   conventional cell a,b,c,alpha,beta,gamma\n")
        f.write(" ".join(_fl(v) for v in c.primlatt.T.reshape(-1)) + "   primitive vectors (columns) in conventional units\n")
        f.write(_fl(c.lattice_parameter) + "   scale\n")
        f.write("1 1 1 0 0 0 0 0   include FC1..FC8\n")
        f.write("1 1 1 0   invariances\n")
        f.write("0 300   boltzmann weighting\n")
        f.write("1 .false.   OUTCAR files, verbosity\n")
        f.write(f"{len(c.masses)}   natom_type\n")
        f.write(" ".join(_fl(m) for m in c.masses) + "   masses\n")
        f.write(" ".join(c.names) + "   names\n")
        f.write(f"{c.natoms0}   atoms in primitive cell\n")
        f.write("1   shell-range flag\n")
        for _ in range(7):
            f.write(" ".join(["0"] * c.natoms0) + "   shells\n")
        for i in range(c.natoms0):
            f.write(f"{i + 1} {c.types[i]} " + " ".join(_fl(v) for v in c.frac[i]) + "\n")
    with open(os.path.join(dirname, "latdyn.params"), "w") as f:
        f.write(" ".join(str(int(n)) for n in rp.nc) + (" .true." if rp.dos_bs_only else " .false.") + "\n")
        f.write(" ".join(_fl(v) for v in rp.shft) + "   shift\n")
        f.write(" ".join(_fl(v) for v in rp.normald) + "   normal\n")
        f.write(f"{rp.wmesh} {_fl(rp.wmax)}   dos mesh, omega max\n")
        f.write(f"{_fl(rp.width)} {_fl(rp.etaz)}   width, etaz\n")
        f.write((".true." if rp.verbose else ".false.") + "   verbose\n")
        f.write(f"{_fl(rp.tmin)} {_fl(rp.tmax)} {rp.ntemp}   tmin tmax ntemp\n")
        f.write(f"{rp.job}   job\n")
        f.write(f"{rp.ksub_size}   ksub_size\n")
        f.write(f"{rp.v3path}\n")
        f.write(f"{rp.iso}   iso\n")
        f.write(f"{rp.classical}   classical\n")
        f.write(f"{rp.calc_cross} " + " ".join(_fl(v) for v in rp.q_cross) + "   calc_cross q_cross\n")
        f.write(f"{_fl(rp.lmicron)}   lmicron\n")
    b = born if born is not None else Born()
    zeu = np.zeros((c.natoms0, 3, 3)) if b.zeu is None else np.asarray(b.zeu, dtype=float)
    with open(os.path.join(dirname, "dielectric.params"), "w") as f:
        f.write(f"{int(b.born_flag)}   born_flag\n")
        for i in range(3):
            f.write(" ".join(_fl(b.epsil[i][j]) for j in range(3)) + "\n")
        for k in range(c.natoms0):
            for i in range(3):
                f.write(" ".join(_fl(zeu[k, i, j]) for j in range(3)) + "\n")
    with open(os.path.join(dirname, "kpbs.params"), "w") as f:
        f.write("0\n10\n3\nG 0 0 0\nX 0 1 1\nL 0.5 0.5 0.5\n")
    P = c.prim_cart
    with open(os.path.join(dirname, "lat_fc.dat"), "w") as f:
        f.write("  Crystal data: translation vectors of the primitive cell \n")
        for r in P:
            f.write("  " + "  ".join(f"{v:19.10f}" for v in r) + "\n")
        f.write("  Crystal data: atoms in primitive cell: label,type,x,y,z,mass \n")
        f.write(f"  {c.natoms0}\n")
        tau = c.home_positions()
        for i in range(c.natoms0):
            f.write(f"  {i + 1:5d} {c.names[c.types[i] - 1]:2s}  {c.types[i]:5d}"
                    + "".join(f"  {v:19.10f}" for v in tau[i])
                    + f"  {_fl(c.masses[c.types[i] - 1])}\n")
        f.write(" Included ranks of FCs \n")
        f.write("  1 1 1 0 0 0 0 0\n")
        f.write(" Number of shells for each atom in each rank \n")
        for _ in range(7):
            f.write("  " + " ".join(["0"] * c.natoms0) + "\n")
        nt2, nt3 = len(fcs.fc2), len(fcs.fc3)
        f.write(" Crystal data: Number of groups for each rank \n")
        f.write(f"  0 {nt2} {nt3} 0 0 0 0 0\n")
        f.write(" Number of FCs for each rank \n")
        f.write(f"  0 {nt2} {nt3} 0 0 0 0 0\n")
        f.write(" Number of independent FCs for each rank \n")
        f.write(f"  0 {nt2} {nt3} 0 0 0 0 0\n")
        f.write(" maxshells,Neighborshell atoms: i,x,y,z,type_tau,n1,n2,n3 \n")
        f.write(f"  {nsmax} {fcs.natoms} {fcs.natoms}\n")
        for a in range(fcs.natoms):
            p = fcs.atompos[a]
            f.write(f"  {a + 1:5d}" + "".join(f"  {v:19.10f}" for v in p)
                    + f"  {int(fcs.iatomcell0[a]):5d}" + "".join(f"  {int(v):5d}" for v in fcs.iatomcell[a])
                    + f"  {np.linalg.norm(p):9.5f}\n")
    with open(os.path.join(dirname, "fc2.dat"), "w") as f:
        f.write("# RANK 2 tensors :term,group,(iatom,ixyz)_2 d^nU/dx_{i,alpha}^n\n")
        for t, (i, al, j, be, v) in enumerate(fcs.fc2):
            f.write(f"{t + 1:7d} {t + 1:6d}   {int(i):5d} {int(al)}   {int(j):5d} {int(be)}   {float(v)!r}\n")
    with open(os.path.join(dirname, "fc3.dat"), "w") as f:
        f.write("# RANK 3 tensors :term,group,(iatom,ixyz)_3 d^nU/dx_{i,alpha}^n\n")
        for t, (i, al, j, be, k, ga, v) in enumerate(fcs.fc3):
            f.write(f"{t + 1:7d} {t + 1:6d}   {int(i):5d} {int(al)}   {int(j):5d} {int(be)}   {int(k):5d} {int(ga)}   {float(v)!r}\n")


# ---------------------------------------------------------------------------
# readers (used by the host driver to build a context from the reference files)
# ---------------------------------------------------------------------------


def _toks(line: str) -> List[str]:
    return line.replace(",", " ").split()


def _flt(t: str) -> float:
    return float(t.lower().replace("d", "e"))


def _bool(t: str) -> bool:
    t = t.strip(".").lower()
    return t.startswith("t")


@dataclasses.dataclass
class InputSet:
    """Everything the three-phonon path needs from the reference input files."""
    params: RunParams
    latparams: np.ndarray        # (6,)
    primlatt: np.ndarray         # (3,3) columns = primitive vectors, conventional units
    lattice_parameter: float
    masses: np.ndarray           # per primitive atom
    types: np.ndarray
    r123: np.ndarray             # (3,3) rows r1,r2,r3 (Cartesian, from lat_fc.dat)
    atompos: np.ndarray          # (natoms,3)
    iatomcell0: np.ndarray       # (natoms,)
    iatomcell: np.ndarray        # (natoms,3)
    fc2_atoms: np.ndarray        # (nt2,2) 1-based
    fc2_xyz: np.ndarray          # (nt2,2) 1-based
    fc2: np.ndarray              # (nt2,)
    fc3_atoms: np.ndarray        # (nt3,3)
    fc3_xyz: np.ndarray          # (nt3,3)
    fc3: np.ndarray              # (nt3,)
    names: Tuple[str, ...] = ()  # element label per atom type (structure.params)
    born_flag: int = 0           # dielectric.params (read_born)
    epsil: Optional[np.ndarray] = None   # (3,3) epsil(i,j)
    zeu: Optional[np.ndarray] = None     # (natoms0,3,3) charge(i,j) per atom

    @property
    def natoms0(self) -> int:
        return len(self.masses)


def read_inputs(dirname: str) -> InputSet:
    P = lambda n: os.path.join(dirname, n)
    with open(P("latdyn.params")) as f:
        L = f.read().splitlines()
    t = _toks(L[0])
    rp = RunParams(nc=tuple(int(x) for x in t[:3]), dos_bs_only=_bool(t[3]))
    rp.shft = tuple(_flt(x) for x in _toks(L[1])[:3])
    rp.normald = tuple(_flt(x) for x in _toks(L[2])[:3])
    t = _toks(L[3]); rp.wmesh, rp.wmax = int(t[0]), _flt(t[1])
    t = _toks(L[4]); rp.width, rp.etaz = _flt(t[0]), _flt(t[1])
    rp.verbose = _bool(_toks(L[5])[0])
    t = _toks(L[6]); rp.tmin, rp.tmax, rp.ntemp = _flt(t[0]), _flt(t[1]), int(t[2])
    rp.job = int(_toks(L[7])[0])
    rp.ksub_size = int(_toks(L[8])[0])
    rp.v3path = L[9]
    rp.iso = int(_toks(L[10])[0])
    rp.classical = int(_toks(L[11])[0])
    t = _toks(L[12]); rp.calc_cross = int(t[0]); rp.q_cross = tuple(_flt(x) for x in t[1:4])
    rp.lmicron = _flt(_toks(L[13])[0])

    with open(P("structure.params")) as f:
        S = f.read().splitlines()
    latp = np.array([_flt(x) for x in _toks(S[0])[:6]])
    prim = np.array([_flt(x) for x in _toks(S[1])[:9]]).reshape(3, 3).T
    scale = _flt(_toks(S[2])[0])
    ntype = int(_toks(S[7])[0])
    mas = [_flt(x) for x in _toks(S[8])[:ntype]]
    names = tuple(_toks(S[9])[:ntype])
    nat0 = int(_toks(S[10])[0])
    types = np.zeros(nat0, dtype=np.int64)
    for r in range(nat0):
        t = _toks(S[19 + r])
        types[int(t[0]) - 1] = int(t[1])
    masses = np.array([mas[ty - 1] for ty in types])

    with open(P("lat_fc.dat")) as f:
        F = f.read().splitlines()
    k = 0
    while not F[k].startswith("  Crystal data: transl"):
        k += 1
    r123 = np.array([[_flt(x) for x in _toks(F[k + 1 + i])[:3]] for i in range(3)])
    k += 4 + 1
    n1 = int(_toks(F[k])[0])
    if n1 != nat0:
        raise ValueError(f"natoms0 in lat_fc.dat incompatible with structure.params {n1} {nat0}")
    k += 1 + n1
    k += 1
    k += 1  # include_fc
    k += 8
    k += 2  # groups
    k += 1
    nterms = [int(x) for x in _toks(F[k])[:4]]
    k += 3
    k += 1
    natoms2 = int(_toks(F[k])[1])
    k += 1
    atompos = np.zeros((natoms2, 3)); i0 = np.zeros(natoms2, dtype=np.int64)
    ic = np.zeros((natoms2, 3), dtype=np.int64)
    for a in range(natoms2):
        t = _toks(F[k + a])
        i = int(t[0]) - 1
        atompos[i] = [_flt(x) for x in t[1:4]]
        i0[i] = int(t[4]); ic[i] = [int(x) for x in t[5:8]]

    def rdfc(name, rank, expect):
        with open(P(name)) as f:
            lines = [l for l in f.read().splitlines()[1:] if l.strip()]
        if len(lines) != expect:
            raise ValueError(f"READ_FC{rank}: number of lines read={len(lines)} inconsistent with {expect} from lat_fc.dat")
        at = np.zeros((len(lines), rank), dtype=np.int64); xyz = np.zeros((len(lines), rank), dtype=np.int64)
        val = np.zeros(len(lines))
        for n, l in enumerate(lines):
            t = _toks(l)
            for r in range(rank):
                at[n, r] = int(t[2 + 2 * r]); xyz[n, r] = int(t[3 + 2 * r])
            val[n] = _flt(t[2 + 2 * rank])
        return at, xyz, val

    a2, x2, v2 = rdfc("fc2.dat", 2, nterms[1])
    a3, x3, v3 = rdfc("fc3.dat", 3, nterms[2])
    born_flag, epsil, zeu = read_born(P("dielectric.params"), nat0)
    return InputSet(rp, latp, prim, scale, masses, types, r123, atompos, i0, ic, a2, x2, v2, a3, x3, v3, names,
                    born_flag, epsil, zeu)


def read_born(path: str, natoms0: int):
    """read_born (rest.f90:3580-3600): born_flag on record 1, epsil(i,1:3) on records 2-4, then
    charge(i,1:3) of atom k on three records per atom (list-directed: blank records are
    skipped, as a Fortran list-directed read skips them).  A missing file stops the reference
    (status='old'); here it means born_flag = 0."""
    if not os.path.exists(path):
        return 0, np.eye(3), np.zeros((natoms0, 3, 3))
    with open(path) as f:
        recs = [_toks(l) for l in f.read().splitlines()]
    recs = [r for r in recs if r]
    flag = int(recs[0][0])
    epsil = np.array([[_flt(x) for x in recs[1 + i][:3]] for i in range(3)])
    zeu = np.zeros((natoms0, 3, 3))
    for k in range(natoms0):
        for i in range(3):
            r = recs[4 + 3 * k + i] if 4 + 3 * k + i < len(recs) else []
            if len(r) < 3:
                if flag:
                    raise ValueError(f"dielectric.params: charges of atom {k + 1} incomplete")
                continue
            zeu[k, i] = [_flt(x) for x in r[:3]]
    return flag, epsil, zeu
