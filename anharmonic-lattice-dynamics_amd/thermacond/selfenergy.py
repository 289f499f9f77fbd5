"""Band-path self-energy of THERMACOND jobs 2/6/7 on the device.

make_kp_bs follows kpoints.f90:37-173 (kpbs.params: units, points per direction, number of
directions, then ndir+1 named end points in conventional (units = 0) or primitive reciprocal
units); the conventional cell is force_constants_init's (force_constants.f90:56-67).
run_job2 then does what kappa.f90:876-925 does for one temperature: the phonon frequencies on
the path and, through alt3_self_w3, the self-energy of every branch at its own frequency, and
writes lw-selfenergy.dat and lw-selfenergy2.dat with the reference's columns:
  lw-selfenergy.dat   tempk, la, n, k(3), dk, omega, Im/2... as kappa.f90:916-925 writes
                      (isempi/2 = Im(nself+uself), isempi, tse, 1/tse)
  lw-selfenergy2.dat  tempk, dk, omega(1..nb), 1/tse(1..nb)
with isempi = 2 Im(nself + uself) and tse = 1e10 / c / isempi.
"""
from __future__ import annotations

import argparse
import os
import sys
from typing import List, Sequence, Tuple

import numpy as np

C_LIGHT = 2.99792458e+8


def conventional_cell(latparams: Sequence[float]) -> np.ndarray:
    """conv_to_cart (force_constants.f90:56-67): columns are the conventional vectors, with a
    along x and b in the xy plane.  latparams = a, b, c (scaled), alpha, beta, gamma (deg)."""
    a, b, c = latparams[:3]
    al, be, ga = np.radians(latparams[3:6])
    m = np.zeros((3, 3))
    m[0, 0] = a
    m[0, 1] = b * np.cos(ga)
    m[1, 1] = b * np.sin(ga)
    m[0, 2] = c * np.cos(be)
    m[1, 2] = c * (np.cos(al) - np.cos(ga) * np.cos(be)) / np.sin(ga)
    m[2, 2] = np.sqrt(c * c - m[0, 2] ** 2 - m[1, 2] ** 2)
    return m


def reciprocal_2pi(cols: np.ndarray) -> np.ndarray:
    """Rows g_i with g_i . r_j = 2 pi delta_ij for the vectors r_j = cols[:, j]."""
    return 2 * np.pi * np.linalg.inv(cols)


def read_kpbs(text: str) -> Tuple[int, int, List[str], np.ndarray]:
    lines = [l for l in text.splitlines() if l.strip()]
    units = int(lines[0].split()[0])
    nkdir = int(lines[1].split()[0])
    ndir = int(lines[2].split()[0])
    names, pts = [], []
    for l in lines[3:3 + ndir + 1]:
        f = l.split()
        names.append(f[0])
        pts.append([float(x) for x in f[1:4]])
    return units, nkdir, names, np.array(pts)


def make_kp_bs(latparams: Sequence[float], primlatt: np.ndarray, kpbs_text: str):
    """Points of the band path (nkp, 3) Cartesian with 2 pi, their path coordinate dk_bs and the
    end-point names (kpoints.f90:63-146).  primlatt: columns are the primitive vectors in
    conventional units (structure.params line 2 read column by column)."""
    units, nkdir, names, ext = read_kpbs(kpbs_text)
    ndir = len(names) - 1
    conv = conventional_cell(latparams)
    prim = conv @ np.asarray(primlatt).reshape(3, 3)          # prim_to_cart
    gconv = reciprocal_2pi(conv)                             # rows g1conv..g3conv
    g0 = reciprocal_2pi(prim)                                # rows g01..g03
    gen = gconv if units == 0 else g0
    ki = ext[:-1] @ gen
    kf = ext[1:] @ gen
    kp = np.zeros((ndir * nkdir, 3))
    dk = np.zeros(ndir * nkdir)
    n = 0
    for i in range(ndir):
        q = kf[i] - ki[i]
        lk = np.linalg.norm(q) / np.linalg.norm(g0[0])
        for j in range(1, nkdir + 1):
            n += 1
            if n >= 2:
                dk[n - 1] = dk[n - 2] + (lk / (nkdir - 1 + 1e-8) if j >= 2 else 0.0)
            kp[n - 1] = ki[i] + (j - 1) * q / (nkdir - 1 + 1e-8)
    return kp, dk, names


def _f(x) -> str:
    return repr(float(x))


def run_job2(workdir: str, precision: str = "fp64", log=print) -> dict:
    """Job 2 for T = tmin: the band path of kpbs.params, omega and the self-energy of every
    branch, lw-selfenergy.dat and lw-selfenergy2.dat in workdir."""
    from . import capi, refio
    inp = refio.read_inputs(workdir)
    text = open(os.path.join(workdir, "kpbs.params")).read()
    lp = np.array(inp.latparams, dtype=float)
    lp[:3] *= inp.lattice_parameter
    kp, dk, _ = make_kp_bs(lp, np.asarray(inp.primlatt).reshape(3, 3, order="F"), text)
    tempk = float(inp.params.tmin)
    eta = float(inp.params.etaz)
    with capi.Context(inp) as c:
        c.eig(want=False)
        om, ns, us = c.self_w3(kp, tempk, eta)
    ise = 2 * np.imag(ns + us)
    tse = 1e10 / C_LIGHT / ise
    with open(os.path.join(workdir, "lw-selfenergy.dat"), "w") as f1, \
            open(os.path.join(workdir, "lw-selfenergy2.dat"), "w") as f2:
        for n in range(len(kp)):
            for la in range(om.shape[0]):
                f1.write(" ".join([_f(tempk), str(la + 1), str(n + 1), *map(_f, kp[n]), _f(dk[n]), _f(om[la, n]),
                                   _f(ise[la, n] / 2), _f(ise[la, n]), _f(tse[la, n]), _f(1 / tse[la, n])]) + "\n")
            f2.write(" ".join([_f(tempk), _f(dk[n]), *map(_f, om[:, n]), *map(_f, 1 / tse[:, n])]) + "\n")
    log(f"job 2: {len(kp)} band points, {om.shape[0]} branches, T = {tempk} K")
    return dict(kp=kp, dk=dk, omega=om, nself=ns, uself=us)


def main(argv=None):
    ap = argparse.ArgumentParser(description="THERMACOND job 2 (band-path self-energy) on the GPU")
    ap.add_argument("workdir", nargs="?", default=".")
    args = ap.parse_args(argv)
    run_job2(args.workdir)


if __name__ == "__main__":
    sys.exit(main())
