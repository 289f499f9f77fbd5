"""Readers for the raw dumps written by oracle/ref_probe.f90 (the reference THERMACOND
routines run on one process) and for the compressed fixtures made from them."""
from __future__ import annotations

import os

import numpy as np


def load_dump(d: str, ksub: int | None = None) -> dict:
    """Arrays in the reference's column-major layouts (numpy order='F'); indices 1-based as
    dumped (mapibz, mapinv, tet, star_n0, mappos).  ksub: the size of the IBZ block of a
    REF_PROBE_SUBSET run (|V3|^2, P and the block's Q rows T<t>_Qsub/_QNsub/_QUsub)."""
    with open(os.path.join(d, "dims.txt")) as f:
        nk, nibz, nb, nat0, nt2, nt3, nops, npos = (int(x) for x in f.read().split())
    out = dict(nk=nk, nibz=nibz, nb=nb, natoms0=nat0, nt2=nt2, nt3=nt3, nops=nops, npos=npos)

    def r(name, shape, dt=np.float64):
        p = os.path.join(d, name + ".bin")
        if not os.path.exists(p):
            return None
        a = np.fromfile(p, dtype=dt)
        return a.reshape(shape, order="F")

    out["kpc_mesh"] = r("kpc_mesh", (3, nk))
    out["kpc"] = r("kpc", (3, nk))
    out["mapibz"] = r("mapibz", (nk,), np.int32)
    out["mapinv"] = r("mapinv", (nibz,), np.int32)
    out["mappos"] = r("mappos", (nk,), np.int32)
    out["kibz"] = r("kibz", (3, nibz))
    out["wibz"] = r("wibz", (nibz,))
    out["eigenval"] = r("eigenval", (nb, nk))
    out["eigenvec"] = r("eigenvec", (nb, nb, nk), np.complex128)
    out["eigenvec_raw"] = r("eigenvec_raw", (nb, nb, nk), np.complex128)
    out["veloc"] = r("veloc", (3, nb, nk))
    out["op_kmatrix"] = r("op_kmatrix", (3, 3, nops))
    out["tet"] = r("tet", (4, 6 * nk), np.int32)
    out["star_n"] = r("star_n", (nibz,), np.int32)
    out["star_op"] = r("star_op", (48, nibz), np.int32)
    out["star_n0"] = r("star_n0", (48, nibz), np.int32)
    nblk = nibz if ksub is None else ksub
    out["v33s8"] = r("v33s8", (nblk, nk, nb, nb, nb))
    out["P1sm"] = r("P1sm", (nblk, nk, nb, nb, nb))
    out["P2sm"] = r("P2sm", (nblk, nk, nb, nb, nb))
    out["frequency"] = r("frequency", (nk, nb))
    temps = sorted({int(f.split("_")[0][1:]) for f in os.listdir(d) if f.startswith("T") and f.endswith(".bin")})
    out["temps"] = temps
    N = nibz * nb * 3
    for T in temps:
        t = f"T{T}"
        if ksub is not None:
            out[t] = dict(Qsub=r(f"{t}_Qsub", (ksub, nb)), QNsub=r(f"{t}_QNsub", (ksub, nb)),
                          QUsub=r(f"{t}_QUsub", (ksub, nb)))
            continue
        out[t] = dict(
            dist=r(f"{t}_dist", (nk, nb)), Q=r(f"{t}_Q", (nibz, nb)), QN=r(f"{t}_QN", (nibz, nb)),
            QU=r(f"{t}_QU", (nibz, nb)), tauN=r(f"{t}_tauN", (nibz, nb)), tauU=r(f"{t}_tauU", (nibz, nb)),
            tautot=r(f"{t}_tautot", (nibz, nb)), taueff=r(f"{t}_taueff", (nibz, nb)),
            FRTA=r(f"{t}_FRTA", (nibz, nb, 3)), kappaRTA=r(f"{t}_kappaRTA", (nb, 3, 3)),
            CM=r(f"{t}_CM", (N, N)), RHS=r(f"{t}_RHS", (N,)), x=r(f"{t}_x", (N,)),
            kappaBTE=r(f"{t}_kappaBTE", (nb, 3, 3)), taueffd=r(f"{t}_taueffd", (nibz, nb)))
        # iterative BTE (REF_PROBE_ITER): first update F2, kappa after each iteration, final F
        kit = os.path.join(d, f"{t}_kappaIter.bin")
        if os.path.exists(kit):
            k = np.fromfile(kit, dtype=np.float64)
            out[t].update(F2first=r(f"{t}_F2first", (nibz, nb, 3)), Fiter=r(f"{t}_Fiter", (nibz, nb, 3)),
                          kappaIter=k.reshape((nb, 3, 3, k.size // (9 * nb)), order="F"),
                          taueffIter=r(f"{t}_taueffIter", (k.size // (9 * nb),)))
    return out
