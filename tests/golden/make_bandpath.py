"""Writes tests/golden/ge_bandpath.npz from the reference's Ge example (THERMACOND/Ge-example/
selfenergy): the kpbs.params text, the conventional lattice parameters and primitive vectors
of structure.params, and the band path the reference wrote in bs_freqs.dat (point index,
path coordinate dk_bs, k in Cartesian units, 3 decimals).  Data only; run in this container:

    python tests/golden/make_bandpath.py
"""
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = "/root/reference/THERMACOND/Ge-example/selfenergy"


def main():
    kpbs = open(os.path.join(SRC, "kpbs.params")).read()
    with open(os.path.join(SRC, "structure.params")) as f:
        lat = [float(x) for x in f.readline().split()[:6]]
        prim = [float(x) for x in f.readline().split()[:9]]
        scale = float(f.readline().split()[0])
    rows = np.array([[float(x) for x in l.split()[:5]] for l in open(os.path.join(SRC, "bs_freqs.dat")) if l.strip()])
    lat = np.array(lat); lat[:3] *= scale
    np.savez_compressed(os.path.join(HERE, "ge_bandpath.npz"), kpbs=np.array(kpbs), latparams=lat,
                        primlatt=np.array(prim).reshape(3, 3, order="F"), dk=rows[:, 1], kp=rows[:, 2:5])
    print("wrote ge_bandpath.npz", rows.shape)


if __name__ == "__main__":
    main()
