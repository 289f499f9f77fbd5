"""Writes sweep_fc.npz: the Stillinger-Weber Si force constants of the 64 instances of the
batched sweep (seeds 0..63; lattice constant and mesh of each from
thermacond.workloads.sweep_draw), generated by thermacond.models.sw_silicon(a) and packed by
workloads.sweep_fc_pack.  Generation is about 7 s per instance on one core, so the tests and
the benchmark read the pack instead of regenerating it; tests/test_sweep.py regenerates
instance 0 and checks it against the pack.

    python tests/golden/make_sweep.py [workers]
"""
from __future__ import annotations

import os
import sys
from concurrent.futures import ProcessPoolExecutor

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(HERE)), "anharmonic-lattice-dynamics_amd"))

from thermacond import models, workloads  # noqa: E402

SEEDS = range(64)


def _one(seed: int):
    import torch
    torch.set_num_threads(1)
    a, _ = workloads.sweep_draw(seed)
    return models.sw_silicon(a, batched=True)


def main():
    workers = int(sys.argv[1]) if len(sys.argv) > 1 else 4
    with ProcessPoolExecutor(workers) as ex:
        fcs = list(ex.map(_one, SEEDS))
    np.savez_compressed(os.path.join(HERE, "sweep_fc.npz"), **workloads.sweep_fc_pack(list(SEEDS), fcs))


if __name__ == "__main__":
    main()
