"""Writes threshold_table_samples.json: cells of the full threshold table
computed on the GPU (`python bench.py --threshold-sweep --sweep-out DIR`),
chosen with seeded generators among the triplets whose lowest tone is at least
0.3 MHz (short CPU solves), for the CPU check against the oracle.

usage: python tests/golden/make_threshold_samples.py DIR/sweep_0_2777500.npy
"""
import json
import sys
from pathlib import Path

import numpy as np

sys.path.insert(0, str(Path(__file__).resolve().parents[2]))
from bubbledynamics_amd import gilmore as gm  # noqa: E402


def main(piece: str) -> None:
    res = np.load(piece)
    fr = gm.freq_table()
    n_r0 = 100
    assert res.shape == (len(fr) * n_r0,)
    cand = np.nonzero(fr.min(axis=1) >= 3e5)[0]
    cells = []
    for seed in (1, 2, 3, 4):
        rng = np.random.default_rng(seed)
        for c in rng.choice(cand, size=4, replace=False):
            cells += [int(c) * n_r0 + int(r) for r in rng.choice(n_r0, 2, replace=False)]
    out = {"tableau": "Dormand-Prince 1980", "n_r0": n_r0, "cells": cells,
           "gpu_threshold_pa": [float(res[i]) for i in cells]}
    path = Path(__file__).with_name("threshold_table_samples.json")
    path.write_text(json.dumps(out, indent=1) + "\n")
    print(f"wrote {len(cells)} cells to {path}")


if __name__ == "__main__":
    main(sys.argv[1])
