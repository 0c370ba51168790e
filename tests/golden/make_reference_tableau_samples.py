"""Writes reference_tableau_samples.json: cells of the threshold table with the
reference's own tableau (a32 = -9/40, G:220) at its tolerances, as the GPU's
drop-in sweep computed them in a bounded run over the table's high-frequency
quarter (`python bench.py --threshold-sweep --sweep-reference-tableau
--sweep-first 2060000 --sweep-budget 20000 --sweep-deadline 660 --sweep-out
DIR`), eight cells chosen with a seeded generator among the finished ones, for
the CPU check against the oracle.

usage: python tests/golden/make_reference_tableau_samples.py DIR/sweep_partial_2060000.npy
"""
import json
import sys
from pathlib import Path

import numpy as np

FIRST = 2060000


def main(piece: str) -> None:
    part = np.load(piece)
    done = np.flatnonzero(~np.isnan(part))
    rng = np.random.default_rng(11)
    pick = np.sort(rng.choice(done, 8, replace=False))
    out = {"tableau": "reference (a32 = -9/40)", "n_r0": 100, "first": FIRST,
           "cells": [FIRST + int(j) for j in pick], "gpu_threshold_pa": [float(part[j]) for j in pick]}
    path = Path(__file__).with_name("reference_tableau_samples.json")
    path.write_text(json.dumps(out, indent=1) + "\n")
    print(f"wrote {len(pick)} cells to {path}")


if __name__ == "__main__":
    main(sys.argv[1])
