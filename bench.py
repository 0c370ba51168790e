#!/usr/bin/env python3
"""Benchmark: FP32 Westervelt, 192^3 tone burst, one MI355X.

A step is one Westervelt time step over the 192^3 grid (interior 6th-order
update with relaxation and nonlinearity, source plane, Mur faces, sponge,
peak pressure, per-step energy and the focus sample), replayed from a
captured 6-step hipGraph.  Inputs live in HBM before timing starts.

Prints one JSON line.  `--gpus N` with N > 1 exits: this build measures one
GPU (the z-slab sharding across GPUs is described in the design notes).
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

PEAK_HBM = 8.0e12  # B/s, MI355X


def log(msg: str) -> None:
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def cpu_baseline(n: int, budget_s: float) -> dict:
    """FP64 oracle (serial C++ restatement) on the host cores, bounded sample."""
    import numpy as np
    import oracle
    from bubbledynamics_amd import presets
    orc = oracle.load()
    p = presets.tone_burst_192(n)
    cores = min(len(os.sched_getaffinity(0)), 16)
    state = np.zeros((12, n, n, n))
    pmax = np.zeros((n, n, n))
    t0 = time.perf_counter()
    steps = 0
    # warm-up step (page faults), not timed
    orc.wv_run(p, n, n, 0, 1, state, pmax, None, None, (-1, -1, -1), None, None, cores)
    steps = 1
    t0 = time.perf_counter()
    done = 0
    while time.perf_counter() - t0 < budget_s or done < 3:
        orc.wv_run(p, n, n, steps, 1, state, pmax, None, None, (-1, -1, -1), None, None, cores)
        steps += 1
        done += 1
    el = time.perf_counter() - t0
    val = n ** 3 * done / el / 1e6
    return {"value": round(val, 3), "unit": "Mcell-steps/s", "cores": cores, "kind": "port",
            "sample": f"FP64 oracle, same 192^3 workload, {done} steps after 1 warm-up step, "
                      f"{el:.1f} s wall, OpenMP over z-planes"}


def read_traffic(profile: Path) -> float | None:
    """HBM bytes per launch of the interior kernel from the committed counter
    summary (line 'hbm_bytes_per_launch: N'), or None."""
    try:
        for line in profile.read_text().splitlines():
            if line.startswith("hbm_bytes_per_launch:"):
                return float(line.split(":", 1)[1])
    except OSError:
        pass
    return None


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3000)
    ap.add_argument("--warmup", type=int, default=60)
    ap.add_argument("--n", type=int, default=192)
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-fp64", action="store_true")
    ap.add_argument("--calib-copy-mb", type=int, default=0,
                    help="first copy a buffer of this size 3 times (HBM counter calibration)")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU only: the multi-GPU z-slab decomposition is not built "
              "in this version (see the design notes); run with --gpus 1", file=sys.stderr)
        return 2

    import numpy as np
    import torch
    from bubbledynamics_amd import _ffi, presets
    from bubbledynamics_amd.westervelt import WesterveltRun, focal_harmonics

    if not torch.cuda.is_available():
        print("no GPU visible", file=sys.stderr)
        return 3
    _ffi.load()
    if args.calib_copy_mb:
        src = torch.ones(args.calib_copy_mb * 2 ** 20 // 4, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst
    n, K, W = args.n, args.steps, args.warmup
    p = presets.tone_burst_192(n)
    focus = presets.tone_burst_focus(p, n)
    total = W + K

    def one_run(dtype):
        run = WesterveltRun(p, n, n, dtype=dtype, max_steps=total + 1, focus=focus)
        run.run(W)
        torch.cuda.synchronize()
        ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        ev0.record()
        run.run(K)
        ev1.record()
        torch.cuda.synchronize()
        ms = ev0.elapsed_time(ev1)
        return run, ms

    log(f"FP32 run {n}^3: {W} warm-up + {K} timed steps")
    run32, ms = one_run(torch.float32)
    value = n ** 3 * K / (ms / 1e3) / 1e6
    log(f"FP32 {value:.1f} Mcell-steps/s ({ms / K * 1e3:.2f} us/step)")

    run32.finalize()
    torch.cuda.synchronize()
    pmax32 = run32.pmax.clone()
    # interior kernel alone, HIP events on the stream it is launched on (these
    # launches recompute the next step's interior and touch pmax, hence the copy)
    R = 50
    lib = _ffi.load()
    stream = torch.cuda.current_stream().cuda_stream
    import ctypes as C
    for _ in range(5):
        _ffi.check("interior_only", lib.bd_westervelt_plan_interior_only(run32._plan, C.c_void_p(stream)))
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(R):
        _ffi.check("interior_only", lib.bd_westervelt_plan_interior_only(run32._plan, C.c_void_p(stream)))
    e1.record()
    torch.cuda.synchronize()
    k_ms = e0.elapsed_time(e1) / R
    cells = int(lib.bd_westervelt_plan_interior_cells(run32._plan))
    bytes_per_cell = 14 * 4          # 9 reads + 3 writes + pmax read/modify/write, FP32
    alg_bytes = cells * bytes_per_cell
    achieved = alg_bytes / (k_ms / 1e3)
    log(f"interior kernel {k_ms * 1e3:.2f} us -> {achieved / 1e9:.0f} GB/s algorithmic")

    f32 = run32.focus.double().cpu().numpy()[:total]
    energy = run32.energy.cpu().numpy()[:total]
    harm = focal_harmonics(f32, p.dt, p.fr, 3)

    extra = {}
    if not args.no_fp64:
        log("FP64 run (same workload)")
        run64, ms64 = one_run(torch.float64)
        run64.finalize()
        torch.cuda.synchronize()
        f64 = run64.focus.cpu().numpy()[:total]
        rel = float(np.linalg.norm(f32 - f64) / np.linalg.norm(f64))
        extra["fp64"] = {"value": round(n ** 3 * K / (ms64 / 1e3) / 1e6, 1),
                         "unit": "Mcell-steps/s",
                         "focal_rel_l2_fp32_vs_fp64": rel,
                         "pmax_rel_l2_fp32_vs_fp64": float(
                             (torch.linalg.vector_norm(pmax32.double() - run64.pmax) /
                              torch.linalg.vector_norm(run64.pmax)).item())}
        log(f"FP64 {extra['fp64']['value']} Mcell-steps/s; focal rel L2 FP32 vs FP64 = {rel:.2e}")
        del run64

    cpu = None
    if not args.no_cpu:
        log(f"CPU baseline (~{args.cpu_seconds:.0f} s)")
        try:
            cpu = cpu_baseline(n, args.cpu_seconds)
        except Exception as exc:  # the baseline is reported, never required
            log(f"cpu baseline failed: {exc}")

    traffic = read_traffic(ROOT / "profiles" / f"westervelt_march_{n}_pmc.txt")
    out = {
        "metric": "Westervelt Mcell-steps/s & %HBM roofline; Gilmore bubble-steps/s; vs FP64 plain",
        "value": round(value, 2),
        "unit": "Mcell-steps/s",
        "n_gpus": 1,
        "steps": K,
        "warmup": W,
        "ms_per_step": ms / K,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": f"westervelt_tone_burst_{n}", "grid": [n, n, n], "dx_m": p.dxy,
                   "dt_s": p.dt, "relaxation": 2, "source": "cap F#1, 1 MHz, 8 cycles, 1 MPa",
                   "boundary": "Mur + 20-cell sponge", "hipgraph": "6-step"},
        "roofline": {"bound": "hbm", "achieved": round(achieved / 1e9, 1), "peak": PEAK_HBM / 1e9,
                     "unit": "GB/s", "frac": round(achieved / PEAK_HBM, 4),
                     "traffic": traffic,
                     "kernel": "wv_march<float,64,4,1,0>", "kernel_us": round(k_ms * 1e3, 2),
                     "alg_bytes_per_launch": alg_bytes, "cells_per_launch": cells},
        "cpu_baseline": cpu,
        "focus": {"peak_pa": float(np.abs(f32).max()), "harmonics_pa": [float(h) for h in harm],
                  "energy_max_j": float(energy[:, 0].max())},
    }
    out.update(extra)
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
