#!/usr/bin/env python3
"""Benchmark: FP32 Westervelt, 192^3 tone burst, one MI355X.

A step is one Westervelt time step over the 192^3 grid (interior 6th-order
update with relaxation and nonlinearity, source plane, Mur faces, sponge,
peak pressure, per-step energy and the focus sample) -- one kernel launch,
replayed from a pair of 6-step hipGraphs captured when the plan is built.
Inputs live in HBM before timing starts.  After --warmup steps, --reps (>= 5)
runs of exactly --steps steps are timed with HIP events and the median is
reported.  The FP32-vs-FP64 accuracy comes from the whole 3,000-step tone
burst in both precisions, whatever --steps is.  The CPU baseline is the FP64
oracle on the host (the per-GPU core share and one core), with lscpu.

Prints one JSON line.  `--gpus N` with N > 1 exits: one GPU is measured (the
z-slab decomposition across GPUs is not built).

`--threshold-sweep` runs the reference's threshold table (27,775 frequency
triplets x 100 radii) from `--sweep-first` on, in resumable pieces
(`--sweep-deadline`), and `--sweep-assemble` joins the pieces into its table file.

`--reference-scenario` instead runs the reference's own programs at their own
scale and prints one JSON line of stage timings: the heat source of
Westervelt_relax_fft_Pressure_Energy.cu (426 x 426 x 2003 quarter domain,
FP64, 165 settle periods, one water period, 4 harmonic periods, 10 harmonics),
its Q3D file, the quarter -> full expansion of
Westervelt_Temperature_SquareSpiral.cu (851 x 851 x 2001) and that program's
continuous spiral (all 77,474 substeps, or `--shifts`) in FP32 with the fields
stored as T - T0 and a compensated dose; the first `--compare` substeps also
run in FP64 (the reference's precision) to state the FP32 error.
"""
from __future__ import annotations

import argparse
import json

import numpy as np
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


def host_description() -> dict:
    """CPU of the host: visible CPUs, the threads this process may use and lscpu."""
    import subprocess
    info = {"nproc": os.cpu_count(), "affinity_cpus": len(os.sched_getaffinity(0)),
            "omp_num_threads_env": os.environ.get("OMP_NUM_THREADS")}
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=10).stdout
        for line in out.splitlines():
            k, _, v = line.partition(":")
            if k.strip() in ("Model name", "Socket(s)", "Core(s) per socket", "Thread(s) per core", "CPU(s)",
                             "CPU max MHz", "L3 cache"):
                info[k.strip()] = v.strip()
    except (OSError, subprocess.SubprocessError):
        pass
    return info


def cpu_threads() -> int:
    """Threads of the CPU baseline: the host CPUs this process may use, at most
    the per-GPU share the box gives one GPU's jobs (OMP_NUM_THREADS, 16)."""
    share = int(os.environ.get("OMP_NUM_THREADS", "16") or 16)
    return max(1, min(len(os.sched_getaffinity(0)), share))


def cpu_baseline(n: int, budget_s: float) -> dict:
    """FP64 oracle (serial C++ restatement, OpenMP over z-planes) on the host
    cores, on a bounded sample of the same 192^3 workload: all threads of the
    share, then one thread."""
    import numpy as np

    import oracle
    from bubbledynamics_amd import presets
    orc = oracle.load()
    p = presets.tone_burst_192(n)
    state = np.zeros((12, n, n, n))
    pmax = np.zeros((n, n, n))

    def rate(threads, budget, step0):
        steps = step0
        orc.wv_run(p, n, n, steps, 1, state, pmax, None, None, (-1, -1, -1), None, None, threads)  # warm-up
        steps += 1
        t0 = time.perf_counter()
        done = 0
        while time.perf_counter() - t0 < budget or done < 2:
            orc.wv_run(p, n, n, steps, 1, state, pmax, None, None, (-1, -1, -1), None, None, threads)
            steps += 1
            done += 1
        el = time.perf_counter() - t0
        return n ** 3 * done / el / 1e6, done, el, steps

    cores = cpu_threads()
    v_all, d_all, e_all, s = rate(cores, budget_s, 0)
    v_one, d_one, e_one, _ = rate(1, budget_s / 2, s)
    return {"value": round(v_all, 3), "unit": "Mcell-steps/s", "cores": cores, "kind": "port",
            "serial_1core": round(v_one, 3),
            "sample": f"FP64 oracle, same 192^3 workload: {d_all} steps on {cores} threads in {e_all:.1f} s "
                      f"and {d_one} steps on 1 thread in {e_one:.1f} s, each after 1 warm-up step",
            "host": host_description()}


def gilmore_cpu_baseline(budget_s: float) -> dict:
    """FP64 oracle timed on the host: config 0 (RK4 dt = 0.1 ns, 20 us) on its
    first bubbles and config 3 (DP5(4), 20 us) on a 4,096-bubble subset
    stratified in R0; rates in bubble-steps/s."""
    import oracle
    from bubbledynamics_amd import gilmore as gm
    orc = oracle.load()
    p = gm.workload_params()
    cores = cpu_threads()
    R0, amp, ph, fr = gm.config0_inputs(1024, 0)
    k = max(cores, 16)
    t0 = time.perf_counter()
    r0 = orc.gilmore_integrate(p, gm.BD_GM_RK4, R0[:k], amp[:k], ph[:k], fr, 20e-6, 1e-10, nthreads=cores)
    e0 = time.perf_counter() - t0
    idx = gm.config3_subset()
    R0, amp, ph, fr = gm.config3_inputs(1_000_000, 3)
    t0 = time.perf_counter()
    r3 = orc.gilmore_integrate(p, gm.BD_GM_DP45, R0[idx], amp[idx], ph[idx], fr, 20e-6, 1e-13, rtol=1e-8,
                               atol=1e-12, nthreads=cores)
    e3 = time.perf_counter() - t0
    s3 = int(r3["n_acc"].sum())
    return {"unit": "bubble-steps/s", "cores": cores, "kind": "port",
            "rk4_1024": {"value": float(r0["n_acc"].sum()) / e0,
                         "sample": f"first {k} of the 1,024 config-0 bubbles, full 20 us, {e0:.1f} s"},
            "dp45_1e6": {"value": s3 / e3, "accepted_steps": s3,
                         "sample": f"4,096 config-3 bubbles stratified in R0, full 20 us, {e3:.1f} s"},
            "dp45_subset_rmax": r3["Rmax"].tolist(), "dp45_subset_collapsed": (r3["t_collapse"] > 0).tolist()}


def bioheat_cpu_baseline(n: int, budget_s: float) -> dict:
    """FP64 oracle Pennes substeps with the Kahan-compensated dose (T:52-105)
    on the host, on a bounded sample of the coupled 256^3 workload (a uniform
    heat source; the work per cell does not depend on it): all threads of the
    share, then one thread; rates in cell-substeps/s."""
    import oracle
    from bubbledynamics_amd import bioheat as bh
    orc = oracle.load()
    p = bh.pennes_params(dx=2e-4, dt=0.05, substeps=2)
    p.T_off = 0.0
    Ta = np.full((n, n, n), p.T_art)
    Tb = Ta.copy()
    Q = np.full((n, n, n), 1e6)
    TD = np.zeros((n, n, n))
    TDc = np.zeros((n, n, n))

    def rate(threads, budget):
        orc.temperature_square(p, Ta, Tb, Q, TD, n, n, 0, 0, 0, TDc, threads)   # warm-up
        t0 = time.perf_counter()
        done = 0
        while time.perf_counter() - t0 < budget or done < 2:
            orc.temperature_square(p, Ta, Tb, Q, TD, n, n, (done + 1) % 2, 0, 0, TDc, threads)
            done += 1
        el = time.perf_counter() - t0
        return (n - 2) ** 3 * done / el, done, el

    cores = cpu_threads()
    v_all, d_all, e_all = rate(cores, budget_s)
    v_one, d_one, e_one = rate(1, budget_s / 2)
    return {"value": v_all, "unit": "cell-substeps/s", "cores": cores, "kind": "port", "serial_1core": v_one,
            "sample": f"FP64 oracle with the compensated dose, {n}^3: {d_all} substeps on {cores} threads in "
                      f"{e_all:.1f} s and {d_one} on 1 thread in {e_one:.1f} s, each after 1 warm-up substep"}


def adjoint_cpu_baseline(p, n: int, amp, ph, budget_s: float) -> dict:
    """FP64 oracle forward steps of the 128^3 phased-array workload on the
    host (all threads of the share): the CPU path of the gradient is the
    forward solve repeated per central finite difference (two solves per
    parameter), timed on a bounded prefix of steps and projected."""
    import oracle
    orc = oracle.load()
    state = np.zeros((12, n, n, n))
    cores = cpu_threads()
    orc.wv_run(p, n, n, 0, 1, state, None, None, None, (-1, -1, -1), amp, ph, cores)   # warm-up
    t0 = time.perf_counter()
    done = 0
    while time.perf_counter() - t0 < budget_s or done < 2:
        orc.wv_run(p, n, n, 1 + done, 1, state, None, None, None, (-1, -1, -1), amp, ph, cores)
        done += 1
    el = time.perf_counter() - t0
    return {"value": n ** 3 * done / el / 1e6, "unit": "Mcell-steps/s", "cores": cores, "kind": "port",
            "sample": f"FP64 oracle, {n}^3 phased array: {done} forward steps on {cores} threads in {el:.1f} s "
                      "after 1 warm-up step"}


def read_traffic(profile: Path, kernel: str) -> float | None:
    """HBM bytes per launch of the step kernel `kernel` from the committed
    counter summary (line 'hbm_bytes_per_launch <kernel>: N'), or None."""
    try:
        for line in profile.read_text().splitlines():
            if line.startswith(f"hbm_bytes_per_launch {kernel}:"):
                return float(line.rsplit(":", 1)[1])
    except OSError:
        pass
    return None


def _timed(fn):
    import torch
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    out = fn()
    e1.record()
    torch.cuda.synchronize()
    return out, e0.elapsed_time(e1) / 1e3


def gilmore_figures(cpu_budget_s: float | None) -> dict:
    """1e6 bubbles, adaptive DP5(4) (rtol 1e-8, atol 1e-12), 20 us: FP64 in the
    fast arithmetic and in the reference's, and the FP32/FP64 split; 1,024
    bubbles, RK4 dt = 0.1 ns, 20 us, FP64 with R(t) every 1 ns; the FP64 oracle
    on the host for both (cpu_budget_s None: skipped), with each GPU variant's
    Rmax on the stratified 4,096-bubble subset against it (bubbles whose Rmax
    the oracle itself moves by > 1e-9 under a one-ulp input change are counted
    apart: tests/golden/gilmore_conditioning.json)."""
    import numpy as np
    from bubbledynamics_amd import gilmore as gm
    from bubbledynamics_amd._ffi import BD_GM_ARITH_REFERENCE, BD_GM_DP45, BD_GM_RK4, BD_GM_SCHED_SLICED, BD_GM_SPLIT
    p = gm.workload_params()
    pref = gm.workload_params(arithmetic=BD_GM_ARITH_REFERENCE)
    R0, amp, ph, fr = gm.config3_inputs(1_000_000, 3)
    b = gm.BubbleBatch(R0, amp, ph, fr)
    gm.BubbleBatch(R0[:4096], amp[:4096], ph[:4096], fr).run(p, BD_GM_DP45, 2e-6, 1e-13)   # warm-up
    runs = {}
    for name, pp, prec in (("dp45_1e6_fp64", p, 0), ("dp45_1e6_fp64_reference", pref, 0),
                           ("dp45_1e6_split", p, BD_GM_SPLIT)):
        # launch shape from the autotuner (probe: every bubble capped at 500
        # attempts), then the sliced schedule (passes of 128 and 2048
        # attempts, the rest longest-first)
        tune, t_tune = _timed(lambda: b.autotune(pp, BD_GM_DP45, 20e-6, 1e-13, precision=prec, rtol=1e-8,
                                                 atol=1e-12))
        o, t = _timed(lambda: b.run(pp, BD_GM_DP45, 20e-6, 1e-13, precision=prec, rtol=1e-8, atol=1e-12,
                                    split_ratio=0.2, schedule=BD_GM_SCHED_SLICED))
        steps = int(o["n_acc"].sum().item())
        runs[name] = (o, {"bubble_steps_per_s": steps / t, "accepted_steps": steps,
                          "rejected": int(o["n_rej"].sum().item()), "seconds": t,
                          "collapsed_frac": float((o["t_collapse"] > 0).float().mean().item()),
                          "autotuned_shape": tune, "autotune_seconds": t_tune,
                          "schedule": "sliced (128, 2048, rest)"})
        log(f"{name}: {steps / t:.3e} bubble-steps/s")
    _, t_in = _timed(lambda: b.run(p, BD_GM_DP45, 20e-6, 1e-13, rtol=1e-8, atol=1e-12))
    runs["dp45_1e6_fp64"][1]["input_order_bubble_steps_per_s"] = runs["dp45_1e6_fp64"][1]["accepted_steps"] / t_in
    runs["dp45_1e6_fp64"][1]["arithmetic"] = "fast (shared subexpressions, no cancellation, FSAL)"
    runs["dp45_1e6_fp64_reference"][1]["arithmetic"] = "reference (G:40-179 expression order, 7 stages)"
    o64 = runs["dp45_1e6_fp64"][0]
    o32 = runs["dp45_1e6_split"][0]
    rel = (o32["Rmax"] / o64["Rmax"] - 1).abs().cpu().numpy()
    runs["dp45_1e6_split"][1]["rmax_rel_err_vs_fp64"] = {
        "median": float(np.median(rel)), "p90": float(np.quantile(rel, 0.9)),
        "p99": float(np.quantile(rel, 0.99)), "max": float(rel.max())}
    runs["dp45_1e6_split"][1]["collapse_flag_flips"] = float(
        ((o32["t_collapse"] > 0) != (o64["t_collapse"] > 0)).float().mean().item())
    R0c, ampc, phc, frc = gm.config0_inputs(1024, 0)
    bc = gm.BubbleBatch(R0c, ampc, phc, frc)
    oc, tc = _timed(lambda: bc.run(p, BD_GM_RK4, 20e-6, 1e-10, out_every=10, n_out=20001))
    sc = int(oc["n_acc"].sum().item())
    out = {k: v[1] for k, v in runs.items()}
    out.update({
        "rk4_1024_fp64": {"bubble_steps_per_s": sc / tc, "steps": sc, "seconds": tc,
                          "collapsed_frac": float((oc["t_collapse"] > 0).float().mean().item())},
        "unit": "bubble-steps/s", "tableau": "Dormand-Prince 1980"})
    if cpu_budget_s is not None:
        cpu = gilmore_cpu_baseline(cpu_budget_s)
        idx = gm.config3_subset()
        ref = np.array(cpu.pop("dp45_subset_rmax"))
        ref_c = np.array(cpu.pop("dp45_subset_collapsed"))
        cond = json.loads((ROOT / "tests" / "golden" / "gilmore_conditioning.json").read_text())
        ill = set(cond["config3_subset_ulp"]["positions"])
        well = np.array([i not in ill for i in range(len(idx))])
        for name, (o, rec) in runs.items():
            got = o["Rmax"].cpu().numpy()[idx]
            got_c = (o["t_collapse"] > 0).cpu().numpy()[idx]
            e = np.abs(got / ref - 1)
            rec["rmax_rel_err_vs_oracle_subset4096"] = {
                "median": float(np.median(e)), "p99": float(np.quantile(e, 0.99)), "max": float(e.max()),
                "well_conditioned": int(well.sum()), "well_conditioned_max": float(e[well].max()),
                "well_conditioned_over_1e-6": int(np.sum(e[well] > 1e-6)),
                "collapse_flag_flips": int(np.sum(got_c != ref_c))}
        cpu["rk4_1024"]["gpu_x"] = (sc / tc) / cpu["rk4_1024"]["value"]
        cpu["dp45_1e6"]["gpu_x"] = runs["dp45_1e6_fp64"][1]["bubble_steps_per_s"] / cpu["dp45_1e6"]["value"]
        cpu["dp45_1e6"]["gpu_reference_arithmetic_x"] = (runs["dp45_1e6_fp64_reference"][1]["bubble_steps_per_s"] /
                                                         cpu["dp45_1e6"]["value"])
        out["cpu_baseline"] = cpu
    return out


def adjoint_figures(cpu_budget_s: float | None = None) -> dict:
    """128^3, 64-element array, 1,500 steps, FP32; checkpoints every 50 steps and
    every level kept; cost (forward + backward phase wall time) as a multiple of
    the forward solve.  The segment recompute overlaps the adjoint sweep.  With
    cpu_budget_s, the FP64 oracle's forward rate on the host and the projected
    host time of the central-difference gradient the configuration checks (8
    elements, amplitude and phase: 32 forward solves) beside the GPU gradient of
    all 64 elements."""
    import math
    import torch
    from bubbledynamics_amd import presets
    from bubbledynamics_amd.adjoint import adjoint_gradient, objective
    from bubbledynamics_amd.gilmore import seeded_uniforms
    n, steps, ne = 128, 1500, 8
    p = presets.phased_array(n, n_el=ne, dx=2e-4)
    u = seeded_uniforms(4, 2 * ne * ne)
    ph, amp = 2 * math.pi * u[:ne * ne], 0.5e6 + 0.5e6 * u[ne * ne:]
    obj = objective(n, n, steps, (n // 2, n // 2, n // 2), 5.0)
    adjoint_gradient(p, n, 60, amp, ph, objective(n, n, 60, (n // 2,) * 3, 5.0), 50, torch.float32)  # warm-up
    out = {}
    for K in (50, steps):
        J, gA, gP, tm = adjoint_gradient(p, n, steps, amp, ph, obj, K, torch.float32)
        out[f"checkpoint_{K}"] = {"forward_ms": tm[0], "recompute_outside_sweep_ms": tm[1], "sweep_ms": tm[2],
                                  "cost_x_forward": (tm[0] + tm[1] + tm[2]) / tm[0], "J": J,
                                  "grad_amp_norm": float(np.linalg.norm(gA))}
    if cpu_budget_s is not None:
        cpu = adjoint_cpu_baseline(p, n, amp, ph, cpu_budget_s)
        fwd_s = n ** 3 * steps / (cpu["value"] * 1e6)
        grad_gpu_s = sum(out["checkpoint_50"][k] for k in ("forward_ms", "recompute_outside_sweep_ms",
                                                           "sweep_ms")) / 1e3
        cpu["forward_solve_s_projected"] = fwd_s
        cpu["fd_gradient_8_elements_s_projected"] = 32 * fwd_s
        cpu["gpu_gradient_64_elements_s"] = grad_gpu_s
        cpu["gpu_forward_x"] = fwd_s / (out["checkpoint_50"]["forward_ms"] / 1e3)
        out["cpu_baseline"] = cpu
    return out


def coupled_figures(cpu_budget_s: float | None = None) -> dict:
    """256^3 CW Westervelt -> Q = 2 alpha I (5 Np/m/MHz) -> Pennes, 49-point spiral, 1 s dwell."""
    import torch
    from bubbledynamics_amd import workloads as wl
    n = 256
    Q, pw, tq = wl.heat_source(n)
    run, secs, nsub = wl.pennes_spiral(Q, n)
    cells = (n - 2) ** 3
    rate = cells * nsub / secs
    T = run.temperature()
    # substeps run in pairs: a pair reads T, Q, TD, TD_comp and writes T'', TD,
    # TD_comp once (7 FP32 words per cell), 14 B per cell-substep; the
    # measured traffic per pair is in the committed counter summary
    traffic = read_traffic(ROOT / "profiles" / "bioheat_pair_256_trace.txt", "bh_march2")
    pair_cells = 2 * cells
    out = {"westervelt_steps": tq["steps"], "westervelt_s": tq["settle_s"] + tq["window_s"],
            "q_max_w_m3": float(Q.max().item()), "bioheat_substeps": nsub, "bioheat_s": secs,
            "bioheat_mcell_steps_per_s": rate / 1e6,
            "bioheat_alg_bytes_per_cell_substep": 14,
            "bioheat_hbm_frac_14B": rate * 14 / PEAK_HBM,
            "bioheat_actual_gb_s": None if traffic is None else rate / pair_cells * traffic / 1e9,
            "bioheat_single_substep_equiv_gb_s_28B": rate * 28 / 1e9,
            "peak_rise_k": float((T.max() - run.p.T_art).item()),
            "max_cem43_min": float(run.dose().max().item() / 60.0)}
    if cpu_budget_s is not None:
        del run, T
        torch.cuda.empty_cache()
        cpu = bioheat_cpu_baseline(n, cpu_budget_s)
        cpu["gpu_x"] = rate / cpu["value"]
        cpu["gpu_x_serial"] = rate / cpu["serial_1core"]
        out["bioheat_cpu_baseline"] = cpu
    return out


def bioheat_shapes_figures() -> dict:
    """FP32 Pennes substeps with the compensated dose at several grid widths:
    pairs against the two region shapes of the three-substep kernel (forced
    through BD_BH_FUSE / BD_BH_VARIANT3, which the library reads on every
    call), and the choice bd_bioheat_run measures by itself; then FP64 pairs
    as the reference-precision thermal program runs them (plain dose, stored
    temperatures); us per substep, best of three runs of 60 substeps after a
    6-substep warm-up."""
    import torch
    from bubbledynamics_amd import bioheat as bh
    p = bh.pennes_params(dx=2e-4, dt=0.05, substeps=2)
    out = {}
    for n, nz in ((256, 256), (426, 300), (851, 200), (1024, 100)):
        Q = torch.zeros((nz, n, n), dtype=torch.float32, device="cuda")
        Q[nz // 2, n // 2, n // 2] = 1e9
        row = {}
        for name, env in (("pairs", {"BD_BH_FUSE": "2"}), ("triples_64x32", {"BD_BH_VARIANT3": "0"}),
                          ("triples_128x16", {"BD_BH_VARIANT3": "1"}), ("measured_choice", {}), ("pairs_f64", {})):
            saved = {k: os.environ.pop(k, None) for k in ("BD_BH_FUSE", "BD_BH_VARIANT3")}
            os.environ.update(env)
            try:
                if name.endswith("_f64"):
                    run = bh.BioheatRun(p, n, nz, Q.double(), p.T_art, dtype=torch.float64, kahan=False,
                                        offset=False)
                else:
                    run = bh.BioheatRun(p, n, nz, Q, p.T_art, dtype=torch.float32, kahan=True, offset=True)
                run.step(0, 0, 6)
                torch.cuda.synchronize()
                best = 1e9
                for _ in range(3):
                    t0 = time.perf_counter()
                    run.step(0, 0, 60)
                    torch.cuda.synchronize()
                    best = min(best, time.perf_counter() - t0)
                row[name] = round(best / 60 * 1e6, 2)
                del run
            finally:
                for k in ("BD_BH_FUSE", "BD_BH_VARIANT3"):
                    os.environ.pop(k, None)
                    if saved[k] is not None:
                        os.environ[k] = saved[k]
        out[f"{n}x{n}x{nz}_us_per_substep"] = row
        del Q
        torch.cuda.empty_cache()
        log(f"bioheat shapes {n}x{n}x{nz}: {row}")
    return out


def reference_scale(shifts: int, compare: int = -1, out_dir: str | None = None, thermal_f64: bool = False,
                    npy_dir: str | None = None) -> dict:
    """The reference's heat-source and thermal programs at their own grids."""
    import ctypes as C
    import tempfile
    import torch
    from bubbledynamics_amd import _ffi, bioheat as bh, workloads as wl
    from bubbledynamics_amd._ffi import BD_F64
    out = {}
    t0 = time.perf_counter()
    stamps = {}

    def progress(stage, step):
        now = time.perf_counter() - t0
        stamps[stage] = now
        log(f"{stage} {step} steps, {now:.1f} s")

    Q, p, info = wl.reference_scenario(full_scale=True, settle_periods=165, periods=4, n_harm=10,
                                       progress=progress)
    nz, nxy = Q.shape[0], Q.shape[1]
    steps = info["settle"] + info["water"] + info["harmonics"]
    cells = nxy * nxy * nz
    t_w = time.perf_counter() - t0
    iz_t = info["iz_tissue"]
    qt = Q[iz_t:]
    k = int(torch.argmax(qt).item())
    out["westervelt"] = {
        "grid": [nxy, nxy, nz], "dtype": "f64", "steps": steps, "seconds": round(t_w, 2),
        "settle_s": round(stamps["settle"], 2),
        "mcell_steps_per_s": round(cells * steps / t_w / 1e6, 1),
        "q_tissue_max_w_m3": float(qt.max().item()),
        "q_tissue_argmax_izyx": [iz_t + k // (nxy * nxy), (k // nxy) % nxy, k % nxy],
        "q_water_max_w_m3": float(Q[:iz_t].max().item()),
        "finite": bool(torch.isfinite(Q).all().item())}
    log(f"westervelt: {out['westervelt']}")
    st = C.c_void_p(torch.cuda.current_stream().cuda_stream)
    with tempfile.TemporaryDirectory() as d:
        path = os.path.join(d, "Q3D_reference.bin")
        t1 = time.perf_counter()
        _ffi.call("bd_q3d_write", path.encode(), BD_F64, C.c_void_p(Q.data_ptr()), nxy, nz, 2, st)
        t_write = time.perf_counter() - t1
        size = os.path.getsize(path)
        del Q, qt
        torch.cuda.empty_cache()
        bp, nb, nzb = bh.reference_params()
        assert nb == 2 * (nxy - 1) + 1 and nzb == nz - 2
        Qf = torch.empty((nzb, nb, nb), dtype=torch.float64, device="cuda")
        t1 = time.perf_counter()
        _ffi.call("bd_q3d_read_expand", path.encode(), BD_F64, C.c_void_p(Qf.data_ptr()), nb, nzb, st)
        t_read = time.perf_counter() - t1
    out["q3d"] = {"bytes": size, "write_s": round(t_write, 2), "read_expand_s": round(t_read, 2)}
    log(f"q3d: {out['q3d']}")
    sx, sy = bh.spiral_reference()
    n = len(sx) if shifts <= 0 else min(shifts, len(sx))

    def runs_of(lo, hi):
        # consecutive equal shifts run as one library call
        out_runs, i = [], lo
        while i < hi:
            j = i
            while j + 1 < hi and sx[j + 1] == sx[i] and sy[j + 1] == sy[i]:
                j += 1
            out_runs.append((int(sx[i]), int(sy[i]), j - i + 1))
            i = j + 1
        return out_runs

    def advance(run, lo, hi, label):
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        done, mark = lo, lo
        for a_, b_, m in runs_of(lo, hi):
            run.step(a_, b_, m)
            done += m
            if done - mark >= 4000:
                mark = done
                torch.cuda.synchronize()
                log(f"{label}: {done}/{hi} substeps, {time.perf_counter() - t1:.1f} s")
        torch.cuda.synchronize()
        return time.perf_counter() - t1

    bcells = (nb - 2) * (nb - 2) * (nzb - 2)
    T0_tissue = 23.6 + 273.15                                  # T:149
    ends = bh.spiral_reference_line_ends()
    assert ends[-1] == len(sx)
    # FP64 drop-in precision on the first `compare` substeps (default: the first spiral line)
    m_cmp = 0 if thermal_f64 else min(ends[0] if compare < 0 else compare, n)
    ref64 = None
    if m_cmp > 0:
        run64 = bh.BioheatRun(bp, nb, nzb, Qf, 0.0, dtype=torch.float64, kahan=False, offset=False)
        run64.set_reference_initial()
        t64 = advance(run64, 0, m_cmp, "FP64")
        ref64 = (run64.temperature(), run64.dose())
        out["bioheat_fp64"] = {"substeps": m_cmp, "seconds": round(t64, 2),
                               "mcell_steps_per_s": round(bcells * m_cmp / t64 / 1e6, 1)}
        log(f"bioheat_fp64: {out['bioheat_fp64']}")
        del run64
        torch.cuda.empty_cache()
    # FP32 with fields stored as T - T0 and a compensated dose (or FP64 as the
    # reference), all substeps
    tag = "FP64" if thermal_f64 else "FP32"
    if thermal_f64:
        run = bh.BioheatRun(bp, nb, nzb, Qf, 0.0, dtype=torch.float64, kahan=False, offset=False)
    else:
        p32 = bp.copy()
        p32.T_art = T0_tissue                  # the storage offset (no perfusion in this program)
        run = bh.BioheatRun(p32, nb, nzb, Qf, 0.0, dtype=torch.float32, kahan=True, offset=True)
    del Qf
    torch.cuda.empty_cache()
    run.set_reference_initial()
    t_b = advance(run, 0, m_cmp, tag)
    if ref64 is not None:
        T64, D64 = ref64
        dT = (run.temperature() - T64).abs().max().item()
        sel = D64 > 1e-3 * D64.max()
        drel = ((run.dose()[sel] / D64[sel]) - 1).abs().max().item()
        out["bioheat_fp64"].update({"fp32_max_abs_dT_k": dT, "fp32_rise_k": float((T64.max() - T0_tissue).item()),
                                    "fp32_dose_max_rel_err": drel})
        log(f"FP32 vs FP64 after {m_cmp} substeps: max |dT| {dT:.2e} K, dose rel err {drel:.2e}")
        del T64, D64, ref64, sel
        torch.cuda.empty_cache()
    for d_ in (out_dir, npy_dir):
        if d_:
            os.makedirs(d_, exist_ok=True)
    from bubbledynamics_amd._ffi import BD_SLICE_XY, BD_SLICE_XZ

    def dump(name, t, off):
        # float64 copy of a slice for analysis (the text slices keep 6 digits)
        np.save(os.path.join(npy_dir, name), (t.double() + off).cpu().numpy())
    # the reference's T slices after spiral lines 1, 5, 9, 13 (T:411-563)
    done = m_cmp
    for k, line in enumerate((1, 5, 9, 13)):
        stop = min(ends[line - 1], n)
        if stop > done:
            t_b += advance(run, done, stop, tag)
            done = stop
        if out_dir and done == ends[line - 1]:
            run.write_slice(os.path.join(out_dir, f"T_XZ_loop{k}.txt"), BD_SLICE_XZ, 425, 1200, 1701, 325, 526)
            run.write_slice(os.path.join(out_dir, f"T_XY_loop{k}_z15.txt"), BD_SLICE_XY, 1500, 325, 526, 325, 526)
        if npy_dir and done == ends[line - 1]:
            t = run.T[run.cur]
            dump(f"T_XZ_loop{k}.npy", t[1200:1701, 425, 325:526], run.p.T_off)
            dump(f"T_XY_loop{k}_z15.npy", t[1500, 325:526, 325:526], run.p.T_off)
    if n > done:
        t_b += advance(run, done, n, tag)
    T = run.temperature()
    out["bioheat"] = {"grid": [nb, nb, nzb], "dtype": "f64" if thermal_f64 else "f32 (T - T0, Kahan dose)",
                      "substeps": n,
                      "of_reference": len(sx), "calls": len(runs_of(0, n)), "seconds": round(t_b, 2),
                      "mcell_steps_per_s": round(bcells * n / t_b / 1e6, 1),
                      "t_max_c": float(T.max().item() - 273.15),
                      "dose_max_min": float(run.dose().max().item() / 60.0),
                      "finite": bool(torch.isfinite(T).all().item())}
    if out_dir:
        # the reference's end-of-run dose slices (T:573-634)
        run.write_slice(os.path.join(out_dir, "TD_XZ.txt"), BD_SLICE_XZ, 425, 1200, 1601, 325, 527, "TD")
        for iz, tag in ((1500, "z15"), (1350, "z13_5"), (1250, "z12_5")):
            run.write_slice(os.path.join(out_dir, f"TD_XY_{tag}.txt"), BD_SLICE_XY, iz, 325, 526, 325, 527, "TD")
        out["bioheat"]["slices"] = out_dir
    if npy_dir:
        D = run.dose()
        dump("TD_XZ.npy", D[1200:1601, 425, 325:527], 0.0)
        for iz, tag in ((1500, "z15"), (1350, "z13_5"), (1250, "z12_5")):
            dump(f"TD_XY_{tag}.npy", D[iz, 325:526, 325:527], 0.0)
        del D
    log(f"bioheat: {out['bioheat']}")
    return out


def threshold_table(first: int, deadline: float, out_dir: str, reference_tableau: bool) -> dict:
    """The reference's threshold table (G:498-606: 27,775 triplets x 100 radii)
    from cell `first` on, stopping to claim cells after `deadline` seconds; the
    done cells go to out_dir/sweep_<first>_<end>.npy, the table file is written
    by --sweep-assemble once the pieces cover it."""
    from bubbledynamics_amd import gilmore as gm
    p = gm.reference_params()
    p.tableau = 0 if reference_tableau else 1
    fr = gm.freq_table()
    n_r0 = 100
    total = len(fr) * n_r0
    out, end, secs = gm.threshold_sweep_range(p, fr, n_r0, cell0=first, deadline_s=deadline or None, poll_s=30.0,
                                              log=log)
    part = out[first:end]
    os.makedirs(out_dir, exist_ok=True)
    np.save(os.path.join(out_dir, f"sweep_{first}_{end}.npy"), part)
    with open(os.path.join(out_dir, f"sweep_{first}_{end}.json"), "w") as f:
        json.dump({"first": first, "end": end, "seconds": secs}, f)
    found = part[(part > 0) & (part < 1e7)]
    return {"workload": "threshold_sweep", "tableau": "reference" if reference_tableau else "Dormand-Prince 1980",
            "cells": total, "first": first, "done_end": end, "seconds": round(secs, 1),
            "cells_per_s": round((end - first) / secs, 1) if secs > 0 else None,
            "no_collapse_1e7": int(np.sum(part == 1e7)), "unset_minus1": int(np.sum(part == -1.0)),
            "threshold_mpa_median": float(np.median(found) / 1e6) if found.size else None}


def threshold_table_bounded(first: int, deadline: float, out_dir: str, reference_tableau: bool,
                            budget: int) -> dict:
    """The same table from cell `first` on with the resumable sweep: launches
    of at most `budget` DP5(4) attempts per lane, every launch's finished cells
    saved (out_dir/sweep_partial_<first>.npy, NaN where a cell is not done),
    and no launch started after `deadline` seconds, so the run ends within one
    launch of the deadline whatever its cells cost."""
    import torch
    from bubbledynamics_amd import gilmore as gm
    p = gm.reference_params()
    p.tableau = 0 if reference_tableau else 1
    fr = gm.freq_table()
    n_r0 = 100
    total = len(fr) * n_r0
    os.makedirs(out_dir, exist_ok=True)
    sw = gm.ThresholdSweep(p, fr, n_r0, cell0=first)
    t0 = time.perf_counter()
    pr = {}
    while True:
        sw.advance(budget)
        torch.cuda.synchronize()
        pr = sw.progress()
        el = time.perf_counter() - t0
        part = sw.results[first:].cpu().numpy()
        np.save(os.path.join(out_dir, f"sweep_partial_{first}.npy"), part)
        done = int(np.sum(~np.isnan(part)))
        log(f"sweep: launch {sw.advances}, {done} of {total - first} cells done, {pr['suspended']} suspended, "
            f"{el:.0f} s")
        if pr["all_done"] or el > deadline > 0:
            break
    secs = time.perf_counter() - t0
    sw.close()
    found = part[(part > 0) & (part < 1e7)]
    info = {"workload": "threshold_sweep_bounded", "tableau": "reference" if reference_tableau else
            "Dormand-Prince 1980", "cells": total, "first": first, "done": done, "all_done": bool(pr["all_done"]),
            "launches": sw.advances, "budget_attempts_per_lane": budget, "seconds": round(secs, 1),
            "no_collapse_1e7": int(np.sum(part == 1e7)), "capped_minus2": int(np.sum(part == -2.0)),
            "unset_minus1": int(np.sum(part == -1.0)),
            "threshold_mpa_median": float(np.median(found) / 1e6) if found.size else None}
    with open(os.path.join(out_dir, f"sweep_partial_{first}.json"), "w") as f:
        json.dump(info, f)
    return info


def assemble_table(out_dir: str, path: str) -> dict:
    """Joins the sweep pieces of out_dir into the reference's table file."""
    import glob
    from bubbledynamics_amd import gilmore as gm
    fr = gm.freq_table()
    n_r0 = 100
    res = np.full(len(fr) * n_r0, np.nan)
    for f in sorted(glob.glob(os.path.join(out_dir, "sweep_*_*.npy"))):
        a, b = (int(x) for x in os.path.basename(f)[6:-4].split("_"))
        res[a:b] = np.load(f)
    seconds = 0.0
    for f in glob.glob(os.path.join(out_dir, "sweep_*_*.json")):
        with open(f) as fh:
            seconds += json.load(fh)["seconds"]
    missing = int(np.isnan(res).sum())
    if missing == 0:
        p = gm.reference_params()
        gm.write_sweep_table(path, fr, res, n_r0, p.minR0, p.stepR0, seconds)
    return {"workload": "threshold_sweep_assemble", "cells": len(res), "missing": missing,
            "table": path if missing == 0 else None}


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3000)
    ap.add_argument("--warmup", type=int, default=60)
    ap.add_argument("--n", type=int, default=192)
    ap.add_argument("--reps", type=int, default=5, help="timed runs of --steps steps (median reported; >= 5)")
    ap.add_argument("--cpu-seconds", type=float, default=8.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-fp64", action="store_true")
    ap.add_argument("--no-extras", action="store_true",
                    help="skip the Gilmore, adjoint and coupled bioheat figures")
    ap.add_argument("--figure", choices=("gilmore", "adjoint", "coupled", "bioheat_shapes"), default=None,
                    help="run only that auxiliary workload and print its JSON (for profiling)")
    ap.add_argument("--calib-copy-mb", type=int, default=0,
                    help="first copy a buffer of this size 3 times (HBM counter calibration)")
    ap.add_argument("--reference-scenario", action="store_true",
                    help="run the reference's programs at their own scale instead (one JSON line)")
    ap.add_argument("--shifts", type=int, default=0,
                    help="spiral substeps of the thermal program in --reference-scenario (0: all 77,474)")
    ap.add_argument("--compare", type=int, default=-1,
                    help="substeps also run in FP64 to state the FP32 error in --reference-scenario "
                         "(-1: the first spiral line, 1,802)")
    ap.add_argument("--slices", default="",
                    help="directory for the reference's end-of-run slice files in --reference-scenario")
    ap.add_argument("--slices-npy", default="",
                    help="directory for float64 .npy copies of the same slices in --reference-scenario")
    ap.add_argument("--thermal-dtype", choices=("f32", "f64"), default="f32",
                    help="precision of the thermal program in --reference-scenario")
    ap.add_argument("--threshold-sweep", action="store_true",
                    help="run the reference's threshold table from --sweep-first on (one JSON line)")
    ap.add_argument("--sweep-first", type=int, default=0)
    ap.add_argument("--sweep-deadline", type=float, default=0.0,
                    help="seconds after which the sweep stops claiming cells (0: run to the end); "
                         "cells already claimed still run to their end")
    ap.add_argument("--sweep-out", default="sweep_parts", help="directory of the sweep pieces")
    ap.add_argument("--sweep-budget", type=int, default=0,
                    help="resumable sweep: DP5(4) attempts per lane per launch (0: one launch, claimed cells "
                         "run to their end); the finished cells of every launch are saved")
    ap.add_argument("--sweep-reference-tableau", action="store_true",
                    help="the reference's a32 = -9/40 instead of the published Dormand-Prince tableau")
    ap.add_argument("--sweep-assemble", default="",
                    help="write the table file from the pieces in --sweep-out (no GPU)")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU only: the multi-GPU z-slab decomposition is not built; "
              "run with --gpus 1", file=sys.stderr)
        return 2

    if args.sweep_assemble:
        print(json.dumps(assemble_table(args.sweep_out, args.sweep_assemble)), flush=True)
        return 0

    import numpy as np
    import torch
    from bubbledynamics_amd import _ffi, presets
    from bubbledynamics_amd.westervelt import WesterveltRun, focal_harmonics

    if not torch.cuda.is_available():
        print("no GPU visible", file=sys.stderr)
        return 3
    _ffi.load()
    if args.threshold_sweep and args.sweep_budget > 0:
        res = threshold_table_bounded(args.sweep_first, args.sweep_deadline, args.sweep_out,
                                      args.sweep_reference_tableau, args.sweep_budget)
        print(json.dumps({"n_gpus": 1, **res}), flush=True)
        return 0
    if args.threshold_sweep:
        res = threshold_table(args.sweep_first, args.sweep_deadline, args.sweep_out, args.sweep_reference_tableau)
        print(json.dumps({"n_gpus": 1, **res}), flush=True)
        return 0
    if args.reference_scenario:
        res = reference_scale(args.shifts, args.compare, args.slices or None, args.thermal_dtype == "f64",
                              args.slices_npy or None)
        print(json.dumps({"workload": "reference_scenario", "n_gpus": 1, **res}), flush=True)
        return 0
    if args.calib_copy_mb:
        src = torch.ones(args.calib_copy_mb * 2 ** 20 // 4, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst
    if args.figure:
        fn = {"gilmore": lambda: gilmore_figures(None), "adjoint": adjoint_figures,
              "coupled": coupled_figures, "bioheat_shapes": bioheat_shapes_figures}[args.figure]
        print(json.dumps({"figure": args.figure, "n_gpus": 1, **fn()}), flush=True)
        return 0
    n, K, W = args.n, args.steps, args.warmup
    reps = max(5, args.reps)
    p = presets.tone_burst_192(n)
    focus = presets.tone_burst_focus(p, n)

    def timed_runs(dtype):
        # the plan captures both 6-step graphs when it is built, so the timed
        # region holds kernel launches and graph replays only; every timed run
        # is the workload itself, K steps of the burst from rest (the state,
        # peak pressure and records are cleared and the plan sought back to
        # step 0 before it, outside the timed region)
        run = WesterveltRun(p, n, n, dtype=dtype, max_steps=max(W, K) + 1, focus=focus)
        run.run(W)
        torch.cuda.synchronize()
        times = []
        for _ in range(reps):
            for t in (run.state, run.pmax, run.energy, run.focus):
                t.zero_()
            run.seek(0)
            ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            torch.cuda.synchronize()
            ev0.record()
            run.run(K)
            ev1.record()
            torch.cuda.synchronize()
            times.append(ev0.elapsed_time(ev1))
        return run, times

    log(f"FP32 run {n}^3: {W} warm-up steps, then {reps} timed runs of {K} steps, each from rest")
    run32, times = timed_runs(torch.float32)
    ms = float(np.median(times))
    value = n ** 3 * K / (ms / 1e3) / 1e6
    log(f"FP32 {value:.1f} Mcell-steps/s ({ms / K * 1e3:.2f} us/step; runs {[round(t, 3) for t in times]} ms)")
    shape = run32.shape
    # The timed region launches one step kernel per step (eager until the ring
    # phase is 0 mod 6, then graph replays) and, at the end of each run, the
    # fold of the levels after the last 6-step pmax window; its time per step
    # bounds the step kernel's average duration from above and the roofline
    # uses it.  For comparison, the same kernel relaunched alone on the stream
    # it runs on (these launches recompute the next step and touch pmax).
    step_ms = ms / K
    R = 50
    lib = _ffi.load()
    stream = torch.cuda.current_stream().cuda_stream
    import ctypes as C
    for _ in range(5):
        _ffi.check("rerun_step", lib.bd_westervelt_plan_rerun_step(run32._plan, C.c_void_p(stream)))
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    # the relaunches queue behind a short device-side wait (a few ms at most),
    # so that they run back to back whatever the host's launch rate
    torch.cuda._sleep(5_000_000)
    e0.record()
    for _ in range(R):
        _ffi.check("rerun_step", lib.bd_westervelt_plan_rerun_step(run32._plan, C.c_void_p(stream)))
    e1.record()
    torch.cuda.synchronize()
    k_ms = e0.elapsed_time(e1) / R
    cells = int(lib.bd_westervelt_plan_interior_cells(run32._plan))
    edge_cells = n ** 3 - cells
    # inside cell: 9 reads + 3 writes + pmax read/modify/write; face / source
    # cell: read p^n, write p^{n+1}, pmax read/modify/write (FP32)
    alg_bytes = cells * 14 * 4 + edge_cells * 4 * 4
    achieved = alg_bytes / (step_ms / 1e3)
    log(f"step kernel {step_ms * 1e3:.2f} us per step in the timed region ({k_ms * 1e3:.2f} us relaunched alone) "
        f"-> {achieved / 1e9:.0f} GB/s algorithmic; shape {shape}")
    kname = "wv_march<float, %d, %d, %d, %d, %d>" % (
        shape["tile_x"], shape["tile_y"], shape["rows"], shape["prefetch"], shape["nontemporal"])
    del run32
    torch.cuda.empty_cache()

    # Accuracy leg, independent of --steps: the whole 3,000-step tone burst
    # (the burst reaches the focus after ~500 steps) in FP32 and in FP64 from
    # rest; focal record, pmax, energy and focal harmonics.
    acc_steps = 3000
    full = {}
    for name, dt_ in (("f32", torch.float32), ("f64", torch.float64)):
        if name == "f64" and args.no_fp64:
            continue
        r = WesterveltRun(p, n, n, dtype=dt_, max_steps=acc_steps, focus=focus)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        r.run(acc_steps)
        r.finalize()
        torch.cuda.synchronize()
        el = time.perf_counter() - t0
        full[name] = {"seconds": el, "focus": r.focus.double().cpu().numpy(), "pmax": r.pmax.double(),
                      "energy": r.energy.cpu().numpy()}
        del r
        log(f"{name} whole {acc_steps}-step run: {el:.3f} s ({n ** 3 * acc_steps / el / 1e6:.0f} Mcell-steps/s)")
    f32 = full["f32"]["focus"]
    harm = focal_harmonics(f32, p.dt, p.fr, 3)
    extra = {"run_3000": {"steps": acc_steps, "f32_seconds": round(full["f32"]["seconds"], 4),
                          "f32_mcell_steps_per_s": round(n ** 3 * acc_steps / full["f32"]["seconds"] / 1e6, 1)}}
    accuracy = {"focal_rel_l2_fp32_vs_fp64": None, "pmax_rel_l2_fp32_vs_fp64": None,
                "record": f"focus sample of every step of the whole {acc_steps}-step run"}
    if "f64" in full:
        f64 = full["f64"]["focus"]
        nf = float(np.linalg.norm(f64))
        accuracy["focal_rel_l2_fp32_vs_fp64"] = float(np.linalg.norm(f32 - f64) / nf) if nf > 0 else None
        p64 = full["f64"]["pmax"]
        accuracy["pmax_rel_l2_fp32_vs_fp64"] = float(
            (torch.linalg.vector_norm(full["f32"]["pmax"] - p64) / torch.linalg.vector_norm(p64)).item())
        extra["run_3000"]["f64_seconds"] = round(full["f64"]["seconds"], 4)
        extra["run_3000"]["f64_mcell_steps_per_s"] = round(n ** 3 * acc_steps / full["f64"]["seconds"] / 1e6, 1)
        extra["run_3000"]["fp32_speedup_vs_fp64"] = round(full["f64"]["seconds"] / full["f32"]["seconds"], 3)
        log(f"FP32 vs FP64 over {acc_steps} steps: focal rel L2 {accuracy['focal_rel_l2_fp32_vs_fp64']}, "
            f"pmax rel L2 {accuracy['pmax_rel_l2_fp32_vs_fp64']:.2e}")
    energy = full["f32"]["energy"]
    focus_out = {"peak_pa": float(np.abs(f32).max()), "harmonics_pa": [float(h) for h in harm],
                 "energy_max_j": float(energy[:, 0].max())}
    del full
    torch.cuda.empty_cache()

    if not args.no_extras:
        gil_cpu = None if args.no_cpu else args.cpu_seconds
        for name, fn in (("gilmore", lambda: gilmore_figures(gil_cpu)), ("adjoint", lambda: adjoint_figures(gil_cpu)),
                         ("coupled_bioheat", lambda: coupled_figures(gil_cpu))):
            try:
                log(f"{name} ...")
                extra[name] = fn()
                log(f"{name}: {json.dumps(extra[name])}")
            except Exception as exc:  # an auxiliary figure never blocks the main line
                log(f"{name} failed: {exc!r}")
                extra[name] = {"error": repr(exc)}

    cpu = None
    if not args.no_cpu:
        log(f"CPU baseline (~{1.5 * args.cpu_seconds:.0f} s)")
        try:
            cpu = cpu_baseline(n, args.cpu_seconds)
            cpu["gpu_x"] = round(value / cpu["value"], 1)
            cpu["gpu_x_serial"] = round(value / cpu["serial_1core"], 1)
        except Exception as exc:  # the baseline is reported, never required
            log(f"cpu baseline failed: {exc}")

    traffic = read_traffic(ROOT / "profiles" / f"westervelt_march_{n}_pmc.txt", kname)
    out = {
        "metric": "Westervelt Mcell-steps/s & %HBM roofline; Gilmore bubble-steps/s; vs FP64 plain",
        "value": round(value, 2),
        "unit": "Mcell-steps/s",
        "n_gpus": 1,
        "steps": K,
        "warmup": W,
        "ms_per_step": step_ms,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "timing": {"repetitions": reps, "statistic": "median", "run_ms": [round(t, 4) for t in times]},
        "config": {"workload": f"westervelt_tone_burst_{n}", "grid": [n, n, n], "dx_m": p.dxy,
                   "dt_s": p.dt, "relaxation": 2, "source": "cap F#1, 1 MHz, 8 cycles, 1 MPa",
                   "boundary": "Mur + 20-cell sponge", "hipgraph": "2 x 6-step, captured before timing"},
        "roofline": {"bound": "hbm", "achieved": round(achieved / 1e9, 1), "peak": PEAK_HBM / 1e9,
                     "unit": "GB/s", "frac": round(achieved / PEAK_HBM, 4),
                     "traffic": traffic,
                     "kernel": kname, "kernel_us": round(step_ms * 1e3, 2),
                     "kernel_us_relaunched_alone": round(k_ms * 1e3, 2),
                     "frac_relaunched_alone": round(alg_bytes / (k_ms / 1e3) / PEAK_HBM, 4),
                     "workgroups": shape["workgroups"],
                     "alg_bytes_per_launch": alg_bytes, "cells_per_launch": n ** 3,
                     "inside_cells_per_launch": cells},
        "cpu_baseline": cpu,
        "accuracy": accuracy,
        "focus": focus_out,
    }
    out.update(extra)
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
