"""The benchmark workloads as host drivers over the C ABI.

* coupled Westervelt -> Pennes (256^3, dx = 0.2 mm): CW cap source to a
  steady state, harmonic accumulators over the last 4 periods, Q = 2 alpha I
  with alpha = 5 Np/m/MHz per harmonic, then the FP32 Pennes stencil along a
  49-point dwell spiral with a Kahan-compensated CEM43 dose;
* batched Gilmore bubbles (RK4 and DP5(4), FP64 and the FP32/FP64 split);
* the adjoint phased-array gradient (128^3, 64 elements, 1,500 steps).
"""
from __future__ import annotations

import ctypes as C
import math
import time

import numpy as np
import torch

from . import _ffi, presets
from . import bioheat as bh
from ._ffi import BD_F32, BD_F64, WesterveltParams
from .westervelt import WesterveltRun


def cw_cap_full(n: int = 256, dx: float = 2e-4, p0: float = 0.25e6) -> WesterveltParams:
    """Full-domain CW cap (F-number 1, focal length 0.8 of the depth) in the
    water-like relaxing medium of the 192^3 workload, Mur faces behind a sponge."""
    p = presets.tone_burst_192(n)
    p.dxy = p.dz = dx
    p.dt = 0.3 * dx / p.c_inf
    p.full_domain = 1
    p.xy_min = -(n - 1) / 2 * dx
    p.z_min = -2 * dx
    p.d = 0.78 * (n - 2) * dx
    p.burst_cycles = 0.0
    p.burst_t0 = 0.0
    p.p0 = p0
    p.sponge_cells = max(4, round(20 * n / 192))
    p.sponge_alpha = 0.015 * 20 / p.sponge_cells
    return p


def heat_source(n: int = 256, dx: float = 2e-4, periods_settle: int = 42, periods_window: int = 4,
                n_harm: int = 5, abs_per_mhz: float = 5.0, dtype=torch.float32):
    """Q = sum_n 2 alpha_n I_n (W/m^3) from the steady CW field.  Returns
    (Q tensor (n,n,n) on the GPU, params, timings)."""
    p = cw_cap_full(n, dx)
    spp = round(1.0 / (p.fr * p.dt))                     # samples per period (25 at CFL 0.3)
    assert abs(spp * p.fr * p.dt - 1.0) < 1e-9
    n_settle = periods_settle * spp
    n_time = periods_window * spp
    run = WesterveltRun(p, n, n, dtype=dtype, max_steps=n_settle + n_time, pmax=False, energy=False)
    code = BD_F32 if dtype == torch.float32 else BD_F64
    H = torch.zeros((4 * n_harm, n, n, n), dtype=dtype, device="cuda")
    st = C.c_void_p(torch.cuda.current_stream().cuda_stream)
    lib = _ffi.load()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    run.run(n_settle)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    for i in range(n_time):
        run.run(1)
        _ffi.check("harmonics", lib.bd_harmonics_accumulate(code, C.c_void_p(H.data_ptr()),
                                                            C.c_void_p(run.pressure(0).data_ptr()), n, n, 0, i,
                                                            n_time, periods_window, n_harm, 1, st))
    Q = torch.zeros((n, n, n), dtype=dtype, device="cuda")
    _ffi.check("harmonics_q", lib.bd_harmonics_q(C.byref(p), code, C.c_void_p(Q.data_ptr()),
                                                C.c_void_p(H.data_ptr()), n, n, 0, n_time, n_harm,
                                                abs_per_mhz, 1, st))
    torch.cuda.synchronize()
    t2 = time.perf_counter()
    del H, run
    return Q, p, {"settle_s": t1 - t0, "window_s": t2 - t1, "steps": n_settle + n_time}


def pennes_spiral(Q: torch.Tensor, n: int, dx: float = 2e-4, dwell_s: float = 1.0, dt: float = 0.05,
                  substeps: int = 2, n_side: int = 7, side: float = 8e-3, dtype=torch.float32):
    """Dwell spiral: n_side^2 points, dwell_s each, dt split into `substeps`.
    Returns (BioheatRun, seconds, substeps run)."""
    p = bh.pennes_params(dx=dx, dt=dt, substeps=substeps)
    sx, sy = bh.spiral_dwell(n_side, side, dx)
    run = bh.BioheatRun(p, n, n, Q, p.T_art, dtype=dtype, kahan=True, offset=True)
    per_point = round(dwell_s / dt) * substeps
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for k in range(len(sx)):
        run.step(int(sx[k]), int(sy[k]), per_point)
    torch.cuda.synchronize()
    return run, time.perf_counter() - t0, per_point * len(sx)


def reference_scenario(nxy: int = 48, nz: int = 64, settle_periods: int = 6, periods: int = 4,
                       n_harm: int = 10, steps_per_period: int = 35, dtype=torch.float64,
                       q3d_path: str | None = None, full_scale: bool = False, progress=None,
                       chunk: int = 1000):
    """The reference's own heat-source computation (W:556-790) on a reduced
    grid: two-layer water/muscle medium, quarter domain with mirrors, Mur
    faces, cap source with a hole.

    1. run `settle_periods` periods (the reference: 165, W:417) and keep the
       state (W:600-613);
    2. one more period accumulating the water heat source (W:617-660);
    3. from the kept state, `periods` periods accumulating the harmonic bins
       of the tissue slab, then Q = sum_n 2 (7.2 n f/MHz) I_n there
       (W:663-755; running DFT bins in place of the slab re-runs and FFT);
    4. optionally write the Q3D file (W:778-790).
    dt is set so that a period is a whole number of steps (the reference's
    0.01/f gives 100).  full_scale=True takes the reference's own grid and
    constants instead (W:356-417: 426 x 426 x 2003, dt = 0.01/f, 100 steps a
    period; nxy, nz and steps_per_period are then ignored).  progress(stage,
    step) is called after every `chunk` settle steps and at the end of each
    stage.  Returns (Q (nz, nxy, nxy) on the GPU, params, steps)."""
    if full_scale:
        p, nxy, nz = presets.reference_westervelt()
        steps_per_period = round(1.0 / (p.fr * p.dt))
    else:
        p = presets.reference_scaled(nxy, nz)
        p.dt = 1.0 / (p.fr * steps_per_period)

    def note(stage, step):
        if progress is not None:
            torch.cuda.synchronize()
            progress(stage, step)
    n_settle = settle_periods * steps_per_period
    n_time = periods * steps_per_period
    total = n_settle + max(steps_per_period, n_time)
    run = WesterveltRun(p, nxy, nz, dtype=dtype, max_steps=total, pmax=False, energy=False)
    code = BD_F32 if dtype == torch.float32 else BD_F64
    lib = _ffi.load()
    st = C.c_void_p(torch.cuda.current_stream().cuda_stream)
    for s0 in range(0, n_settle, chunk):
        run.run(min(chunk, n_settle - s0))
        note("settle", min(s0 + chunk, n_settle))
    snapshot = run.state.clone()
    Q = torch.zeros((nz, nxy, nxy), dtype=dtype, device="cuda")
    for _ in range(steps_per_period):
        run.run(1)
        # Q += ((p^{n+1} - p^{n-1}) / 2dt)^2 dt in water (W:642)
        _ffi.check("acoustic_power_water", lib.bd_acoustic_power_water(
            C.byref(p), code, C.c_void_p(Q.data_ptr()), C.c_void_p(run.pressure(0).data_ptr()),
            C.c_void_p(run.pressure(2).data_ptr()), nxy, nz, 0, st))
    _ffi.check("acoustic_power_water", lib.bd_acoustic_power_water(
        C.byref(p), code, C.c_void_p(Q.data_ptr()), C.c_void_p(run.pressure(0).data_ptr()),
        C.c_void_p(run.pressure(2).data_ptr()), nxy, nz, 1, st))
    note("water", steps_per_period)
    # restart from the kept state (W:674-703)
    run.state.copy_(snapshot)
    del snapshot
    run.seek(n_settle)
    iz_t = int(lib.bd_westervelt_first_tissue_plane(C.byref(p), nz))
    H = torch.zeros((2 * n_harm, nz - iz_t, nxy, nxy), dtype=dtype, device="cuda")
    for i in range(n_time):
        run.run(1)
        _ffi.check("harmonics", lib.bd_harmonics_accumulate(code, C.c_void_p(H.data_ptr()),
                                                            C.c_void_p(run.pressure(0).data_ptr()), nxy, nz, iz_t,
                                                            i, n_time, periods, n_harm, 0, st))
    note("harmonics", n_time)
    _ffi.check("harmonics_q", lib.bd_harmonics_q(C.byref(p), code, C.c_void_p(Q.data_ptr()),
                                                C.c_void_p(H.data_ptr()), nxy, nz, iz_t, n_time, n_harm, 7.2, 0,
                                                st))
    if q3d_path is not None:
        _ffi.check("q3d_write", lib.bd_q3d_write(str(q3d_path).encode(), code, C.c_void_p(Q.data_ptr()), nxy, nz,
                                                2, st))
    torch.cuda.synchronize()
    return Q, p, {"settle": n_settle, "water": steps_per_period, "harmonics": n_time, "iz_tissue": iz_t}
