"""Host-side mirror of the reference's Gilmore driver (G:433-631) and the
batched bubble workloads, over the C ABI.

Device memory comes from torch tensors; the integration itself is the HIP
library's persistent work-queue kernel.
"""
from __future__ import annotations

import ctypes as C
import math

import numpy as np
import torch

from . import _ffi
from ._ffi import (BD_GM_ARITH_FAST, BD_GM_ARITH_REFERENCE, BD_GM_DP45, BD_GM_FP64, BD_GM_RK4,  # noqa: F401
                   BD_GM_SCHED_INPUT, BD_GM_SCHED_SLICED, BD_GM_SPLIT, BD_GM_SWEEP_CAPPED, BD_GM_TAB_DP1980,
                   BD_GM_TAB_REFERENCE, GilmoreBatch, GilmoreParams, call)


def reference_params() -> GilmoreParams:
    """Liver constants of G:440-460, reference tableau, controller and arithmetic."""
    p = GilmoreParams()
    _ffi.load().bd_gilmore_reference_params(C.byref(p))
    return p


def workload_params(published_tableau: bool = True, arithmetic: int = BD_GM_ARITH_FAST) -> GilmoreParams:
    """Medium of the batched workloads: rho=1060, c=1590, sigma=0.056,
    mu=0.005, B=3.3e8, n=7, kappa=1.4; P0, G and lambda from the reference
    (G:441, G:444, G:446).  The fast arithmetic by default; pass
    BD_GM_ARITH_REFERENCE for the reference's expression order."""
    p = reference_params()
    p.arithmetic = arithmetic
    p.rho = 1060.0
    p.C_inf = 1590.0
    p.S2 = 2 * 0.056
    p.mu4 = 4 * 0.005
    p.B = 3.3e8
    p.n = 7.0
    p.gamma1 = 1.4
    p.gamma3 = 4.2
    p.tableau = BD_GM_TAB_DP1980 if published_tableau else BD_GM_TAB_REFERENCE
    return p


def freq_table(min_fr=0.02e6, max_fr=5e6, n_fr=55) -> np.ndarray:
    """Frequency triplets of G:465-469, G:498-520 (27,775 combos for 55 tones)."""
    step_log = (math.log10(max_fr) - math.log10(min_fr)) / (n_fr - 1.0)
    f = [10.0 ** (math.log10(min_fr) + i * step_log) for i in range(n_fr)]
    out = []
    for i in range(n_fr):
        first = True
        for ii in range(i, n_fr):
            iii = ii if first else ii + 1
            for k in range(iii, n_fr):
                out.append((f[i], f[ii], f[k]))
            first = False
    return np.array(out)


def _vp(t):
    return C.c_void_p(0 if t is None else t.data_ptr())


def _stream():
    return C.c_void_p(torch.cuda.current_stream().cuda_stream)


def threshold_sweep(params: GilmoreParams, fr_combo: np.ndarray, n_r0: int) -> np.ndarray:
    """Drop-in run of DPRK_solver_threshold_fix_tissue (G:323-429)."""
    fr = torch.as_tensor(np.ascontiguousarray(fr_combo, dtype=np.float64).ravel(), device="cuda")
    n_combo = fr.numel() // 3
    res = torch.empty(n_combo * n_r0, dtype=torch.float64, device="cuda")
    call("bd_gilmore_threshold_sweep", C.byref(params), _vp(res), _vp(fr), n_combo, n_r0, 0, _stream())
    torch.cuda.synchronize()
    return res.cpu().numpy()


def threshold_sweep_range(params: GilmoreParams, fr_combo: np.ndarray, n_r0: int, cell0: int = 0,
                          cell_end: int | None = None, deadline_s: float | None = None, poll_s: float = 20.0,
                          log=None, results: np.ndarray | None = None) -> tuple[np.ndarray, int, float]:
    """The reference's sweep (G:498-573) over the cells [cell0, cell_end) of
    idx = combo * n_r0 + i_r0 in one persistent launch.  Every `poll_s`
    seconds the cells claimed and completed, as the kernel reports them, go to
    `log`; after `deadline_s` the kernel is asked to stop claiming, and the
    cells it had claimed are finished.  Returns (results with the done cells
    filled in, end of the done prefix, seconds)."""
    import time
    fr_np = np.ascontiguousarray(fr_combo, dtype=np.float64).reshape(-1)
    n_combo = fr_np.size // 3
    total = n_combo * n_r0
    cell_end = total if cell_end is None else cell_end
    out = np.full(total, np.nan) if results is None else results
    lib = _ffi.load()
    flags = lib.bd_gilmore_sweep_flags()
    if not flags:
        raise RuntimeError("bd_gilmore_sweep_flags: host-mapped allocation failed")
    fr = torch.as_tensor(fr_np, device="cuda")
    res = torch.full((total,), float("nan"), dtype=torch.float64, device="cuda")
    flags[2] = 0
    t0 = time.perf_counter()
    call("bd_gilmore_threshold_sweep_range", C.byref(params), _vp(res), _vp(fr), n_combo, n_r0, cell0, cell_end,
         _stream())
    stopped = False
    last = 0.0
    while flags[3] == 0:            # set by the library's last kernel of the launch
        time.sleep(0.25)
        el = time.perf_counter() - t0
        if log and el - last >= poll_s and flags[3] == 0:
            last = el
            log(f"sweep: {flags[0]} cells claimed, {flags[1]} completed of {cell_end - cell0}, {el:.0f} s")
        if deadline_s is not None and el > deadline_s and not stopped:
            flags[2] = 1
            stopped = True
            if log:
                log("sweep: stop requested, finishing the claimed cells")
    torch.cuda.synchronize()
    done = int(flags[1])
    flags[2] = 0
    out[cell0:cell0 + done] = res[cell0:cell0 + done].cpu().numpy()
    return out, cell0 + done, time.perf_counter() - t0


class ThresholdSweep:
    """Resumable run of the drop-in sweep over cells [cell0, cell_end)
    (bd_gilmore_sweep_*): each advance() is one launch in which every lane
    stops after `budget` DP5(4) attempts, saving unfinished cells for the
    next advance.  results (device, float64, NaN until a cell is done) is
    indexed like the drop-in's output."""

    def __init__(self, params: GilmoreParams, fr_combo: np.ndarray, n_r0: int, cell0: int = 0,
                 cell_end: int | None = None):
        fr_np = np.ascontiguousarray(fr_combo, dtype=np.float64).reshape(-1)
        self.n_combo = fr_np.size // 3
        self.total = self.n_combo * n_r0
        self.cell0 = cell0
        self.cell_end = self.total if cell_end is None else cell_end
        self.fr = torch.as_tensor(fr_np, device="cuda")
        self.results = torch.full((self.total,), float("nan"), dtype=torch.float64, device="cuda")
        self._h = C.c_void_p()
        call("bd_gilmore_sweep_create", C.byref(params), _vp(self.fr), self.n_combo, n_r0, cell0, self.cell_end,
             C.byref(self._h))
        self.advances = 0

    def advance(self, budget: int) -> None:
        call("bd_gilmore_sweep_advance", self._h, _vp(self.results), int(budget), _stream())
        self.advances += 1

    def progress(self) -> dict:
        out = (C.c_longlong * 5)()
        call("bd_gilmore_sweep_progress", self._h, out)
        return {"claimed": out[0], "completed": out[1], "suspended": out[2], "all_done": bool(out[3]),
                "launch_ended": bool(out[4])}

    def stop_word(self):
        return _ffi.load().bd_gilmore_sweep_stop_word(self._h)

    def run(self, budget: int, max_advances: int = 1_000_000, log=None) -> np.ndarray:
        """Advances until every cell of the range is done; returns the results."""
        while self.advances < max_advances:
            self.advance(budget)
            torch.cuda.synchronize()
            pr = self.progress()
            if log:
                log(f"sweep advance {self.advances}: {pr}")
            if pr["all_done"]:
                break
        return self.results.cpu().numpy()

    def close(self) -> None:
        if self._h:
            _ffi.load().bd_gilmore_sweep_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def write_sweep_table(path: str, fr_combo: np.ndarray, results: np.ndarray, n_r0: int, min_r0: float,
                      step_r0: float, seconds: float) -> None:
    """Threshold table in the reference's format (G:580-606)."""
    fr = np.ascontiguousarray(fr_combo, dtype=np.float64).reshape(-1)
    res = np.ascontiguousarray(results, dtype=np.float64).reshape(-1)
    dp = C.POINTER(C.c_double)
    call("bd_gilmore_sweep_write", str(path).encode(), fr.ctypes.data_as(dp), len(fr) // 3,
         res.ctypes.data_as(dp), n_r0, min_r0, step_r0, seconds)


class BubbleBatch:
    """A batch of bubbles with per-bubble R0 and three-tone drive amplitudes and
    phases on shared frequencies; outputs live on the GPU."""

    def __init__(self, R0, amp, phase, freq):
        self.n = len(R0)
        self.R0 = torch.as_tensor(np.ascontiguousarray(R0, dtype=np.float64), device="cuda")
        self.amp = torch.as_tensor(np.ascontiguousarray(amp, dtype=np.float64).reshape(-1), device="cuda")
        self.phase = torch.as_tensor(np.ascontiguousarray(phase, dtype=np.float64).reshape(-1), device="cuda")
        self.freq = tuple(float(f) for f in freq)
        self.workspace = torch.zeros(int(_ffi.load().bd_gilmore_workspace_bytes()), dtype=torch.uint8,
                                     device="cuda")

    def _batch(self, method, t_end, dt, precision, rtol, atol, split_ratio, out_every, n_out, max_steps,
               block, chunk, waves, schedule=0, slice_=0):
        n = self.n
        dev = "cuda"
        out = {
            "Rmax": torch.empty(n, dtype=torch.float64, device=dev),
            "t_collapse": torch.empty(n, dtype=torch.float64, device=dev),
            "n_acc": torch.empty(n, dtype=torch.int64, device=dev),
            "n_rej": torch.empty(n, dtype=torch.int64, device=dev),
            "R_out": torch.empty((n, n_out), dtype=torch.float64, device=dev) if n_out else None,
        }
        b = GilmoreBatch()
        b.method, b.precision, b.n = method, precision, n
        b.out_every, b.n_out = out_every, n_out
        b.block, b.chunk, b.waves = block, chunk, waves
        b.schedule = schedule
        b.slice = slice_
        for k in range(3):
            b.freq[k] = self.freq[k]
        b.t_end, b.dt, b.rtol, b.atol, b.split_ratio = t_end, dt, rtol, atol, split_ratio
        b.max_steps = max_steps
        b.R0, b.amp, b.phase = self.R0.data_ptr(), self.amp.data_ptr(), self.phase.data_ptr()
        b.R_out = out["R_out"].data_ptr() if n_out else None
        b.Rmax, b.t_collapse = out["Rmax"].data_ptr(), out["t_collapse"].data_ptr()
        b.n_acc, b.n_rej = out["n_acc"].data_ptr(), out["n_rej"].data_ptr()
        return b, out

    def run(self, params: GilmoreParams, method: int, t_end: float, dt: float, *, precision=BD_GM_FP64,
            rtol=1e-8, atol=1e-12, split_ratio=0.2, out_every=0, n_out=0, max_steps=0, block=0,
            chunk=0, waves=0, schedule=BD_GM_SCHED_INPUT, slice_attempts=0) -> dict:
        """Integrates every bubble; block / chunk / waves left at 0 take the
        autotuned shape for this method and precision (or the default);
        schedule=BD_GM_SCHED_SLICED starts the predicted-longest bubbles first."""
        b, out = self._batch(method, t_end, dt, precision, rtol, atol, split_ratio, out_every, n_out,
                             max_steps, block, chunk, waves, schedule, slice_attempts)
        call("bd_gilmore_integrate", C.byref(params), C.byref(b), _vp(self.workspace), _stream())
        return out

    def autotune(self, params: GilmoreParams, method: int, t_end: float, dt: float, *, precision=BD_GM_FP64,
                 rtol=1e-8, atol=1e-12, split_ratio=0.2, sample=0, probe_steps=0) -> dict:
        """Times the candidate launch shapes on this batch's first `sample`
        bubbles and records the fastest for later runs; returns it."""
        b, _ = self._batch(method, t_end, dt, precision, rtol, atol, split_ratio, 0, 0, 0, 0, 0, 0)
        blk, ch, wv, rate = C.c_int32(), C.c_int32(), C.c_int32(), C.c_double()
        call("bd_gilmore_autotune", C.byref(params), C.byref(b), int(sample), int(probe_steps), _stream(),
             C.byref(blk), C.byref(ch), C.byref(wv), C.byref(rate))
        return {"block": blk.value, "chunk": ch.value, "waves": wv.value, "probe_rate": rate.value}


def seeded_uniforms(seed: int, n: int) -> np.ndarray:
    """SplitMix64 -> 53-bit uniforms in [0, 1), the seeded inputs of the workloads."""
    out = np.empty(n)
    s = seed & 0xFFFFFFFFFFFFFFFF
    M = 0xFFFFFFFFFFFFFFFF
    for i in range(n):
        s = (s + 0x9E3779B97F4A7C15) & M
        z = s
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & M
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & M
        z = z ^ (z >> 31)
        out[i] = (z >> 11) * (1.0 / 9007199254740992.0)
    return out


def config0_inputs(n: int = 1024, seed: int = 0):
    """1,024 bubbles: R0 ~ U(1,5) um, A_k ~ U(0.5,3) MPa, phi_k ~ U(0,2pi); f = (1,2,3) MHz.
    Draw order per bubble: R0, A1..A3, phi1..phi3."""
    u = seeded_uniforms(seed, 7 * n).reshape(n, 7)
    R0 = 1e-6 + 4e-6 * u[:, 0]
    amp = 0.5e6 + 2.5e6 * u[:, 1:4]
    ph = 2 * math.pi * u[:, 4:7]
    return R0, amp, ph, (1e6, 2e6, 3e6)


def config3_inputs(n: int = 1_000_000, seed: int = 3):
    """1e6 bubbles: R0 log-uniform 0.5-20 um, A_k ~ U(0.5,8) MPa, phi_k ~ U(0,2pi)."""
    u = _splitmix_vec(seed, 7 * n).reshape(n, 7)
    R0 = np.exp(math.log(0.5e-6) + u[:, 0] * math.log(40.0))
    amp = 0.5e6 + 7.5e6 * u[:, 1:4]
    ph = 2 * math.pi * u[:, 4:7]
    return R0, amp, ph, (1e6, 2e6, 3e6)


def config3_subset(n: int = 1_000_000, k: int = 4096) -> np.ndarray:
    """Indices (increasing) of k of the n config-3 bubbles, stratified in R0:
    the bubbles at ranks (j + 1/2) n / k of the R0 order."""
    R0, _, _, _ = config3_inputs(n, 3)
    order = np.argsort(R0, kind="stable")
    return np.sort(order[((np.arange(k) + 0.5) * n / k).astype(np.int64)])


def _splitmix_vec(seed: int, n: int) -> np.ndarray:
    """Vectorised SplitMix64 (identical stream to seeded_uniforms)."""
    with np.errstate(over="ignore"):
        idx = np.arange(1, n + 1, dtype=np.uint64)
        z = np.uint64(seed) + idx * np.uint64(0x9E3779B97F4A7C15)
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        z = z ^ (z >> np.uint64(31))
    return (z >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)
