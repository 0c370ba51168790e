"""Host-side mirror of the reference's Westervelt driver (W:556-585).

``WesterveltRun`` owns the device buffers (torch tensors, used only as device
memory) and drives the C-ABI plan: a 6-level pressure ring and 3-level
relaxation rings replace the reference's nine device-to-device copies per
step, and repeated steps replay a captured 6-step hipGraph.
"""
from __future__ import annotations

import ctypes as C
import math

import numpy as np
import torch

from . import _ffi
from ._ffi import BD_F32, BD_F64, WesterveltParams, call


def _dtype_code(dtype: torch.dtype) -> int:
    if dtype == torch.float32:
        return BD_F32
    if dtype == torch.float64:
        return BD_F64
    raise ValueError(f"unsupported dtype {dtype}")


def _ptr(t: torch.Tensor | None) -> C.c_void_p:
    return C.c_void_p(0 if t is None else t.data_ptr())


def _stream() -> C.c_void_p:
    return C.c_void_p(torch.cuda.current_stream().cuda_stream)


class WesterveltRun:
    """One Westervelt run on the current GPU.

    state: (12, Nz, Nxy, Nxy) tensor -- P ring slots 0..5 (p^m in slot m mod 6),
    R1 ring slots 6..8 and R2 ring slots 9..11 (R^m in slot m mod 3).
    """

    def __init__(self, params: WesterveltParams, nxy: int, nz: int, *, dtype=torch.float32,
                 max_steps: int = 1, focus: tuple[int, int, int] | None = None,
                 pmax: bool = True, energy: bool = True, elem_amp=None, elem_phase=None,
                 device: str | torch.device = "cuda", prepare_graphs: bool = True):
        _ffi.load()
        self.params = params.copy()
        self.nxy, self.nz, self.dtype = nxy, nz, dtype
        self.max_steps = max_steps
        self.device = torch.device(device)
        n = int(_ffi.load().bd_westervelt_state_elems(nxy, nz))
        assert n == 12 * nxy * nxy * nz
        self.state = torch.zeros((12, nz, nxy, nxy), dtype=dtype, device=self.device)
        self.pmax = torch.zeros((nz, nxy, nxy), dtype=dtype, device=self.device) if pmax else None
        self.energy = torch.zeros((max_steps, 2), dtype=torch.float64, device=self.device) if energy else None
        self.focus = torch.zeros(max_steps, dtype=torch.float64, device=self.device) if focus else None
        self.elem_amp = self.elem_phase = None
        if elem_amp is not None:
            self.elem_amp = torch.as_tensor(elem_amp, dtype=torch.float64, device=self.device).contiguous()
            self.elem_phase = torch.as_tensor(elem_phase, dtype=torch.float64, device=self.device).contiguous()
        fx, fy, fz = focus if focus else (-1, -1, -1)
        plan = C.c_void_p()
        call("bd_westervelt_plan_create", C.byref(self.params), _dtype_code(dtype), nxy, nz,
             max_steps, fx, fy, fz, C.byref(plan))
        self._plan = plan
        call("bd_westervelt_plan_bind", plan, _ptr(self.state), _ptr(self.pmax),
             _ptr(self.energy), _ptr(self.focus), _ptr(self.elem_amp), _ptr(self.elem_phase))
        if prepare_graphs:
            call("bd_westervelt_plan_prepare", plan)

    # -- driving -----------------------------------------------------------
    def run(self, n_steps: int, use_graph: bool = True) -> None:
        call("bd_westervelt_plan_run", self._plan, n_steps, 1 if use_graph else 0, _stream())

    def finalize(self) -> None:
        call("bd_westervelt_plan_finalize", self._plan, _stream())

    @property
    def steps(self) -> int:
        return int(_ffi.load().bd_westervelt_plan_steps(self._plan))

    @property
    def shape(self) -> dict:
        """Launch shape of the step kernel (tile, rows per thread, prefetch,
        non-temporal history loads, workgroups per step)."""
        v = [C.c_int() for _ in range(6)]
        call("bd_westervelt_plan_shape", self._plan, *[C.byref(x) for x in v])
        keys = ("tile_x", "tile_y", "rows", "prefetch", "nontemporal", "workgroups")
        return {k: x.value for k, x in zip(keys, v)}

    def seek(self, step: int) -> None:
        """The state block holds the levels of `step` (e.g. a restored copy)."""
        call("bd_westervelt_plan_seek", self._plan, int(step))

    def pressure(self, level: int = 0) -> torch.Tensor:
        """p^{n-level} after n steps (a view into the state block)."""
        return self.state[_ffi.load().bd_westervelt_plan_slot(self._plan, 0, level)]

    def relax(self, which: int, level: int = 0) -> torch.Tensor:
        return self.state[_ffi.load().bd_westervelt_plan_slot(self._plan, which, level)]

    def close(self) -> None:
        if getattr(self, "_plan", None):
            _ffi.load().bd_westervelt_plan_destroy(self._plan)
            self._plan = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def focal_harmonics(series: np.ndarray, dt: float, fr: float, n_harm: int = 3,
                    t0: float = 0.0) -> np.ndarray:
    """Amplitudes of harmonics n*fr of a focus record by direct DFT at the
    exact frequencies: |2/N sum_j p_j exp(-i 2 pi n fr t_j)| over the record.
    Evaluated in FP64 with compensated (Kahan) accumulation."""
    t = t0 + dt * np.arange(1, len(series) + 1)
    out = np.empty(n_harm)
    for n in range(1, n_harm + 1):
        ph = 2.0 * math.pi * n * fr * t
        re = _kahan_sum(series * np.cos(ph))
        im = _kahan_sum(-series * np.sin(ph))
        out[n - 1] = 2.0 * math.hypot(re, im) / len(series)
    return out


def _kahan_sum(x: np.ndarray) -> float:
    s = 0.0
    c = 0.0
    for v in x.tolist():
        y = v - c
        t = s + y
        c = (t - s) - y
        s = t
    return s
