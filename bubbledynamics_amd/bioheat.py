"""Host-side mirror of the reference's bioheat driver (T:109-683) over the C ABI.

Parameters follow the reference's coefficient precompute (T:125-173,
T:212-224); the Pennes workload adds perfusion and runs the dwell spiral with
FP32 fields stored as T - T_art and a Kahan-compensated CEM43 dose.
"""
from __future__ import annotations

import ctypes as C
import math

import numpy as np
import torch

from . import _ffi
from ._ffi import BD_F32, BD_F64, BioheatParams, call


def reference_params() -> tuple[BioheatParams, int, int]:
    """Reference run constants (T:115-173, T:212-224).  Returns (params, Nxy, Nz)."""
    p = BioheatParams()
    rho_w, c_w, K_w = 1000.0, 4200.0, 0.6          # T:125-128
    rho_t, c_t, K_t = 1055.0, 3700.0, 0.51         # T:132-135
    k_w = K_w / (rho_w * c_w)
    k_t = K_t / (rho_t * c_t)
    p.water_tissue = 0.12                          # T:139
    xy_min, xy_max, p.dxy = -0.085, 0.085, 0.0002  # T:156-157
    nxy = 1 + round((xy_max - xy_min) / p.dxy)     # T:158
    z_max, p.dz = 0.2, 0.0001                      # T:161-162
    nz = 1 + round(z_max / p.dz)                   # T:163
    p.dt_T = 0.001                                 # T:166
    p.kdt_dxy = p.dt_T * k_t / (p.dxy * p.dxy)     # T:212-217
    p.kdt_dz = p.dt_T * k_t / (p.dz * p.dz)
    p.kdt_dxy_water = p.dt_T * k_w / (p.dxy * p.dxy)
    p.kdt_dz_water = p.dt_T * k_w / (p.dz * p.dz)
    p.dt_rho_ct = p.dt_T / (rho_t * c_t)
    p.dt_rho_ct_water = p.dt_T / (rho_w * c_w)
    return p, nxy, nz


def pennes_params(dx: float = 2e-4, dt: float = 0.05, substeps: int = 2, k: float = 0.52,
                  rho_c: float = 3.6e6, w_b: float = 0.5, c_b: float = 3617.0,
                  T_art_c: float = 37.0) -> BioheatParams:
    """Homogeneous liver Pennes workload: k=0.52 W/m/K, rho c=3.6e6 J/m^3/K,
    w_b=0.5 kg/m^3/s (c_b = 3617 J/kg/K, blood), T_art = 37 C, isotropic dx.
    dt = 0.05 s is split into `substeps` explicit substeps: the 7-point
    scheme needs k dt/(rho c dx^2) <= 1/6 (0.1806 at 0.05 s; 0.0903 at 0.025 s)."""
    p = BioheatParams()
    h = dt / substeps
    kappa = k / rho_c
    p.water_tissue = -1.0                          # every plane is tissue
    p.dxy = p.dz = dx
    p.dt_T = h
    p.kdt_dxy = p.kdt_dz = p.kdt_dxy_water = p.kdt_dz_water = kappa * h / (dx * dx)
    p.dt_rho_ct = p.dt_rho_ct_water = h / rho_c
    p.perf = w_b * c_b * h / rho_c
    p.T_art = T_art_c + 273.15
    return p


_SPIRAL = (2e-3, 3e-3, 0.05e-3, 0.15, 1e-3, 2e-4, -0.085 + 425 * 2e-4)   # speed, scale, speed_init, t_acc, dt, dxy, start


def spiral_reference_line_ends(lines: int = 13) -> list[int]:
    """Number of substeps done when each line of the reference's spiral ends
    (the loop of T:345-408: line k has increase*Nt_scale + 2*Nt_acc steps,
    increase = 1 + k//2, the first line one fewer)."""
    speed, scale, _, t_acc, dt = _SPIRAL[:5]
    nt_scale = 1 + int(scale / (speed * dt))                 # T:167
    nt_acc = 1 + int(t_acc / dt)                             # T:170
    ends, total = [], 0
    for k in range(lines):
        total += (1 + k // 2) * nt_scale + 2 * nt_acc - (1 if k == 0 else 0)
        ends.append(total)
    return ends


def spiral_reference(lines: int = 13) -> tuple[np.ndarray, np.ndarray]:
    """Shift sequence of the reference's continuous spiral (T:141-173, T:323-409)."""
    lib = _ffi.load()
    args = _SPIRAL + (lines,)
    n = lib.bd_spiral_reference(*args, None, None, 0)
    sx = (C.c_int * n)(); sy = (C.c_int * n)()
    lib.bd_spiral_reference(*args, sx, sy, n)
    return np.frombuffer(sx, dtype=np.int32).copy(), np.frombuffer(sy, dtype=np.int32).copy()


def spiral_dwell(n_side: int = 7, side: float = 8e-3, dxy: float = 2e-4) -> tuple[np.ndarray, np.ndarray]:
    """n_side^2 dwell points on a side x side square, as integer cell shifts."""
    sx = (C.c_int * (n_side * n_side))(); sy = (C.c_int * (n_side * n_side))()
    call("bd_spiral_dwell", n_side, side / (n_side - 1), dxy, sx, sy)
    return np.frombuffer(sx, dtype=np.int32).copy(), np.frombuffer(sy, dtype=np.int32).copy()


def _vp(t):
    return C.c_void_p(0 if t is None else t.data_ptr())


class BioheatRun:
    """Temperature (ping-pong), dose and compensation fields on the GPU.

    With offset=True the fields hold T - T_art (FP32 resolves small rises);
    temperature() always returns T in kelvin."""

    def __init__(self, params: BioheatParams, nxy: int, nz: int, Q: torch.Tensor, T0_kelvin: float,
                 dtype=torch.float32, kahan: bool = True, offset: bool = True):
        _ffi.load()
        self.p = params.copy()
        self.p.T_off = params.T_art if offset else 0.0
        self.nxy, self.nz, self.dtype = nxy, nz, dtype
        self.code = BD_F32 if dtype == torch.float32 else BD_F64
        shape = (nz, nxy, nxy)
        self.T = [torch.full(shape, T0_kelvin - self.p.T_off, dtype=dtype, device="cuda") for _ in range(2)]
        self.TD = torch.zeros(shape, dtype=dtype, device="cuda")
        self.TDc = torch.zeros(shape, dtype=dtype, device="cuda") if kahan else None
        self.Q = Q.to(dtype).contiguous()
        self.cur = 0
        self.steps = 0

    def step(self, sx: int, sy: int, n: int = 1) -> None:
        """n substeps with the source shifted by (sx, sy) cells (launched by the library)."""
        st = C.c_void_p(torch.cuda.current_stream().cuda_stream)
        which = C.c_int(self.cur)
        rc = _ffi.load().bd_bioheat_run(C.byref(self.p), self.code, _vp(self.T[0]), _vp(self.T[1]), _vp(self.Q),
                                        _vp(self.TD), _vp(self.TDc), self.nxy, self.nz, sx, sy, n, C.byref(which),
                                        st)
        _ffi.check("bd_bioheat_run", rc)
        self.cur = which.value
        self.steps += n

    def set_reference_initial(self, T0_tissue_c: float = 23.6, T0_water_c: float = 22.7) -> None:
        """The reference's initial state (T:149-150, T:233-257): T0 by region
        (water where iz*dz < water_tissue, T:79) in both buffers, and TD seeded
        with one step of dose at T0, 4^(T0 - 43) dt_T."""
        iz = 0
        while iz < self.nz and iz * self.p.dz < self.p.water_tissue:
            iz += 1
        for t in self.T:
            t[:iz] = T0_water_c + 273.15 - self.p.T_off
            t[iz:] = T0_tissue_c + 273.15 - self.p.T_off
        self.TD[:iz] = 4.0 ** (T0_water_c - 43.0) * self.p.dt_T
        self.TD[iz:] = 4.0 ** (T0_tissue_c - 43.0) * self.p.dt_T
        if self.TDc is not None:
            self.TDc.zero_()

    def temperature(self) -> torch.Tensor:
        return self.T[self.cur].double() + self.p.T_off

    def write_slice(self, path, plane: int, fixed: int, r0: int, r1: int, c0: int, c1: int,
                    field: str = "T") -> None:
        """A slice of T (kelvin) or of the dose in the reference's text format
        (T:412-447); plane BD_SLICE_XZ or BD_SLICE_XY."""
        if field == "T":
            t, off = self.T[self.cur], self.p.T_off
        else:
            t, off = (self.dose().to(self.dtype) if self.TDc is not None else self.TD), 0.0
        t = t.contiguous()
        _ffi.call("bd_write_slice", str(path).encode(), self.code, _vp(t), self.nxy, self.nz, plane, fixed,
                  r0, r1, c0, c1, off, C.c_void_p(torch.cuda.current_stream().cuda_stream))

    def dose(self) -> torch.Tensor:
        d = self.TD.double()
        return d - self.TDc.double() if self.TDc is not None else d
