"""Host-side driver of the adjoint phased-array gradient over the C ABI.

J = -mean_{t, target}(p^2) + 0.1 mean_{t, outside}(p^2) on the interior cells
(the "time-integrated p^2 over the target sphere" taken as a time-and-volume
mean so that both terms are in Pa^2), gradient with respect to the element
amplitudes and phases of the array on plane iz=2.
"""
from __future__ import annotations

import ctypes as C

import numpy as np
import torch

from . import _ffi
from ._ffi import BD_F32, BD_F64, AdjointObjective, WesterveltParams


def objective(n: int, nz: int, n_steps: int, center, radius: float, outside_weight: float = 0.1,
              full_domain: bool = True) -> AdjointObjective:
    lo = 3 if full_domain else 0
    idx = np.arange(n)
    x = idx[(idx >= lo) & (idx <= n - 4)]
    z = np.arange(3, nz - 3)
    d2 = ((z[:, None, None] - center[2]) ** 2 + (x[None, :, None] - center[1]) ** 2 +
          (x[None, None, :] - center[0]) ** 2)
    n_t = int(np.count_nonzero(d2 <= radius * radius))
    n_o = d2.size - n_t
    o = AdjointObjective()
    o.cx, o.cy, o.cz = center
    o.radius = radius
    o.w_target = 1.0 / (n_t * n_steps)
    o.w_outside = outside_weight / (n_o * n_steps)
    return o


def adjoint_gradient(p: WesterveltParams, n: int, n_steps: int, amp, phase, obj: AdjointObjective,
                     checkpoint_every: int = 50, dtype=torch.float64, gradient: bool = True):
    """Returns (J, dJ/dA, dJ/dphi, timings_ms[forward, recompute outside the sweep, adjoint sweep])."""
    amp_d = torch.as_tensor(np.asarray(amp, dtype=np.float64), device="cuda").contiguous()
    ph_d = torch.as_tensor(np.asarray(phase, dtype=np.float64), device="cuda").contiguous()
    ne = amp_d.numel()
    gA = torch.zeros(ne, dtype=torch.float64, device="cuda") if gradient else None
    gP = torch.zeros(ne, dtype=torch.float64, device="cuda") if gradient else None
    J = C.c_double(0.0)
    tm = (C.c_double * 3)()
    code = BD_F32 if dtype == torch.float32 else BD_F64
    st = C.c_void_p(torch.cuda.current_stream().cuda_stream)
    _ffi.call("bd_westervelt_adjoint", C.byref(p), code, n, n, n_steps, checkpoint_every,
              C.c_void_p(amp_d.data_ptr()), C.c_void_p(ph_d.data_ptr()), C.byref(obj), C.byref(J),
              C.c_void_p(gA.data_ptr()) if gradient else None, C.c_void_p(gP.data_ptr()) if gradient else None,
              tm, st)
    return (J.value, gA.cpu().numpy() if gradient else None, gP.cpu().numpy() if gradient else None,
            list(tm))
