"""ctypes binding of the C ABI in include/bubbledynamics.h.

The shared library is built in-tree by ``bubbledynamics_amd.build``.  Loading
fails loudly when it is missing: there is no fallback implementation.
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

LIB_PATH = Path(__file__).resolve().parent / "libbubbledynamics.so"

BD_OK = 0
BD_ERR_INVALID = 1
BD_ERR_UNSTABLE = 2
BD_ERR_ALLOC = 3
BD_ERR_HIP = 4
BD_ERR_NONFINITE = 5
BD_ERR_UNSUPPORTED = 6
BD_ERR_IO = 7
BD_SLICE_XZ = 0
BD_SLICE_XY = 1
BD_F32 = 0
BD_F64 = 1
BD_BC_MUR = 0
BD_BC_ZERO = 1
BD_SRC_CAP = 0
BD_SRC_ARRAY = 1

_ERRS = {1: "invalid argument", 2: "unstable parameters", 3: "allocation failed",
         4: "HIP runtime error", 5: "non-finite result", 6: "unsupported option"}


class BDError(RuntimeError):
    def __init__(self, fn: str, code: int):
        super().__init__(f"{fn} returned {code} ({_ERRS.get(code, 'unknown')})")
        self.code = code


class WesterveltParams(C.Structure):
    """Mirror of bd_westervelt_params (W:16-52 plus extensions)."""
    _fields_ = [(n, C.c_double) for n in (
        "rho", "c_inf", "delta", "beta", "abs_water",
        "rho_tissue", "c_inf_tissue", "beta_tissue", "abs_tissue",
        "delta_tissue", "c1", "c2", "tau1", "tau2",
        "fr", "d", "alpha", "alpha_hole", "p0",
        "water_tissue",
        "xy_min", "dxy", "z_min", "dz", "dt")] + [
        ("full_domain", C.c_int32), ("boundary", C.c_int32), ("source", C.c_int32),
        ("sponge_cells", C.c_int32), ("sponge_alpha", C.c_double),
        ("burst_cycles", C.c_double), ("burst_t0", C.c_double),
        ("arr_nx", C.c_int32), ("arr_ny", C.c_int32),
        ("arr_x0", C.c_double), ("arr_y0", C.c_double), ("arr_pitch", C.c_double),
        ("reserved", C.c_double * 8),
    ]

    def copy(self) -> "WesterveltParams":
        o = WesterveltParams()
        C.pointer(o)[0] = self
        return o

    def as_dict(self) -> dict:
        return {n: getattr(self, n) for n, _ in self._fields_ if n != "reserved"}


class GilmoreParams(C.Structure):
    """Mirror of bd_gilmore_params (G:13-33 plus integrator controls)."""
    _fields_ = [(n, C.c_double) for n in (
        "rho", "P0", "S2", "C_inf", "lambda_", "n", "B", "gamma1", "gamma3", "mu4", "G4_3")] + [
        ("triple_fr", C.c_int32), ("tableau", C.c_int32),
        ("minA", C.c_double), ("stepA", C.c_double), ("minR0", C.c_double), ("stepR0", C.c_double),
        ("abs_eps", C.c_double), ("rel_eps", C.c_double), ("h0", C.c_double),
        ("max_solve_steps", C.c_double), ("arithmetic", C.c_int32), ("pad0", C.c_int32),
        ("reserved", C.c_double * 6),
    ]

    def copy(self) -> "GilmoreParams":
        o = GilmoreParams()
        C.pointer(o)[0] = self
        return o


class GilmoreBatch(C.Structure):
    """Mirror of bd_gilmore_batch."""
    _fields_ = [("method", C.c_int32), ("precision", C.c_int32), ("n", C.c_int32),
                ("out_every", C.c_int32), ("n_out", C.c_int32), ("block", C.c_int32),
                ("chunk", C.c_int32), ("waves", C.c_int32),
                ("freq", C.c_double * 3), ("t_end", C.c_double), ("dt", C.c_double),
                ("rtol", C.c_double), ("atol", C.c_double), ("split_ratio", C.c_double),
                ("max_steps", C.c_longlong),
                ("R0", C.c_void_p), ("amp", C.c_void_p), ("phase", C.c_void_p),
                ("R_out", C.c_void_p), ("Rmax", C.c_void_p), ("t_collapse", C.c_void_p),
                ("n_acc", C.c_void_p), ("n_rej", C.c_void_p), ("schedule", C.c_int32), ("slice", C.c_int32),
                ("reserved", C.c_double * 3)]


class BioheatParams(C.Structure):
    """Mirror of bd_bioheat_params (T:30-45 thermal block plus Pennes extension)."""
    _fields_ = [(n, C.c_double) for n in (
        "water_tissue", "dxy", "dz", "dt_T", "kdt_dxy", "kdt_dz", "kdt_dxy_water", "kdt_dz_water",
        "dt_rho_ct", "dt_rho_ct_water", "perf", "T_art", "T_off")] + [("reserved", C.c_double * 8)]

    def copy(self) -> "BioheatParams":
        o = BioheatParams()
        C.pointer(o)[0] = self
        return o


class AdjointObjective(C.Structure):
    _fields_ = [("cx", C.c_int32), ("cy", C.c_int32), ("cz", C.c_int32), ("pad0", C.c_int32),
                ("radius", C.c_double), ("w_target", C.c_double), ("w_outside", C.c_double),
                ("reserved", C.c_double * 4)]


BD_GM_TAB_REFERENCE = 0
BD_GM_TAB_DP1980 = 1
BD_GM_RK4 = 0
BD_GM_DP45 = 1
BD_GM_FP64 = 0
BD_GM_SPLIT = 1
BD_GM_SCHED_INPUT = 0
BD_GM_SCHED_SLICED = 1
BD_GM_ARITH_REFERENCE = 0
BD_GM_ARITH_FAST = 1
BD_GM_SWEEP_CAPPED = -2.0

_VP = C.c_void_p
_I = C.c_int
_D = C.c_double
_IP = C.POINTER(C.c_int)
_PP = C.POINTER(WesterveltParams)

# name -> (restype, argtypes); entries added by the other modules' binding tables
SIGNATURES: dict[str, tuple] = {
    "bd_version": (C.c_char_p, []),
    "bd_device_count": (_I, []),
    "bd_westervelt_space6": (_I, [_PP, _I] + [_VP] * 12 + [_I, _I, _VP]),
    "bd_boundary_condition_source": (_I, [_PP, _I, _VP, _D, _I, _I, _VP, _VP, _VP]),
    "bd_boundary_condition_space6": (_I, [_PP, _I, _VP, _VP, _D, _I, _I, _VP]),
    "bd_acoustic_power_water": (_I, [_PP, _I, _VP, _VP, _VP, _I, _I, _I, _VP]),
    "bd_harmonics_accumulate": (_I, [_I, _VP, _VP, _I, _I, _I, _I, _I, _I, _I, _I, _VP]),
    "bd_harmonics_q": (_I, [_PP, _I, _VP, _VP, _I, _I, _I, _I, _I, _D, _I, _VP]),
    "bd_westervelt_state_elems": (C.c_size_t, [_I, _I]),
    "bd_westervelt_plan_create": (_I, [_PP, _I, _I, _I, _I, _I, _I, _I, C.POINTER(_VP)]),
    "bd_westervelt_plan_bind": (_I, [_VP, _VP, _VP, _VP, _VP, _VP, _VP]),
    "bd_westervelt_plan_run": (_I, [_VP, _I, _I, _VP]),
    "bd_westervelt_plan_prepare": (_I, [_VP]),
    "bd_westervelt_plan_finalize": (_I, [_VP, _VP]),
    "bd_westervelt_plan_steps": (_I, [_VP]),
    "bd_westervelt_plan_seek": (_I, [_VP, _I]),
    "bd_westervelt_first_tissue_plane": (_I, [C.POINTER(WesterveltParams), _I]),
    "bd_westervelt_plan_slot": (_I, [_VP, _I, _I]),
    "bd_westervelt_plan_destroy": (_I, [_VP]),
    "bd_westervelt_plan_rerun_step": (_I, [_VP, _VP]),
    "bd_westervelt_plan_interior_cells": (C.c_longlong, [_VP]),
    "bd_westervelt_plan_shape": (_I, [_VP, _IP, _IP, _IP, _IP, _IP, _IP]),
    "bd_westervelt_adjoint": (_I, [_PP, _I, _I, _I, _I, _I, _VP, _VP, C.POINTER(AdjointObjective),
                                   C.POINTER(_D), _VP, _VP, C.POINTER(_D), _VP]),
    "bd_temperature_square": (_I, [C.POINTER(BioheatParams), _I, _VP, _VP, _VP, _VP, _I, _I, _I, _I, _I, _VP]),
    "bd_bioheat_step": (_I, [C.POINTER(BioheatParams), _I, _VP, _VP, _VP, _VP, _VP, _I, _I, _I, _I, _VP]),
    "bd_write_slice": (_I, [C.c_char_p, _I, _VP, _I, _I, _I, _I, _I, _I, _I, _I, C.c_double, _VP]),
    "bd_q3d_write": (_I, [C.c_char_p, _I, _VP, _I, _I, _I, _VP]),
    "bd_q3d_read_expand": (_I, [C.c_char_p, _I, _VP, _I, _I, _VP]),
    "bd_mirror_expand": (_I, [_I, _VP, _VP, _I, _I, _VP]),
    "bd_bioheat_run": (_I, [C.POINTER(BioheatParams), _I, _VP, _VP, _VP, _VP, _VP, _I, _I, _I, _I, _I,
                            C.POINTER(C.c_int), _VP]),
    "bd_spiral_reference": (_I, [_D, _D, _D, _D, _D, _D, _D, _I, C.POINTER(_I), C.POINTER(_I), _I]),
    "bd_spiral_dwell": (_I, [_I, _D, _D, C.POINTER(_I), C.POINTER(_I)]),
    "bd_gilmore_reference_params": (None, [C.POINTER(GilmoreParams)]),
    "bd_gilmore_threshold_sweep": (_I, [C.POINTER(GilmoreParams), _VP, _VP, _I, _I, _I, _VP]),
    "bd_gilmore_threshold_sweep_noalloc": (_I, [C.POINTER(GilmoreParams), _VP, _VP, _I, _I, _I, _VP]),
    "bd_gilmore_threshold_sweep_range": (_I, [C.POINTER(GilmoreParams), _VP, _VP, _I, _I, _I, _I, _VP]),
    "bd_gilmore_sweep_flags": (_IP, []),
    "bd_gilmore_sweep_create": (_I, [C.POINTER(GilmoreParams), _VP, _I, _I, _I, _I, C.POINTER(_VP)]),
    "bd_gilmore_sweep_advance": (_I, [_VP, _VP, C.c_longlong, _VP]),
    "bd_gilmore_sweep_progress": (_I, [_VP, C.POINTER(C.c_longlong)]),
    "bd_gilmore_sweep_stop_word": (_IP, [_VP]),
    "bd_gilmore_sweep_destroy": (_I, [_VP]),
    "bd_gilmore_sweep_write": (_I, [C.c_char_p, C.POINTER(C.c_double), _I, C.POINTER(C.c_double), _I,
                                    C.c_double, C.c_double, C.c_double]),
    "bd_gilmore_workspace_bytes": (C.c_size_t, []),
    "bd_gilmore_integrate": (_I, [C.POINTER(GilmoreParams), C.POINTER(GilmoreBatch), _VP, _VP]),
    "bd_gilmore_autotune": (_I, [C.POINTER(GilmoreParams), C.POINTER(GilmoreBatch), C.c_int, C.c_longlong, _VP,
                                 C.POINTER(C.c_int32), C.POINTER(C.c_int32), C.POINTER(C.c_int32),
                                 C.POINTER(C.c_double)]),
}

_lib = None


def load() -> C.CDLL:
    """Load the HIP library; raise if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not LIB_PATH.exists():
        raise RuntimeError(
            f"{LIB_PATH} is missing: build it with `python -m bubbledynamics_amd.build` "
            "(there is no CPU fallback)")
    lib = C.CDLL(str(LIB_PATH))
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def check(fn: str, rc: int) -> None:
    if rc != BD_OK:
        raise BDError(fn, rc)


def call(fn: str, *args) -> int:
    rc = getattr(load(), fn)(*args)
    check(fn, rc)
    return rc
