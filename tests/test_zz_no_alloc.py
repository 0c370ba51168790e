"""The allocation-free threshold sweep (bd_gilmore_threshold_sweep_noalloc)
against the drop-in bd_gilmore_threshold_sweep, through the C ABI: bitwise
equal results on two calls in a row, the second with a smaller input (the
static counters keep the first call's contents and must be cleared)."""
import ctypes as C
import json
from pathlib import Path

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

from bubbledynamics_amd import _ffi, gilmore as gm  # noqa: E402
from bubbledynamics_amd._ffi import BD_ERR_INVALID, BD_GM_ARITH_FAST, BD_GM_ARITH_REFERENCE  # noqa: E402

FIX = json.loads(Path(__file__).with_name("golden").joinpath("sweep_cells.json").read_text())


def _sweep(entry, p, fr, n_r0):
    frd = torch.as_tensor(np.ascontiguousarray(fr, dtype=np.float64).ravel(), device="cuda")
    n_combo = frd.numel() // 3
    res = torch.full((n_combo * n_r0,), float("nan"), dtype=torch.float64, device="cuda")
    rc = getattr(_ffi.load(), entry)(C.byref(p), C.c_void_p(res.data_ptr()), C.c_void_p(frd.data_ptr()),
                                     n_combo, n_r0, 0, C.c_void_p(torch.cuda.current_stream().cuda_stream))
    _ffi.check(entry, rc)
    torch.cuda.synchronize()
    return res.cpu().numpy()


def _params(arith):
    s = FIX["settings"]["dp1980"]
    p = gm.reference_params()
    p.tableau, p.arithmetic = s["tableau"], arith
    p.minR0, p.stepR0 = FIX["min_r0"], FIX["step_r0"]
    return p


def test_noalloc_sweep_equals_drop_in_on_two_calls():
    fr = np.array(FIX["fr_combo"])
    n_r0 = FIX["n_r0"]
    # call 1: 24 cells, fast arithmetic
    p = _params(BD_GM_ARITH_FAST)
    print("noalloc call 1: 6 combos x 4 radii", flush=True)
    a = _sweep("bd_gilmore_threshold_sweep_noalloc", p, fr, n_r0)
    b = _sweep("bd_gilmore_threshold_sweep", p, fr, n_r0)
    np.testing.assert_array_equal(a, b)
    assert np.all(a > 0)
    # call 2: a smaller input (3 combos x 2 radii) in the reference's arithmetic
    p = _params(BD_GM_ARITH_REFERENCE)
    print("noalloc call 2: 3 combos x 2 radii", flush=True)
    a = _sweep("bd_gilmore_threshold_sweep_noalloc", p, fr[:3], 2)
    b = _sweep("bd_gilmore_threshold_sweep", p, fr[:3], 2)
    np.testing.assert_array_equal(a, b)
    expect = np.array(FIX["settings"]["dp1980"]["threshold_pa"]).reshape(6, 4)[:3, :2].ravel()
    np.testing.assert_array_equal(a, expect)


def test_noalloc_sweep_rejects_sizes_beyond_its_maximum():
    p = _params(BD_GM_ARITH_FAST)
    frd = torch.zeros(3 * 65536, dtype=torch.float64, device="cuda")
    res = torch.zeros(1, dtype=torch.float64, device="cuda")
    # 65,536 combos x 65,535 radii > 2^31 - 2^24 - 1 cells: rejected before any launch
    rc = _ffi.load().bd_gilmore_threshold_sweep_noalloc(C.byref(p), C.c_void_p(res.data_ptr()),
                                                       C.c_void_p(frd.data_ptr()), 65536, 65535, 0, None)
    assert rc == BD_ERR_INVALID
