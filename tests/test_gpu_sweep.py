"""GPU drop-in threshold sweep (DPRK_solver_threshold_fix_tissue, G:323-429)
against the FP64 oracle, through the C ABI.

Expected values: tests/golden/sweep_cells.json, written by the oracle
(tests/golden/make_sweep_cells.py; the CPU suite re-checks the fixture).  In
the reference's arithmetic the GPU sweep must give exactly the oracle's
thresholds; the resumable sweep must give exactly the drop-in's.
"""
import json
from pathlib import Path

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

from bubbledynamics_amd import gilmore as gm  # noqa: E402
from bubbledynamics_amd._ffi import BD_GM_ARITH_FAST, BD_GM_ARITH_REFERENCE  # noqa: E402

FIX = json.loads(Path(__file__).with_name("golden").joinpath("sweep_cells.json").read_text())


def setting(name, arithmetic=BD_GM_ARITH_REFERENCE):
    s = FIX["settings"][name]
    p = gm.reference_params()
    assert p.arithmetic == BD_GM_ARITH_REFERENCE       # the drop-in default
    p.arithmetic = arithmetic
    p.tableau = s["tableau"]
    p.abs_eps = p.rel_eps = s["eps"]
    p.max_solve_steps = s["max_solve_steps"]
    p.minR0, p.stepR0 = FIX["min_r0"], FIX["step_r0"]
    return p, np.array(FIX["fr_combo"]), FIX["n_r0"], np.array(s["threshold_pa"])


@pytest.mark.parametrize("name", ["dp1980", "reference", "capped"])
def test_drop_in_sweep_equals_oracle(name):
    p, fr, n_r0, expect = setting(name)
    print(f"sweep {name}: {len(expect)} cells", flush=True)
    got = gm.threshold_sweep(p, fr, n_r0)
    print("GPU   ", got.tolist(), "\noracle", expect.tolist())
    np.testing.assert_array_equal(got, expect)
    if name == "capped":
        assert (got == gm.BD_GM_SWEEP_CAPPED).any() and (got > 0).any()


@pytest.mark.parametrize("name", ["dp1980", "reference"])
def test_resumable_sweep_equals_drop_in(name):
    # launches of 1,500 attempts per lane: cells are suspended and continued
    # many times, and every result equals the unsliced sweep's
    p, fr, n_r0, expect = setting(name)
    sw = gm.ThresholdSweep(p, fr, n_r0)
    sw.advance(1500)
    torch.cuda.synchronize()
    first = sw.progress()
    print(f"{name}: after one advance {first}", flush=True)
    assert first["launch_ended"] and first["suspended"] > 0
    got = sw.run(1500)
    print(f"{name}: {sw.advances} advances", flush=True)
    sw.close()
    assert sw.advances > 3
    np.testing.assert_array_equal(got, expect)


def test_resumable_sweep_sub_range_and_stop_word():
    # a stop request makes an advance claim nothing and suspend what it holds;
    # cleared, the sweep goes on; a sub-range leaves the other cells untouched
    p, fr, n_r0, expect = setting("dp1980")
    sw = gm.ThresholdSweep(p, fr, n_r0, cell0=5, cell_end=19)
    stop = sw.stop_word()
    stop[0] = 1
    try:
        sw.advance(0)
        torch.cuda.synchronize()
        pr = sw.progress()
        assert pr["launch_ended"] and pr["claimed"] == 0 and pr["completed"] == 0
    finally:
        stop[0] = 0
    got = sw.run(4000)
    sw.close()
    np.testing.assert_array_equal(got[5:19], expect[5:19])
    assert np.isnan(got[:5]).all() and np.isnan(got[19:]).all()


def test_fast_arithmetic_sweep_near_oracle():
    # the fast arithmetic (shared subexpressions, cancellation-free
    # I^((n-1)/n) - 1, FSAL) is not bit-identical to the reference's: its
    # thresholds may differ by the 1 kPa resolution on cells whose collapse
    # is marginal.  The measured agreement is printed; the bound is 1 kPa.
    p, fr, n_r0, expect = setting("dp1980", BD_GM_ARITH_FAST)
    got = gm.threshold_sweep(p, fr, n_r0)
    same = int(np.sum(got == expect))
    print(f"fast arithmetic: {same}/{len(expect)} cells equal, max |diff| {np.abs(got - expect).max():.0f} Pa")
    assert np.all(np.abs(got - expect) <= 1e3)
