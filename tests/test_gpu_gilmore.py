"""GPU parity of the Gilmore kernels against the FP64 oracle (through the C ABI)."""
import math

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

from bubbledynamics_amd import gilmore as gm  # noqa: E402
from bubbledynamics_amd._ffi import BD_GM_DP45, BD_GM_FP64, BD_GM_RK4, BD_GM_SPLIT  # noqa: E402


def host(t):
    return t.cpu().numpy()


def test_threshold_sweep_range_prefix_and_stop():
    # the range sweep gives the drop-in's values on the whole table and on a
    # sub-range, with the exact completed count; a stop request set before the
    # launch makes it claim nothing (the kernel exits at its first claim)
    import ctypes as C
    import time
    from bubbledynamics_amd import _ffi
    # (the fast arithmetic: this test is about the range and stop mechanics)
    p = gm.reference_params()
    p.tableau = 1
    p.arithmetic = gm.BD_GM_ARITH_FAST
    p.minR0, p.stepR0 = 1e-6, 1e-6
    fr = np.array([[1e6, 1e6, 1e6], [1e6, 2e6, 3e6], [2e6, 3e6, 5e6]])
    n_r0 = 4
    full = gm.threshold_sweep(p, fr, n_r0)
    out, end, _ = gm.threshold_sweep_range(p, fr, n_r0, poll_s=0.5)
    assert end == 12
    np.testing.assert_array_equal(out, full)
    out, end, _ = gm.threshold_sweep_range(p, fr, n_r0, cell0=5, cell_end=11, poll_s=0.5)
    assert end == 11
    np.testing.assert_array_equal(out[5:11], full[5:11])
    assert np.isnan(out[:5]).all() and np.isnan(out[11:]).all()
    lib = _ffi.load()
    flags = lib.bd_gilmore_sweep_flags()
    res = torch.full((12,), float("nan"), dtype=torch.float64, device="cuda")
    frd = torch.as_tensor(fr.reshape(-1), device="cuda")
    flags[2] = 1
    try:
        _ffi.call("bd_gilmore_threshold_sweep_range", C.byref(p), C.c_void_p(res.data_ptr()),
                  C.c_void_p(frd.data_ptr()), 3, n_r0, 0, 12, C.c_void_p(torch.cuda.current_stream().cuda_stream))
        t0 = time.perf_counter()
        while flags[3] == 0 and time.perf_counter() - t0 < 60.0:
            time.sleep(0.05)
        torch.cuda.synchronize()
        assert flags[3] == 1 and flags[1] == 0
        assert torch.isnan(res).all()
    finally:
        flags[2] = 0


def _batch(n, seed, t_end_us):
    R0, amp, ph, fr = gm.config0_inputs(n, seed)
    return R0, amp, ph, fr, t_end_us * 1e-6


def test_rk4_batch_matches_oracle(orc):
    p = gm.workload_params()
    R0, amp, ph, fr, T = _batch(64, 0, 2.0)
    b = gm.BubbleBatch(R0, amp, ph, fr)
    out = b.run(p, BD_GM_RK4, T, 1e-10, out_every=10, n_out=2001)
    torch.cuda.synchronize()
    ref = orc.gilmore_integrate(p, BD_GM_RK4, R0, amp, ph, fr, T, 1e-10, out_every=10, n_out=2001, nthreads=8)
    np.testing.assert_allclose(host(out["Rmax"]), ref["Rmax"], rtol=1e-9)
    np.testing.assert_array_equal(host(out["n_acc"]), ref["n_acc"])
    tc, tcr = host(out["t_collapse"]), ref["t_collapse"]
    np.testing.assert_allclose(tc, tcr, rtol=1e-12)
    Ro, Rr = host(out["R_out"]), ref["R_out"]
    np.testing.assert_array_equal(np.isnan(Ro), np.isnan(Rr))
    fin = ~np.isnan(Rr)
    np.testing.assert_allclose(Ro[fin], Rr[fin], rtol=1e-8)
    assert (tcr > 0).any() and (tcr < 0).any()       # both collapsing and surviving bubbles


@pytest.mark.parametrize("tab", [0, 1])
def test_dp45_batch_matches_oracle(orc, tab):
    p = gm.workload_params()
    p.tableau = tab
    R0, amp, ph, fr = gm.config3_inputs(48, 3)
    T = 3e-6
    b = gm.BubbleBatch(R0, amp, ph, fr)
    out = b.run(p, BD_GM_DP45, T, 1e-13, rtol=1e-8, atol=1e-12)
    torch.cuda.synchronize()
    ref = orc.gilmore_integrate(p, BD_GM_DP45, R0, amp, ph, fr, T, 1e-13, rtol=1e-8, atol=1e-12, nthreads=8)
    rel = np.abs(host(out["Rmax"]) / ref["Rmax"] - 1)
    print("DP45 Rmax rel err max", rel.max(), "median", np.median(rel))
    assert np.median(rel) <= 1e-9
    assert rel.max() <= 1e-6
    acc, acc_r = host(out["n_acc"]), ref["n_acc"]
    assert np.all(np.abs(acc - acc_r) <= np.maximum(3, 0.01 * acc_r))
    np.testing.assert_array_equal(host(out["t_collapse"]) > 0, ref["t_collapse"] > 0)


@pytest.mark.parametrize("tab", [0, 1])
def test_dp45_batch_reference_arithmetic_matches_oracle(orc, tab):
    # the reference's expression order with the first stage recomputed: only
    # libm rounding separates the GPU from the oracle
    from bubbledynamics_amd._ffi import BD_GM_ARITH_REFERENCE
    p = gm.workload_params(arithmetic=BD_GM_ARITH_REFERENCE)
    p.tableau = tab
    R0, amp, ph, fr = gm.config3_inputs(48, 3)
    T = 3e-6
    b = gm.BubbleBatch(R0, amp, ph, fr)
    out = b.run(p, BD_GM_DP45, T, 1e-13, rtol=1e-8, atol=1e-12)
    torch.cuda.synchronize()
    ref = orc.gilmore_integrate(p, BD_GM_DP45, R0, amp, ph, fr, T, 1e-13, rtol=1e-8, atol=1e-12, nthreads=8)
    rel = np.abs(host(out["Rmax"]) / ref["Rmax"] - 1)
    acc, acc_r = host(out["n_acc"]), ref["n_acc"]
    same = np.mean(acc == acc_r)
    print(f"DP45 reference arithmetic: Rmax rel err max {rel.max():.3e} median {np.median(rel):.3e}; "
          f"equal step counts {same:.3f}")
    assert rel.max() <= 1e-8
    assert np.all(np.abs(acc - acc_r) <= np.maximum(3, 0.01 * acc_r))
    np.testing.assert_array_equal(host(out["t_collapse"]) > 0, ref["t_collapse"] > 0)


def test_split_precision_rmax_error(orc):
    # FP32 right-hand sides while R/R0 >= 0.2 (per wavefront), FP64 below, on
    # the stratified config-3 subset over 20 us: Rmax against the FP64 oracle
    # and against the all-FP64 kernel.  rtol 1e-8 is below FP32's epsilon and
    # atol 1e-12 on dR/dt (SI) is far below the FP32 rounding of d2R/dt2 near
    # rest, so the split does not reach the 1e-6 target: the bound here is
    # what it achieves on the bubbles the oracle can pin (one-ulp inputs move
    # Rmax by <= 1e-9), and all-FP64 is the configuration's answer.
    import json
    from pathlib import Path
    p = gm.workload_params()
    idx = gm.config3_subset()
    R0, amp, ph, fr = gm.config3_inputs(1_000_000, 3)
    R0, amp, ph = R0[idx], amp[idx], ph[idx]
    T = 20e-6
    b = gm.BubbleBatch(R0, amp, ph, fr)
    o64 = b.run(p, BD_GM_DP45, T, 1e-13, rtol=1e-8, atol=1e-12)
    ref = orc.gilmore_integrate(p, BD_GM_DP45, R0, amp, ph, fr, T, 1e-13, rtol=1e-8, atol=1e-12, nthreads=16)
    fix = json.loads((Path(__file__).parent / "golden" / "gilmore_conditioning.json").read_text())
    ill = set(fix["config3_subset_ulp"]["positions"])
    well = np.array([i not in ill for i in range(len(idx))])
    # input order, and the sliced schedule (its first two passes run the
    # kernel's build for two wavefronts per SIMD)
    for schedule in (gm.BD_GM_SCHED_INPUT, gm.BD_GM_SCHED_SLICED):
        o32 = b.run(p, BD_GM_DP45, T, 1e-13, precision=BD_GM_SPLIT, rtol=1e-8, atol=1e-12, schedule=schedule)
        torch.cuda.synchronize()
        rel = np.abs(host(o32["Rmax"]) / ref["Rmax"] - 1)
        rel64 = np.abs(host(o32["Rmax"]) / host(o64["Rmax"]) - 1)
        flips = host(o32["t_collapse"] > 0) != (ref["t_collapse"] > 0)
        print(f"split (schedule {schedule}) vs oracle: median {np.median(rel):.3e} p99 {np.quantile(rel, 0.99):.3e} "
              f"max {rel.max():.3e}; well-conditioned max {rel[well].max():.3e}, > 1e-6: "
              f"{int(np.sum(rel[well] > 1e-6))}/{well.sum()}; vs all-FP64 median {np.median(rel64):.3e}; collapse "
              f"flips {int(flips.sum())}; accepted steps split {host(o32['n_acc']).sum()} FP64 "
              f"{host(o64['n_acc']).sum()}")
        keep = well & ~flips
        print(f"split: well-conditioned bubbles without a collapse flip: max {rel[keep].max():.3e}, "
              f"p99 {np.quantile(rel[keep], 0.99):.3e}")
        # measured: median 3.3e-6, 14 flips in 4,096; the bounds guard that level
        assert np.median(rel) <= 1e-5
        assert np.quantile(rel[keep], 0.99) <= 1e-2
        assert np.mean(flips) <= 0.01


def test_queue_scheduling_does_not_change_results():
    p = gm.workload_params()
    R0, amp, ph, fr = gm.config3_inputs(512, 3)
    b = gm.BubbleBatch(R0, amp, ph, fr)
    base = b.run(p, BD_GM_DP45, 4e-6, 1e-13)
    for block, chunk in ((64, 1), (128, 16), (256, 64)):
        o = b.run(p, BD_GM_DP45, 4e-6, 1e-13, block=block, chunk=chunk)
        torch.cuda.synchronize()
        assert torch.equal(o["Rmax"], base["Rmax"])
        assert torch.equal(o["n_acc"], base["n_acc"])


def test_autotune_picks_a_candidate_and_shape_does_not_change_results():
    # the autotuner measures launch shapes on a capped probe and records the
    # fastest; runs with the tuned shape and with explicit waves match bitwise
    p = gm.workload_params()
    R0, amp, ph, fr = gm.config3_inputs(4096, 3)
    b = gm.BubbleBatch(R0, amp, ph, fr)
    base = b.run(p, BD_GM_DP45, 4e-6, 1e-13, block=256, chunk=1, waves=8)
    t = b.autotune(p, BD_GM_DP45, 4e-6, 1e-13, sample=4096, probe_steps=300)
    assert t["block"] in (64, 128, 256) and t["chunk"] in (1, 4) and t["waves"] in (2, 4, 8)
    assert t["probe_rate"] > 0
    tuned = b.run(p, BD_GM_DP45, 4e-6, 1e-13)
    for waves in (2, 4):
        o = b.run(p, BD_GM_DP45, 4e-6, 1e-13, block=128, waves=waves)
        torch.cuda.synchronize()
        assert torch.equal(o["Rmax"], base["Rmax"]) and torch.equal(o["n_acc"], base["n_acc"])
    torch.cuda.synchronize()
    assert torch.equal(tuned["Rmax"], base["Rmax"])
    assert torch.equal(tuned["t_collapse"], base["t_collapse"])
    assert torch.equal(tuned["n_rej"], base["n_rej"])


@pytest.mark.parametrize("method,arith", [(BD_GM_DP45, "fast"), (BD_GM_DP45, "reference"), (BD_GM_RK4, "fast")],
                         ids=["dp45", "dp45-reference", "rk4"])
def test_sliced_schedule_gives_the_same_results(method, arith):
    # slicing, suspending and the cost-ordered second pass change only which
    # lane takes which bubble when (and, for the reference arithmetic, which
    # register build of the kernel runs a pass): every output matches the
    # input order bitwise.  (The split is left out: it chooses FP32 or FP64
    # per wavefront, so its results depend on which bubbles share one.)
    from bubbledynamics_amd._ffi import BD_GM_ARITH_FAST, BD_GM_ARITH_REFERENCE, BD_GM_SCHED_SLICED
    p = gm.workload_params(arithmetic=BD_GM_ARITH_REFERENCE if arith == "reference" else BD_GM_ARITH_FAST)
    R0, amp, ph, fr = gm.config3_inputs(6000, 3)
    b = gm.BubbleBatch(R0, amp, ph, fr)
    dt = 1e-13 if method == BD_GM_DP45 else 1e-10
    kw = dict(out_every=10, n_out=41) if method == BD_GM_RK4 else {}
    base = b.run(p, method, 4e-6, dt, **kw)
    o = b.run(p, method, 4e-6, dt, schedule=BD_GM_SCHED_SLICED, **kw)
    torch.cuda.synchronize()
    for k in ("Rmax", "t_collapse", "n_acc", "n_rej"):
        assert torch.equal(o[k], base[k]), k
    if method == BD_GM_RK4:
        assert torch.equal(torch.nan_to_num(o["R_out"], nan=-1.0), torch.nan_to_num(base["R_out"], nan=-1.0))
