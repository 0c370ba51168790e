"""GPU parity of the Westervelt kernels against the FP64 oracle (through the C ABI).

Tolerances: FP64 <= 1e-12 relative to the field's max (device sin/exp/sqrt may
differ from glibc by an ulp); FP32 single step <= 2e-6 of the field max;
multi-step FP32 runs are held to the relative L2 error stated per test.
"""
import ctypes as C
import math

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

from bubbledynamics_amd import _ffi, presets  # noqa: E402
from bubbledynamics_amd._ffi import BD_BC_MUR, BD_BC_ZERO, BD_F32, BD_F64, BD_SRC_ARRAY  # noqa: E402
from bubbledynamics_amd.westervelt import WesterveltRun  # noqa: E402


def dev(a, dtype):
    return torch.as_tensor(np.ascontiguousarray(a), dtype=dtype, device="cuda")


def host(t):
    return t.detach().double().cpu().numpy()


def vp(t):
    return C.c_void_p(t.data_ptr())


def smooth_random(rng, shape, scale=1e5):
    a = rng.standard_normal(shape)
    # light smoothing keeps the fields wave-like
    for ax in range(3):
        a = 0.5 * a + 0.25 * (np.roll(a, 1, ax) + np.roll(a, -1, ax))
    return a * scale


def rel_err(a, b):
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
@pytest.mark.parametrize("full", [0, 1])
def test_space6_dropin_parity(orc, dtype, full):
    rng = np.random.default_rng(7 + full)
    n, nz = 40, 48
    p = presets.reference_scaled(n, nz)
    p.full_domain = full
    p.sponge_cells = 6 if full else 0
    p.sponge_alpha = 0.05
    shp = (nz, n, n)
    lv = [smooth_random(rng, shp) for _ in range(5)]
    Rh = [smooth_random(rng, shp, 1e-3) for _ in range(4)]
    # oracle
    Pn = np.zeros(shp); R1 = np.zeros(shp); R2 = np.zeros(shp)
    orc.wv_space6(p, n, nz, Pn, *lv, R1, Rh[0], Rh[1], R2, Rh[2], Rh[3])
    # device (same inputs rounded to dtype; the oracle then sees the rounded values too)
    dl = [dev(a, dtype) for a in lv]
    dR = [dev(a, dtype) for a in Rh]
    lv_r = [host(t) for t in dl]
    Rh_r = [host(t) for t in dR]
    Pn_r = np.zeros(shp); R1_r = np.zeros(shp); R2_r = np.zeros(shp)
    orc.wv_space6(p, n, nz, Pn_r, *lv_r, R1_r, Rh_r[0], Rh_r[1], R2_r, Rh_r[2], Rh_r[3])
    dPn = torch.zeros(shp, dtype=dtype, device="cuda")
    dR1 = torch.zeros(shp, dtype=dtype, device="cuda")
    dR2 = torch.zeros(shp, dtype=dtype, device="cuda")
    code = BD_F64 if dtype == torch.float64 else BD_F32
    _ffi.call("bd_westervelt_space6", C.byref(p), code, vp(dPn), *map(vp, dl), vp(dR1), vp(dR[0]),
              vp(dR[1]), vp(dR2), vp(dR[2]), vp(dR[3]), n, nz, None)
    torch.cuda.synchronize()
    lo = 3 if full else 0
    sl = (slice(3, nz - 3), slice(lo, n - 3), slice(lo, n - 3))
    izt = orc.first_tissue_plane(p, nz)
    tol = 1e-13 if dtype == torch.float64 else 2e-6
    assert rel_err(host(dPn)[sl], Pn_r[sl]) <= tol
    slt = (slice(max(3, izt), nz - 3), sl[1], sl[2])
    assert rel_err(host(dR1)[slt], R1_r[slt]) <= tol * 10
    assert rel_err(host(dR2)[slt], R2_r[slt]) <= tol * 10
    if dtype == torch.float64:
        assert rel_err(host(dPn)[sl], Pn[sl]) <= 1e-13


@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
@pytest.mark.parametrize("case", ["mirror_mur", "full_mur", "full_zero_array"])
def test_source_and_abc_dropin_parity(orc, dtype, case):
    rng = np.random.default_rng(11)
    n, nz = 36, 40
    if case == "full_zero_array":
        p = presets.phased_array(n, n_el=4, dx=2e-4)
        amp = rng.uniform(0.5e6, 1e6, 16)
        ph = rng.uniform(0, 2 * math.pi, 16)
    else:
        p = presets.reference_scaled(n, nz)
        p.full_domain = 1 if case == "full_mur" else 0
        if p.full_domain:
            p.xy_min = -(n - 1) / 2 * p.dxy
        amp = ph = None
    shp = (nz, n, n)
    P = smooth_random(rng, shp)
    Pm1 = smooth_random(rng, shp)
    code = BD_F64 if dtype == torch.float64 else BD_F32
    dP, dPm1 = dev(P, dtype), dev(Pm1, dtype)
    P_r, Pm1_r = host(dP), host(dPm1)
    t = 1.234e-6
    orc.wv_source(p, P_r, t, n, nz, amp, ph)
    orc.wv_abc(p, P_r, Pm1_r, t, n, nz)
    damp = dev(amp, torch.float64) if amp is not None else None
    dph = dev(ph, torch.float64) if ph is not None else None
    _ffi.call("bd_boundary_condition_source", C.byref(p), code, vp(dP), t, n, nz,
              vp(damp) if damp is not None else None, vp(dph) if dph is not None else None, None)
    _ffi.call("bd_boundary_condition_space6", C.byref(p), code, vp(dP), vp(dPm1), t, n, nz, None)
    torch.cuda.synchronize()
    tol = 1e-13 if dtype == torch.float64 else 2e-6
    assert rel_err(host(dP), P_r) <= tol


def _oracle_run(orc, p, n, nz, steps, focus, amp=None, ph=None):
    state = np.zeros((12, nz, n, n))
    pmax = np.zeros((nz, n, n))
    energy = np.zeros(2 * steps)
    fs = np.zeros(steps)
    orc.wv_run(p, n, nz, 0, steps, state, pmax, energy, fs, focus, amp, ph, 0)
    return state, pmax, energy.reshape(steps, 2), fs


@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
def test_plan_run_parity_reference_physics(orc, dtype):
    # the reference's two-layer water/muscle medium, cap with hole, Mur faces
    n, nz, steps = 40, 56, 150
    p = presets.reference_scaled(n, nz)
    focus = (0, 0, nz - 10)
    st, pmax, en, fs = _oracle_run(orc, p, n, nz, steps, focus)
    run = WesterveltRun(p, n, nz, dtype=dtype, max_steps=steps, focus=focus)
    run.run(steps, use_graph=True)
    run.finalize()
    torch.cuda.synchronize()
    slot = steps % 6
    ref = st[slot]
    got = host(run.pressure(0))
    if dtype == torch.float64:
        tol_f, tol_e = 1e-11, 1e-11
    else:
        tol_f, tol_e = 2e-4, 5e-4
    assert np.linalg.norm(got - ref) / np.linalg.norm(ref) <= tol_f
    assert np.linalg.norm(host(run.pmax) - pmax) / np.linalg.norm(pmax) <= tol_f
    e = host(run.energy)
    assert np.abs(e[:, 0] - en[:, 0]).max() / np.abs(en[:, 0]).max() <= tol_e
    f = host(run.focus)
    assert np.linalg.norm(f - fs) / np.linalg.norm(fs) <= tol_f * 5
    # relaxation state (tissue half)
    r1 = host(run.relax(1, 0))
    r1_ref = st[6 + (steps - 1) % 3]
    assert np.linalg.norm(r1 - r1_ref) / np.linalg.norm(r1_ref) <= tol_f * 10


def test_graph_replay_equals_eager_bitwise():
    # (static tile: a tile measured per plan could differ between the plans and
    # regroup the energy partials)
    n, nz, steps = 40, 48, 40
    p = presets.reference_scaled(n, nz)
    a = WesterveltRun(p, n, nz, dtype=torch.float32, max_steps=steps, focus=(0, 0, 30), prepare_graphs=False)
    b = WesterveltRun(p, n, nz, dtype=torch.float32, max_steps=steps, focus=(0, 0, 30), prepare_graphs=False)
    a.run(steps, use_graph=False)
    b.run(5, use_graph=True)      # eager head, then graph replays, then eager tail
    b.run(steps - 5, use_graph=True)
    a.finalize(); b.finalize()
    torch.cuda.synchronize()
    assert torch.equal(a.state, b.state)
    assert torch.equal(a.pmax, b.pmax)
    assert torch.equal(a.energy, b.energy)
    assert torch.equal(a.focus, b.focus)
    # determinism: a second identical run gives identical bits
    c = WesterveltRun(p, n, nz, dtype=torch.float32, max_steps=steps, focus=(0, 0, 30), prepare_graphs=False)
    c.run(steps)
    c.finalize()
    torch.cuda.synchronize()
    assert torch.equal(a.state, c.state) and torch.equal(a.energy, c.energy)


@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
def test_phased_array_full_domain_parity(orc, dtype):
    n, steps = 48, 120
    p = presets.phased_array(n, n_el=4, dx=2e-4)
    rng = np.random.default_rng(4)
    amp = rng.uniform(0.5e6, 1e6, 16)
    ph = rng.uniform(0, 2 * math.pi, 16)
    focus = (n // 2, n // 2, n // 2)
    st, pmax, en, fs = _oracle_run(orc, p, n, n, steps, focus, amp, ph)
    run = WesterveltRun(p, n, n, dtype=dtype, max_steps=steps, focus=focus, elem_amp=amp, elem_phase=ph)
    run.run(steps)
    run.finalize()
    torch.cuda.synchronize()
    ref = st[steps % 6]
    tol = 1e-11 if dtype == torch.float64 else 2e-4
    assert np.linalg.norm(host(run.pressure(0)) - ref) / np.linalg.norm(ref) <= tol
    assert np.linalg.norm(host(run.focus) - fs) / np.linalg.norm(fs) <= tol * 5


@pytest.mark.parametrize("n,full", [(67, 0), (70, 1), (40, 1), (44, 0)])
def test_plan_faces_at_tile_edges(orc, n, full):
    # Mur faces of a run: quarter (67) and full (70) domains
    # put the last inside column and row on the last lane / row of a 64 x 4
    # tile, so the corner layers come from neighbouring tiles' halo; FP64
    # state and pmax must match the oracle's face order everywhere.
    nz, steps = 24, 90
    p = presets.reference_scaled(n, nz)
    p.d = 0.8 * (n - 4) * p.dxy
    p.alpha, p.alpha_hole = math.asin(0.9), 0.0
    if full:
        p.full_domain = 1
        p.xy_min = -(n - 1) / 2 * p.dxy
    focus = (n // 2, n // 2, nz // 2)
    st, pmax, en, fs = _oracle_run(orc, p, n, nz, steps, focus)
    run = WesterveltRun(p, n, nz, dtype=torch.float64, max_steps=steps, focus=focus)
    run.run(steps, use_graph=True)
    run.finalize()
    torch.cuda.synchronize()
    ref = st[steps % 6]
    got = host(run.pressure(0))
    face = np.zeros((nz, n, n), bool)
    face[:, :, n - 3:] = face[:, n - 3:, :] = True
    if full:
        face[:, :, :3] = face[:, :3, :] = True
    assert np.abs(ref[face]).max() > 1e-3 * np.abs(ref).max()     # the faces are exercised
    scale = np.abs(ref).max()
    assert np.abs(got - ref).max() <= 1e-11 * scale
    assert np.abs(host(run.pmax) - pmax).max() <= 1e-11 * np.abs(pmax).max()
    e = host(run.energy)
    assert np.abs(e[:, 0] - en[:, 0]).max() <= 1e-11 * np.abs(en[:, 0]).max()


# tile shapes the step kernel is instantiated for: (columns, thread rows,
# rows per thread, plane prefetch), selected with BD_WV_VARIANT
TILE_SHAPES = [(64, 4, 1, 0), (64, 4, 1, 1), (64, 4, 2, 1), (64, 8, 1, 1), (64, 8, 2, 1),
               (64, 4, 4, 1), (64, 8, 2, 0), (32, 8, 2, 1), (128, 4, 1, 0), (64, 4, 2, 2),
               (64, 4, 1, 2), (64, 8, 1, 2), (128, 4, 1, 2), (64, 4, 1, 3)]
_TILE_ORACLE = {}


def _tile_case(orc, n, full, nz=24, steps=60):
    key = (n, full, nz, steps)
    if key not in _TILE_ORACLE:
        p = presets.reference_scaled(n, nz)
        p.d = 0.8 * (n - 4) * p.dxy
        p.alpha, p.alpha_hole = math.asin(0.9), 0.0
        if full:
            p.full_domain = 1
            p.xy_min = -(n - 1) / 2 * p.dxy
        focus = (n // 2, n // 2, nz // 2)
        _TILE_ORACLE[key] = (p, focus, _oracle_run(orc, p, n, nz, steps, focus) if orc else None)
    return _TILE_ORACLE[key]


def _run_shape(monkeypatch, variant, p, n, nz, steps, focus, dtype):
    if variant is None:
        monkeypatch.delenv("BD_WV_VARIANT", raising=False)
    else:
        monkeypatch.setenv("BD_WV_VARIANT", str(variant))
    run = WesterveltRun(p, n, nz, dtype=dtype, max_steps=steps, focus=focus)
    monkeypatch.delenv("BD_WV_VARIANT", raising=False)
    sh = run.shape
    if variant is not None:
        assert (sh["tile_x"], sh["tile_y"], sh["rows"], sh["prefetch"]) == TILE_SHAPES[variant]
    run.run(steps, use_graph=True)
    run.finalize()
    torch.cuda.synchronize()
    return run


@pytest.mark.parametrize("variant", range(len(TILE_SHAPES)))
@pytest.mark.parametrize("n,full", [(67, 0), (70, 1)])
def test_every_tile_shape_matches_oracle(orc, monkeypatch, variant, n, full):
    # every instantiated tile shape (not only the default one) against the FP64
    # oracle on grids whose last inside row and column end a 64 x 4 tile
    nz, steps = 24, 60
    p, focus, (st, pmax, en, fs) = _tile_case(orc, n, full, nz, steps)
    run = _run_shape(monkeypatch, variant, p, n, nz, steps, focus, torch.float64)
    ref = st[steps % 6]
    assert np.abs(host(run.pressure(0)) - ref).max() <= 1e-11 * np.abs(ref).max()
    assert np.abs(host(run.pmax) - pmax).max() <= 1e-11 * np.abs(pmax).max()
    e = host(run.energy)
    assert np.abs(e[:, 0] - en[:, 0]).max() <= 1e-11 * np.abs(en[:, 0]).max()
    assert np.abs(host(run.focus) - fs).max() <= 1e-11 * np.abs(fs).max()


@pytest.mark.parametrize("variant", range(1, len(TILE_SHAPES)))
def test_tile_shapes_fp32_bitwise_equal(monkeypatch, variant):
    # the per-cell arithmetic does not depend on the tile: FP32 fields, pmax and
    # focus record are bit-identical across shapes; the per-step energies sum
    # per-thread Kahan-compensated FP32 partials (each within 2 u = 1.2e-7 of
    # its exact sum) whose grouping follows the tile
    n, nz, steps = 67, 40, 120
    p, focus, _ = _tile_case(None, n, 0, nz, steps)
    a = _run_shape(monkeypatch, 0, p, n, nz, steps, focus, torch.float32)
    b = _run_shape(monkeypatch, variant, p, n, nz, steps, focus, torch.float32)
    assert np.abs(host(a.pressure(0))).max() > 0
    for lvl in range(5):
        assert torch.equal(a.pressure(lvl), b.pressure(lvl))
    for which in (1, 2):
        assert torch.equal(a.relax(which, 0), b.relax(which, 0))
    assert torch.equal(a.pmax, b.pmax)
    assert torch.equal(a.focus, b.focus)
    ea, eb = host(a.energy), host(b.energy)
    for k in range(2):
        assert np.abs(ea[:, k] - eb[:, k]).max() <= 1e-6 * np.abs(ea[:, k]).max()


@pytest.mark.parametrize("nt", [0, 1, 2])
@pytest.mark.parametrize("variant,dtype", [(13, torch.float32), (10, torch.float32), (2, torch.float32),
                                           (0, torch.float32), (0, torch.float64), (10, torch.float64)])
def test_history_cache_policies_bitwise_equal(monkeypatch, variant, dtype, nt):
    # the history loads' cache policy (cached / every level non-temporal /
    # only p^{n-3}, p^{n-4}, R^{n-2} non-temporal, forced with BD_WV_NT) does
    # not touch the arithmetic: fields, pmax and focus equal the default's
    # (FP64 variant 0 issues every load of a plane, through zero-sized
    # descriptors where the branchy tiles skip them)
    n, nz, steps = 67, 40, 120
    p, focus, _ = _tile_case(None, n, 0, nz, steps)
    a = _run_shape(monkeypatch, variant, p, n, nz, steps, focus, dtype)
    monkeypatch.setenv("BD_WV_NT", str(nt))
    b = _run_shape(monkeypatch, variant, p, n, nz, steps, focus, dtype)
    monkeypatch.delenv("BD_WV_NT", raising=False)
    assert b.shape["nontemporal"] == nt and a.shape["nontemporal"] == 0
    assert np.abs(host(a.pressure(0))).max() > 0
    assert torch.equal(a.state, b.state) and torch.equal(a.pmax, b.pmax) and torch.equal(a.focus, b.focus)


def test_default_shape_rule(monkeypatch):
    # without the measured choice of bd_westervelt_plan_prepare: one row per
    # thread with three planes of loads in flight for FP32, none for FP64; the
    # history loads are non-temporal once the 12-field state exceeds 256 MB
    monkeypatch.delenv("BD_WV_VARIANT", raising=False)
    monkeypatch.delenv("BD_WV_BLOCKS", raising=False)
    expect = {(192, torch.float32): (1, 3, 1), (128, torch.float32): (1, 3, 0),
              (160, torch.float32): (1, 3, 0), (192, torch.float64): (1, 0, 1)}
    for (n, dt), (rows, pf, nt) in expect.items():
        p = presets.tone_burst_192(n)
        run = WesterveltRun(p, n, n, dtype=dt, max_steps=1, pmax=False, energy=False, prepare_graphs=False)
        sh = run.shape
        assert (sh["rows"], sh["prefetch"], sh["nontemporal"]) == (rows, pf, nt), (n, dt, sh)
        assert sh["tile_x"] == 64 and 768 <= sh["workgroups"] <= 1024
        run.close()
        del run


@pytest.mark.parametrize("dtype", [torch.float32, torch.float64])
def test_prepare_chooses_a_measured_tile(monkeypatch, dtype):
    # bd_westervelt_plan_prepare times the candidate tiles on the bound state
    # before the first step and keeps the fastest; its side effects are
    # overwritten by the first step, so fields, pmax and focus equal those of
    # the static tile bitwise (energy to the partials' grouping)
    monkeypatch.delenv("BD_WV_VARIANT", raising=False)
    monkeypatch.delenv("BD_WV_BLOCKS", raising=False)
    cands = {torch.float32: (13, 10, 2, 0), torch.float64: (0, 10)}[dtype]
    n, nz, steps = 67, 40, 60
    p, focus, _ = _tile_case(None, n, 0, nz, steps)
    a = WesterveltRun(p, n, nz, dtype=dtype, max_steps=steps, focus=focus)             # measured tile
    b = WesterveltRun(p, n, nz, dtype=dtype, max_steps=steps, focus=focus, prepare_graphs=False)
    sh = a.shape
    assert (sh["tile_x"], sh["tile_y"], sh["rows"], sh["prefetch"]) in [TILE_SHAPES[c] for c in cands]
    assert sh["nontemporal"] in (0, 2)          # a state the last-level cache holds: cached or partial
    a.run(steps)
    b.run(steps)
    a.finalize(); b.finalize()
    torch.cuda.synchronize()
    assert np.abs(host(a.pressure(0))).max() > 0
    assert torch.equal(a.state, b.state) and torch.equal(a.pmax, b.pmax) and torch.equal(a.focus, b.focus)
    ea, eb = host(a.energy), host(b.energy)
    assert np.abs(ea - eb).max() <= 1e-6 * np.abs(eb).max()


@pytest.mark.parametrize("dtype", [torch.float32, torch.float64])
def test_pmax_exact_after_every_run(dtype):
    # the interior peak pressure is updated once every 6 steps over the six
    # levels a step holds, plus a fold at the end of each run; after every run
    # it must equal the running maximum formed here level by level, also after
    # a restart from a kept state with pmax cleared
    n, nz = 44, 40
    p = presets.reference_scaled(n, nz)
    run = WesterveltRun(p, n, nz, dtype=dtype, max_steps=200, energy=False)
    ref_run = WesterveltRun(p, n, nz, dtype=dtype, max_steps=200, energy=False)
    ref = torch.zeros_like(ref_run.pmax)
    kept = None
    for m in (7, 5, 13, 1, 6, 11, 24, 3):
        run.run(m)
        for _ in range(m):
            ref_run.run(1, use_graph=False)
            ref = torch.maximum(ref, ref_run.pressure(0))
        if run.steps == 26:
            kept = run.state.clone()
        torch.cuda.synchronize()
        assert torch.equal(run.pressure(0), ref_run.pressure(0))
        assert torch.equal(run.pmax, ref), run.steps
    assert torch.count_nonzero(ref) > 0.5 * ref.numel()
    # restart at step 26 with pmax cleared: only the levels produced after it count
    run.state.copy_(kept)
    run.seek(26)
    run.pmax.zero_()
    ref_run.state.copy_(kept)
    ref_run.seek(26)
    ref = torch.zeros_like(ref)
    for m in (4, 9, 17):
        run.run(m)
        for _ in range(m):
            ref_run.run(1, use_graph=False)
            ref = torch.maximum(ref, ref_run.pressure(0))
        torch.cuda.synchronize()
        assert torch.equal(run.pmax, ref), run.steps


def test_tone_burst_small_fp32_vs_fp64(orc):
    # the benchmark workload's physics at 64^3: FP32 focal pressure against the
    # FP64 oracle, relative L2 over the whole record
    n = 64
    p = presets.tone_burst_192(n)
    focus = presets.tone_burst_focus(p, n)
    steps = 600
    st, pmax, en, fs = _oracle_run(orc, p, n, n, steps, focus)
    run = WesterveltRun(p, n, n, dtype=torch.float32, max_steps=steps, focus=focus)
    run.run(steps)
    run.finalize()
    torch.cuda.synchronize()
    f = host(run.focus)
    err = np.linalg.norm(f - fs) / np.linalg.norm(fs)
    print(f"tone burst 64^3 FP32 focal rel L2 = {err:.3e}")
    assert np.abs(fs).max() > 1e5       # the burst reached the focus
    assert err <= 1e-4


def test_invalid_arguments_rejected():
    p = presets.reference_scaled(40, 48)
    q = p.copy()
    q.dt = 10 * q.dt   # Courant number far beyond the leapfrog limit
    plan = C.c_void_p()
    rc = _ffi.load().bd_westervelt_plan_create(C.byref(q), BD_F32, 40, 48, 10, -1, -1, -1, C.byref(plan))
    assert rc == _ffi.BD_ERR_UNSTABLE
    rc = _ffi.load().bd_westervelt_plan_create(C.byref(p), 7, 40, 48, 10, -1, -1, -1, C.byref(plan))
    assert rc == _ffi.BD_ERR_INVALID
    rc = _ffi.load().bd_westervelt_plan_create(C.byref(p), BD_F32, 40, 48, 10, 39, 0, 10, C.byref(plan))
    assert rc == _ffi.BD_ERR_INVALID
