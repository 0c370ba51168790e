"""Pins of the FP64 Gilmore oracle (CPU only): equilibrium, limits, derivative
consistency, linear natural frequency, the DP5(4) tableau, the frequency
table and the threshold-search semantics."""
import math

import numpy as np
import pytest


@pytest.fixture(scope="module")
def g(orc):
    return orc.gilmore_reference_params()


def test_reference_constants(g):
    # G:440-460
    assert g.B == pytest.approx(1549.0 ** 2 * 1100.0 / 7.0 - 1.013e5, rel=1e-15)
    assert g.B == pytest.approx(3.7695e8, rel=1e-4)
    assert g.gamma3 == pytest.approx(4.2)
    assert g.G4_3 == pytest.approx(0.04e6 * 4 / 3)


def test_rest_state_is_equilibrium(orc, g):
    # R=R0, dR=tau=q=0, no drive: H=0, C=c_inf, all derivatives 0
    for R0 in (1e-7, 2e-6, 1e-5):
        out = orc.gilmore_rhs(g, R0, [R0, 0.0, 0.0, 0.0], 0.0, [1e6, 2e6, 3e6], amp=[0, 0, 0], phase=[0, 0, 0])
        assert out[1] == pytest.approx(0.0, abs=1e-3)       # m/s^2 (terms ~1e12 cancel)
        assert out[2] == 0.0 and out[3] == 0.0
        assert out[4] == pytest.approx(0.0, abs=1e-9)
        assert out[6] == pytest.approx(g.C_inf, rel=1e-15)


def test_enthalpy_derivative_matches_finite_difference(orc, g):
    # G:132-157 is the time derivative of G:121-128 along the trajectory
    R0 = 3e-6
    amp, ph, fr = [0.4e6, 0.3e6, 0.2e6], [0.3, 1.1, 2.0], [1e6, 2e6, 3e6]
    y = np.array([2.1e-6, -35.0, 1.2e4, -3e3])
    t = 0.37e-6
    d = orc.gilmore_rhs(g, R0, y, t, fr, amp, ph)
    eps = 1e-12
    f = lambda s: orc.gilmore_rhs(g, R0, y + s * d[:4], t + s, fr, amp, ph)[4]
    fd = (f(eps) - f(-eps)) / (2 * eps)
    assert d[5] == pytest.approx(fd, rel=1e-5)


def test_rayleigh_plesset_limit(orc, g):
    # c_inf, B -> infinity: Gilmore -> R Rdd + 1.5 Rd^2 = (p_in - 2s/R - P0 - p + 3q)/rho.
    # c_inf = 1e6 keeps the compressible corrections ~1e-5 while (u-v)/v ~ 1e-9 is still
    # resolved by I^((n-1)/n) - 1 in FP64 (at c_inf = 1e9 it would cancel completely)
    # (a slow stress relaxation keeps tau' -- which enters through dH/dt * R/C -- small)
    q = g.copy()
    q.C_inf = 1e6
    q.lambda_ = 1e-6
    q.B = q.C_inf ** 2 * q.rho / q.n - q.P0
    R0 = 2e-6
    y = np.array([1.7e-6, 12.0, 5e3, -2e3])
    t = 0.21e-6
    amp, ph, fr = [0.5e6, 0, 0], [0.0, 0, 0], [1e6, 2e6, 3e6]
    out = orc.gilmore_rhs(q, R0, y, t, fr, amp, ph)
    R, Rd, tau, qq = y
    p = amp[0] * math.sin(2 * math.pi * fr[0] * t)
    p_in = (q.P0 + q.S2 / R0) * (R0 / R) ** q.gamma3
    rdd = (-1.5 * Rd * Rd + (p_in - q.S2 / R + tau - q.P0 - p) / q.rho + (3 * qq - tau) / q.rho) / R
    assert out[1] == pytest.approx(rdd, rel=5e-5)
    # at the liver's c_inf the compressible terms are visible
    ql = g.copy()
    ql.lambda_ = 1e-6
    out_l = orc.gilmore_rhs(ql, R0, y, t, fr, amp, ph)
    assert abs(out_l[1] / rdd - 1) > 1e-3


def test_linear_natural_frequency(orc, g):
    # free small oscillation: omega_d^2 = omega0^2 - beta^2,
    # omega0^2 = (3 gamma (P0 + 2s/R0) - 2s/R0 + 4G) / (rho R0^2), beta = 2 mu / (rho R0^2)
    R0 = 5e-6
    t, y, h = 0.0, np.array([R0 * (1 + 1e-4), 0.0, 0.0, 0.0]), 1e-12
    ts, rs = [t], [y[0] - R0]
    while t < 6e-6:
        t, y, h, _, _ = orc.dprk_step(g, R0, [1e6, 1e6, 1e6], 0.0, 1.0, t, y, min(h, 2e-9))
        ts.append(t); rs.append(y[0] - R0)
    ts, r = np.array(ts), np.array(rs)
    # zero crossings (linear interpolation)
    s = np.sign(r)
    idx = np.where(s[:-1] * s[1:] < 0)[0]
    assert len(idx) >= 6
    tz = ts[idx] - r[idx] * (ts[idx + 1] - ts[idx]) / (r[idx + 1] - r[idx])
    period = 2 * np.mean(np.diff(tz[1:]))
    G = g.G4_3 * 3 / 4
    mu = g.mu4 / 4
    w0sq = (g.gamma3 * (g.P0 + g.S2 / R0) - g.S2 / R0 + 4 * G) / (g.rho * R0 ** 2)
    beta = 2 * mu / (g.rho * R0 ** 2)
    wd = math.sqrt(w0sq - beta ** 2)
    assert 2 * math.pi / period == pytest.approx(wd, rel=0.01)
    # without the elastic term the frequency differs by far more than the tolerance
    assert abs(math.sqrt(w0sq - 4 * G / (g.rho * R0 ** 2) - beta ** 2) / wd - 1) > 0.05


def test_dp45_tableau_order_conditions(orc):
    # Dormand & Prince (1980): b gives order 5, bb order 4, c_i = sum_j a_ij, FSAL row
    a, b, bb, c = orc.dp_tableau(1)
    np.testing.assert_allclose(a.sum(axis=1), c, atol=1e-15)
    for k in range(5):
        assert b @ c ** k == pytest.approx(1.0 / (k + 1), rel=1e-14)
    for k in range(4):
        assert bb @ c ** k == pytest.approx(1.0 / (k + 1), rel=1e-14)
    # fifth-order conditions involving the matrix
    assert b @ (a @ c) == pytest.approx(1 / 6, rel=1e-14)
    assert b @ (c * (a @ c)) == pytest.approx(1 / 8, rel=1e-14)
    assert b @ (a @ (c ** 2)) == pytest.approx(1 / 12, rel=1e-14)
    assert b @ (a @ (a @ c)) == pytest.approx(1 / 24, rel=1e-14)
    assert b @ (c * (a @ (a @ c))) == pytest.approx(1 / 30, rel=1e-14)
    np.testing.assert_allclose(a[6], b, atol=0)              # FSAL: last stage row = b
    assert bb @ c ** 4 != pytest.approx(1 / 5, rel=1e-6)     # bb is genuinely 4th order


def test_reference_tableau_differs_only_in_a32(orc):
    # G:220 writes a32 = -9/40; the published coefficient is +9/40, so the
    # reference's third stage is not at c3 = 0.3 (row sum -0.15)
    a_ref, b_ref, bb_ref, c_ref = orc.dp_tableau(0)
    a, b, bb, c = orc.dp_tableau(1)
    d = a_ref - a
    assert np.count_nonzero(d) == 1 and d[2, 1] == pytest.approx(-18 / 40)
    np.testing.assert_array_equal(b_ref, b)
    np.testing.assert_array_equal(bb_ref, bb)
    assert a_ref[2].sum() == pytest.approx(-0.15)


def test_dp45_converges_to_rk4_reference(orc, g):
    # the same driven trajectory from the adaptive DP5(4) and from fine fixed-step
    # RK4 agree to the resolution of Rmax sampled at step end points (~1e-7); a
    # loose tolerance is visibly worse.  The reference tableau (a32 = -9/40) is
    # far less accurate at the same tolerance and needs many more steps.
    g = g.copy()
    g.tableau = 1
    R0 = [2.5e-6]
    amp, ph, fr = [[0.15e6, 0.1e6, 0.05e6]], [[0.2, 0.7, 1.3]], [1e6, 2e6, 3e6]
    T = 2e-6
    ref = orc.gilmore_integrate(g, 0, R0, amp, ph, fr, T, 1e-11)
    err = lambda r: abs(r["Rmax"][0] / ref["Rmax"][0] - 1)
    tight = orc.gilmore_integrate(g, 1, R0, amp, ph, fr, T, 1e-13, rtol=1e-10, atol=1e-14)
    loose = orc.gilmore_integrate(g, 1, R0, amp, ph, fr, T, 1e-13, rtol=1e-4, atol=1e-8)
    assert err(tight) < 3e-7
    assert err(loose) > 5 * err(tight)
    assert tight["n_acc"][0] > loose["n_acc"][0]
    gr = g.copy()
    gr.tableau = 0
    pub = orc.gilmore_integrate(g, 1, R0, amp, ph, fr, T, 1e-13, rtol=1e-8, atol=1e-12)
    refq = orc.gilmore_integrate(gr, 1, R0, amp, ph, fr, T, 1e-13, rtol=1e-8, atol=1e-12)
    assert err(refq) > 50 * err(pub)
    assert refq["n_acc"][0] > 5 * pub["n_acc"][0]


def test_frequency_table(orc):
    # G:465-469, G:498-520: 55 log-spaced tones, combos (i,i,k>=i) and (i<j<k)
    fr = orc.freq_table()
    assert fr.shape == (27775, 3)
    single = np.sum((fr[:, 0] == fr[:, 1]) & (fr[:, 1] == fr[:, 2]))
    dual = np.sum(((fr[:, 0] == fr[:, 1]) | (fr[:, 1] == fr[:, 2])) & ~((fr[:, 0] == fr[:, 1]) & (fr[:, 1] == fr[:, 2])))
    assert (single, dual, len(fr) - single - dual) == (55, 1485, 26235)
    tones = np.unique(fr[:, 0])
    assert tones[0] == pytest.approx(0.02e6) and np.unique(fr).max() == pytest.approx(5e6)
    np.testing.assert_allclose(np.diff(np.log10(np.unique(fr))), (math.log10(5e6) - math.log10(0.02e6)) / 54, rtol=1e-9)
    assert np.all(fr[:, 0] <= fr[:, 1]) and np.all(fr[:, 1] <= fr[:, 2])


def _collapses(orc, g, R0, fr, A):
    # one fresh solve (h reset to h0) at amplitude A, as in stages 2-3 (G:393-399)
    c, n = orc.gilmore_solve(g, R0, fr, A, 5.0 / min(fr), g.h0)
    assert n < g.max_solve_steps
    return c


def test_threshold_is_a_threshold(orc, g):
    # G:371-426: the returned amplitude collapses and 1 kPa less does not
    # (single 1 MHz tone, R0 = 2 um: one sweep cell, ~10 s of oracle time)
    fr = np.array([[1e6, 1e6, 1e6]])
    n_r0, i = 100, 19
    res = orc.gilmore_threshold(g, fr, n_r0, i, i + 1)
    A = res[i]
    R0 = g.minR0 + i * g.stepR0
    assert 1e5 <= A < 1e7
    assert A == pytest.approx(round(A / 1e3) * 1e3, abs=1e-6)
    assert _collapses(orc, g, R0, fr[0], A)
    assert not _collapses(orc, g, R0, fr[0], A - 1e3)


def test_full_threshold_table_samples_match_oracle(orc):
    # 32 cells of the full 27,775 x 100 threshold table as the GPU computed it
    # (fixture written by tests/golden/make_threshold_samples.py from a
    # `bench.py --threshold-sweep` run): the oracle's search gives the same
    # amplitude, bit for bit
    import json
    from pathlib import Path
    from bubbledynamics_amd import gilmore as gm
    fx = json.loads((Path(__file__).parent / "golden" / "threshold_table_samples.json").read_text())
    p = gm.reference_params()
    p.tableau = 1
    fr = gm.freq_table()
    n_r0 = fx["n_r0"]
    for idx, want in zip(fx["cells"], fx["gpu_threshold_pa"]):
        got = orc.gilmore_threshold(p, fr, n_r0, idx, idx + 1, nthreads=1)[idx]
        assert got == want, (idx, fr[idx // n_r0], got, want)


def test_sweep_fixture_is_the_oracle(orc):
    # tests/golden/sweep_cells.json (the GPU sweep's expected values) is what
    # the oracle computes: the capped setting in full, two cells of the others
    import json
    from pathlib import Path
    fix = json.loads((Path(__file__).parent / "golden" / "sweep_cells.json").read_text())
    fr = np.array(fix["fr_combo"])
    n_r0 = fix["n_r0"]
    for name, cells in (("capped", range(len(fr) * n_r0)), ("dp1980", (4, 21)), ("reference", (9,))):
        s = fix["settings"][name]
        p = orc.gilmore_reference_params()
        p.tableau, p.max_solve_steps = s["tableau"], s["max_solve_steps"]
        p.abs_eps = p.rel_eps = s["eps"]
        p.minR0, p.stepR0 = fix["min_r0"], fix["step_r0"]
        for i in cells:
            got = orc.gilmore_threshold(p, fr, n_r0, i, i + 1, nthreads=1)[i]
            assert got == s["threshold_pa"][i], (name, i)
    cap = fix["settings"]["capped"]["threshold_pa"]
    assert -2.0 in cap and any(v > 0 for v in cap)


def test_capped_solves_give_the_sentinel(orc):
    # a search whose solve meets max_solve_steps undecided returns -2 (the
    # reference has no cap); with a cap the solves never reach, the result
    # is the ordinary threshold
    p = orc.gilmore_reference_params()
    p.tableau = 1
    p.minR0, p.stepR0 = 1e-6, 1e-6
    fr = np.array([[5e6, 5e6, 5e6]])
    free = orc.gilmore_threshold(p, fr, 1, 0, 1, nthreads=1)[0]
    assert free > 0
    p.max_solve_steps = 10
    assert orc.gilmore_threshold(p, fr, 1, 0, 1, nthreads=1)[0] == -2.0


def test_conditioning_fixture_spot_check(orc):
    # tests/golden/gilmore_conditioning.json: a listed config-3 bubble moves by
    # more than 1e-9 under a one-ulp R0 or amplitude change, an unlisted one
    # does not
    import json
    from pathlib import Path
    from bubbledynamics_amd import gilmore as gm
    fix = json.loads((Path(__file__).parent / "golden" / "gilmore_conditioning.json").read_text())
    lst = fix["config3_subset_ulp"]
    idx = gm.config3_subset()
    R0, amp, ph, fr = gm.config3_inputs(1_000_000, 3)
    p = gm.workload_params(arithmetic=gm.BD_GM_ARITH_REFERENCE)
    u = 2.0 ** -52
    worst = lst["positions"][int(np.argmax(lst["rmax_spread"]))]
    calm = next(i for i in range(len(idx)) if i not in set(lst["positions"]))

    def spread(pos):
        i = idx[pos:pos + 1]
        run = lambda r, a: orc.gilmore_integrate(p, gm.BD_GM_DP45, r, a, ph[i], fr, 20e-6, 1e-13, rtol=1e-8,
                                                 atol=1e-12)["Rmax"][0]
        b = run(R0[i], amp[i])
        return max(abs(run(R0[i] * (1 + u), amp[i]) / b - 1), abs(run(R0[i], amp[i] * (1 + u)) / b - 1))
    assert spread(worst) > 1e-9
    assert spread(calm) <= 1e-9


def test_reference_tableau_table_samples_match_oracle(orc):
    # cells of the threshold table with the reference's own tableau and
    # tolerances as the GPU's drop-in sweep computed them (fixture written by
    # tests/golden/make_reference_tableau_samples.py from a bounded
    # `bench.py --threshold-sweep --sweep-reference-tableau` run); two of the
    # eight are recomputed here (about 7 s each on one core): the oracle's
    # search gives the same amplitude, bit for bit
    import json
    from pathlib import Path
    from bubbledynamics_amd import gilmore as gm
    fx = json.loads((Path(__file__).parent / "golden" / "reference_tableau_samples.json").read_text())
    p = gm.reference_params()
    assert p.tableau == 0
    fr = gm.freq_table()
    n_r0 = fx["n_r0"]
    for idx, want in list(zip(fx["cells"], fx["gpu_threshold_pa"]))[::4]:
        got = orc.gilmore_threshold(p, fr, n_r0, idx, idx + 1, nthreads=1)[idx]
        assert got == want, (idx, fr[idx // n_r0], got, want)
