"""Pins of the FP64 Westervelt oracle against closed forms (CPU only).

Each test is chosen so that a plausible transcription mistake in the oracle
(a wrong stencil weight, sign, index or time level) fails it.
"""
import math

import numpy as np
import pytest

from bubbledynamics_amd import presets
from bubbledynamics_amd._ffi import BD_BC_ZERO, WesterveltParams


def lossless(n=24, c=1500.0, dx=1e-4, cfl=0.3, full=1):
    p = WesterveltParams()
    p.rho = p.rho_tissue = 1000.0
    p.c_inf = p.c_inf_tissue = c
    p.water_tissue = 1e9          # all planes in the non-relaxing branch
    p.tau1 = p.tau2 = 1e-9
    p.dxy = p.dz = dx
    p.dt = cfl * dx / c
    p.full_domain = full
    p.xy_min = 0.0
    p.z_min = -2 * dx
    p.boundary = BD_BC_ZERO
    p.fr = 1e6
    p.d = 1.0            # cap of zero aperture and p0 = 0: no source
    return p


def fields(n, nz=None, k=12):
    nz = nz or n
    return [np.zeros((nz, n, n)) for _ in range(k)]


def step(orc, p, n, levels, nz=None, R=None):
    nz = nz or n
    Pn = np.zeros((nz, n, n))
    R = R or [np.zeros((nz, n, n)) for _ in range(6)]
    orc.wv_space6(p, n, nz, Pn, *levels, R[0], R[1], R[2], R[3], R[4], R[5])
    return Pn, R


def interior(a, lo=3):
    return a[3:-3, lo:-3, lo:-3]


@pytest.mark.parametrize("deg", [2, 4, 6, 7])
def test_laplacian_exact_on_polynomials(orc, deg):
    # L (W:203-205) is the 6th-order central second difference: exact up to degree 7
    n = 20
    p = lossless(n)
    idx = np.arange(n, dtype=np.float64) - 9.5
    z, y, x = np.meshgrid(idx, idx, idx, indexing="ij")
    f = 1e-3 * (x ** deg + 0.5 * y ** deg - 0.25 * z ** deg)
    Pn, _ = step(orc, p, n, [f] * 5)
    dd = deg * (deg - 1) * 1e-3
    lap = dd * (x ** (deg - 2) + 0.5 * y ** (deg - 2) - 0.25 * z ** (deg - 2))  # per unit h^2
    kap = (p.c_inf * p.dt / p.dxy) ** 2
    np.testing.assert_allclose(interior(Pn - f), interior(kap * lap), rtol=1e-9, atol=1e-9 * np.abs(f).max())


def test_laplacian_not_exact_at_degree_8(orc):
    n = 20
    p = lossless(n)
    idx = np.arange(n, dtype=np.float64) - 9.5
    z, y, x = np.meshgrid(idx, idx, idx, indexing="ij")
    f = 1e-6 * x ** 8
    Pn, _ = step(orc, p, n, [f] * 5)
    kap = (p.c_inf * p.dt / p.dxy) ** 2
    err = interior(Pn - f) - interior(kap * 56e-6 * x ** 6)
    # leading error term h^6 f^(8) * (-2/560)... non-zero: the stencil is 6th order, not higher
    assert np.abs(err).max() > 1e-9


def test_laplacian_nyquist_symbol(orc):
    n = 20
    p = lossless(n)
    i = np.arange(n)
    f = np.broadcast_to(((-1.0) ** i)[None, None, :], (n, n, n)).copy()
    Pn, _ = step(orc, p, n, [f] * 5)
    kap = (p.c_inf * p.dt / p.dxy) ** 2
    np.testing.assert_allclose(interior(Pn - f), interior(kap * (-1088.0 / 180.0) * f), rtol=1e-12)


def test_plane_wave_discrete_dispersion(orc):
    # leapfrog + L6: cos(k z - w n dt) is an exact discrete solution when
    # sin^2(w dt/2) = (c dt/h)^2 |sigma(kh)| / 4, sigma(t) = (-490+540cos t-54cos 2t+4cos 3t)/180
    n = 24
    p = lossless(n, full=0)
    kh = 2 * math.pi / 9.0
    sig = (-490 + 540 * math.cos(kh) - 54 * math.cos(2 * kh) + 4 * math.cos(3 * kh)) / 180
    cfl = p.c_inf * p.dt / p.dz
    wdt = 2 * math.asin(math.sqrt(cfl ** 2 * abs(sig) / 4))
    zi = np.arange(n, dtype=np.float64)[:, None, None] * np.ones((1, n, n))
    lev = [np.cos(kh * zi + wdt * m) for m in range(5)]  # p^{n-m}
    Pn, _ = step(orc, p, n, lev)
    expect = np.cos(kh * zi - wdt)
    np.testing.assert_allclose(interior(Pn, 0), interior(expect, 0), atol=1e-12)


def test_time_differences_and_nonlinear_term(orc):
    # Uniform field, history from p(t) = a + b t + g t^2 + e t^3: D3 = 2 dt^3 p''' (exact for cubics),
    # N1 + N2 = dt^2 (p^2)''/2 exact for quadratics (W:207-209)
    n = 16
    p = lossless(n)
    p.delta = 2e-6
    dt = p.dt
    a, b, g, e = 1.0e5, 3.0e11, -2.0e18, 0.0
    tn = 5 * dt
    times = [tn - m * dt for m in range(5)]
    val = [a + b * t + g * t * t + e * t ** 3 for t in times]
    lev = [np.full((n, n, n), v) for v in val]
    Pn, _ = step(orc, p, n, lev)
    base = 2 * val[0] - val[1]
    dlt = p.delta / (2 * dt * p.c_inf ** 2)
    D3 = 0.0  # quadratic: p''' = 0
    pt = b + 2 * g * tn
    ptt = 2 * g
    expect_nl_zero = base + dlt * D3
    np.testing.assert_allclose(interior(Pn), expect_nl_zero, rtol=1e-12)
    # now switch beta on
    p.beta = 3.5
    Pn, _ = step(orc, p, n, lev)
    bet = 2 * p.beta / (p.rho * p.c_inf ** 2)
    expect = base + bet * dt * dt * (pt * pt + val[0] * ptt)
    np.testing.assert_allclose(interior(Pn), expect, rtol=1e-11)
    # cubic history: D3 = 2 dt^3 * 6 e, the delta term alone
    p.beta = 0.0
    e = 4.0e24
    val = [a + b * t + g * t * t + e * t ** 3 for t in times]
    lev = [np.full((n, n, n), v) for v in val]
    Pn, _ = step(orc, p, n, lev)
    expect = 2 * val[0] - val[1] + dlt * (2 * dt ** 3 * 6 * e)
    np.testing.assert_allclose(interior(Pn), expect, rtol=1e-11)


def test_relaxation_fixed_point(orc):
    # W:226-228: constant D3 drives R_nu to R* = 2 c_nu tau_nu (D3/(2 dt^3)) / c^3
    n = 12
    p = lossless(n)
    p.water_tissue = -1e9        # every plane relaxes
    p.c1, p.tau1, p.c2, p.tau2 = 1.5, 1e-9, 3.0, 1e-8
    dt = p.dt
    e = 1.0e24
    times = [5 * dt - m * dt for m in range(5)]
    lev = [np.full((n, n, n), e * t ** 3) for t in times]
    R = [np.zeros((n, n, n)) for _ in range(6)]
    D3_2dt = 6 * e
    for _ in range(400):
        Pn = np.zeros((n, n, n))
        r1, r1m1, r1m2, r2, r2m1, r2m2 = R
        orc.wv_space6(p, n, n, Pn, *lev, r1, r1m1, r1m2, r2, r2m1, r2m2)
        R = [r1m2, r1, r1m1, r2m2, r2, r2m1]   # rotate: new m1 = r1, new m2 = old m1
    r1 = R[1][5, 5, 5]
    r2 = R[4][5, 5, 5]
    assert r1 == pytest.approx(2 * p.c1 * p.tau1 * D3_2dt / p.c_inf_tissue ** 3, rel=1e-10)
    assert r2 == pytest.approx(2 * p.c2 * p.tau2 * D3_2dt / p.c_inf_tissue ** 3, rel=1e-6)


def test_energy_invariant_conserved(orc):
    # lossless homogeneous leapfrog: ||p^{n+1}-p^n||^2 - <p^{n+1}, (c dt)^2 Lh p^n> is constant
    n = 40
    p = lossless(n)
    c = np.arange(n) - (n - 1) / 2
    z, y, x = np.meshgrid(c, c, c, indexing="ij")
    r2 = (x * x + y * y + z * z) / 36.0
    bump = np.where(r2 < 1, (1 - r2) ** 4, 0.0) * 1e5
    state = np.zeros((12, n, n, n))
    state[0] = bump          # p^0
    state[5] = bump          # p^-1
    state[4] = bump; state[3] = bump; state[2] = bump
    energy = np.zeros(2 * 30)
    orc.wv_run(p, n, n, 0, 30, state, energy=energy)
    e1 = energy[1::2]
    assert np.all(np.isfinite(e1))
    assert np.abs(e1 - e1[0]).max() <= 1e-11 * abs(e1[0])


def test_source_delays_focus_in_phase(orc):
    # W:78-82: a ring at radius r that fires at t_e = t_a - sqrt(d^2+r^2)/c reaches
    # the focus at t_a; every ring must then carry the same phase omega (t_a - d/c)
    p, nxy, nz = presets.reference_westervelt()
    n = 440
    t_a = 130.0e-6
    xs = p.xy_min + np.arange(n) * p.dxy
    rmin, rmax = p.d * math.tan(p.alpha_hole), p.d * math.tan(p.alpha)
    seen = []
    for ix, iy in [(104, 0), (150, 30), (200, 100), (300, 140), (420, 50), (60, 80)]:
        r = math.hypot(xs[ix], xs[iy])
        t_e = t_a - math.sqrt(p.d ** 2 + r * r) / p.c_inf
        P = np.zeros((3, n, n))
        orc.wv_source(p, P, t_e, n, 3)
        v = P[2, iy, ix]
        if rmin <= r <= rmax:
            var = math.sqrt(1 + (r / p.d) ** 2)
            seen.append(v * var / p.p0)
        else:
            assert v == 0.0
    assert len(seen) >= 4
    np.testing.assert_allclose(seen, math.sin(2 * math.pi * p.fr * (t_a - p.d / p.c_inf)), atol=1e-9)


def test_reference_preset_derived_constants(orc):
    # numbers derived from the reference's own host code (W:356-417)
    p, nxy, nz = presets.reference_westervelt()
    assert (nxy, nz) == (426, 2003)
    assert p.p0 == pytest.approx(117.9e3, rel=1e-3)
    # W:372 evaluates to 4.445e-6 (frequency-independent for alpha ~ f^2); its comment
    # "about 4.6257e-6" belongs to another parameter set
    assert p.delta == pytest.approx(4.4455e-6, rel=1e-4)
    assert p.d * math.tan(p.alpha) == pytest.approx(87.37e-3, rel=1e-3)
    assert p.d * math.tan(p.alpha_hole) == pytest.approx(20.69e-3, rel=1e-3)
    assert orc.first_tissue_plane(p, nz) == 1202         # W:673
    assert p.dt == pytest.approx(7.981e-9, rel=1e-3)


def test_config_water_delta():
    p = presets.tone_burst_192()
    assert p.delta == pytest.approx(4.445e-6, rel=1e-3)
    assert p.dt == pytest.approx(3e-8, rel=1e-12)


def test_mur_face_low_reflection(orc):
    # a beam travelling +z meets the z-max Mur face (W:108-113); the energy left
    # after the pulse has passed the face must be a small fraction of the incident
    n, nz = 24, 72
    p = lossless(n, full=0)
    p.boundary = 0
    zc = np.arange(nz)[:, None, None]
    yy = np.arange(n)[None, :, None]
    xx = np.arange(n)[None, None, :]
    kz = 2 * math.pi / 10
    env = lambda zz: np.exp(-((zz - 40) / 5.0) ** 2)
    beam = np.exp(-(xx ** 2 + yy ** 2) / 60.0)
    cfl = p.c_inf * p.dt / p.dz
    state = np.zeros((12, nz, n, n))
    for m in range(5):
        zz = zc + m * cfl          # earlier levels sit further back
        state[(-m) % 6] = env(zz) * np.cos(kz * zz) * beam
    energy = np.zeros(2 * 400)
    orc.wv_run(p, n, nz, 0, 400, state, energy=energy)
    e0 = energy[0::2]
    # incident energy (first steps) vs energy after the pulse has left (last steps)
    assert e0[-1] < 0.02 * e0[:20].max()


def test_water_power_plane_wave(orc):
    # W:242-265: one period of p = A sin(w t) gives Q = 2 alpha I, I = A^2/(2 rho c),
    # up to the central-difference factor (sin(w dt)/(w dt))^2
    p, nxy, nz = presets.reference_westervelt()
    n, nzz = 4, 4
    p.water_tissue = 1.0
    A = 2.0e5
    w = 2 * math.pi * p.fr
    Q = np.zeros((nzz, n, n))
    steps = 100
    for j in range(steps):
        t = (j + 1) * p.dt
        Pn = np.full((nzz, n, n), A * math.sin(w * t))
        Pm1 = np.full((nzz, n, n), A * math.sin(w * (t - 2 * p.dt)))
        orc.acoustic_power_water(p, Q, Pn, Pm1, n, nzz, 0)
    orc.acoustic_power_water(p, Q, Pn, Pm1, n, nzz, 1)
    I = A * A / (2 * p.rho * p.c_inf)
    fac = (math.sin(w * p.dt) / (w * p.dt)) ** 2
    np.testing.assert_allclose(Q, 2 * p.abs_water * I * fac, rtol=1e-9)


def test_harmonics_match_numpy_fft(orc):
    # W:324-345 with the FFT bins replaced by running DFT accumulators
    p, _, _ = presets.reference_westervelt()
    n, nz, izt = 3, 5, 2
    N_time, periods, n_harm = 400, 4, 10
    rng = np.random.default_rng(0)
    amps = rng.uniform(0, 1e5, size=(n_harm, nz - izt, n, n))
    phs = rng.uniform(0, 2 * math.pi, size=(n_harm, nz - izt, n, n))
    H = np.zeros((2 * n_harm, nz - izt, n, n))
    series = np.zeros((N_time, nz - izt, n, n))
    for j in range(N_time):
        v = sum(amps[k] * np.cos(2 * math.pi * (k + 1) * periods * j / N_time + phs[k]) for k in range(n_harm))
        series[j] = v
        full = np.zeros((nz, n, n))
        full[izt:] = v
        orc.harmonics_accumulate(H, full, n, nz, izt, j, N_time, periods, n_harm)
    X = np.fft.fft(series, axis=0)
    for k in range(n_harm):
        b = (k + 1) * periods
        np.testing.assert_allclose(H[2 * k], X[b].real, rtol=1e-9, atol=1e-3)
        np.testing.assert_allclose(H[2 * k + 1], X[b].imag, rtol=1e-9, atol=1e-3)
        # |X_k|/N = A/2 for a pure cosine
        np.testing.assert_allclose(np.abs(X[b]) / N_time, amps[k] / 2, rtol=1e-9)
    Q = np.zeros((nz, n, n))
    orc.harmonics_q(p, Q, H, n, nz, izt, N_time, n_harm, 7.2)
    expect = sum(2 * (7.2 * (k + 1) * p.fr / 1e6) * 2 * (amps[k] / 2) ** 2 / (p.rho_tissue * p.c_inf_tissue)
                 for k in range(n_harm))
    np.testing.assert_allclose(Q[izt:], expect, rtol=1e-9)


def _nyquist_companion(p):
    """One step of the linear scheme (beta = 0) for the grid's highest mode
    (-1)^(x+y+z), whose 6th-order symbol is -1088/180 per axis, as a 9 x 9
    matrix on (p^n..p^{n-4}, R1^{n-1}, R1^{n-2}, R2^{n-1}, R2^{n-2}): the
    leapfrog, D3 (W:207) and the relaxation recurrence (W:226-230) written from
    the parameters, independently of the oracle's code."""
    dt, c = p.dt, p.c_inf_tissue
    kL = ((c * dt / p.dxy) ** 2 * 2 + (c * dt / p.dz) ** 2) * (-1088.0 / 180.0)
    dlt = p.delta_tissue / (2.0 * dt * c * c)
    D3 = np.array([5.0, -18.0, 24.0, -14.0, 3.0, 0, 0, 0, 0])
    rows = []
    for cv, tv, k in ((p.c1, p.tau1, 5), (p.c2, p.tau2, 7)):
        hist = np.zeros(9)
        hist[k], hist[k + 1] = 4.0, -1.0
        rows.append((2.0 * dt * 2.0 * cv * tv / ((2.0 * dt + 3.0 * tv) * c ** 3)) * D3 / (2.0 * dt ** 3)
                    + tv / (2.0 * dt + 3.0 * tv) * hist)
    M = np.zeros((9, 9))
    M[0] = np.array([2.0 + kL, -1.0, 0, 0, 0, 0, 0, 0, 0]) + dlt * D3 + (c * dt) ** 2 * (rows[0] + rows[1])
    M[1, 0] = M[2, 1] = M[3, 2] = M[4, 3] = 1.0
    M[5], M[6, 5], M[7], M[8, 7] = rows[0], 1.0, rows[1], 1.0
    return M


def test_relaxation_scheme_amplifies_the_grid_nyquist_mode(orc):
    # The relaxation discretisation of W:226-230 at the 192^3 tone-burst
    # config (CFL 0.3, dt = 30 tau_1 = 3 tau_2) amplifies the grid's highest
    # mode by 1.00103 per step, so content the burst leaves there grows ~6e5-fold
    # in 13,000 steps (both precisions end non-finite after ~14,000 steps;
    # the configured run is 3,000); without relaxation the scheme is neutral.
    # The oracle's one-step map on that mode equals the companion matrix.
    n, nz = 16, 16
    p = presets.tone_burst_192(n)
    p.beta = p.beta_tissue = 0.0
    p.sponge_cells = 0
    M = _nyquist_companion(p)
    ev, V = np.linalg.eig(M)
    i = int(np.argmax(np.abs(ev)))
    assert abs(abs(ev[i]) - 1.001025) < 2e-6
    q = p.copy()
    q.c1 = q.c2 = 0.0
    assert np.abs(np.linalg.eigvals(_nyquist_companion(q))).max() < 1.0 + 1e-9
    v = np.real(V[:, i])
    v /= np.abs(v).max()
    z, y, x = np.meshgrid(np.arange(nz), np.arange(n), np.arange(n), indexing="ij")
    s = np.where((x + y + z) % 2 == 0, 1.0, -1.0)
    lv = [v[k] * s for k in range(5)]
    R1m1, R1m2, R2m1, R2m2 = (v[k] * s for k in (5, 6, 7, 8))
    Pn, R1, R2 = np.zeros((nz, n, n)), np.zeros((nz, n, n)), np.zeros((nz, n, n))
    orc.wv_space6(p, n, nz, Pn, *lv, R1, R1m1, R1m2, R2, R2m1, R2m2)
    w = M @ v
    sl = (slice(3, nz - 3), slice(0, n - 3), slice(0, n - 3))
    for got, k in ((Pn, 0), (R1, 5), (R2, 7)):
        np.testing.assert_allclose(got[sl], (w[k] * s)[sl], rtol=1e-11, atol=1e-11 * abs(w[k]))
