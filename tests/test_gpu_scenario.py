"""The reference's own heat-source pipeline (W:556-790) at a reduced size:
settle, water heat source over one period, restart from the kept state,
harmonic bins of the tissue slab, Q, Q3D file -- against the same sequence
run with the FP64 oracle."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

from bubbledynamics_amd import workloads as wl  # noqa: E402


def test_reference_scenario_matches_oracle(orc, tmp_path):
    n, nz, spp, settle, periods, n_harm = 24, 44, 35, 2, 2, 10
    path = tmp_path / "Q3D.bin"
    Q, p, info = wl.reference_scenario(n, nz, settle_periods=settle, periods=periods, n_harm=n_harm,
                                       steps_per_period=spp, dtype=torch.float64, q3d_path=path)
    n_settle, n_time, iz_t = info["settle"], info["harmonics"], info["iz_tissue"]
    assert 2 < iz_t < nz - 4 and iz_t == orc.first_tissue_plane(p, nz)
    st = np.zeros((12, nz, n, n))
    orc.wv_run(p, n, nz, 0, n_settle, st)
    snap = st.copy()
    Qr = np.zeros((nz, n, n))
    for k in range(spp):
        orc.wv_run(p, n, nz, n_settle + k, 1, st)
        m = n_settle + k + 1
        orc.acoustic_power_water(p, Qr, st[m % 6], st[(m - 2) % 6], n, nz, False)
    orc.acoustic_power_water(p, Qr, st[m % 6], st[(m - 2) % 6], n, nz, True)
    st = snap
    H = np.zeros((2 * n_harm, nz - iz_t, n, n))
    for i in range(n_time):
        orc.wv_run(p, n, nz, n_settle + i, 1, st)
        orc.harmonics_accumulate(H, st[(n_settle + i + 1) % 6], n, nz, iz_t, i, n_time, periods, n_harm)
    orc.harmonics_q(p, Qr, H, n, nz, iz_t, n_time, n_harm)
    got = Q.cpu().numpy()
    water, tissue = Qr[:iz_t], Qr[iz_t:]
    assert np.abs(water).max() > 0 and np.abs(tissue).max() > 0
    np.testing.assert_allclose(got[:iz_t], water, rtol=1e-10, atol=1e-10 * np.abs(water).max())
    np.testing.assert_allclose(got[iz_t:], tissue, rtol=1e-10, atol=1e-10 * np.abs(tissue).max())
    np.testing.assert_array_equal(np.fromfile(path, dtype="<f8"), got[2:].reshape(-1))


def test_reference_scenario_progress_reports_every_stage():
    # the progress hook sees the settle chunks and the end of each stage
    seen = []
    wl.reference_scenario(24, 44, settle_periods=2, periods=1, n_harm=2, steps_per_period=35,
                          dtype=torch.float64, progress=lambda stage, step: seen.append((stage, step)), chunk=30)
    assert seen == [("settle", 30), ("settle", 60), ("settle", 70), ("water", 35), ("harmonics", 35)]


def test_coupled_workload_drivers():
    # the coupled Westervelt -> Pennes drivers at 48^3: a focal heat source
    # (FP32 within 1e-3 of FP64 in relative L2) and a temperature rise and dose
    # where it sits
    n = 48
    Q32, p, info = wl.heat_source(n, periods_settle=6, periods_window=2, n_harm=3, dtype=torch.float32)
    Q64, _, _ = wl.heat_source(n, periods_settle=6, periods_window=2, n_harm=3, dtype=torch.float64)
    assert info["steps"] == 8 * 25
    q32, q64 = Q32.double().cpu().numpy(), Q64.cpu().numpy()
    assert np.isfinite(q64).all() and q64.min() >= 0.0 and q64.max() > 0.0
    assert np.linalg.norm(q32 - q64) / np.linalg.norm(q64) <= 1e-3
    k = np.unravel_index(np.argmax(q64), q64.shape)
    assert abs(k[1] - (n - 1) / 2) <= 2 and abs(k[2] - (n - 1) / 2) <= 2     # on the axis
    run, secs, nsub = wl.pennes_spiral(Q32, n, dwell_s=0.1, n_side=3, side=4 * 2e-4)
    T = run.temperature().cpu().numpy()
    assert nsub == 9 * 4
    assert T.max() > run.p.T_art and np.isfinite(T).all()
    d = run.dose().cpu().numpy()
    assert np.unravel_index(np.argmax(T), T.shape)[0] == pytest.approx(k[0], abs=4)
    assert d.max() > 0.0
