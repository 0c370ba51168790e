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
