"""Adjoint phased-array gradient against FP64 central finite differences."""
import math

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

from bubbledynamics_amd import presets  # noqa: E402
from bubbledynamics_amd.adjoint import adjoint_gradient, objective  # noqa: E402
from bubbledynamics_amd.gilmore import seeded_uniforms  # noqa: E402


def setup(n=40, steps=120, n_el=4, seed=4):
    p = presets.phased_array(n, n_el=n_el, dx=2e-4)
    u = seeded_uniforms(seed, 2 * n_el * n_el)
    ph = 2 * math.pi * u[: n_el * n_el]
    amp = 0.5e6 + 0.5e6 * u[n_el * n_el:]
    c = (n // 2, n // 2, n // 2)
    obj = objective(n, n, steps, c, 2.5)
    return p, amp, ph, obj


def test_adjoint_matches_central_fd_fp64():
    from autograd_adjoint import WesterveltObjective
    n, steps = 40, 120
    p, amp, ph, obj = setup(n, steps)
    a = torch.tensor(amp, requires_grad=True)
    f = torch.tensor(ph, requires_grad=True)
    J = WesterveltObjective.apply(a, f, p, n, steps, obj, 30)
    J.backward()
    gA, gP = a.grad.numpy(), f.grad.numpy()
    sel = np.argsort(seeded_uniforms(40, len(amp)))[:8]      # 8 elements, seed 40
    eps = 1e-6
    fdA, fdP = [], []
    for e in sel:
        ap = amp.copy(); am = amp.copy()
        ap[e] *= 1 + eps; am[e] *= 1 - eps
        Jp = adjoint_gradient(p, n, steps, ap, ph, obj, 1000, torch.float64, gradient=False)[0]
        Jm = adjoint_gradient(p, n, steps, am, ph, obj, 1000, torch.float64, gradient=False)[0]
        fdA.append((Jp - Jm) / (2 * eps * amp[e]))
        pp = ph.copy(); pm = ph.copy()
        pp[e] += eps; pm[e] -= eps
        Jp = adjoint_gradient(p, n, steps, amp, pp, obj, 1000, torch.float64, gradient=False)[0]
        Jm = adjoint_gradient(p, n, steps, amp, pm, obj, 1000, torch.float64, gradient=False)[0]
        fdP.append((Jp - Jm) / (2 * eps))
    fdA, fdP = np.array(fdA), np.array(fdP)
    errA = np.abs(gA[sel] - fdA) / np.abs(fdA).max()
    errP = np.abs(gP[sel] - fdP) / np.abs(fdP).max()
    print(f"J={J.item():.6e}  max rel err dA {errA.max():.2e}  dphi {errP.max():.2e}")
    print("adjoint dA", gA[sel], "\nFD      dA", fdA)
    assert np.abs(fdA).max() > 0 and np.abs(fdP).max() > 0
    assert errA.max() <= 1e-5
    assert errP.max() <= 1e-5


def test_checkpoint_interval_does_not_change_gradient():
    n, steps = 32, 80
    p, amp, ph, obj = setup(n, steps)
    ref = adjoint_gradient(p, n, steps, amp, ph, obj, steps)
    for K in (1, 7, 33):
        g = adjoint_gradient(p, n, steps, amp, ph, obj, K)
        assert g[0] == ref[0]
        np.testing.assert_array_equal(g[1], ref[1])
        np.testing.assert_array_equal(g[2], ref[2])


def test_fp32_adjoint_close_to_fp64():
    n, steps = 40, 120
    p, amp, ph, obj = setup(n, steps)
    J64, gA64, gP64, _ = adjoint_gradient(p, n, steps, amp, ph, obj, 30, torch.float64)
    J32, gA32, gP32, _ = adjoint_gradient(p, n, steps, amp, ph, obj, 30, torch.float32)
    eA = np.abs(gA32 - gA64).max() / np.abs(gA64).max()
    eP = np.abs(gP32 - gP64).max() / np.abs(gP64).max()
    print(f"FP32 vs FP64: J {abs(J32 / J64 - 1):.2e}, dA {eA:.2e}, dphi {eP:.2e}")
    assert abs(J32 / J64 - 1) < 1e-4
    assert eA < 1e-3 and eP < 1e-3


@pytest.mark.parametrize("tile", range(5))
def test_every_adjoint_tile_same_gradient(monkeypatch, tile):
    # the backward kernel's tile shapes (64x4, 64x8, 32x8, 128x4, 128x8,
    # selected with BD_ADJ_TILE) give the default tile's gradient up to the
    # summation order of the per-workgroup partials
    n, steps = 40, 80
    p, amp, ph, obj = setup(n, steps)
    monkeypatch.delenv("BD_ADJ_TILE", raising=False)
    J0, gA0, gP0, _ = adjoint_gradient(p, n, steps, amp, ph, obj, 20, torch.float64)
    monkeypatch.setenv("BD_ADJ_TILE", str(tile))
    J1, gA1, gP1, _ = adjoint_gradient(p, n, steps, amp, ph, obj, 20, torch.float64)
    assert J1 == J0
    assert np.abs(gA1 - gA0).max() <= 1e-12 * np.abs(gA0).max()
    assert np.abs(gP1 - gP0).max() <= 1e-12 * np.abs(gP0).max()


@pytest.mark.parametrize("env", [{"BD_ADJ_ZB": "0"}, {"BD_ADJ_ZCHUNK": "7"}, {"BD_ADJ_ZCHUNK": "40"},
                                 {"BD_ADJ_TILE": "1", "BD_ADJ_ZB": "0"}])
def test_border_stores_and_z_chunks_do_not_change_gradient(monkeypatch, env):
    # the zero stores of the border columns write the zeros those cells hold,
    # and the z-chunk length only moves the per-workgroup boundaries of the
    # backward march: the gradient is bit-identical (n = 67 leaves a partly
    # filled last tile in x; 7-plane chunks a short last chunk)
    n, steps = 67, 60
    p, amp, ph, obj = setup(n, steps)
    for k in ("BD_ADJ_ZB", "BD_ADJ_ZCHUNK", "BD_ADJ_TILE"):
        monkeypatch.delenv(k, raising=False)
    J0, gA0, gP0, _ = adjoint_gradient(p, n, steps, amp, ph, obj, 20, torch.float32)
    for k, v in env.items():
        monkeypatch.setenv(k, v)
    J1, gA1, gP1, _ = adjoint_gradient(p, n, steps, amp, ph, obj, 20, torch.float32)
    assert J1 == J0
    np.testing.assert_array_equal(gA1, gA0)
    np.testing.assert_array_equal(gP1, gP0)
