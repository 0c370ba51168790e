"""GPU parity of the bioheat kernel against the FP64 oracle (through the C ABI)."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

from bubbledynamics_amd import _ffi  # noqa: E402
from bubbledynamics_amd import bioheat as bh  # noqa: E402
from bubbledynamics_amd._ffi import BD_F32, BD_F64  # noqa: E402


def dev(a, dtype=torch.float64):
    return torch.as_tensor(np.ascontiguousarray(a), dtype=dtype, device="cuda")


def vp(t):
    return C.c_void_p(t.data_ptr())


def two_layer(n, nz):
    # the reference's water/tissue split (T:79) at mid-depth, coarse grid
    p, _, _ = bh.reference_params()
    p.water_tissue = (nz // 2) * p.dz + 1e-9
    return p


@pytest.mark.parametrize("use_p", [1, 0])
def test_temperature_square_dropin_fp64(orc, use_p):
    rng = np.random.default_rng(5)
    n, nz = 40, 36
    p = two_layer(n, nz)
    shp = (nz, n, n)
    Tn = 300.0 + rng.uniform(0, 20, shp)
    Tp = 300.0 + rng.uniform(0, 20, shp)
    Q = rng.uniform(0, 1e8, shp)
    TD = rng.uniform(0, 10, shp)
    dTn, dTp, dQ, dTD = dev(Tn), dev(Tp), dev(Q), dev(TD)
    _ffi.call("bd_temperature_square", C.byref(p), BD_F64, vp(dTn), vp(dTp), vp(dQ), vp(dTD), n, nz,
              use_p, 3, -2, None)
    torch.cuda.synchronize()
    orc.temperature_square(p, Tn, Tp, Q, TD, n, nz, use_p, 3, -2)
    out_ref = Tn if use_p else Tp
    out = (dTn if use_p else dTp).cpu().numpy()
    np.testing.assert_allclose(out, out_ref, rtol=1e-15, atol=0)
    np.testing.assert_allclose(dTD.cpu().numpy(), TD, rtol=1e-13)


def test_pennes_fp32_offset_kahan_vs_oracle(orc):
    # 300 explicit substeps with perfusion and a focal source moved along 6 dwell points:
    # FP32 fields stored as T - T_art and a compensated dose against the FP64 oracle
    n = 48
    p = bh.pennes_params()
    rng = np.random.default_rng(2)
    z, y, x = np.meshgrid(*(np.arange(n) - n / 2,) * 3, indexing="ij")
    Q = 3e7 * np.exp(-(x * x + y * y) / 8.0 - z * z / 30.0)
    T0 = p.T_art
    sx, sy = bh.spiral_dwell(3, 8 * 2e-4, 2e-4)
    run = bh.BioheatRun(p, n, n, dev(Q, torch.float32), T0, dtype=torch.float32, kahan=True, offset=True)
    plain = bh.BioheatRun(p, n, n, dev(Q, torch.float32), T0, dtype=torch.float32, kahan=False, offset=False)
    A = np.full((n, n, n), T0); B = A.copy(); TD = np.zeros_like(A); TDc = np.zeros_like(A)
    use_p = True
    for k in range(6):
        run.step(int(sx[k]), int(sy[k]), 50)
        plain.step(int(sx[k]), int(sy[k]), 50)
        for _ in range(50):
            orc.temperature_square(p, A, B, Q, TD, n, n, use_p, int(sx[k]), int(sy[k]), TDc=TDc)
            use_p = not use_p
    torch.cuda.synchronize()
    T_ref = A if not use_p else B
    dT = run.temperature().cpu().numpy() - T_ref
    rise = (T_ref - T0).max()
    dose_ref = TD - TDc
    print(f"peak rise {rise:.2f} K; FP32(offset) max |dT| {np.abs(dT).max():.2e} K; "
          f"FP32(kelvin) max |dT| {np.abs(plain.temperature().cpu().numpy() - T_ref).max():.2e} K")
    assert rise > 5.0
    assert np.abs(dT).max() <= 2e-4 * rise
    d = run.dose().cpu().numpy()
    sel = dose_ref > 1e-3 * dose_ref.max()
    rel = np.abs(d[sel] / dose_ref[sel] - 1)
    print(f"dose rel err max {rel.max():.2e}")
    assert rel.max() <= 1e-3


def _random_state(p, n, nz, dtype, kahan, offset, seed):
    rng = np.random.default_rng(seed)
    Q = dev(rng.uniform(0.0, 5e8, (nz, n, n)), dtype)
    run = bh.BioheatRun(p, n, nz, Q, p.T_art, dtype=dtype, kahan=kahan, offset=offset)
    T = dev(p.T_art - run.p.T_off + rng.uniform(-5.0, 30.0, (nz, n, n)), dtype)
    run.T[0].copy_(T)
    run.T[1].copy_(T)                      # both buffers hold the same boundary values
    run.TD.copy_(dev(rng.uniform(0.0, 2.0, (nz, n, n)), dtype))
    if kahan:
        run.TDc.copy_(dev(rng.uniform(-1e-7, 1e-7, (nz, n, n)), dtype))
    return run


@pytest.mark.parametrize("dtype,kahan,offset", [(torch.float32, True, True), (torch.float32, False, False),
                                                (torch.float64, False, False), (torch.float64, True, True)])
@pytest.mark.parametrize("n,nz", [(37, 29), (70, 41), (131, 20)])
def test_fused_substep_pairs_bitwise_equal(monkeypatch, dtype, kahan, offset, n, nz):
    # bd_bioheat_run launches two substeps at a time; the temperatures, dose
    # and compensation must be the bits of the same substeps run one at a time
    # (two-layer medium, perfusion, shifted source, odd substep counts, grids
    # that end inside a 62 x 14 / 62 x 6 output tile and inside a z-chunk)
    p = two_layer(n, nz)
    p.T_art = 310.15
    p.perf = 2e-3
    a = _random_state(p, n, nz, dtype, kahan, offset, 11)
    b = _random_state(p, n, nz, dtype, kahan, offset, 11)
    for sx, sy, m in ((3, -2, 5), (0, 0, 4), (-7, 5, 7), (40, 1, 2)):
        monkeypatch.delenv("BD_BH_FUSE", raising=False)
        a.step(sx, sy, m)
        monkeypatch.setenv("BD_BH_FUSE", "0")
        b.step(sx, sy, m)
    monkeypatch.delenv("BD_BH_FUSE", raising=False)
    torch.cuda.synchronize()
    assert not torch.equal(a.T[a.cur], b.T[1 - b.cur])     # the fields moved
    assert torch.equal(a.T[a.cur], b.T[b.cur])
    assert torch.equal(a.TD, b.TD)
    if kahan:
        assert torch.equal(a.TDc, b.TDc)


@pytest.mark.parametrize("dtype,kahan,variant", [(torch.float32, True, v) for v in range(6)] +
                         [(torch.float32, False, 2), (torch.float64, True, 3), (torch.float64, True, 4),
                          (torch.float64, True, 5), (torch.float64, False, 5)])
def test_every_pair_region_bitwise_equal(monkeypatch, dtype, kahan, variant):
    # each region shape of the pair kernel (64 / 128 / 256 columns, four or
    # two rows per thread) against single substeps: a grid that ends inside a
    # region in x and y, two media, perfusion, shifted source
    n, nz = 131, 29
    p = two_layer(n, nz)
    p.T_art = 310.15
    p.perf = 2e-3
    a = _random_state(p, n, nz, dtype, kahan, kahan, 5)
    b = _random_state(p, n, nz, dtype, kahan, kahan, 5)
    for sx, sy, m in ((2, -3, 6), (-9, 4, 3)):
        monkeypatch.setenv("BD_BH_VARIANT", str(variant))
        monkeypatch.setenv("BD_BH_FUSE", "2")
        a.step(sx, sy, m)
        monkeypatch.delenv("BD_BH_VARIANT", raising=False)
        monkeypatch.setenv("BD_BH_FUSE", "0")
        b.step(sx, sy, m)
    monkeypatch.delenv("BD_BH_FUSE", raising=False)
    torch.cuda.synchronize()
    assert torch.equal(a.T[a.cur], b.T[b.cur])
    assert torch.equal(a.TD, b.TD)
    if kahan:
        assert torch.equal(a.TDc, b.TDc)


@pytest.mark.parametrize("variant", range(2))
@pytest.mark.parametrize("n,nz,kahan", [(131, 29, True), (70, 41, False), (37, 13, True)])
def test_every_triple_region_bitwise_equal(monkeypatch, variant, n, nz, kahan):
    # each region shape of the FP32 triple kernel (64 x 32 and 128 x 16
    # regions, 60 x 28 and 124 x 12 output) against single substeps: grids
    # that end inside a region in x and y and inside a z-chunk, two media,
    # perfusion, shifted source, substep counts that leave a pair or a single
    # substep after the triples
    p = two_layer(n, nz)
    p.T_art = 310.15
    p.perf = 2e-3
    a = _random_state(p, n, nz, torch.float32, kahan, kahan, 7)
    b = _random_state(p, n, nz, torch.float32, kahan, kahan, 7)
    for sx, sy, m in ((2, -3, 9), (-9, 4, 5), (0, 0, 4), (5, 5, 3)):
        monkeypatch.setenv("BD_BH_VARIANT3", str(variant))
        monkeypatch.delenv("BD_BH_FUSE", raising=False)
        a.step(sx, sy, m)
        monkeypatch.delenv("BD_BH_VARIANT3", raising=False)
        monkeypatch.setenv("BD_BH_FUSE", "0")
        b.step(sx, sy, m)
    monkeypatch.delenv("BD_BH_FUSE", raising=False)
    torch.cuda.synchronize()
    assert not torch.equal(a.T[a.cur], a.T[1 - a.cur])
    assert torch.equal(a.T[a.cur], b.T[b.cur])
    assert torch.equal(a.TD, b.TD)
    if kahan:
        assert torch.equal(a.TDc, b.TDc)


@pytest.mark.parametrize("kahan", [True, False])
def test_measured_choice_equals_single_substeps(monkeypatch, kahan):
    # an FP32 call with at least 11 substeps left and no choice kept yet for
    # its grid times a triple of each region shape and a pair (real substeps
    # of the run) and keeps the fastest for the later calls: whatever it picks,
    # the fields are the bits of the same substeps run one at a time
    n, nz = 53, 23
    p = two_layer(n, nz)
    p.T_art = 310.15
    p.perf = 2e-3
    a = _random_state(p, n, nz, torch.float32, kahan, True, 5)
    b = _random_state(p, n, nz, torch.float32, kahan, True, 5)
    for k in ("BD_BH_FUSE", "BD_BH_VARIANT3"):
        monkeypatch.delenv(k, raising=False)
    for sx, sy, m in ((2, -1, 31), (0, 3, 17), (-4, 0, 5)):
        monkeypatch.delenv("BD_BH_FUSE", raising=False)
        a.step(sx, sy, m)
        monkeypatch.setenv("BD_BH_FUSE", "0")
        b.step(sx, sy, m)
    monkeypatch.delenv("BD_BH_FUSE", raising=False)
    torch.cuda.synchronize()
    assert torch.equal(a.T[a.cur], b.T[b.cur])
    assert torch.equal(a.TD, b.TD)
    if kahan:
        assert torch.equal(a.TDc, b.TDc)


@pytest.mark.parametrize("dtype", [torch.float32, torch.float64])
def test_differing_boundary_shells_paired_equals_single(monkeypatch, dtype):
    # the two fields start with different boundary shells: bd_bioheat_run
    # holds the current field's shell fixed, so paired and single substeps
    # give the same temperatures (and the shell of the result is the
    # current field's)
    n, nz = 70, 41
    p = two_layer(n, nz)
    p.T_art = 310.15
    p.perf = 2e-3
    runs = []
    for fuse in ("1", "0"):
        r = _random_state(p, n, nz, dtype, True, True, 9)
        r.T[1].add_(dev(np.random.default_rng(2).uniform(-3.0, 3.0, (nz, n, n)), dtype))
        monkeypatch.setenv("BD_BH_FUSE", fuse)
        for sx, sy, m in ((1, 2, 5), (0, -4, 4)):
            r.step(sx, sy, m)
        runs.append(r)
    monkeypatch.delenv("BD_BH_FUSE", raising=False)
    torch.cuda.synchronize()
    a, b = runs
    ta, tb = a.T[a.cur], b.T[b.cur]
    assert torch.equal(ta, tb)
    assert torch.equal(a.TD, b.TD) and torch.equal(a.TDc, b.TDc)
    shell = torch.ones((nz, n, n), dtype=torch.bool, device=ta.device)
    shell[1:-1, 1:-1, 1:-1] = False
    t0 = _random_state(p, n, nz, dtype, True, True, 9).T[0]
    assert torch.equal(ta[shell], t0[shell])


def test_offset_storage_is_exact_change_of_variables(orc):
    # FP64: storing T - T_art gives the same temperatures as kelvin storage
    n = 24
    p = bh.pennes_params()
    Q = np.zeros((n, n, n)); Q[12, 12, 12] = 5e8
    a = bh.BioheatRun(p, n, n, dev(Q), p.T_art + 1.0, dtype=torch.float64, kahan=False, offset=True)
    b = bh.BioheatRun(p, n, n, dev(Q), p.T_art + 1.0, dtype=torch.float64, kahan=False, offset=False)
    a.step(1, 1, 40)
    b.step(1, 1, 40)
    torch.cuda.synchronize()
    np.testing.assert_allclose(a.temperature().cpu().numpy(), b.temperature().cpu().numpy(), rtol=1e-13)


def test_unstable_step_rejected():
    p = bh.pennes_params(substeps=1)
    t = torch.zeros(8 ** 3, dtype=torch.float32, device="cuda")
    rc = _ffi.load().bd_bioheat_step(C.byref(p), BD_F32, vp(t), vp(t.clone()), vp(t), vp(t), None, 8, 8, 0, 0, None)
    assert rc == _ffi.BD_ERR_UNSTABLE


@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
def test_slice_writer_reference_format(tmp_path, dtype):
    # T:412-447 layout: header, rows of '%g;' values; XZ at fixed iy, XY at fixed iz
    from bubbledynamics_amd._ffi import BD_SLICE_XY, BD_SLICE_XZ
    n, nz = 12, 9
    f = (torch.arange(nz * n * n, dtype=torch.float64) / 7.0 + 300.0).reshape(nz, n, n)
    d = f.to(dtype).to("cuda")
    host = d.double().cpu().numpy()
    code = BD_F64 if dtype == torch.float64 else BD_F32
    st = C.c_void_p(torch.cuda.current_stream().cuda_stream)
    for plane, fixed, hdr in ((BD_SLICE_XZ, 5, "Z=rows;X=columns"), (BD_SLICE_XY, 4, "Y=rows;X=columns")):
        path = tmp_path / f"s{plane}.dat"
        _ffi.call("bd_write_slice", str(path).encode(), code, vp(d), n, nz, plane, fixed, 2, 7, 3, 9, 0.5, st)
        lines = path.read_text().split("\n")
        assert lines[0] == hdr and len(lines) == 1 + 5
        for k, r in enumerate(range(2, 7)):
            row = host[r, fixed, 3:9] if plane == BD_SLICE_XZ else host[fixed, r, 3:9]
            assert lines[1 + k] == "".join(f"{v + 0.5:g};" for v in row)
    rc = _ffi.load().bd_write_slice(str(tmp_path / "x.dat").encode(), code, vp(d), n, nz, BD_SLICE_XY, nz, 0, 1, 0, 1,
                                    0.0, st)
    assert rc == _ffi.BD_ERR_INVALID
