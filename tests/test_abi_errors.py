"""Error contract of the C ABI, exercised on the host: every case here is
rejected by argument checks that run before any HIP call or pointer use, so
the library answers with its status code on a machine without a GPU."""
import ctypes as C

import numpy as np
import pytest

from bubbledynamics_amd import _ffi, presets
from bubbledynamics_amd import bioheat as bh
from bubbledynamics_amd._ffi import (BD_ERR_INVALID, BD_ERR_IO, BD_ERR_UNSTABLE, BD_ERR_UNSUPPORTED, BD_F32,
                                     BD_F64, BD_SLICE_XZ, BD_SRC_ARRAY)

NULL = None


@pytest.fixture(scope="module")
def lib():
    return _ffi.load()


def test_westervelt_parameter_checks(lib):
    p = presets.reference_scaled(40, 48)
    nulls = [NULL] * 12
    assert lib.bd_westervelt_space6(C.byref(p), 7, *nulls, 40, 48, NULL) == BD_ERR_INVALID       # dtype
    assert lib.bd_westervelt_space6(C.byref(p), BD_F32, *nulls, 9, 48, NULL) == BD_ERR_INVALID   # grid
    assert lib.bd_westervelt_space6(C.byref(p), BD_F32, *nulls, 40, 48, NULL) == BD_ERR_INVALID  # pointers
    assert lib.bd_westervelt_space6(None, BD_F32, *nulls, 40, 48, NULL) == BD_ERR_INVALID
    q = p.copy()
    q.dt *= 10                                   # far beyond the leapfrog limit
    assert lib.bd_westervelt_space6(C.byref(q), BD_F64, *nulls, 40, 48, NULL) == BD_ERR_UNSTABLE
    q = p.copy()
    q.boundary = 5
    assert lib.bd_westervelt_space6(C.byref(q), BD_F64, *nulls, 40, 48, NULL) == BD_ERR_INVALID
    q = p.copy()
    q.sponge_cells = 300
    assert lib.bd_westervelt_space6(C.byref(q), BD_F64, *nulls, 40, 48, NULL) == BD_ERR_INVALID
    # the array source needs element amplitudes and phases
    q = p.copy()
    q.source = BD_SRC_ARRAY
    assert lib.bd_boundary_condition_source(C.byref(q), BD_F64, NULL, 0.0, 40, 48, NULL, NULL, NULL) == BD_ERR_INVALID
    assert lib.bd_boundary_condition_space6(C.byref(p), BD_F64, NULL, NULL, 0.0, 40, 48, NULL) == BD_ERR_INVALID
    assert lib.bd_acoustic_power_water(C.byref(p), BD_F64, NULL, NULL, NULL, 40, 48, 0, NULL) == BD_ERR_INVALID


def test_westervelt_plane_size_limit(lib):
    # the step kernel forms 32-bit byte offsets over 4 planes from a plane's
    # descriptor: 4 * Nxy^2 * sizeof(T) must stay below 4 GiB
    p = presets.reference_scaled(40, 48)
    plan = C.c_void_p()
    for dtype, nxy in ((BD_F32, 16384), (BD_F64, 11586)):
        assert 4 * nxy * nxy * (4 if dtype == BD_F32 else 8) > 0xFFFFFFFF
        q = p.copy()
        q.dt = p.dt * 40 / nxy       # keep the step stable at the finer spacing
        q.dxy = p.dxy * 40 / nxy
        assert lib.bd_westervelt_plan_create(C.byref(q), dtype, nxy, 48, 10, -1, -1, -1, C.byref(plan)) \
            == BD_ERR_INVALID


def test_harmonic_bin_checks(lib):
    h = (C.c_double * 4)()
    args = dict(dtype=BD_F64, H=C.cast(h, C.c_void_p), P=C.cast(h, C.c_void_p))

    def call(nxy=8, nz=8, izt=2, i_time=0, n_time=4, periods=1, n_harm=2):
        return lib.bd_harmonics_accumulate(args["dtype"], args["H"], args["P"], nxy, nz, izt, i_time, n_time,
                                           periods, n_harm, 0, NULL)
    assert call(n_harm=65) == BD_ERR_INVALID
    assert call(n_harm=0) == BD_ERR_INVALID
    assert call(i_time=4) == BD_ERR_INVALID
    assert call(izt=8) == BD_ERR_INVALID
    assert call(periods=0) == BD_ERR_INVALID
    assert lib.bd_harmonics_accumulate(BD_F64, NULL, NULL, 8, 8, 2, 0, 4, 1, 2, 0, NULL) == BD_ERR_INVALID


def test_plan_checks(lib):
    p = presets.reference_scaled(40, 48)
    plan = C.c_void_p()
    assert lib.bd_westervelt_plan_create(C.byref(p), BD_F32, 40, 48, 0, -1, -1, -1, C.byref(plan)) == BD_ERR_INVALID
    assert lib.bd_westervelt_plan_create(C.byref(p), BD_F32, 40, 48, 10, 37, 5, 10, C.byref(plan)) == BD_ERR_INVALID
    assert lib.bd_westervelt_plan_create(C.byref(p), BD_F32, 40, 48, 10, 5, 5, 45, C.byref(plan)) == BD_ERR_INVALID
    assert plan.value is None
    assert lib.bd_westervelt_plan_create(C.byref(p), BD_F32, 40, 48, 10, -1, -1, -1, NULL) == BD_ERR_INVALID
    assert lib.bd_westervelt_plan_run(NULL, 1, 1, NULL) == BD_ERR_INVALID
    assert lib.bd_westervelt_plan_shape(NULL, NULL, NULL, NULL, NULL, NULL, NULL) == BD_ERR_INVALID
    assert lib.bd_westervelt_plan_steps(NULL) == -1


def test_adjoint_checks(lib):
    p = presets.phased_array(32, n_el=4, dx=2e-4)
    amp = (C.c_double * 16)()
    obj = _ffi.AdjointObjective()
    J = C.c_double()
    args = (C.cast(amp, C.POINTER(C.c_double)), C.cast(amp, C.POINTER(C.c_double)), C.byref(obj), C.byref(J),
            NULL, NULL, NULL, NULL)
    assert lib.bd_westervelt_adjoint(C.byref(p), BD_F32, 32, 32, 0, 10, *args) == BD_ERR_INVALID     # no steps
    q = p.copy()
    q.full_domain = 0
    assert lib.bd_westervelt_adjoint(C.byref(q), BD_F32, 32, 32, 10, 5, *args) == BD_ERR_UNSUPPORTED


def test_bioheat_checks(lib):
    p, _, _ = bh.reference_params()
    nulls = (NULL, NULL, NULL, NULL)
    q = p.copy()
    q.kdt_dxy = 0.3                              # 1 - 4 kxy - 2 kz < 0: the explicit scheme diverges
    assert lib.bd_temperature_square(C.byref(q), BD_F64, *nulls, 20, 20, 1, 0, 0, NULL) == BD_ERR_UNSTABLE
    assert lib.bd_temperature_square(C.byref(p), BD_F64, *nulls, 20, 20, 1, 0, 0, NULL) == BD_ERR_INVALID
    assert lib.bd_temperature_square(C.byref(p), 3, *nulls, 20, 20, 1, 0, 0, NULL) == BD_ERR_INVALID
    assert lib.bd_temperature_square(C.byref(p), BD_F64, *nulls, 2, 20, 1, 0, 0, NULL) == BD_ERR_INVALID
    buf = (C.c_double * 8)()
    a = C.cast(buf, C.c_void_p)
    b = C.c_void_p(a.value + 8)
    which = C.c_int(0)
    # the two fields must differ, `which` must be 0 or 1, the count non-negative
    assert lib.bd_bioheat_run(C.byref(p), BD_F64, a, a, a, a, NULL, 20, 20, 0, 0, 2, C.byref(which), NULL) == \
        BD_ERR_INVALID
    which.value = 2
    assert lib.bd_bioheat_run(C.byref(p), BD_F64, a, b, a, a, NULL, 20, 20, 0, 0, 2, C.byref(which), NULL) == \
        BD_ERR_INVALID
    which.value = 0
    assert lib.bd_bioheat_run(C.byref(p), BD_F64, a, b, a, a, NULL, 20, 20, 0, 0, -1, C.byref(which), NULL) == \
        BD_ERR_INVALID
    assert lib.bd_bioheat_run(C.byref(p), BD_F64, a, b, a, a, NULL, 20, 20, 0, 0, 2, NULL, NULL) == BD_ERR_INVALID
    # slices: plane kind and ranges
    assert lib.bd_write_slice(b"x.txt", BD_F64, a, 20, 20, 5, 3, 0, 4, 0, 4, 0.0, NULL) == BD_ERR_INVALID
    assert lib.bd_write_slice(b"x.txt", BD_F64, a, 20, 20, BD_SLICE_XZ, 3, 4, 4, 0, 4, 0.0, NULL) == BD_ERR_INVALID
    assert lib.bd_write_slice(b"x.txt", BD_F64, a, 20, 20, BD_SLICE_XZ, 20, 0, 4, 0, 4, 0.0, NULL) == BD_ERR_INVALID


def test_gilmore_checks(lib):
    g = _ffi.GilmoreParams()
    lib.bd_gilmore_reference_params(C.byref(g))
    arr = (C.c_double * 4)()
    ll = (C.c_longlong * 4)()
    dp = C.cast(arr, C.c_void_p)

    def batch(**kw):
        b = _ffi.GilmoreBatch()
        b.method, b.precision, b.n = _ffi.BD_GM_DP45, _ffi.BD_GM_FP64, 4
        b.t_end, b.dt, b.rtol, b.atol = 1e-6, 1e-10, 1e-8, 1e-12
        b.R0 = b.amp = b.phase = dp
        b.Rmax = b.t_collapse = dp
        b.n_acc = b.n_rej = C.cast(ll, C.c_void_p)
        for k, v in kw.items():
            setattr(b, k, v)
        return b
    ws = C.cast(arr, C.c_void_p)
    for bad in (dict(n=0), dict(method=7), dict(precision=3), dict(dt=0.0), dict(rtol=0.0), dict(block=-1)):
        assert lib.bd_gilmore_integrate(C.byref(g), C.byref(batch(**bad)), ws, NULL) == BD_ERR_INVALID, bad
    assert lib.bd_gilmore_integrate(C.byref(g), C.byref(batch()), NULL, NULL) == BD_ERR_INVALID          # workspace
    assert lib.bd_gilmore_integrate(C.byref(g), C.byref(batch(schedule=4)), ws, NULL) == BD_ERR_INVALID
    assert lib.bd_gilmore_integrate(C.byref(g), C.byref(batch(slice=-1)), ws, NULL) == BD_ERR_INVALID
    g2 = _ffi.GilmoreParams()
    lib.bd_gilmore_reference_params(C.byref(g2))
    g2.tableau = 9
    assert lib.bd_gilmore_integrate(C.byref(g2), C.byref(batch()), ws, NULL) == BD_ERR_INVALID
    out = C.c_int32()
    assert lib.bd_gilmore_autotune(C.byref(g), C.byref(batch()), -1, 100, NULL, C.byref(out), C.byref(out),
                                   C.byref(out), NULL) == BD_ERR_INVALID
    assert lib.bd_gilmore_threshold_sweep(C.byref(g), dp, dp, 10, 0, 1, NULL) == BD_ERR_INVALID
    assert lib.bd_gilmore_threshold_sweep(C.byref(g), dp, dp, 1 << 20, 1 << 12, 1, NULL) == BD_ERR_INVALID
    assert lib.bd_gilmore_threshold_sweep(C.byref(g), NULL, dp, 10, 10, 1, NULL) == BD_ERR_INVALID


def test_q3d_checks(lib, tmp_path):
    buf = (C.c_double * 8)()
    a = C.cast(buf, C.c_void_p)
    missing = str(tmp_path / "none.bin").encode()
    assert lib.bd_q3d_read_expand(missing, BD_F64, a, 9, 4, NULL) == BD_ERR_IO
    short = tmp_path / "short.bin"
    np.zeros(10).tofile(short)                   # fewer values than 5 x 5 x 4
    assert lib.bd_q3d_read_expand(str(short).encode(), BD_F64, a, 9, 4, NULL) == BD_ERR_IO
    assert lib.bd_q3d_read_expand(missing, 5, a, 9, 4, NULL) == BD_ERR_INVALID
    assert lib.bd_q3d_write(b"q.bin", BD_F64, a, 9, 4, 4, NULL) == BD_ERR_INVALID                  # iz0 >= Nz
    assert lib.bd_mirror_expand(BD_F64, a, a, 9, 4, NULL) == BD_ERR_INVALID                        # in place
    assert lib.bd_mirror_expand(BD_F64, a, NULL, 9, 4, NULL) == BD_ERR_INVALID
