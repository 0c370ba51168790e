"""The Westervelt -> bioheat hand-off file (Q3D) and the quarter -> full
mirror expansion, against the byte layout of W:778-790 and the index map of
T:281-299 restated in numpy."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

from bubbledynamics_amd import _ffi  # noqa: E402
from bubbledynamics_amd._ffi import BD_ERR_IO, BD_F32, BD_F64  # noqa: E402


def vp(t):
    return C.c_void_p(t.data_ptr())


def expand_ref(Qq, nxy_full):
    # T:285-294: full (ix, iy) reads quarter (|ix - half|, |iy - half|)
    half = nxy_full // 2
    idx = np.abs(np.arange(nxy_full) - half)
    return Qq[:, idx][:, :, idx]


@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
def test_q3d_write_layout_and_read_expand(tmp_path, dtype):
    rng = np.random.default_rng(7)
    nq, nz_w = 21, 30                       # quarter-domain Westervelt grid
    Q = rng.uniform(0.0, 1e6, (nz_w, nq, nq))
    code = BD_F64 if dtype == torch.float64 else BD_F32
    dQ = torch.tensor(Q, dtype=dtype, device="cuda")
    Qd = dQ.double().cpu().numpy()          # the values as stored on the device
    path = tmp_path / "Q3D.bin"
    st = C.c_void_p(torch.cuda.current_stream().cuda_stream)
    _ffi.call("bd_q3d_write", str(path).encode(), code, vp(dQ), nq, nz_w, 2, st)
    raw = np.fromfile(path, dtype="<f8")
    np.testing.assert_array_equal(raw, Qd[2:].reshape(-1))      # z from iz = 2, then y, x fastest
    nfull, nz_t = 2 * (nq - 1) + 1, nz_w - 2
    full = torch.empty((nz_t, nfull, nfull), dtype=dtype, device="cuda")
    _ffi.call("bd_q3d_read_expand", str(path).encode(), code, vp(full), nfull, nz_t, st)
    ref = expand_ref(Qd[2:], nfull)
    np.testing.assert_array_equal(full.double().cpu().numpy(), ref.astype(np.float32 if code == BD_F32 else np.float64))
    # a file shorter than the grid asks for is an I/O error
    rc = _ffi.load().bd_q3d_read_expand(str(path).encode(), code, vp(full), nfull, nz_t + 1, st)
    assert rc == BD_ERR_IO
    rc = _ffi.load().bd_q3d_read_expand(str(tmp_path / "missing.bin").encode(), code, vp(full), nfull, nz_t, st)
    assert rc == BD_ERR_IO


def test_mirror_expand_device():
    rng = np.random.default_rng(8)
    nfull, nz = 33, 12
    nq = nfull // 2 + 1
    Qq = rng.standard_normal((nz, nq, nq))
    dq = torch.tensor(Qq, device="cuda")
    full = torch.empty((nz, nfull, nfull), dtype=torch.float64, device="cuda")
    _ffi.call("bd_mirror_expand", BD_F64, vp(dq), vp(full), nfull, nz,
              C.c_void_p(torch.cuda.current_stream().cuda_stream))
    np.testing.assert_array_equal(full.cpu().numpy(), expand_ref(Qq, nfull))
    # the expanded field is mirror-symmetric about the centre cell
    f = full.cpu().numpy()
    np.testing.assert_array_equal(f, f[:, ::-1, :])
    np.testing.assert_array_equal(f, f[:, :, ::-1])
