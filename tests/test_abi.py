"""The C-ABI library loads without a GPU and exports every function that
include/bubbledynamics.h declares (no compute calls)."""
import ctypes as C
import re
from pathlib import Path

ROOT = Path(__file__).resolve().parent.parent


def declared_functions():
    text = (ROOT / "include" / "bubbledynamics.h").read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = re.findall(r"^\s*(?:const\s+)?[A-Za-z_][\w\s\*]*?\b(bd_\w+)\s*\(", text, flags=re.M)
    return sorted(set(names))


def test_header_declares_entry_points():
    names = declared_functions()
    for n in ("bd_westervelt_space6", "bd_boundary_condition_source", "bd_boundary_condition_space6",
              "bd_acoustic_power_water", "bd_gilmore_threshold_sweep", "bd_gilmore_integrate"):
        assert n in names
    assert len(names) >= 20


def test_library_exports_every_declared_symbol():
    from bubbledynamics_amd import _ffi
    lib = C.CDLL(str(_ffi.LIB_PATH))
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing


def test_ffi_signature_table_matches_header():
    from bubbledynamics_amd import _ffi
    declared = set(declared_functions())
    assert set(_ffi.SIGNATURES) <= declared
    lib = _ffi.load()            # binds every signature: fails on a missing export
    assert lib.bd_version().startswith(b"bubbledynamics")


def test_struct_sizes_match_c_layout():
    # the ctypes mirrors must have the C sizes (x86-64 SysV alignment)
    from bubbledynamics_amd._ffi import GilmoreBatch, GilmoreParams, WesterveltParams
    assert C.sizeof(WesterveltParams) == 25 * 8 + 4 * 4 + 3 * 8 + 2 * 4 + 3 * 8 + 8 * 8
    assert C.sizeof(GilmoreParams) == 11 * 8 + 2 * 4 + 4 * 8 + 4 * 8 + 7 * 8
    assert C.sizeof(GilmoreBatch) == 8 * 4 + 3 * 8 + 5 * 8 + 8 + 8 * 8 + 4 * 8


def test_c_host_programs_link_and_print_usage():
    # the C host programs link against the library and the HIP runtime and
    # start without a GPU (usage text, exit status 2)
    import subprocess
    host = ROOT / "bubbledynamics_amd" / "host"
    for name in ("gilmore_threshold", "westervelt_dropin", "temperature_spiral"):
        exe = host / f"bd_{name}"
        assert exe.exists(), f"{exe} is not built (python -m bubbledynamics_amd.build)"
        r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=60)
        assert r.returncode == 2 and "usage:" in r.stderr, (name, r.returncode, r.stderr)
