"""The C host programs (bubbledynamics_amd/host/*.c), which drive the library
through the C ABI the way the reference's main() programs drive their
kernels, run on the GPU and agree with the library's other paths."""
import re
import subprocess
from pathlib import Path

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

HOST = Path(__file__).resolve().parent.parent / "bubbledynamics_amd" / "host"


def _run(name, *args, timeout=240):
    exe = HOST / f"bd_{name}"
    assert exe.exists(), f"{exe} is not built (python -m bubbledynamics_amd.build)"
    return subprocess.run([str(exe), *map(str, args)], capture_output=True, text=True, timeout=timeout)


@pytest.mark.parametrize("dtype", ["f64", "f32"])
def test_c_dropin_step_loop_equals_plan(dtype):
    # the reference's loop -- interior, t += dt, source, Mur faces, nine
    # device copies per step -- through the drop-in calls, against the plan
    # (pointer ring, fused kernel, 6-step graphs) on the same parameters:
    # two-layer medium with relaxation, quarter domain, 60 steps; the same
    # bits (the fused faces keep the drop-in's order)
    r = _run("westervelt_dropin", 40, 48, 60, dtype)
    assert r.returncode == 0, r.stdout + r.stderr
    m = re.search(r"max\|p\| (\S+) max\|dp\| (\S+) finite (\d)", r.stdout)
    amax, dmax = float(m.group(1)), float(m.group(2))
    print(r.stdout.strip())
    assert m.group(3) == "1" and amax > 1e4
    assert dmax == 0.0


def test_c_threshold_table_equals_python_sweep(tmp_path):
    # the reference Gilmore program's flow in C (constants, 55-tone triplet
    # table, sweep launch, table file) on the three highest-frequency
    # triplets, four radii, the reference's arithmetic and tableau at 1e-6:
    # the same table, value for value, as the drop-in called from Python
    torch = pytest.importorskip("torch")
    from bubbledynamics_amd import gilmore as gm
    first, n_combo, n_r0, tol = 27772, 3, 4, 1e-6
    out_c = tmp_path / "c.csv"
    r = _run("gilmore_threshold", out_c, first, n_combo, n_r0, 0, tol)
    assert r.returncode == 0, r.stdout + r.stderr
    p = gm.reference_params()
    p.abs_eps = p.rel_eps = tol
    fr = gm.freq_table()[first:first + n_combo]
    assert len(gm.freq_table()) == 27775
    res = gm.threshold_sweep(p, fr, n_r0)
    assert np.all(np.isfinite(res))
    out_py = tmp_path / "py.csv"
    gm.write_sweep_table(str(out_py), fr, res, n_r0, p.minR0, p.stepR0, 0.0)
    a = out_c.read_text().splitlines()
    b = out_py.read_text().splitlines()
    assert len(a) == len(b) == n_combo + 2
    assert a[:-1] == b[:-1]                    # header and every row; the last line is the run time
    assert a[-1].startswith("time")
    torch.cuda.synchronize()


def test_c_thermal_program_drop_in_against_run(tmp_path):
    # the reference thermal program's flow in C on a 101 x 101 x 121 grid:
    # Q handed over through a Q3D file with the mirror expansion, initial
    # state by region, the spiral's first 400 shifts with one drop-in
    # Temperature_square per time step, the end-of-run slices; then the same
    # substeps through bd_bioheat_run (runs of equal shift, pairs per launch)
    r = _run("temperature_spiral", 101, 121, 400, tmp_path)
    assert r.returncode == 0, r.stdout + r.stderr
    print(r.stdout.strip())
    m = re.search(r"max\|T\| (\S+) max\|dT\| (\S+) max\|TD\| (\S+) max\|dTD\| (\S+) finite (\d)", r.stdout)
    tmax, dT, dmax, dTD = (float(m.group(k)) for k in range(1, 5))
    assert m.group(5) == "1"
    assert tmax > 273.15 + 23.6 + 1.0               # the focus heated by more than 1 K
    assert dT <= 1e-12 * tmax and dTD <= 1e-12 * dmax
    for name in ("Q3D_spiral.bin", "T_xz.txt", "TD_xz.txt"):
        assert (tmp_path / name).stat().st_size > 0
    rows = (tmp_path / "T_xz.txt").read_text().splitlines()
    assert rows[0].startswith("Z=") and len(rows) == 1 + 121
