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
