"""Seeded synthetic inputs: SplitMix64 known answers and distributions."""
import math

import numpy as np

from bubbledynamics_amd import gilmore as gm


def test_splitmix64_known_answer():
    # SplitMix64 seeded with 0: first output 0xE220A8397B1DCDAF (Steele, Lea & Flood 2014)
    u = gm.seeded_uniforms(0, 2)
    assert u[0] == (0xE220A8397B1DCDAF >> 11) / 2.0 ** 53
    v = gm._splitmix_vec(0, 1000)
    np.testing.assert_array_equal(v, gm.seeded_uniforms(0, 1000))


def test_workload_distributions():
    R0, amp, ph, fr = gm.config0_inputs(1024, 0)
    assert R0.min() >= 1e-6 and R0.max() < 5e-6
    assert amp.min() >= 0.5e6 and amp.max() < 3e6
    assert ph.min() >= 0 and ph.max() < 2 * math.pi
    R0, amp, ph, fr = gm.config3_inputs(100000, 3)
    assert R0.min() >= 0.5e-6 and R0.max() < 20e-6
    # log-uniform: log R0 is uniform
    lr = (np.log(R0) - math.log(0.5e-6)) / math.log(40)
    assert abs(lr.mean() - 0.5) < 0.01
    assert amp.min() >= 0.5e6 and amp.max() < 8e6


def test_frequency_table_host_matches_oracle(orc):
    np.testing.assert_array_equal(gm.freq_table(), orc.freq_table())


def test_sweep_table_format(tmp_path):
    # the reference's table layout (G:580-606), numbers in the default stream
    # format; known strings for the reference R0 grid and tone table
    from bubbledynamics_amd import gilmore as gm
    fr = gm.freq_table()[[0, 27774]]
    res = np.array([[676000.0, 1e7, -1.0], [1878000.0, 1562000.0, 2.5e-3]])
    path = tmp_path / "sweep.dat"
    gm.write_sweep_table(path, fr, res, 3, 1e-7, 1e-7, 12.5)
    lines = path.read_text().split("\n")
    assert lines[0] == "fr1;fr2;fr3 --> R0;1e-07;2e-07;3e-07"
    assert lines[1] == "20000;20000;20000;676000;1e+07;-1"
    assert lines[2] == "5e+06;5e+06;5e+06;1.878e+06;1.562e+06;0.0025"
    assert lines[3] == "time (s);12.5"
    assert len(lines) == 4
