"""The command line (esthera_pf) on the GPU: its per-configuration lines are
the library driver's result for the same configuration (pf_generic.c:814
format), its sweeps run the reference's loop order, and -T prints one
print_estimate line per trace row."""
import os
import subprocess

import numpy as np
import pytest

import esthera_amd as ea

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))
ARM = os.path.join(HERE, "golden", "robotarm_input5.txt")
UNI = os.path.join(HERE, "golden", "unicycle_input.txt")


def cli(*args):
    r = subprocess.run([ea.CLI_PATH, *args], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    return r.stdout.strip().splitlines()


def test_cli_line_equals_driver(engine):
    lines = cli("-m", "64", "-N", "16", "-t", "1", ARM)
    assert lines[0] == "N m r t e"
    N, m, r, t, e = lines[1].split()
    assert (int(N), int(m), float(r), int(t)) == (16, 64, 1.0, 1)
    err, steps = ea.particle_filter(64, 16, 1, 1.0, ARM)
    assert e == f"{err:.16f}"


def test_cli_sweep_order_and_resamplers(engine):
    lines = cli("-m", "64:128", "-N", "4:8", "-t", "0:1", "-R", "systematic", ARM)
    got = [tuple(x.split()[:4]) for x in lines[1:]]
    want = [(str(G), str(m), "1.000000", str(t)) for t in (0, 1) for G in (4, 8) for m in (64, 128)]
    assert got == want
    errs = np.array([float(x.split()[4]) for x in lines[1:]])
    assert np.all(np.isfinite(errs)) and np.all(errs > 0)
    lines = cli("-m", "128", "-N", "4", "-A", "-r", "0.0:0.02", ARM)   # adaptive alias, step 0.01
    assert [x.split()[2] for x in lines[1:]] == ["0.000000", "0.010000", "0.020000"]


def test_cli_per_step_trace(engine):
    lines = cli("-T", "-m", "64", "-N", "8", ARM)
    assert lines[0] == "ex ey ew ee ax ay"
    rows = np.array([[float(v) for v in x.split()] for x in lines[1:]])
    assert rows.shape == (100, 6) and np.all(np.isfinite(rows))
    truth = np.loadtxt(ARM)
    np.testing.assert_allclose(rows[:, 4:6], truth[:, 13:15], atol=1e-6)
    lines = cli("-T", "-M", "unicycle10", "-m", "128", "-N", "8", UNI)
    rows = np.array([[float(v) for v in x.split()] for x in lines[1:]])
    err, steps = ea.unicycle10_filter(128, 8, 1, 1.0, UNI)
    assert rows.shape[0] == steps
    np.testing.assert_allclose(rows[:, 3].mean(), err, rtol=1e-5)


def test_cli_camera_trace(engine, tmp_path):
    ctrl, meas, truth = ea.robotarm5_camera_scenario(4, 30, 0.1)
    path = str(tmp_path / "cam.txt")
    ea.write_robotarm_trace(path, meas, ctrl, np.full(30, 0.1, np.float32), truth[1:])
    lines = cli("-M", "robotarm5_camera", "-m", "256", "-N", "8", path)
    e = float(lines[1].split()[4])
    assert np.isfinite(e)
