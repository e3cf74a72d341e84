"""The Fortran host driver (shark_amd/fortran, ISO_C_BINDING) calling the C-ABI gives the same
bits as the Python host calling it, and as the oracle, from the same initial state."""
import os
import subprocess

import numpy as np
import pytest

import pyoracle
from shark_amd import gpu, setups, statefile

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRIVER = os.path.join(ROOT, "shark_amd", "bin", "shark_gpu_driver")


def run_driver(args, tmp_path):
    if not os.path.exists(DRIVER):
        pytest.fail("Fortran driver not built (make -C shark_amd/fortran)")
    out = str(tmp_path / "result.bin")
    env = dict(os.environ, SHARK_OUTPUT_DIR=str(tmp_path))
    r = subprocess.run([DRIVER] + args + [out], capture_output=True, text=True, timeout=120, env=env)
    print(r.stdout, r.stderr)
    assert r.returncode == 0, r.stderr
    return statefile.read_result(out)


def test_fortran_dustywave_matches_python_host_and_oracle(tmp_path):
    u0, u, steps, time = run_driver(["dustywave", "1000"], tmp_path)
    assert steps == 1000
    cfg, geom, state = setups.dustywave()
    state = dict(state, u=np.ascontiguousarray(u0))
    g = gpu.Step(cfg, geom, state)
    g.run(1000)
    up = g.download()
    o = pyoracle.Oracle(cfg, geom, state)
    o.run(1000)
    assert np.array_equal(u, up)
    assert np.array_equal(u, o.u())
    assert time == pytest.approx(o.time, rel=0, abs=0)
    # the Fortran driver's output(1) through the binding: rho and the dust velocity over active cells
    act = np.zeros(u.shape[1], bool)
    act[2:-2] = True
    rho = np.fromfile(str(tmp_path / "output_00001" / "rho"))
    vd = np.fromfile(str(tmp_path / "output_00001" / "vd"))
    assert np.array_equal(rho, np.maximum(u[0, act], cfg.smallr))
    assert np.array_equal(vd, u[6, act] / u[5, act])
    assert "time        =" in open(str(tmp_path / "output_00001" / "info.dat")).read()


def test_fortran_state_file_disk_matches_python_host(tmp_path):
    cfg, geom, state = setups.disk(nr=64, nphi=96, nd=4)
    path = str(tmp_path / "disk.state")
    statefile.write_state(path, cfg, geom, state, nsteps=10)
    u0, u, steps, _ = run_driver(["state", path], tmp_path)
    assert steps == 10
    assert np.array_equal(u0, state["u"])
    g = gpu.Step(cfg, geom, state)
    g.run(10)
    assert np.array_equal(u, g.download())
