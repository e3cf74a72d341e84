"""Output snapshots and restart dumps from the device state.

shark_gpu_write_output after the fixture's steps on the GPU writes the same files, byte for byte,
as the reference's own output(iout) (tests/golden/ref_output_<case>.npz).  A run continued from a
shark_gpu_save_state dump in a fresh context takes exactly the steps of the uninterrupted run."""
import os
import tempfile

import numpy as np
import pytest

from refcases import load_ref
from shark_amd import gpu
from test_output import OUT_CASES, compare_dir, load_expected

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("case", OUT_CASES)
def test_gpu_output_matches_reference_files(case):
    cfg, geom, state, final, dts, _ = load_ref(case)
    m, files = load_expected(case)
    print(f"output {case}: {len(dts)} steps", flush=True)
    g = gpu.Step(cfg, geom, state)
    try:
        n, _ = g.run(len(dts))
        assert n == len(dts)
        assert g.time == m["time"]
        with tempfile.TemporaryDirectory() as td:
            g.write_output(m["iout"], m["units"], m["gridspace"], base_dir=td)
            compare_dir(os.path.join(td, f"output_{m['iout']:05d}"), files)
    finally:
        g.close()


@pytest.mark.parametrize("case", ["disk_2d", "si_2d"])
def test_restart_dump_continues_bit_exact(case):
    cfg, geom, state, _, _, _ = load_ref(case)
    g = gpu.Step(cfg, geom, state)
    try:
        ref = [g.step() for _ in range(10)]
        u_ref, t_ref = g.download(), g.time
    finally:
        g.close()
    with tempfile.TemporaryDirectory() as td:
        path = os.path.join(td, "dump.bin")
        g = gpu.Step(cfg, geom, state)
        try:
            first = [g.step() for _ in range(4)]
            g.save_state(path)
        finally:
            g.close()
        g = gpu.Step(cfg, geom, state)      # a fresh run of the same setup, then the dump
        try:
            g.load_state(path)
            rest = [g.step() for _ in range(6)]
            u, t = g.download(), g.time
        finally:
            g.close()
    assert first + rest == ref
    assert t == t_ref
    assert np.array_equal(u, u_ref)


def test_load_state_rejects_other_configuration():
    cfg, geom, state, _, _, _ = load_ref("disk_2d")
    cfg2, geom2, state2, _, _, _ = load_ref("si_2d")
    with tempfile.TemporaryDirectory() as td:
        path = os.path.join(td, "dump.bin")
        g = gpu.Step(cfg, geom, state)
        try:
            g.save_state(path)
        finally:
            g.close()
        g = gpu.Step(cfg2, geom2, state2)
        try:
            with pytest.raises(gpu.SharkGPUError, match="not this configuration"):
                g.load_state(path)
            with pytest.raises(gpu.SharkGPUError, match="cannot open"):
                g.load_state(os.path.join(td, "missing.bin"))
        finally:
            g.close()
