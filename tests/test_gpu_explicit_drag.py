"""Explicit drag in the force (drag = .false. with force_kick, SI setup.f90:355-362): every
species' drag term enters the gas and dust forces of update_force_setup instead of the implicit
drag step.  The oracle is bit-exact with the flang-built reference on two shearing-box runs
(tests/test_oracle_ref.py, ref_si_2d_explicit_drag*); here the GPU is bit-exact with the oracle
on a 128 x 128 box with 4 species over 30 steps, single steps and graph replay alike."""
import numpy as np
import pytest

from refcases import active_mask
from shark_amd import setups

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("by_run", [False, True])
def test_gpu_explicit_drag_matches_oracle(by_run):
    import pyoracle
    from shark_amd import gpu
    cfg, g, st = setups.si_box(n=128, nd=4, st_min=0.1, st_max=1.0)
    cfg.drag = 0
    print(f"shearing box 128x128 nd=4, explicit drag, 30 steps ({'run' if by_run else 'step'}), GPU vs oracle",
          flush=True)
    o = pyoracle.Oracle(cfg, g, st)
    dts_o = [o.step() for _ in range(30)]
    s = gpu.Step(cfg, g, st)
    try:
        if by_run:
            n, t = s.run(30)
            assert n == 30 and t == o.time
        else:
            dts = [s.step() for _ in range(30)]
            np.testing.assert_array_equal(dts, dts_o)
        u = s.download()
    finally:
        s.close()
    m = active_mask(cfg)
    print(f"  state bit-exact {np.array_equal(u[:, m], o.u()[:, m])}", flush=True)
    np.testing.assert_array_equal(u[:, m], o.u()[:, m])
