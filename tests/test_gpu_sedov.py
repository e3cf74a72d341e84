"""GPU Sedov blast: bit-exact with the oracle at 192 x 192, and at 2048 x 2048 symmetric to rounding
and conservative over the periodic box."""
import numpy as np
import pytest

from shark_amd import cfg as C
from shark_amd import setups
from test_sedov import check_blast

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("solver,slope", [(C.SOLVER_HLLC, 3), (C.SOLVER_LLF, 1)])
def test_gpu_sedov_matches_oracle(solver, slope):
    import pyoracle
    from shark_amd import gpu
    cfg, g, st = setups.sedov(n=192, solver=solver, slope_type=slope)
    print(f"sedov 192^2 solver={solver} slope={slope}", flush=True)
    o = pyoracle.Oracle(cfg, g, st)
    dts_o = [o.step() for _ in range(40)]
    s = gpu.Step(cfg, g, st)
    dts = [s.step() for _ in range(40)]
    u = s.download()
    s.close()
    np.testing.assert_array_equal(dts, dts_o)
    np.testing.assert_array_equal(u, o.u())


def test_gpu_sedov_large_symmetric_conservative():
    from shark_amd import gpu
    n = 2048
    cfg, g, st = setups.sedov(n=n)
    print("sedov 2048^2, 200 steps", flush=True)
    s = gpu.Step(cfg, g, st)
    s.step()
    u0 = s.download()
    s.run(200)
    u = s.download()
    s.close()
    check_blast(u0, u, n)
