"""GPU adiabatic gas with dust on the Kelvin-Helmholtz setup: bit-exact with the oracle (drag,
uniform Epstein stopping times from the step's sound speed, drag heating), and at 2048 x 2048 the
gas mass, the species masses above the floor and the total momentum conserved over a periodic box."""
import numpy as np
import pytest

from refcases import active_mask
from shark_amd import cfg as C
from shark_amd import setups
from test_kh import conserved

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("nd,solver,solverdust,slope", [(1, C.SOLVER_HLLC, 0, 3), (4, C.SOLVER_HLL, 2, 1),
                                                        (8, C.SOLVER_LLF, 3, 2)])
def test_gpu_kh_matches_oracle(nd, solver, solverdust, slope):
    import pyoracle
    from shark_amd import gpu
    cfg, g, st = setups.kh(nx=192, ny=160, ndust=nd, solver=solver, solverdust=solverdust,
                           slope_type=slope)
    print(f"kh 192x160 nd={nd} solver={solver}/{solverdust} slope={slope}", flush=True)
    o = pyoracle.Oracle(cfg, g, st)
    dts_o = [o.step() for _ in range(25)]
    s = gpu.Step(cfg, g, st)
    dts = [s.step() for _ in range(25)]
    u = s.download()
    s.close()
    np.testing.assert_array_equal(dts, dts_o)
    m = active_mask(cfg)
    np.testing.assert_array_equal(u[:, m], o.u()[:, m])


def test_gpu_kh_large_conserves():
    from shark_amd import gpu
    nd = 4
    cfg, g, st = setups.kh(nx=2048, ny=2048, ndust=nd)
    print("kh 2048x2048 nd=4, 60 steps", flush=True)
    m = active_mask(cfg)
    s = gpu.Step(cfg, g, st)
    u0 = s.download()
    s.run(60)
    u = s.download()
    s.close()
    M0, D0, X0, Y0 = conserved(u0, nd, m)
    M, D, X, Y = conserved(u, nd, m)
    assert abs(M - M0) <= 1e-12 * M0
    above_floor = u0[5:5 + nd, m].min(axis=1) > 1e-9 * u0[0, m].min()
    np.testing.assert_allclose(D[above_floor], D0[above_floor], rtol=1e-10)
    scale = np.abs(u0[1, m]).sum()
    assert abs(X - X0) <= 1e-12 * scale and abs(Y - Y0) <= 1e-12 * scale
    assert np.isfinite(u[:, m]).all()
