"""GPU cartesian disk: bit-exact with the oracle (zero-gradient boundaries, softened point mass
and planet, buffer zones, alpha viscosity, Epstein stopping times), for gas only and with dust;
and the polar disk with alpha viscosity, with and without FARGO."""
import numpy as np
import pytest

from refcases import active_mask
from shark_amd import cfg as C
from shark_amd import setups

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("nd,visc,mplanet,solver,slope", [(0, 1, 0.0, C.SOLVER_HLL, 1), (1, 1, 0.0, C.SOLVER_HLLC, 3),
                                                          (2, 0, 0.0, C.SOLVER_LLF, 2), (1, 1, 3e-3, C.SOLVER_HLL, 1)])
def test_gpu_cdisk_matches_oracle(nd, visc, mplanet, solver, slope):
    import pyoracle
    from shark_amd import gpu
    cfg, g, st = setups.cdisk(nx=160, ny=136, ndust=nd, viscosity=visc, mplanet=mplanet, solver=solver,
                              slope_type=slope, smooth_r=1.0)
    print(f"cdisk 160x136 nd={nd} visc={visc} mplanet={mplanet} solver={solver} slope={slope}", flush=True)
    o = pyoracle.Oracle(cfg, g, st)
    dts_o = [o.step() for _ in range(30)]
    s = gpu.Step(cfg, g, st)
    dts = [s.step() for _ in range(30)]
    u = s.download()
    s.close()
    np.testing.assert_array_equal(dts, dts_o)
    np.testing.assert_array_equal(u, o.u())


@pytest.mark.parametrize("fargo,nd", [(False, 2), (True, 4)])
def test_gpu_viscous_polar_disk_matches_oracle(fargo, nd):
    import pyoracle
    from shark_amd import gpu
    cfg, g, st = setups.disk(nr=128, nphi=192, nd=nd, fargo=fargo)
    cfg.viscosity = 1
    cfg.alpha_visc = 5e-2
    print(f"viscous polar disk 128x192 nd={nd} fargo={fargo}", flush=True)
    o = pyoracle.Oracle(cfg, g, st)
    dts_o = [o.step() for _ in range(25)]
    s = gpu.Step(cfg, g, st)
    dts = [s.step() for _ in range(25)]
    u = s.download()
    s.close()
    np.testing.assert_array_equal(dts, dts_o)
    m = active_mask(cfg)
    np.testing.assert_array_equal(u[:, m], o.u()[:, m])
