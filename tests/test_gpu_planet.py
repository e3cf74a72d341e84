"""The polar disk's planet (disk_polar/setup.f90:323-378: the softened potential of a planet on
a circular orbit, the reference's ghost values of apply_boundaries_phi, its centred gradient in
every fluid's force) on the GPU.  The oracle is bit-exact with the flang-built reference on the
planet fixtures (tests/test_oracle_ref.py, ref_disk_2d_planet*); here the GPU is bit-exact with the
oracle on a larger disk, with and without FARGO and dust, over 30 steps."""
import numpy as np
import pytest

from refcases import active_mask
from shark_amd import setups

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("fargo,nd", [(False, 4), (True, 4), (False, 0)])
def test_gpu_planet_matches_oracle(fargo, nd):
    import pyoracle
    from shark_amd import gpu
    cfg, g, st = setups.disk(nr=128, nphi=192, nd=nd, fargo=fargo)
    cfg.mplanet, cfg.rplanet, cfg.a_smooth_pl = 3e-3, 1.3, 0.05
    print(f"polar disk 128x192 nd={nd} fargo={fargo} with a planet: 30 steps, GPU vs oracle", flush=True)
    o = pyoracle.Oracle(cfg, g, st)
    dts_o = [o.step() for _ in range(30)]
    s = gpu.Step(cfg, g, st)
    try:
        dts = [s.step() for _ in range(30)]
        u = s.download()
    finally:
        s.close()
    m = active_mask(cfg)
    print(f"  dt bit-exact {dts == dts_o}; state bit-exact {np.array_equal(u[:, m], o.u()[:, m])}", flush=True)
    np.testing.assert_array_equal(dts, dts_o)
    np.testing.assert_array_equal(u[:, m], o.u()[:, m])
