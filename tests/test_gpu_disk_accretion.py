"""GPU accretion through the outer boundary of the polar disk: bit-exact with the oracle, with
and without FARGO."""
import numpy as np
import pytest

from refcases import active_mask
from shark_amd import setups

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("fargo,nd", [(False, 2), (True, 4)])
def test_gpu_accreting_disk_matches_oracle(fargo, nd):
    import pyoracle
    from shark_amd import gpu
    cfg, g, st = setups.disk(nr=128, nphi=192, nd=nd, fargo=fargo)
    cfg.accretion = 1
    cfg.m_acc = 1e-6
    cfg.box_l = 2.1
    cfg.dust2gas = 0.01
    print(f"accreting polar disk 128x192 nd={nd} fargo={fargo}", flush=True)
    o = pyoracle.Oracle(cfg, g, st)
    dts_o = [o.step() for _ in range(25)]
    s = gpu.Step(cfg, g, st)
    dts = [s.step() for _ in range(25)]
    u = s.download()
    s.close()
    np.testing.assert_array_equal(dts, dts_o)
    m = active_mask(cfg)
    np.testing.assert_array_equal(u[:, m], o.u()[:, m])
