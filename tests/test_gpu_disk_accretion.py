"""GPU accretion through the outer boundary of the polar disk: bit-exact with the oracle, with
and without FARGO."""
import numpy as np
import pytest

from refcases import active_mask
from shark_amd import setups

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("fargo,nd,phi_mom", [(False, 2, 2.0), (True, 4, 2.0), (True, 4, 1.0),
                                             (False, 2, 1.0)])
def test_gpu_accreting_disk_matches_oracle(fargo, nd, phi_mom):
    """phi_mom < 2 feeds only part of the outer edge; with FARGO the rows left out are remapped
    from the column as the boundaries before the remap left it (fargo_scheme.f90:15)."""
    import pyoracle
    from shark_amd import gpu
    cfg, g, st = setups.disk(nr=128, nphi=192, nd=nd, fargo=fargo)
    cfg.accretion = 1
    cfg.m_acc = 1e-6
    cfg.box_l = 2.1
    cfg.dust2gas = 0.01
    cfg.phi_mom = phi_mom
    print(f"accreting polar disk 128x192 nd={nd} fargo={fargo} phi_mom={phi_mom}", flush=True)
    o = pyoracle.Oracle(cfg, g, st)
    dts_o = [o.step() for _ in range(25)]
    s = gpu.Step(cfg, g, st)
    dts = [s.step() for _ in range(25)]
    u = s.download()
    s.close()
    np.testing.assert_array_equal(dts, dts_o)
    m = active_mask(cfg)
    np.testing.assert_array_equal(u[:, m], o.u()[:, m])
