"""Cartesian disk (src/setups/disk): softened point-mass gravity with an optional planet,
zero-gradient boundaries (boundary_types.f90:48-100, source cells taken as flat indices as the
reference does), buffer zones reset every step, alpha viscosity (source_terms.f90:99-110) and the
disk's Epstein stopping times.

The oracle is pinned bit for bit on four flang-built reference runs (ref_cdisk_*.npz, in
test_oracle_ref.py).  Here: the Python setup builds the reference's initial state, sound speeds,
viscosity coefficients and grain sizes bit for bit, and the viscous fixtures fail without the
viscous terms (so they pin them)."""
import numpy as np
import pytest

from refcases import GOLDEN, load_ref
from shark_amd import setups

CASES = [("cdisk_visc_n1", dict(nx=32, ny=32, ndust=1, smooth_r=0.5, alpha_visc=1e-2)),
         ("cdisk_n2_s3", dict(nx=40, ny=32, ndust=2, smooth_r=2.0, viscosity=0)),
         ("cdisk_visc_gas", dict(nx=32, ny=24, ndust=0, alpha_visc=5e-2)),
         ("cdisk_planet_n1", dict(nx=32, ny=32, ndust=1, smooth_r=1.0, mplanet=3e-3))]


@pytest.mark.parametrize("case,kw", CASES)
def test_setup_matches_reference_initial_state(case, kw):
    import os
    cfg, geom, state, _, _, m = load_ref(case)
    u = setups.units_cdisk()
    for k in ("unit_d", "unit_l", "unit_t", "unit_v", "unit_p", "unit_dcol"):
        assert u[k] == m[k], k
    c2, g2, s2 = setups.cdisk(**kw)
    for f in ("sigma_r0", "r0_disk", "mstar", "mplanet", "rplanet", "smooth_r", "box_l", "box_l_y", "smallr",
              "smallp", "hoverr", "gamma"):
        assert getattr(c2, f) == getattr(cfg, f), f
    np.testing.assert_array_equal(s2["u"], state["u"])
    np.testing.assert_array_equal(s2["cs"], state["cs"])
    eta_ref = np.load(os.path.join(GOLDEN, f"ref_{case}.npz"))["eta_visc"]
    np.testing.assert_array_equal(s2["eta_visc"], eta_ref)
    if cfg.ndust:
        np.testing.assert_array_equal(s2["sdust"], state["sdust"])


@pytest.mark.parametrize("case", ["cdisk_visc_n1", "cdisk_visc_gas", "disk_2d_visc", "disk_2d_visc_gas"])
def test_viscous_fixtures_pin_the_viscous_terms(case):
    """Also the polar disk's (source_terms.f90:111-116)."""
    import pyoracle
    cfg, geom, state, final, dts, _ = load_ref(case)
    cfg.viscosity = 0
    o = pyoracle.Oracle(cfg, geom, state)
    for _ in range(len(dts)):
        o.step()
    assert not np.array_equal(o.u(), final)
