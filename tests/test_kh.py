"""Kelvin-Helmholtz setup (src/setups/KH/setup.f90): adiabatic gas with dust, uniform-medium Epstein
stopping times from the step's sound speed, implicit drag and the drag heating of the gas
(source_terms.f90:88-95).

The oracle is pinned bit for bit on the flang-built reference's KH runs (ref_kh_*.npz, in
test_oracle_ref.py); dropping the heating term fails both fixtures.  Here: the Python setup
builds the reference's initial state and grain sizes bit for bit (one species, and the MRN
distribution over four), and the oracle conserves the gas mass, the mass of every species above the
density floor and the total momentum over a periodic box."""
import numpy as np
import pytest

from refcases import load_ref
from shark_amd import setups


@pytest.mark.parametrize("case,kw", [("kh_n1_hllc", dict(nx=32, ny=32, ndust=1, rho_diffuse=0.5e-18)),
                                     ("kh_n4_hll", dict(nx=24, ny=32, ndust=4, rho_diffuse=1.0e-18))])
def test_setup_matches_reference_initial_state(case, kw):
    cfg, geom, state, _, _, m = load_ref(case)
    u = setups.units_kh()
    for k in ("unit_d", "unit_l", "unit_t", "unit_v", "unit_p", "unit_dcol"):
        assert u[k] == m[k], k
    c2, g2, s2 = setups.kh(**kw)
    assert (c2.smallr, c2.smallp, c2.rhograin) == (cfg.smallr, cfg.smallp, cfg.rhograin)
    np.testing.assert_array_equal(s2["u"], state["u"])
    np.testing.assert_array_equal(s2["sdust"], state["sdust"])
    for k in ("dx", "surf", "vol", "position"):
        np.testing.assert_array_equal(np.asarray(g2[k]).reshape(-1), np.asarray(geom[k]).reshape(-1))


def conserved(u, nd, m):
    """gas mass, species masses, total x and y momentum over the active cells"""
    return (u[0, m].sum(), u[5:5 + nd][:, m].sum(axis=1),
            u[1, m].sum() + u[5 + nd:5 + 2 * nd][:, m].sum(),
            u[2, m].sum() + u[5 + 2 * nd:5 + 3 * nd][:, m].sum())


@pytest.mark.parametrize("nd", [1, 4])
def test_oracle_conserves_mass_and_momentum(nd):
    import pyoracle
    from refcases import active_mask
    cfg, g, st = setups.kh(nx=48, ny=48, ndust=nd)
    o = pyoracle.Oracle(cfg, g, st)
    m = active_mask(cfg)
    u0 = o.u()
    o.run(120, 1e300)
    u = o.u()
    M0, D0, X0, Y0 = conserved(u0, nd, m)
    M, D, X, Y = conserved(u, nd, m)
    assert abs(M - M0) <= 1e-14 * M0
    above_floor = st["u"][5:5 + nd, m].min(axis=1) > 1e-9 * st["u"][0, m].min()
    assert above_floor.any()
    np.testing.assert_allclose(D[above_floor], D0[above_floor], rtol=1e-13)
    scale = np.abs(u0[1, m]).sum()
    assert abs(X - X0) <= 1e-14 * scale and abs(Y - Y0) <= 1e-14 * scale
    # the shear layer heats: drag moves the at-rest dust and the heating raises the gas energy
    assert u[4, m].sum() > u0[4, m].sum()
