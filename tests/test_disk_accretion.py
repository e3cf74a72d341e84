"""Accretion through the outer boundary of the polar disk (disk_polar/boundaries.f90:63-94): the
reference test suite's DISK_EQ and DISK_S7-1 accretion namelists (ref_suite_*_accrete.npz) pin the
oracle bit for bit (test_oracle_ref.py); without the inflow the same runs differ, so the fixtures
pin it."""
import numpy as np
import pytest

from refcases import load_ref


@pytest.mark.parametrize("case", ["suite_disk_eq_accrete", "suite_s71_accrete"])
def test_accretion_fixtures_pin_the_inflow(case):
    import pyoracle
    cfg, geom, state, final, dts, _ = load_ref(case)
    assert cfg.accretion == 1
    cfg.accretion = 0
    o = pyoracle.Oracle(cfg, geom, state)
    for _ in range(len(dts)):
        o.step()
    assert not np.array_equal(o.u(), final)
