"""Pin the CPU oracle to the flang-built reference (fixtures from tests/golden/make_ref_fixtures.py).

Every case runs the same initial state for the same number of steps through the oracle and
compares the full u_prim (ghosts included) and the dt sequence.  The bar is bit-exact: the
oracle restates the reference in its own expression order and both are FP64 without
contraction (the disk relaxation's exp is glibc's on both sides)."""
import numpy as np
import pytest

import pyoracle
from refcases import REF_CASES, load_ref


@pytest.mark.parametrize("name", REF_CASES)
def test_oracle_matches_reference(name):
    cfg, geom, state, final, dts, meta = load_ref(name)
    o = pyoracle.Oracle(cfg, geom, state)
    got_dt = [o.step() for _ in range(len(dts))]
    np.testing.assert_array_equal(np.array(got_dt), dts)
    np.testing.assert_array_equal(o.u(), final)
