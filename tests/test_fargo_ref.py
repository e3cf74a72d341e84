"""The oracle's FARGO remap against the flang-built reference's fargo_scheme
(src/solver/fargo_scheme.f90:2-81), built with its undeclared `one` set to 1 and its implicit
reals in double precision (tests/golden/make_ref_fixtures.py FARGO_CASES).

* Integer shifts (ref_fargo_int_*): before every step each ring gets a shift velocity with
  |vk dt| an exact multiple of dy, both signs, up to ny - 1 cells, so the reference's remap is a
  pure permutation whatever its other defects; six full steps (hydro, FARGO, drag, in-loop hook),
  every cell and every dt bit for bit.
* The setup's own velocity (ref_fargo_frac_*), fractional shifts with all three limiters: one
  step, bit for bit on every active cell whose remap reads only the slopes the reference sets
  (it leaves the last two columns and rows unset, :20-31).
"""
import numpy as np
import pytest

import pyoracle
from refcases import FARGO_REF_CASES, load_ref, run_fargo_fixture


def test_fargo_fixtures_present():
    assert {"fargo_int_disk", "fargo_int_disk_s3", "fargo_frac_disk", "fargo_frac_disk_s2",
            "fargo_frac_disk_s3"} <= set(FARGO_REF_CASES)


@pytest.mark.parametrize("name", FARGO_REF_CASES)
def test_oracle_fargo_matches_reference(name):
    dts, u, final, mask = run_fargo_fixture(
        name, lambda c, g, s: pyoracle.Oracle(c, g, s), lambda o: o.u())
    np.testing.assert_array_equal(dts, load_ref(name)[4])
    cfg = load_ref(name)[0]
    if name.startswith("fargo_frac"):
        # most active cells are compared: all but the last two columns and rows' worth
        assert mask.sum() >= 0.5 * cfg.nx * cfg.ny
    bad = np.argwhere(u[:, mask] != final[:, mask])
    assert bad.size == 0, f"{len(bad)} values differ, first {bad[:5]}"
