"""Sedov blast (src/setups/sedov): the Python setup builds the reference's initial state bit for
bit, the oracle is pinned on the reference's runs (ref_sedov_*.npz, test_oracle_ref.py), and over a
periodic box the blast keeps the symmetries of its square (x/y transposition, mirror in x) to
rounding (the x and y flux differences are added in a fixed order, so a transposed flow rounds
differently) while conserving mass, momentum and total energy."""
import numpy as np
import pytest

from refcases import active_mask, load_ref
from shark_amd import cfg as C
from shark_amd import setups


@pytest.mark.parametrize("case,kw", [("sedov_hllc_s3", dict(n=32, cfl=0.5, slope_type=3)),
                                     ("sedov_llf_s1", dict(n=40, cfl=0.8, slope_type=1, solver=C.SOLVER_LLF))])
def test_setup_matches_reference_initial_state(case, kw):
    cfg, geom, state, _, _, _ = load_ref(case)
    c2, g2, s2 = setups.sedov(**kw)
    np.testing.assert_array_equal(s2["u"], state["u"])
    for k in ("dx", "surf", "vol", "position"):
        np.testing.assert_array_equal(np.asarray(g2[k]).reshape(-1), np.asarray(geom[k]).reshape(-1))


def check_blast(u0, u, n):
    a = u.reshape(5, n + 4, n + 4)[:, 2:n + 2, 2:n + 2]
    scale = np.abs(a).reshape(5, -1).max(axis=1)[:, None, None]
    t = np.swapaxes(a, 1, 2)[[0, 2, 1, 3, 4]]
    assert (np.abs(t - a) <= 1e-12 * scale).all()       # x <-> y
    mir = a[:, :, ::-1].copy()
    mir[1] = -mir[1]
    assert (np.abs(mir - a) <= 1e-12 * scale).all()     # x -> -x
    m = np.zeros((n + 4, n + 4), bool)
    m[2:n + 2, 2:n + 2] = True
    m = m.reshape(-1)
    for v, tol in ((0, 1e-13), (4, 1e-13)):
        assert abs(u[v, m].sum() - u0[v, m].sum()) <= tol * abs(u0[v, m].sum()), v
    assert np.abs(u[1, m].sum()) <= 1e-13 * np.abs(u[1, m]).sum()


@pytest.mark.parametrize("solver,slope", [(C.SOLVER_HLLC, 3), (C.SOLVER_HLL, 2), (C.SOLVER_LLF, 1)])
def test_oracle_blast_symmetric_and_conservative(solver, slope):
    import pyoracle
    n = 48
    cfg, g, st = setups.sedov(n=n, solver=solver, slope_type=slope)
    o = pyoracle.Oracle(cfg, g, st)
    o.step()   # the first step fills the ghost cells
    u0 = o.u()
    o.run(60, 1e300)
    check_blast(u0, o.u(), n)
