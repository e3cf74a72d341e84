"""Sod shock tube (SOD setup, src/setups/SOD/setup.f90:1-43 and boundaries.f90:10-80) against the
exact Riemann solution at t = 0.245 that the reference's test suite holds
(test/test-suite/Sod-tube/solution.dat, columns x, u, rho, P, e over 1024 cells; kept here as the
data fixture tests/golden/sod_exact_t0245.npz).  The suite's run (sod.nml) is CFL 0.8, slope_type 3,
gamma 1.4, 1024 cells, tend 0.245.

The oracle, pinned bit for bit on the flang-built reference's Sod runs (tests/golden/ref_sod_*.npz),
must also lie close to the exact solution in L1: a wrong sign in the pressure-gradient predictor or
a dropped energy flux term moves the contact and the shock by many cells and fails these bounds."""
import os

import numpy as np
import pytest

from refcases import GOLDEN
from shark_amd import cfg as C
from shark_amd import setups

GAMMA = 1.4
# L1 distances of the oracle (and the GPU path) from the exact solution at 1024 cells
L1_BOUND = {"rho": 1.5e-3, "u": 2e-3, "P": 1e-3}


def exact():
    z = np.load(os.path.join(GOLDEN, "sod_exact_t0245.npz"))
    return {k: z[k] for k in ("x", "u", "rho", "P", "e")}


def prims_1d(u, nx):
    act = slice(2, nx + 2)
    rho = u[0, act]
    vx = u[1, act] / rho
    P = (GAMMA - 1.0) * (u[4, act] - 0.5 * rho * vx * vx)
    return rho, vx, P


def l1_errors(u, nx):
    ex = exact()
    rho, vx, P = prims_1d(u, nx)
    return {"rho": np.abs(rho - ex["rho"]).mean(), "u": np.abs(vx - ex["u"]).mean(),
            "P": np.abs(P - ex["P"]).mean()}


def test_exact_fixture_is_a_sod_solution():
    """The fixture's undisturbed states and the specific internal energy e = P/((gamma-1) rho)."""
    ex = exact()
    assert ex["x"].size == 1024
    np.testing.assert_allclose(ex["e"], ex["P"] / ((GAMMA - 1.0) * ex["rho"]), rtol=2e-3)
    left, right = ex["x"] < 0.2, ex["x"] > 0.95
    np.testing.assert_allclose(ex["rho"][left], 1.0, rtol=1e-3)
    np.testing.assert_allclose(ex["P"][left], 1.0, rtol=1e-3)
    np.testing.assert_allclose(ex["rho"][right], 0.125, rtol=1e-3)
    np.testing.assert_allclose(ex["P"][right], 0.1, rtol=1e-3)


def test_setup_matches_reference_initial_state():
    """setups.sod builds the reference's initial u_prim bit for bit (the ref_sod fixtures' IC)."""
    from refcases import load_ref
    cfg, geom, state, _, _, _ = load_ref("sod_hll_s3")
    c2, g2, s2 = setups.sod(nx=cfg.nx, cfl=cfg.cfl, slope_type=cfg.slope_type, solver=cfg.solver)
    np.testing.assert_array_equal(s2["u"], state["u"])


def test_oracle_against_exact_solution():
    import pyoracle
    cfg, geom, st = setups.sod()
    o = pyoracle.Oracle(cfg, geom, st)
    o.run(100000, 0.245)
    assert 0.245 <= o.time < 0.2456
    err = l1_errors(o.u(), cfg.nx)
    for k, b in L1_BOUND.items():
        assert err[k] < b, (k, err[k])
    # the planted mistakes named above are caught: drop the pressure jump (both sides at p_l)
    cfg2, geom2, st2 = setups.sod(p_r=1.0)
    o2 = pyoracle.Oracle(cfg2, geom2, st2)
    o2.run(100000, 0.245)
    assert l1_errors(o2.u(), cfg2.nx)["P"] > 10 * L1_BOUND["P"]


@pytest.mark.parametrize("direction", [0, 1])
def test_sod_2d_tube_is_uniform_across(direction):
    """A 2D tube along x or y (oracle): every row across the tube is the same bit for bit, the
    transverse momentum stays exactly zero, and the profile is close to the 1D tube's (its dt
    differs: the courant condition sums over both directions, courant.f90).  The cells are square:
    the reference's cartesian 2D grid sets surf(:,1) = dx(:,1) and surf(:,2) = dx(:,2)
    (grid.f90:132-133), which is the face area only when dx = dy."""
    import pyoracle
    n = 64
    cfg, geom, st = setups.sod(nx=n)
    o = pyoracle.Oracle(cfg, geom, st)
    o.run(100000, 0.1)
    ref = o.u()[:, 2:n + 2]
    nx, ny = (n, 4) if direction == 0 else (4, n)
    cfg2, geom2, st2 = setups.sod(nx=nx, ny=ny, shock_dir=direction, box_l=nx / n, box_l_y=ny / n)
    o2 = pyoracle.Oracle(cfg2, geom2, st2)
    o2.run(100000, 0.1)
    u = o2.u().reshape(-1, ny + 4, nx + 4)[:, 2:ny + 2, 2:nx + 2]
    if direction == 1:
        u = np.swapaxes(u, 1, 2)[[0, 2, 1, 3, 4]]   # rows across the tube; its momentum is rho*vy
    assert np.all(u[2] == 0.0)
    for row in range(1, u.shape[1]):
        np.testing.assert_array_equal(u[:, row, :], u[:, 0, :])
    assert np.abs(u[0, 0, :] - ref[0]).mean() < 5e-3


def tube(direction, flip, n=128, w=8, slope=2, solver=C.SOLVER_HLL):
    """The reference's Sod2d tests (test/Sod2d/sod_{x,y}[_up].nml): a tube along x or y, with the
    dense gas on the left, or on the right (_up)."""
    kw = dict(rho_l=0.125, rho_r=1.0, p_l=0.1, p_r=1.0) if flip else {}
    if direction == 0:
        return setups.sod(nx=n, ny=w, shock_dir=0, box_l=1.0, box_l_y=w / n, slope_type=slope, solver=solver, **kw)
    return setups.sod(nx=w, ny=n, shock_dir=1, box_l=w / n, box_l_y=1.0, slope_type=slope, solver=solver, **kw)


def along_x(u, direction, n=128, w=8):
    """Active cells as (var, across, along) with the tube's velocity in slot 1."""
    if direction == 0:
        return u.reshape(5, w + 4, n + 4)[:, 2:w + 2, 2:n + 2]
    return np.swapaxes(u.reshape(5, n + 4, w + 4)[:, 2:n + 2, 2:w + 2], 1, 2)[[0, 2, 1, 3, 4]]


def mirrored(u):
    m = u[:, :, ::-1].copy()
    m[1] = -m[1]
    return m


@pytest.mark.parametrize("slope,solver", [(1, C.SOLVER_LLF), (2, C.SOLVER_HLL), (3, C.SOLVER_HLLC)])
def test_sod2d_symmetries(slope, solver):
    """Oracle: the tube along y is the tube along x transposed, and the tube with the dense gas on
    the right is the mirror image (velocity reversed), bit for bit, to t = 0.245."""
    import pyoracle
    runs = {}
    for d in (0, 1):
        for f in (False, True):
            cfg, g, st = tube(d, f, slope=slope, solver=solver)
            o = pyoracle.Oracle(cfg, g, st)
            o.run(100000, 0.245)
            runs[d, f] = (o.time, along_x(o.u(), d))
    t0, u0 = runs[0, False]
    for (d, f), (t, u) in runs.items():
        assert t == t0
        np.testing.assert_array_equal(mirrored(u) if f else u, u0)
