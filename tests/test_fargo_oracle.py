"""The oracle's FARGO remap (the code its step runs; src/solver/fargo_scheme.f90:17-77) pinned
by what the mathematics of the scheme fixes, independently of the formula as written.

The reference remap replaces each cell of a ring (all phi at fixed r) by the average, over the
cell shifted back by |vk dt| / dy cells, of the limited piecewise-linear reconstruction of the
ring (slopes d_j = limit(u_j - u_{j-1}, u_{j+1} - u_j), slope_limiters.f90:5-37).  From that:
  * an integer shift (eps = 0) is an exact cyclic permutation, in the direction
    jshft = iy - nshft - 1 (+1 for the right neighbour) for vk > 0 and iy + nshft for vk < 0
    (fargo_scheme.f90:48-67), bit for bit, whatever the slopes (their weight 1/4 - (1/2)^2 is 0);
  * a constant ring stays constant (to rounding of u*eps + u*(1-eps));
  * ring sums are conserved (the slope terms cancel pairwise), for every limiter and both signs;
  * a linear ring away from the periodic seam is translated exactly by the fractional shift;
  * in general, each cell equals the exact integral of the reconstruction over the shifted cell,
    computed here piece by piece over the cells it overlaps.
The reference leaves `one` and `eps` undeclared (implicit REAL(4), :47, :57, :69) and the edge
slopes unset (:20-31); the oracle takes one = 1.0d0, FP64 eps and periodic slopes for every
active cell, and these checks hold for that decision.
"""
import os
import sys

import numpy as np
import pytest

sys.path.insert(0, os.path.join(os.path.dirname(__file__), "..", "oracle"))
import pyoracle  # noqa: E402

NX, NY, NV = 6, 48, 5


def grid(dy_col, r_col=None):
    """Per-cell dx(:,2) and radii_c of a (NX+4) x (NY+4) grid from per-column values."""
    nxm, nym = NX + 4, NY + 4
    dx2 = np.tile(np.asarray(dy_col, float), nym)
    r = np.ones(nxm * nym) if r_col is None else np.tile(np.asarray(r_col, float), nym)
    assert dx2.size == nxm * nym
    return dx2, r


def per_cell(col_vals):
    return np.tile(np.asarray(col_vals, float), NY + 4)


def rings(u):
    """(nvar, ny, nx) view of the active block of a (nvar, ncells) state."""
    return u.reshape(u.shape[0], NY + 4, NX + 4)[:, 2:NY + 2, 2:NX + 2]


def state_from_rings(a, fill=np.nan):
    """(nvar, ncells) state whose active block is a (nvar, ny, nx); other cells hold `fill`
    (NaN: the remap must not read them, ghost rows are taken as periodic copies)."""
    u = np.full((a.shape[0], NY + 4, NX + 4), fill)
    u[:, 2:NY + 2, 2:NX + 2] = a
    return u.reshape(a.shape[0], -1)


def test_integer_shift_is_exact_cyclic_permutation():
    # dy = 2^-5, dt = 2^-7: vk = 4 m gives |vk dt| = m dy exactly, so nshft = m and eps = 0
    dy, dt = 2.0 ** -5, 2.0 ** -7
    shifts = [(1, +1), (5, +1), (NY - 1, +1), (3, -1), (NY - 1, -1), (0, +1)]
    vk = np.array([0.0, 0.0] + [s * m * dy / dt for m, s in shifts] + [0.0, 0.0])
    dx2, r = grid(np.full(NX + 4, dy))
    rng = np.random.default_rng(1)
    a = rng.uniform(0.5, 2.0, (NV, NY, NX)) * rng.choice([-1.0, 1.0], (NV, NY, NX))
    for slope in (1, 2, 3):
        out = rings(pyoracle.fargo_remap(state_from_rings(a), NX, NY, slope, dt, dx2, r, per_cell(vk)))
        for i, (m, s) in enumerate(shifts):
            # vk > 0: cell iy takes cell iy - m (the pattern moves to larger phi); vk < 0: iy + m
            want = np.roll(a[:, :, i], s * m, axis=1)
            assert np.array_equal(out[:, :, i], want), (slope, m, s)


def test_shift_beyond_one_turn_wraps():
    # the reference subtracts ny once (fargo_scheme.f90:50, 63); the oracle wraps modulo ny, the
    # same for shifts below one turn and a whole number of turns more beyond
    dy, dt = 2.0 ** -5, 2.0 ** -7
    ms = [NY, NY + 3, 2 * NY + 7, NY + 1, 3 * NY, 1]
    sg = [+1, +1, -1, -1, +1, -1]
    vk = np.array([0.0, 0.0] + [s * m * dy / dt for m, s in zip(ms, sg)] + [0.0, 0.0])
    dx2, r = grid(np.full(NX + 4, dy))
    a = np.random.default_rng(2).uniform(1.0, 2.0, (NV, NY, NX))
    out = rings(pyoracle.fargo_remap(state_from_rings(a), NX, NY, 1, dt, dx2, r, per_cell(vk)))
    for i, (m, s) in enumerate(zip(ms, sg)):
        assert np.array_equal(out[:, :, i], np.roll(a[:, :, i], s * m, axis=1)), (m, s)


def random_case(seed):
    rng = np.random.default_rng(seed)
    dphi = 2 * np.pi / NY
    r_col = np.concatenate([[0.4, 0.45], np.linspace(0.5, 2.5, NX), [2.55, 2.6]])
    dx2, r = grid(np.full(NX + 4, dphi), r_col)
    dt = 1.7e-2
    # shifts of 0.1 to 9.6 cells per step, different on each ring
    cells = rng.uniform(0.1, 9.6, NX + 4)
    vk = cells * dphi * r_col / dt
    return rng, dx2, r, r_col, dt, vk


@pytest.mark.parametrize("slope", [1, 2, 3])
@pytest.mark.parametrize("sign", [1.0, -1.0])
def test_constant_ring_stays_constant(slope, sign):
    rng, dx2, r, _, dt, vk = random_case(3)
    c = np.array([1.0, -0.37, 2.5e-3, 7.1, 1e-9])
    a = np.broadcast_to(c[:, None, None], (NV, NY, NX)).copy()
    out = rings(pyoracle.fargo_remap(state_from_rings(a), NX, NY, slope, dt, dx2, r, per_cell(sign * vk)))
    # u*eps + u*(1-eps) rounds twice: within two ulps of u
    tol = 2 * np.finfo(float).eps * np.abs(c)[:, None, None]
    assert np.all(np.abs(out - a) <= tol)


@pytest.mark.parametrize("slope", [1, 2, 3])
@pytest.mark.parametrize("sign", [1.0, -1.0])
def test_ring_sums_are_conserved(slope, sign):
    rng, dx2, r, _, dt, vk = random_case(4)
    a = rng.uniform(0.1, 3.0, (NV, NY, NX)) + np.sin(np.arange(NY) * 0.7)[None, :, None]
    out = rings(pyoracle.fargo_remap(state_from_rings(a), NX, NY, slope, dt, dx2, r, per_cell(sign * vk)))
    s0, s1 = a.sum(axis=1), out.sum(axis=1)
    assert np.all(np.abs(s1 - s0) <= 1e-14 * np.abs(a).sum(axis=1))
    assert not np.array_equal(out, a)


@pytest.mark.parametrize("slope", [1, 2, 3])
@pytest.mark.parametrize("sign", [1.0, -1.0])
def test_linear_ring_is_translated_by_the_fractional_shift(slope, sign):
    _, dx2, r, r_col, dt, vk = random_case(5)
    j = np.arange(NY, dtype=float)
    a0 = np.array([1.0, -2.0, 0.5, 3.0, 10.0])
    b = np.array([0.25, -0.125, 0.03, 1.0, -0.5])
    a = (a0[:, None] + b[:, None] * j[None, :])[:, :, None] * np.ones((1, 1, NX))
    out = rings(pyoracle.fargo_remap(state_from_rings(a), NX, NY, slope, dt, dx2, r, per_cell(sign * vk)))
    for i in range(NX):
        ic = i + 2
        s = abs(vk[ic] * dt) / (dx2[ic] * r_col[ic])       # shift in cells
        n = int(s)
        for iy in range(NY):
            # source cells j0-1 .. j1+1 of the remap must not straddle the periodic seam
            j0 = iy - n - 1 if sign > 0 else iy + n
            if j0 - 1 < 0 or j0 + 2 > NY - 1:
                continue
            want = a0 + b * (iy - sign * s)
            np.testing.assert_allclose(out[:, iy, i], want, rtol=0, atol=1e-12 * np.abs(a).max())


def pl_average(u, d, c):
    """Average over [c - 1/2, c + 1/2] (cell units, cell j centred at j, periodic in len(u)) of
    the reconstruction u_j + d_j (x - j), integrated piece by piece over the cells it meets."""
    n = len(u)
    lo, hi = c - 0.5, c + 0.5
    total = 0.0
    x = lo
    while x < hi - 1e-15:
        j = int(np.floor(x + 0.5))
        edge = min(j + 0.5, hi)
        mid = 0.5 * (x + edge)
        jj = j % n
        total += (edge - x) * (u[jj] + d[jj] * (mid - j))
        x = edge
    return total


@pytest.mark.parametrize("slope", [1, 2, 3])
@pytest.mark.parametrize("sign", [1.0, -1.0])
def test_remap_is_the_exact_average_of_the_limited_reconstruction(slope, sign):
    rng, dx2, r, r_col, dt, vk = random_case(6)
    phi = np.arange(NY) * 2 * np.pi / NY
    a = np.empty((NV, NY, NX))
    for v in range(NV):
        for i in range(NX):
            k = rng.integers(1, 5, 3)
            a[v, :, i] = 1.5 + (rng.uniform(-1, 1, 3)[:, None] * np.sin(np.outer(k, phi) + rng.uniform(0, 6, 3)[:, None])).sum(0)
    out = rings(pyoracle.fargo_remap(state_from_rings(a), NX, NY, slope, dt, dx2, r, per_cell(sign * vk)))
    L = pyoracle.lib()
    for i in range(NX):
        ic = i + 2
        s = abs(vk[ic] * dt) / (dx2[ic] * r_col[ic])
        for v in range(NV):
            u = a[v, :, i]
            d = np.array([L.oracle_slope_limit(slope, u[j] - u[j - 1], u[(j + 1) % NY] - u[j])
                          for j in range(NY)])
            want = np.array([pl_average(u, d, iy - sign * s) for iy in range(NY)])
            np.testing.assert_allclose(out[v, :, i], want, rtol=0, atol=1e-12 * np.abs(u).max())


def test_slope_limiters_closed_forms():
    """slope_limiters.f90:5-37: minmod (1), harmonic mean (2), monotonised central (3)."""
    L = pyoracle.lib()
    lim = L.oracle_slope_limit
    for t in (1, 2, 3):
        assert lim(t, 1.0, -2.0) == 0.0 and lim(t, -3.0, 0.5) == 0.0   # extrema: no slope
        assert lim(t, 0.75, 0.75) == 0.75                               # smooth: the slope itself
    assert lim(1, 1.0, 3.0) == 1.0 and lim(1, -4.0, -2.0) == -2.0      # smaller magnitude
    assert lim(2, 1.0, 3.0) == 1.5                                      # 2ab/(a+b)
    assert lim(3, 1.0, 3.0) == 1.5 and lim(3, 1.0, 1.2) == 1.1         # min(1.5 min|.|, |mean|)
    assert lim(3, -1.0, -5.0) == -1.5
