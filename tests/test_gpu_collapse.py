"""Collapse_1D on the GPU (spherical 1D grid, barotropic EOS, 1D self-gravity, collapse
boundaries): against the flang-built reference's one-step fixtures on every active cell the
reference leaves finite (its first active cell turns NaN, tests/test_collapse_oracle.py), against
the oracle over hundreds of steps, and on a 8192-cell grid through mass conservation.

The barotropic EOS evaluates pow() in every cell and face (the GPU's pow against the host's): the
runs agree bit for bit, states and dt sequences alike."""
import numpy as np
import pytest

import pyoracle
import refcases
from shark_amd import gpu
from shark_amd import setups as S

pytestmark = pytest.mark.gpu


def total_mass(cfg, geom, u):
    act = refcases.active_mask(cfg)
    rho = u[0] + u[5:5 + cfg.ndust].sum(axis=0)
    return float((rho * geom["vol"])[act].sum())


@pytest.mark.parametrize("name", ["collapse_gas_1", "collapse_dust_1"])
def test_gpu_collapse_one_step_matches_the_reference(name):
    cfg, geom, state, final, dts, _ = refcases.load_ref(name)
    print(f"{name}: one GPU step against the reference", flush=True)
    g = gpu.Step(cfg, geom, state)
    try:
        dt = g.step()
        u = g.download()
    finally:
        g.close()
    act = refcases.active_mask(cfg)
    cmp = act & np.isfinite(final).all(axis=0)
    rel = refcases.rel_l1(u[:, cmp], final[:, cmp])
    print(f"  dt {dt!r} ref {dts[0]!r}; bit-exact {np.array_equal(u[:, cmp], final[:, cmp])}; "
          f"max rel-L1 {rel.max():.2e}", flush=True)
    assert dt == dts[0]
    assert np.array_equal(u[:, cmp], final[:, cmp])
    assert np.isfinite(u).all()


@pytest.mark.parametrize("nd,nx,nsteps", [(0, 128, 300), (4, 96, 200), (8, 200, 100)])
def test_gpu_collapse_matches_the_oracle(nd, nx, nsteps):
    cfg, geom, st = S.collapse(nx=nx, ndust=nd)
    print(f"collapse nx={nx} nd={nd}: {nsteps} steps, GPU vs oracle", flush=True)
    o = pyoracle.Oracle(cfg, geom, st)
    g = gpu.Step(cfg, geom, st)
    try:
        dg, do = [], []
        for _ in range(nsteps):
            dg.append(g.step())
            do.append(o.step())
        u = g.download()
    finally:
        g.close()
    uo = o.u()
    act = refcases.active_mask(cfg)
    rel = refcases.rel_l1(u, uo, act)
    print(f"  bit-exact {np.array_equal(u[:, act], uo[:, act])} dt bit-exact {dg == do}; "
          f"max rel-L1 {rel.max():.2e}", flush=True)
    assert dg == do
    assert np.array_equal(u[:, act], uo[:, act])


def test_gpu_collapse_large_grid_conserves_mass():
    """8192 cells with 8 dust species for 500 steps: finite everywhere, mass conserved to
    round-off (closed grid), the enclosed mass's single-chain sum spanning eight 1024-cell tiles."""
    cfg, geom, st = S.collapse(nx=8192, ndust=8)
    print("collapse nx=8192 nd=8: 500 steps", flush=True)
    g = gpu.Step(cfg, geom, st)
    try:
        u0 = g.download()
        n, _ = g.run(500)
        u = g.download()
    finally:
        g.close()
    assert n == 500
    assert np.isfinite(u).all()
    assert abs(total_mass(cfg, geom, u) / total_mass(cfg, geom, u0) - 1.0) < 1e-12
