"""Dust growth inside the step (src/solver/solve.f90:112): the post pass splits around a growth
call on the step's own dust densities.  FP64 growth in the step is bit-exact with the oracle's
step + growth; the learned step in the step leaves the gas untouched and stays within its error
of the plain FP64 growth."""
import numpy as np
import pytest

import pyoracle
from shark_amd import gpu, growth as G, setups, surrogate as S

pytestmark = pytest.mark.gpu


def active(cfg):
    nxm, nym = cfg.nx + 4, cfg.ny + 4
    m = np.zeros((nym, nxm), bool)
    m[2:-2, 2:-2] = True
    return m.reshape(-1)


def test_fp64_growth_in_step_bit_exact():
    cfg, geom, state, p = setups.disk_growth(nr=32, nphi=48, nbins=16)
    nc = (cfg.nx + 4) * (cfg.ny + 4)
    o = pyoracle.Oracle(cfg, geom, state)
    o.set_growth(p)
    g = gpu.Step(cfg, geom, state)
    gr = G.Growth(p, nc)
    try:
        g.attach_growth(gr, G.FP64)
        for i in range(3):
            dg = g.step()
            do = o.step()
            assert dg == do, (i, dg, do)
        ug, uo = g.download(), o.u()
        act = active(cfg)
        assert np.array_equal(ug[:, act], uo[:, act])
        # the growth actually moved mass between bins
        nd = cfg.ndust
        assert not np.allclose(ug[5:5 + nd, act], state["u"][5:5 + nd, act], rtol=1e-3)
    finally:
        g.close()
        gr.close()


def test_learned_growth_in_step():
    cfg, geom, state, p = setups.disk_growth(nr=1024, nphi=16, nbins=64)
    nc = (cfg.nx + 4) * (cfg.ny + 4)
    net, norm, _ = S.load(S.DATA + "/surrogate_disk64.npz")
    out = {}
    try:
        for mode in (G.FP64, G.SURROGATE):
            g = gpu.Step(cfg, geom, state)
            gr = G.Growth(p, nc)
            try:
                if mode == G.SURROGATE:
                    gr.set_surrogate(net, norm)
                g.attach_growth(gr, mode)
                g.set_timing(True)
                g.run(1)
                out[mode] = (g.download(), g.kernel_time("growth")[0], gr.surrogate_stats())
            finally:
                g.close()
                gr.close()
    finally:
        net.close()
    (uf, tf, _), (us, ts, (nfb, _)) = out[G.FP64], out[G.SURROGATE]
    act = active(cfg)
    nd = cfg.ndust
    gas = [0, 1, 2, 3, 4]
    assert np.array_equal(uf[gas], us[gas])                        # growth does not touch the gas
    rf, rs = uf[5:5 + nd][:, act], us[5:5 + nd][:, act]
    # each species keeps its velocity across the growth call: momentum / density agrees between
    # the two modes to rounding (the momenta themselves scale with the new densities)
    for c in range(3):
        mf = uf[5 + nd + c * nd:5 + nd + (c + 1) * nd][:, act]
        ms = us[5 + nd + c * nd:5 + nd + (c + 1) * nd][:, act]
        ok = (rf > 0) & (rs > 0)
        np.testing.assert_allclose(ms[ok] / rs[ok], mf[ok] / rf[ok], rtol=1e-13, atol=0)
    err = S.mass_l1(rs, rf)
    mass = np.abs(rs.sum(0) / rf.sum(0) - 1)
    print(f"learned step in the step: {ts:.2f} ms vs FP64 {tf:.2f} ms; L1 mean {err.mean():.2e} "
          f"max {err.max():.2e}; fallback {nfb}; mass {mass.max():.1e}")
    assert err.mean() < 5e-3
    assert mass.max() < 1e-12
    assert tf > 10 * ts
