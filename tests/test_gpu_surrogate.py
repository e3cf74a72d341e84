"""Learned growth step: trained here on the GPU from plain-solver samples (small set, so the
test runs in seconds), then compared with the plain FP64 coagulation on held-out cells.

The ablation: the same growth call with the network (mode SURROGATE) and without it (mode FP64,
the plain coagulation that is also its fallback and its reference).  The network must be >= 10x
faster, beat the no-change prediction by a wide margin, conserve dust mass to 1e-12, and send the
cells it cannot handle (outside its training envelope, or a failed mass residual) to the
fallback."""
import numpy as np
import pytest

import pyoracle
from shark_amd import growth as G
from shark_amd import surrogate as S

pytestmark = pytest.mark.gpu

NB = 64


@pytest.fixture(scope="module")
def trained():
    p = G.disk_params(nbins=NB, max_substeps=200000)
    t = pyoracle.growth_tables(p)
    sd = np.array(t.sdust[:NB])
    X, Y, _ = S.sample_states(p, sd, nr=128, nphi=128, seed=1, dt_lo=1e-6, dt_hi=1e-5, age_max=1e-4)
    norm = S.norm_of(X, Y)
    net, hist = S.train(X, Y, norm, epochs=(25, 10), lrs=(2e-3, 3e-4), batch=1024, max_batch=16384)
    print("train loss", hist[0], "->", hist[-1])
    yield p, sd, net, norm, hist
    net.close()


def test_training_converges(trained):
    _, _, _, _, hist = trained
    assert hist[-1] < 0.2 * hist[0]


def test_surrogate_vs_plain_fp64(trained):
    p, sd, net, norm, _ = trained
    X, Y, st = S.sample_states(p, sd, nr=32, nphi=16, seed=7, dt_lo=1e-6, dt_hi=1e-5, age_max=1e-4,
                               gen_mode=G.FP64)
    n = st["sigma"].size
    g = G.Growth(p, n)
    try:
        g.set_surrogate(net, norm, resid_max=1e30)    # the network alone (no residual fallback)
        g.set_cell_dt(st["dt"])
        g.upload(st["rhod"], st["sigma"], st["cs"], st["r"])
        g.run(G.SURROGATE, 0.0)
        out, nsub = g.download()
        ms_s = g.kernel_ms()
        nfb, raw = g.surrogate_stats()
        g.upload(st["rhod"], st["sigma"], st["cs"], st["r"])
        g.run(G.FP64, 0.0)
        ref, _ = g.download()
        ms_p = g.kernel_ms()
    finally:
        g.close()
    assert np.array_equal(ref, st["rhod_next"])           # the reference is the plain step itself
    err = S.mass_l1(out, ref)
    base = S.mass_l1(st["rhod"], ref)                     # no-change prediction
    mass = np.abs(out.sum(axis=0) / st["rhod"].sum(axis=0) - 1)
    print(f"surrogate {ms_s:.2f} ms vs plain fp64 {ms_p:.2f} ms ({ms_p / ms_s:.0f}x); "
          f"L1 err mean {err.mean():.2e} max {err.max():.2e}; no-change {base.mean():.2e}; "
          f"fallback {nfb}/{n}; raw residual {raw:.2e}; mass {mass.max():.1e}")
    assert ms_p >= 10 * ms_s
    assert err.mean() <= 0.1 * base.mean()
    assert mass.max() <= 1e-12
    assert (nsub[nsub > 0].size == nfb)


def test_fallback_cells_are_exactly_plain(trained):
    """A training envelope shrunk a thousand-fold puts every input outside it: every cell falls
    back and the result is exactly the plain step."""
    p, sd, net, norm, _ = trained
    rhod, sigma, cs, r = G.disk_state(p, nr=8, nphi=8, bins=sd, seed=3)
    dt = 1e-6
    bad = norm.copy()
    ni = NB + 5
    bad[ni:2 * ni] *= 1e-3
    g = G.Growth(p, sigma.size)
    try:
        g.set_surrogate(net, bad)
        g.upload(rhod, sigma, cs, r)
        g.run(G.SURROGATE, dt)
        out, _ = g.download()
        nfb, _ = g.surrogate_stats()
        g.upload(rhod, sigma, cs, r)
        g.run(G.FP64, dt)
        ref, _ = g.download()
    finally:
        g.close()
    assert nfb == sigma.size
    assert np.array_equal(out, ref)


def raw_residuals(net, norm, X, rhod):
    """|sum_k share_k exp(change_k) - 1| of the network's raw prediction (before normalisation),
    evaluated as the learned step evaluates it (hi + lo activations)."""
    ni, no = X.shape[1], NB + 1
    Xs, _ = S.standardise(X, np.zeros((X.shape[0], no), np.float32), norm)
    chg = net.forward(Xs, hilo=True) * norm[2 * ni + no:] + norm[2 * ni:2 * ni + no]
    sh = (rhod / rhod.sum(0)).T
    return np.abs((sh * np.exp(chg[:, :NB].astype(np.float64))).sum(1) - 1.0)


def test_raw_mass_residual_fallback(trained):
    """Coagulation conserves mass, so a consistent prediction has a zero raw mass residual; cells
    above the threshold run the plain step and equal it exactly, the others keep the network's
    answer.  With the median residual as threshold about half the cells fall back."""
    p, sd, net, norm, _ = trained
    X, _, st = S.sample_states(p, sd, nr=32, nphi=16, seed=11, dt_lo=1e-6, dt_hi=1e-5, age_max=1e-4,
                               gen_mode=G.FP64)
    n = st["sigma"].size
    raw = raw_residuals(net, norm, X, st["rhod"])
    th = float(np.median(raw))
    g = G.Growth(p, n)
    try:
        g.set_cell_dt(st["dt"])
        g.set_surrogate(net, norm, resid_max=th)
        g.upload(st["rhod"], st["sigma"], st["cs"], st["r"])
        g.run(G.SURROGATE, 0.0)
        out, nsub = g.download()
        nfb, _ = g.surrogate_stats()
        g.set_surrogate(net, norm, resid_max=1e30)
        g.upload(st["rhod"], st["sigma"], st["cs"], st["r"])
        g.run(G.SURROGATE, 0.0)
        net_only, _ = g.download()
    finally:
        g.close()
    fb = nsub > 0                                   # fallback cells ran plain sub-steps
    print(f"threshold {th:.2e}: {nfb}/{n} fell back, host count {(raw > th).sum()}", flush=True)
    assert fb.sum() == nfb
    assert abs(nfb - (raw > th).sum()) <= 0.02 * n  # the device's FP32 sum rounds near the threshold
    assert 0.3 * n < nfb < 0.7 * n
    assert np.array_equal(out[:, fb], st["rhod_next"][:, fb])
    assert np.array_equal(out[:, ~fb], net_only[:, ~fb])
