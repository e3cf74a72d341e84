"""Learned growth-and-charging step: training set from the plain solver, training on the GPU,
weights on disk, and attachment to a growth state.

Samples come from polar-disk cells (the disk workload's radial grid and surface density): an
MRN size distribution aged by the plain coagulation for a random time per cell, then advanced by
the plain coagulation over a random interval dt.  Inputs are the growth module's own features of
the aged state (shark_growth_features); targets are the change of the natural log of each bin's
share of the dust mass and log10(-<Z>) of the advanced state (charge solve of
shark_growth_charge).  Trajectory samples (trajectory_states) start instead from the states the
1024 x 2048 disk reaches when it is stepped with the learned step itself.
"""
import json
import os

import numpy as np

from . import growth as G
from . import mlp as M

DATA = os.path.join(os.path.dirname(os.path.abspath(__file__)), "data")


def sample_states(params, tables_sdust, nr, nphi, seed, age_max=1e-2, dt_lo=1e-5, dt_hi=1e-3,
                  gen_mode=G.MFMA_FP32, perturb=0.0, scale=(1e-3, 0.05), scale_small=(1e-2, 1.0),
                  small_share=1e-5):
    """(X raw features, Y raw targets, state dict) for nr*nphi cells.  With perturb > 0 that
    fraction of the aged cells is moved off the coagulation's own states before the interval: each
    bin's density times exp(s z) (z standard normal per bin; s per cell log-uniform in `scale`, or
    in `scale_small` for bins holding less than small_share of the cell's dust mass), the cell's
    dust mass kept.  The targets of such cells are still the plain step's, so the network learns
    how the coagulation pulls a perturbed distribution back: the bins above the fragmentation
    barrier hold ~1e-8 of the mass in a fast gain-loss balance that a learned change of log-share
    would otherwise leave to drift (noise injection against the compounding of a learned step's
    errors over many steps; Sanchez-Gonzalez et al. 2020, Learning to Simulate Complex Physics with
    Graph Networks, sec. 3)."""
    rng = np.random.default_rng(seed)
    rhod, sigma, cs, r = G.disk_state(params, nr=nr, nphi=nphi, bins=tables_sdust, seed=seed)
    n = sigma.size
    age = np.where(rng.random(n) < 0.15, 0.0, 10.0 ** rng.uniform(-7, np.log10(age_max), n))
    dt = 10.0 ** rng.uniform(np.log10(dt_lo), np.log10(dt_hi), n)
    g = G.Growth(params, n)
    try:
        g.upload(rhod, sigma, cs, r)
        g.set_cell_dt(age)
        g.run(gen_mode, 0.0)
        s1, _ = g.download()
    finally:
        g.close()
    if perturb > 0.0:
        pick = rng.random(n) < perturb
        sg = np.where(pick, 10.0 ** rng.uniform(np.log10(scale[0]), np.log10(scale[1]), n), 0.0)
        ss = np.where(pick, 10.0 ** rng.uniform(np.log10(scale_small[0]), np.log10(scale_small[1]), n), 0.0)
        small = s1 < small_share * s1.sum(axis=0)[None, :]
        z = rng.standard_normal(s1.shape)
        s1p = s1 * np.exp(np.where(small, ss[None, :], sg[None, :]) * z)
        s1 = np.ascontiguousarray(s1p * (s1.sum(axis=0) / s1p.sum(axis=0))[None, :])
    return advance_states(params, s1, sigma, cs, r, dt, gen_mode)


def advance_states(params, rhod, sigma, cs, r, dt, gen_mode=G.MFMA_FP32):
    """(X raw features, Y raw targets, state dict) of cells with dust densities rhod (nbins, n),
    gas surface density sigma, sound speed cs and radius r, advanced by the plain coagulation over
    the per-cell intervals dt."""
    g = G.Growth(params, sigma.size)
    try:
        g.upload(rhod, sigma, cs, r)
        g.set_cell_dt(dt)
        X = g.features(0.0)
        g.run(gen_mode, 0.0)
        s2, nsub = g.download()
        g.upload(s2, sigma, cs, r)
        zm, _ = g.charge()
    finally:
        g.close()
    sh1 = rhod / rhod.sum(axis=0)
    sh2 = s2 / s2.sum(axis=0)
    Y = np.concatenate([(np.log(sh2) - np.log(sh1)).T, np.log10(-zm)[:, None]], axis=1).astype(np.float32)
    state = dict(rhod=rhod, sigma=sigma, cs=cs, r=r, dt=dt, rhod_next=s2, nsub=nsub,
                 share_next=sh2.T.astype(np.float32))
    return X, Y, state


def trajectory_states(steps, every, cells, seed, dt_lo=1e-4, dt_hi=1e-3, gen_mode=G.MFMA_FP32,
                      weights=None, nr=1024, nphi=2048, nbins=64):
    """Samples along the disk's own learned trajectory: the nr x nphi polar disk with nbins dust
    species that are the growth bins (setups.disk_growth) is stepped with the learned growth step
    of `weights` inside every step and its fallback off; after every `every` steps, `cells` random
    active cells of the state reached are advanced by the plain coagulation over random intervals
    (log-uniform in [dt_lo, dt_hi]).  Yields (step, X, Y, state) per snapshot.  Such states carry
    what the hydro step and the learned step's own errors do to the size distributions, which the
    aged box states of sample_states do not.  The trajectory ends early, after a message, at the
    first step in which a cell falls back to the plain coagulation (an input outside the trained
    envelope or a non-finite output: the learned step no longer drives every cell)."""
    from . import gpu, setups
    cfg, geom, state, p = setups.disk_growth(nr=nr, nphi=nphi, nbins=nbins)
    nxm, nym = cfg.nx + 4, cfg.ny + 4
    c = np.arange(nxm * nym)
    ix, iy = c % nxm, c // nxm
    act = np.flatnonzero((ix >= 2) & (ix < nxm - 2) & (iy >= 2) & (iy < nym - 2))
    rng = np.random.default_rng(seed)
    cs_all = np.asarray(state["cs"], np.float64)
    r_all = np.asarray(geom["radii_c"], np.float64)
    step = gpu.Step(cfg, geom, state)
    gr = G.Growth(p, nxm * nym)
    net = Learned(weights)
    try:
        net.attach(gr, resid_max=1e30)
        step.attach_growth(gr, G.SURROGATE)
        for i in range(1, steps + 1):
            step.step()
            nfb, _ = gr.surrogate_stats()
            if nfb:
                print(f"trajectory: {nfb} cells fell back to the plain coagulation at step {i}; "
                      f"the trajectory ends there", flush=True)
                return
            if i % every:
                continue
            u = step.download()
            idx = np.sort(rng.choice(act, size=cells, replace=False))
            rhod = np.ascontiguousarray(u[5:5 + nbins, idx])
            sigma = u[0, idx].copy()
            del u
            dt = 10.0 ** rng.uniform(np.log10(dt_lo), np.log10(dt_hi), cells)
            X, Y, st = advance_states(p, rhod, sigma, cs_all[idx].copy(), r_all[idx].copy(), dt, gen_mode)
            yield i, X, Y, st
    finally:
        step.close()
        gr.close()
        net.close()


def mass_weights(share_next, norm, n_in, floor=1e-6):
    """Per-element loss weights that make the squared error of the standardised outputs a
    mass-weighted error of the log-shares: bin k of a sample gets (share_k + floor) * s_k^2, with
    s_k the output's standard deviation (the change of log-share per standard unit), normalised to
    mean 1 over the bins; the charge output keeps weight 1."""
    nb = share_next.shape[1]
    so = norm[2 * n_in + nb + 1:][:nb].astype(np.float64)
    w = np.empty((share_next.shape[0], nb + 1), dtype=np.float32)
    wb = (share_next.astype(np.float64) + floor) * (so * so)[None, :]
    w[:, :nb] = wb / wb.mean()
    w[:, nb] = 1.0
    return w


def norm_of(X, Y):
    mi, si = X.mean(axis=0), X.std(axis=0)
    mo, so = Y.mean(axis=0), Y.std(axis=0)
    si = np.where(si > 1e-6, si, 1.0)
    so = np.where(so > 1e-12, so, 1.0)
    return np.concatenate([mi, si, mo, so]).astype(np.float32)


def standardise(X, Y, norm):
    ni, no = X.shape[1], Y.shape[1]
    mi, si = norm[:ni], norm[ni:2 * ni]
    mo, so = norm[2 * ni:2 * ni + no], norm[2 * ni + no:]
    return ((X - mi) / si).astype(np.float32), ((Y - mo) / so).astype(np.float32)


def train(X, Y, norm, hidden=256, layers=4, epochs=(30, 20, 10), lrs=(2e-3, 5e-4, 1e-4),
          batch=4096, seed=0, max_batch=65536, progress_every=5, verbose=False, W=None, init=None):
    """Adam on the GPU (W: per-element loss weights or None; init: starting parameters, else a
    seeded initialisation); returns (MLP, loss history)."""
    Xs, Ys = standardise(X, Y, norm)
    cfg = M.make_cfg(n_in=X.shape[1], n_hidden=hidden, n_layers=layers, n_out=Y.shape[1],
                     max_batch=max_batch)
    net = M.MLP(cfg, M.init_params(cfg, seed) if init is None else init)
    net.set_data(Xs, Ys, W)
    del Xs, Ys
    hist = []
    for i, (e, lr) in enumerate(zip(epochs, lrs)):
        done = 0
        while done < e:                      # a few epochs per call, reporting as it goes
            n = min(progress_every, e - done)
            hist += list(net.train(n, batch, lr, seed=seed + 101 * i + 7 * done))
            done += n
            if verbose:
                print(f"  lr {lr:g}: epoch {len(hist)} loss {hist[-1]:.4e}", flush=True)
    return net, np.array(hist)


def save(path, net, norm, meta, net2=None, out_norm2=None):
    """Weights, normalisation and metadata; with net2, the residual network and its output
    normalisation (out_mean2, out_std2) as well."""
    extra = {}
    if net2 is not None:
        extra = dict(params2=net2.get_params(), norm2=np.asarray(out_norm2, np.float32),
                     dims2=np.array([net2.cfg.n_in, net2.cfg.n_hidden, net2.cfg.n_layers, net2.cfg.n_out]))
    np.savez_compressed(path, params=net.get_params(), norm=norm,
                        dims=np.array([net.cfg.n_in, net.cfg.n_hidden, net.cfg.n_layers, net.cfg.n_out]),
                        meta=json.dumps(meta), **extra)


def load(path, max_batch=65536):
    z = np.load(path, allow_pickle=False)
    n_in, n_hidden, n_layers, n_out = [int(v) for v in z["dims"]]
    cfg = M.make_cfg(n_in=n_in, n_hidden=n_hidden, n_layers=n_layers, n_out=n_out, max_batch=max_batch)
    net = M.MLP(cfg, z["params"])
    return net, z["norm"], json.loads(str(z["meta"]))


def load_residual(path, max_batch=65536):
    """The residual network of a weights file and its output normalisation, or (None, None)."""
    z = np.load(path, allow_pickle=False)
    if "params2" not in z:
        return None, None
    n_in, n_hidden, n_layers, n_out = [int(v) for v in z["dims2"]]
    cfg = M.make_cfg(n_in=n_in, n_hidden=n_hidden, n_layers=n_layers, n_out=n_out, max_batch=max_batch)
    return M.MLP(cfg, z["params2"]), z["norm2"]


class Learned:
    """The learned step's networks from a weights file, attached to growth states with their
    calibrated fallback threshold."""

    def __init__(self, path=None, max_batch=65536):
        path = path or os.path.join(DATA, "surrogate_disk64.npz")
        self.net, self.norm, self.meta = load(path, max_batch)
        self.net2, self.norm2 = load_residual(path, max_batch)

    def attach(self, g, resid_max=None):
        g.set_surrogate(self.net, self.norm, resid_max=resid_max if resid_max is not None else
                        globals()["resid_max"](self.meta))
        if self.net2 is not None:
            g.set_surrogate_residual(self.net2, self.norm2)

    def close(self):
        for n in (self.net, self.net2):
            if n is not None:
                n.close()


RESID_MAX_DEFAULT = 3e-3


def resid_max(meta):
    """The raw mass-residual threshold of the learned step's fallback calibrated for these weights
    at training time (meta["resid_max"]), or the default."""
    return float(meta.get("resid_max", RESID_MAX_DEFAULT))


def mass_l1(a, ref):
    """Mass-weighted size-distribution L1 error per cell: sum_k |a_k - ref_k| / sum_k ref_k."""
    return np.abs(a - ref).sum(axis=0) / ref.sum(axis=0)
