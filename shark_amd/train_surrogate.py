"""Train the learned growth-and-charging step on the GPU and write its weights.

    python -m shark_amd.train_surrogate --chunks 16 --out shark_amd/data/surrogate_disk64.npz

Samples: `chunks` patches of nr x nphi cells of the 64-bin polar disk (the 1024 x 2048 disk's
radial grid, a different seed per patch), each cell aged by the plain coagulation for a random
time (0 for 10% of cells, else log-uniform in [1e-7, age_max]) and advanced over a random interval
log-uniform in [dt_lo, dt_hi] (the range of the disk's hydro step).  Targets come from the plain
coagulation, its FP32 MFMA contraction by default (--gen fp64: the FP64 pair loop; the two agree
to a mass-weighted L1 of 2e-8 on these states).  With --res-epochs a second (residual) network is
trained on what the first one misses.  A held-out patch advanced by the plain FP64 coagulation
gives the errors and the fallback threshold on the raw mass residual (a margin above the largest
residual of held-out cells).
"""
import argparse
import json
import os
import sys
import time

import numpy as np

from . import growth as G
from . import surrogate as S


def tables_sdust(p):
    """Bin centre sizes (code units) of the growth bins, as build_tables defines them."""
    nb = p.nbins
    zeta = (p.smax / p.smin) ** (1.0 / nb)
    k = np.arange(nb)
    am = p.smin * zeta ** k / p.unit_l
    ap = p.smin * zeta ** (k + 1) / p.unit_l
    s = np.sqrt(am * ap)
    return s + (p.ice_mantle / p.unit_l if p.kernel_type == 0 else 0.0)


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--chunks", type=int, default=16)
    ap.add_argument("--nr", type=int, default=1024)
    ap.add_argument("--nphi", type=int, default=256)
    ap.add_argument("--dt-lo", type=float, default=1e-4)
    ap.add_argument("--dt-hi", type=float, default=1e-3)
    ap.add_argument("--age-max", type=float, default=1.0)
    ap.add_argument("--epochs", type=str, default="24,12,6")
    ap.add_argument("--lrs", type=str, default="2e-3,4e-4,8e-5")
    ap.add_argument("--res-epochs", type=str, default="",
                    help="epochs of a residual network trained on what the first one misses (empty: none)")
    ap.add_argument("--res-lrs", type=str, default="1e-3,2e-4,4e-5")
    ap.add_argument("--batch", type=int, default=4096)
    ap.add_argument("--eval-cells", type=int, default=8192)
    ap.add_argument("--resid-margin", type=float, default=3.0,
                    help="the fallback threshold on the raw mass residual is this factor times the largest "
                         "residual of the held-out cells: a guard against inconsistent predictions (one "
                         "fallback cell costs about a whole plain step: its sub-steps run in sequence)")
    ap.add_argument("--gen", choices=("fp64", "mfma_fp32"), default="mfma_fp32",
                    help="generator of the training targets: the FP64 plain coagulation or its FP32 "
                         "MFMA contraction (mass-weighted L1 within 2e-8 of FP64 on these states)")
    ap.add_argument("--out", default=os.path.join(S.DATA, "surrogate_disk64.npz"))
    ap.add_argument("--seed", type=int, default=20261202)
    ap.add_argument("--weighted", action="store_true",
                    help="mass-weighted loss (surrogate.mass_weights) instead of plain MSE")
    a = ap.parse_args(argv)

    p = G.disk_params(nbins=64, max_substeps=1000000)
    sd = tables_sdust(p)
    gen = G.FP64 if a.gen == "fp64" else G.MFMA_FP32
    t0 = time.time()
    Xs, Ys, Ss = [], [], []
    for c in range(a.chunks):
        X, Y, st = S.sample_states(p, sd, a.nr, a.nphi, seed=a.seed + c, age_max=a.age_max,
                                   dt_lo=a.dt_lo, dt_hi=a.dt_hi, gen_mode=gen)
        Xs.append(X)
        Ys.append(Y)
        if a.weighted:
            Ss.append(st["share_next"])
        print(f"chunk {c}: {X.shape[0]} samples, sub-steps {st['nsub'].min()}..{st['nsub'].max()} "
              f"({time.time() - t0:.0f} s)", flush=True)
    X = np.concatenate(Xs)
    Y = np.concatenate(Ys)
    del Xs, Ys
    share = np.concatenate(Ss) if a.weighted else None
    del Ss
    norm = S.norm_of(X, Y)
    W = S.mass_weights(share, norm, X.shape[1]) if a.weighted else None
    t1 = time.time()
    epochs = [int(e) for e in a.epochs.split(",")]
    lrs = [float(v) for v in a.lrs.split(",")]
    net, hist = S.train(X, Y, norm, epochs=epochs, lrs=lrs, batch=a.batch, seed=a.seed, verbose=True, W=W)
    del W
    print(f"trained on {X.shape[0]} samples in {time.time() - t1:.0f} s: loss {hist[0]:.3e} -> {hist[-1]:.3e}",
          flush=True)
    net2 = norm2 = None
    hist2 = None
    if a.res_epochs:
        # residual stage: a second network learns Y minus the first network's prediction
        ni, no = X.shape[1], Y.shape[1]
        Xst, _ = S.standardise(X, Y[:1], norm)
        pred = net.forward(Xst) * norm[2 * ni + no:] + norm[2 * ni:2 * ni + no]
        R = (Y - pred).astype(np.float32)
        del pred, Xst
        mo2, so2 = R.mean(axis=0), R.std(axis=0)
        so2 = np.where(so2 > 1e-12, so2, 1.0)
        norm2_full = np.concatenate([norm[:2 * ni], mo2, so2]).astype(np.float32)
        W2 = S.mass_weights(share, norm2_full, ni) if a.weighted else None
        t2 = time.time()
        net2, hist2 = S.train(X, R, norm2_full, epochs=[int(e) for e in a.res_epochs.split(",")],
                              lrs=[float(v) for v in a.res_lrs.split(",")], batch=a.batch, seed=a.seed + 1,
                              verbose=True, W=W2)
        del W2, R
        norm2 = np.concatenate([mo2, so2]).astype(np.float32)
        print(f"residual network in {time.time() - t2:.0f} s: loss {hist2[0]:.3e} -> {hist2[-1]:.3e}", flush=True)
    t_train = time.time() - t1
    del X, Y, share

    # held-out cells advanced by the plain FP64 coagulation: errors of every cell with no fallback,
    # the raw mass residual, and the threshold that sends `fallback_frac` of the cells to the plain step
    nphi_e = max(1, a.eval_cells // a.nr)
    _, _, st = S.sample_states(p, sd, a.nr, nphi_e, seed=a.seed + 7777, age_max=a.age_max,
                               dt_lo=a.dt_lo, dt_hi=a.dt_hi, gen_mode=G.FP64)
    n = st["sigma"].size

    def run(resid_max):
        g = G.Growth(p, n)
        try:
            g.set_surrogate(net, norm, resid_max=resid_max)
            if net2 is not None:
                g.set_surrogate_residual(net2, norm2)
            g.set_cell_dt(st["dt"])
            g.upload(st["rhod"], st["sigma"], st["cs"], st["r"])
            X = g.features(0.0)
            g.run(G.SURROGATE, 0.0)
            out, _ = g.download()
            return out, g.kernel_ms(), g.surrogate_stats(), X
        finally:
            g.close()

    out, _, _, Xe = run(1e30)
    err_all = S.mass_l1(out, st["rhod_next"])
    ni = Xe.shape[1]
    Xst, _ = S.standardise(Xe, np.zeros((n, 65), np.float32), norm)
    no = 65
    chg = net.forward(Xst) * norm[2 * ni + no:] + norm[2 * ni:2 * ni + no]
    if net2 is not None:
        chg = chg + net2.forward(Xst) * norm2[no:] + norm2[:no]
    sh = (st["rhod"] / st["rhod"].sum(0)).T
    raw = np.abs((sh * np.exp(chg[:, :64].astype(np.float64))).sum(1) - 1.0)
    resid_max = float(a.resid_margin * raw.max())
    out, ms_s, (nfb, rawm), _ = run(resid_max)
    err = S.mass_l1(out, st["rhod_next"])
    base = S.mass_l1(st["rhod"], st["rhod_next"])
    meta = dict(samples=a.chunks * a.nr * a.nphi, chunks=a.chunks, generator=a.gen,
                nr=a.nr, nphi=a.nphi, dt_range=[a.dt_lo, a.dt_hi], age_max=a.age_max,
                epochs=epochs, lrs=lrs, batch=a.batch, seed=a.seed, weighted=bool(a.weighted),
                residual_epochs=a.res_epochs, loss_first=float(hist[0]), loss_last=float(hist[-1]),
                residual_loss_last=None if hist2 is None else float(hist2[-1]), train_seconds=t_train,
                resid_max=resid_max, resid_margin=a.resid_margin,
                raw_resid_heldout=dict(mean=float(raw.mean()), p99=float(np.quantile(raw, 0.99)),
                                       max=float(raw.max())),
                eval=dict(cells=n, l1_mean=float(err.mean()), l1_p50=float(np.median(err)),
                          l1_p99=float(np.quantile(err, 0.99)), l1_max=float(err.max()),
                          l1_mean_no_fallback=float(err_all.mean()),
                          no_change_l1_mean=float(base.mean()), fallback=int(nfb),
                          raw_mass_residual_mean=float(rawm), surrogate_ms=ms_s))
    print(json.dumps(meta["eval"]), flush=True)
    os.makedirs(os.path.dirname(os.path.abspath(a.out)), exist_ok=True)
    S.save(a.out, net, norm, meta, net2, norm2)
    print("wrote", a.out)
    net.close()
    if net2 is not None:
        net2.close()
    return 0


if __name__ == "__main__":
    sys.exit(main())
