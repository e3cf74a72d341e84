"""Train the learned growth-and-charging step on the GPU and write its weights.

    python -m shark_amd.train_surrogate --chunks 16 --out shark_amd/data/surrogate_disk64.npz

Samples: `chunks` patches of nr x nphi cells of the 64-bin polar disk (the 1024 x 2048 disk's
radial grid, a different seed per patch), each cell aged by the plain coagulation for a random
time (0 for 10% of cells, else log-uniform in [1e-7, age_max]) and advanced over a random interval
log-uniform in [dt_lo, dt_hi] (the range of the disk's hydro step).  Generated with the FP32 MFMA
contraction (mass-weighted L1 within 2e-8 of the FP64 coagulation on these states).  A held-out
patch is evaluated against the plain FP64 coagulation.
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
    ap.add_argument("--batch", type=int, default=4096)
    ap.add_argument("--eval-cells", type=int, default=8192)
    ap.add_argument("--out", default=os.path.join(S.DATA, "surrogate_disk64.npz"))
    ap.add_argument("--seed", type=int, default=20261202)
    ap.add_argument("--weighted", action="store_true",
                    help="mass-weighted loss (surrogate.mass_weights) instead of plain MSE")
    a = ap.parse_args(argv)

    p = G.disk_params(nbins=64, max_substeps=1000000)
    sd = tables_sdust(p)
    t0 = time.time()
    Xs, Ys, Ss = [], [], []
    for c in range(a.chunks):
        X, Y, st = S.sample_states(p, sd, a.nr, a.nphi, seed=a.seed + c, age_max=a.age_max,
                                   dt_lo=a.dt_lo, dt_hi=a.dt_hi)
        Xs.append(X)
        Ys.append(Y)
        if a.weighted:
            Ss.append(st["share_next"])
        print(f"chunk {c}: {X.shape[0]} samples, sub-steps {st['nsub'].min()}..{st['nsub'].max()} "
              f"({time.time() - t0:.0f} s)", flush=True)
    X = np.concatenate(Xs)
    Y = np.concatenate(Ys)
    del Xs, Ys
    norm = S.norm_of(X, Y)
    W = S.mass_weights(np.concatenate(Ss), norm, X.shape[1]) if a.weighted else None
    del Ss
    t1 = time.time()
    epochs = [int(e) for e in a.epochs.split(",")]
    lrs = [float(v) for v in a.lrs.split(",")]
    net, hist = S.train(X, Y, norm, epochs=epochs, lrs=lrs, batch=a.batch, seed=a.seed, verbose=True, W=W)
    del W
    t_train = time.time() - t1
    print(f"trained on {X.shape[0]} samples in {t_train:.0f} s: loss {hist[0]:.3e} -> {hist[-1]:.3e}",
          flush=True)
    del X, Y

    # held-out evaluation against the plain FP64 coagulation
    nphi_e = max(1, a.eval_cells // a.nr)
    _, _, st = S.sample_states(p, sd, a.nr, nphi_e, seed=a.seed + 7777, age_max=a.age_max,
                               dt_lo=a.dt_lo, dt_hi=a.dt_hi, gen_mode=G.FP64)
    n = st["sigma"].size
    g = G.Growth(p, n)
    try:
        g.set_surrogate(net, norm)
        g.set_cell_dt(st["dt"])
        g.upload(st["rhod"], st["sigma"], st["cs"], st["r"])
        g.run(G.SURROGATE, 0.0)
        out, _ = g.download()
        ms_s = g.kernel_ms()
        nfb, raw = g.surrogate_stats()
    finally:
        g.close()
    err = S.mass_l1(out, st["rhod_next"])
    base = S.mass_l1(st["rhod"], st["rhod_next"])
    meta = dict(samples=a.chunks * a.nr * a.nphi, chunks=a.chunks,
                nr=a.nr, nphi=a.nphi, dt_range=[a.dt_lo, a.dt_hi], age_max=a.age_max,
                epochs=epochs, lrs=lrs, batch=a.batch, seed=a.seed, weighted=bool(a.weighted),
                loss_first=float(hist[0]),
                loss_last=float(hist[-1]), train_seconds=t_train,
                eval=dict(cells=n, l1_mean=float(err.mean()), l1_p50=float(np.median(err)),
                          l1_p99=float(np.quantile(err, 0.99)), l1_max=float(err.max()),
                          no_change_l1_mean=float(base.mean()), fallback=int(nfb),
                          raw_mass_residual_mean=float(raw), surrogate_ms=ms_s))
    print(json.dumps(meta["eval"]), flush=True)
    os.makedirs(os.path.dirname(os.path.abspath(a.out)), exist_ok=True)
    S.save(a.out, net, norm, meta)
    print("wrote", a.out)
    net.close()
    return 0


if __name__ == "__main__":
    sys.exit(main())
