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

With --traj-steps, samples from the disk's own learned trajectory are added (surrogate.
trajectory_states: the 1024 x 2048 disk stepped with the learned step of --traj-weights, every
--traj-every steps --traj-cells cells advanced by the plain coagulation); with --init the networks
start from a weights file and keep its normalisations (a fine-tuning run):

    python -m shark_amd.train_surrogate --chunks 16 --traj-steps 300 --init shark_amd/data/surrogate_disk64.npz \
        --epochs 6,3 --lrs 3e-4,5e-5 --res-epochs 4,2 --res-lrs 2e-4,4e-5 --weighted
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
    ap.add_argument("--perturb", type=float, default=0.0,
                    help="fraction of the aged cells moved off the coagulation's states before the "
                         "interval (surrogate.sample_states)")
    ap.add_argument("--perturb-sigma", default="1e-3,0.05",
                    help="range of the per-cell log-density perturbation scale")
    ap.add_argument("--perturb-sigma-small", default="1e-2,1.0",
                    help="the same for bins holding less than 1e-5 of the cell's dust mass")
    ap.add_argument("--weight-floor", type=float, default=1e-6,
                    help="share added to every bin's loss weight (surrogate.mass_weights)")
    ap.add_argument("--init", default="",
                    help="weights file to start both networks from (its normalisations are kept)")
    ap.add_argument("--traj-steps", type=int, default=0,
                    help="steps of the learned disk trajectory to sample (0: no trajectory samples)")
    ap.add_argument("--traj-every", type=int, default=10)
    ap.add_argument("--traj-cells", type=int, default=131072)
    ap.add_argument("--traj-weights", default="",
                    help="weights of the learned step that drives the trajectory (default: --init, "
                         "else the shipped weights)")
    a = ap.parse_args(argv)

    p = G.disk_params(nbins=64, max_substeps=1000000)
    sd = tables_sdust(p)
    gen = G.FP64 if a.gen == "fp64" else G.MFMA_FP32
    psig = [float(v) for v in a.perturb_sigma.split(",")]
    psig_s = [float(v) for v in a.perturb_sigma_small.split(",")]
    t0 = time.time()
    Xs, Ys, Ss = [], [], []
    for c in range(a.chunks):
        X, Y, st = S.sample_states(p, sd, a.nr, a.nphi, seed=a.seed + c, age_max=a.age_max,
                                   dt_lo=a.dt_lo, dt_hi=a.dt_hi, gen_mode=gen, perturb=a.perturb,
                                   scale=psig, scale_small=psig_s)
        Xs.append(X)
        Ys.append(Y)
        if a.weighted:
            Ss.append(st["share_next"])
        print(f"chunk {c}: {X.shape[0]} samples, sub-steps {st['nsub'].min()}..{st['nsub'].max()} "
              f"({time.time() - t0:.0f} s)", flush=True)
    traj_w = a.traj_weights or a.init or None
    if a.traj_steps:
        for step, X, Y, st in S.trajectory_states(a.traj_steps, a.traj_every, a.traj_cells, seed=a.seed + 5000,
                                                  dt_lo=a.dt_lo, dt_hi=a.dt_hi, gen_mode=gen, weights=traj_w):
            Xs.append(X)
            Ys.append(Y)
            if a.weighted:
                Ss.append(st["share_next"])
            print(f"trajectory step {step}: {X.shape[0]} samples, sub-steps {st['nsub'].min()}..{st['nsub'].max()} "
                  f"({time.time() - t0:.0f} s)", flush=True)
    X = np.concatenate(Xs)
    Y = np.concatenate(Ys)
    del Xs, Ys
    share = np.concatenate(Ss) if a.weighted else None
    del Ss
    init1 = init2 = None
    if a.init:
        n0, norm, _ = S.load(a.init)
        r0, norm2_0 = S.load_residual(a.init)
        init1 = n0.get_params()
        n0.close()
        if r0 is not None:
            init2 = r0.get_params()
            r0.close()
    else:
        norm = S.norm_of(X, Y)
    W = S.mass_weights(share, norm, X.shape[1], floor=a.weight_floor) if a.weighted else None
    t1 = time.time()
    epochs = [int(e) for e in a.epochs.split(",")]
    lrs = [float(v) for v in a.lrs.split(",")]
    net, hist = S.train(X, Y, norm, epochs=epochs, lrs=lrs, batch=a.batch, seed=a.seed, verbose=True, W=W,
                         init=init1)
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
        if init2 is not None:
            mo2, so2 = norm2_0[:no], norm2_0[no:]
        else:
            mo2, so2 = R.mean(axis=0), R.std(axis=0)
            so2 = np.where(so2 > 1e-12, so2, 1.0)
        norm2_full = np.concatenate([norm[:2 * ni], mo2, so2]).astype(np.float32)
        W2 = S.mass_weights(share, norm2_full, ni, floor=a.weight_floor) if a.weighted else None
        t2 = time.time()
        net2, hist2 = S.train(X, R, norm2_full, epochs=[int(e) for e in a.res_epochs.split(",")],
                              lrs=[float(v) for v in a.res_lrs.split(",")], batch=a.batch, seed=a.seed + 1,
                              verbose=True, W=W2, init=init2)
        del W2, R
        norm2 = np.concatenate([mo2, so2]).astype(np.float32)
        print(f"residual network in {time.time() - t2:.0f} s: loss {hist2[0]:.3e} -> {hist2[-1]:.3e}", flush=True)
    t_train = time.time() - t1
    n_samples = X.shape[0]
    del X, Y, share

    # held-out cells advanced by the plain FP64 coagulation (aged disk cells, and with trajectory
    # samples cells of the learned trajectory half an interval before its last snapshot): errors of
    # every cell with no fallback, the raw mass residual with the first network's activations in
    # BF16 hi + lo as the learned step runs it, and the fallback threshold on it
    nphi_e = max(1, a.eval_cells // a.nr)
    sets = {}
    _, _, sets["eval"] = S.sample_states(p, sd, a.nr, nphi_e, seed=a.seed + 7777, age_max=a.age_max,
                                         dt_lo=a.dt_lo, dt_hi=a.dt_hi, gen_mode=G.FP64)
    if a.traj_steps:
        te = max(1, a.traj_steps - a.traj_every // 2)
        for _, _, _, st in S.trajectory_states(te, te, a.eval_cells, seed=a.seed + 8888, dt_lo=a.dt_lo,
                                               dt_hi=a.dt_hi, gen_mode=G.FP64, weights=traj_w):
            sets["eval_trajectory"] = st

    def run(st, resid_max):
        g = G.Growth(p, st["sigma"].size)
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

    no = 65
    raws, errs_all = {}, {}
    for k, st in sets.items():
        out, _, _, Xe = run(st, 1e30)
        errs_all[k] = S.mass_l1(out, st["rhod_next"])
        ni = Xe.shape[1]
        Xst, _ = S.standardise(Xe, np.zeros((Xe.shape[0], no), np.float32), norm)
        chg = net.forward(Xst, hilo=True) * norm[2 * ni + no:] + norm[2 * ni:2 * ni + no]
        if net2 is not None:
            chg = chg + net2.forward(Xst) * norm2[no:] + norm2[:no]
        sh = (st["rhod"] / st["rhod"].sum(0)).T
        raws[k] = np.abs((sh * np.exp(chg[:, :64].astype(np.float64))).sum(1) - 1.0)
    raw = np.concatenate(list(raws.values()))
    resid_max = float(a.resid_margin * raw.max())
    evals = {}
    for k, st in sets.items():
        out, ms_s, (nfb, rawm), _ = run(st, resid_max)
        err = S.mass_l1(out, st["rhod_next"])
        base = S.mass_l1(st["rhod"], st["rhod_next"])
        evals[k] = dict(cells=int(st["sigma"].size), l1_mean=float(err.mean()), l1_p50=float(np.median(err)),
                        l1_p99=float(np.quantile(err, 0.99)), l1_max=float(err.max()),
                        l1_mean_no_fallback=float(errs_all[k].mean()),
                        no_change_l1_mean=float(base.mean()), fallback=int(nfb),
                        raw_mass_residual_mean=float(rawm), surrogate_ms=ms_s,
                        raw_resid=dict(mean=float(raws[k].mean()), p99=float(np.quantile(raws[k], 0.99)),
                                       max=float(raws[k].max())))
    meta = dict(samples=int(n_samples), chunks=a.chunks, generator=a.gen,
                nr=a.nr, nphi=a.nphi, dt_range=[a.dt_lo, a.dt_hi], age_max=a.age_max,
                epochs=epochs, lrs=lrs, batch=a.batch, seed=a.seed, weighted=bool(a.weighted),
                residual_epochs=a.res_epochs, loss_first=float(hist[0]), loss_last=float(hist[-1]),
                residual_loss_last=None if hist2 is None else float(hist2[-1]), train_seconds=t_train,
                resid_max=resid_max, resid_margin=a.resid_margin,
                raw_resid_heldout=dict(mean=float(raw.mean()), p99=float(np.quantile(raw, 0.99)),
                                       max=float(raw.max())),
                init=os.path.basename(a.init) if a.init else None,
                perturb=dict(fraction=a.perturb, scale=psig, scale_small=psig_s) if a.perturb else None,
                weight_floor=a.weight_floor,
                trajectory=dict(steps=a.traj_steps, every=a.traj_every, cells=a.traj_cells,
                                weights=os.path.basename(traj_w) if traj_w else "shipped") if a.traj_steps else None,
                **evals)
    for k in evals:
        print(k, json.dumps(meta[k]), flush=True)
    os.makedirs(os.path.dirname(os.path.abspath(a.out)), exist_ok=True)
    S.save(a.out, net, norm, meta, net2, norm2)
    print("wrote", a.out)
    net.close()
    if net2 is not None:
        net2.close()
    return 0


if __name__ == "__main__":
    sys.exit(main())
