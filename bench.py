#!/usr/bin/env python3
"""Benchmark of the FP64 SHARK step on one MI355X: zone-updates/s (gas + dust fluids).

Default workload: the 2D polar disk with FARGO, Nr=1024 x Nphi=2048, 16 dust species
(17 fluids), per-cell 17-fluid implicit drag — the configuration the 60%-of-HBM and 20x-CPU
targets are stated on.  `--workload si` runs the 1024^2 8-species shearing box, `dustywave`
the 512-cell 1D wave, `c5` the 1024 x 2048 disk with 64 dust bins and the learned
growth-and-charging step inside every step (velocity-keeping density update between drag and the
in-loop hook).

Launch shapes: before the warm-up the library's autotuner (shark_gpu_autotune) times the
candidate hydro strip heights, post segment widths and block orders on the workload's own state
and keeps the fastest; the state is restored, and the chosen shape is reported under "tune".

Timed region: W warm-up steps, then exactly K steps replaying the step's hipGraph (one graph per
step), host synchronised on both sides; then up to 50 more steps with every kernel bracketed by
HIP events on the library's stream, whose average durations feed the roofline.  `--events` times
the event-bracketed launches instead.  One JSON line is printed.

cpu_baseline: the CPU restatement of the same step (oracle/liboracle.so, OpenMP over the
host cores) on the same workload for a bounded number of steps.

growth (second half of the metric, "growth-step speedup at fixed L1 err"): on the 512 x 512
polar disk with 64 mass bins (MRN initial distribution), one growth call per mode over the
same cells and interval: the plain FP64 coagulation (reference order) and the FP32 / BF16
MFMA contractions, each with its sub-step count, time per cell-sub-step, speedup over FP64,
mass-weighted L1 error against it and the relative L1 error of every bin's density field
(l1_per_bin); then the FP64 coagulation's two launch shapes (one lane per cell, one wave per
cell) on 2,048 to 65,536 of the same cells, with their times and bit-equality.  With trained
weights in shark_amd/data, the learned step is timed against the plain FP64 step on cells aged by
the plain solver, at intervals in the disk's hydro-step range, with its L1 errors (per cell and
per bin) and fallback count.  In the c5 workload every timed step's fallback count is summed.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "zone-updates/s (gas+dust) on 1 MI355X; growth-step speedup at fixed L1 err"
HBM_PEAK_GBS = 8000.0
WORKLOADS = {
    "disk": ("disk_1024x2048_n16_fargo", lambda s: s.disk()),
    "si": ("si_1024x1024_n8", lambda s: s.si_box()),
    "dustywave": ("dustywave_512_n1", lambda s: s.dustywave()),
    "c5": ("disk_1024x2048_b64_learned_growth", lambda s: s.disk_growth(nr=1024, nphi=2048, nbins=64)),
}
MLP_DIMS = (69, 256, 256, 256, 65)   # the learned step's layer widths
GAS_SOLVERS = {0: "llf", 1: "hll", 2: "hllc"}                  # SOLVER (bin/Makefile:10)
DUST_SOLVERS = {0: "huang_bai", 1: "llf", 2: "hll", 3: "hllc"}  # SOLVERDUST


def copy_probe(nbytes=1 << 30, reps=10):
    """A plain device copy of known size (the library's streaming probe), for calibrating
    rocprofv3 FETCH_SIZE/WRITE_SIZE."""
    from shark_amd import gpu
    print(json.dumps({"copy_probe_bytes": nbytes, "reps": reps, **gpu.hbm_probe(nbytes, reps)}))


# the reference's step time over the CPU restatement's on the same state and step count, from
# oracle/ref/calibrate.py (flang-built reference, the development host's cores), kept in
# oracle/calibration.json
CALIBRATION = {"disk": "disk_256x512_n16", "si": "si_256x256_n8", "dustywave": "dustywave_512_n1"}


def calibration(key):
    path = os.path.join(ROOT, "oracle", "calibration.json")
    if key not in CALIBRATION or not os.path.exists(path):
        return None
    c = json.load(open(path))
    row = c.get(CALIBRATION[key])
    if not row:
        return None
    threads = {k: v for k, v in row.items() if k.isdigit()}
    most = max(threads, key=int)
    return {"config": CALIBRATION[key], "host": c.get("host"),
            "reference_s_over_port_s": {k: round(v["ratio"], 3) for k, v in threads.items()},
            "all_cores_ratio": threads[most]["ratio"], "all_cores_threads": int(most)}


def cpu_baseline(cfg, geom, state, zones_per_step, steps, key=None):
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle
    cores = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    o = pyoracle.Oracle(cfg, geom, state)
    o.step()  # first touch of the oracle's work arrays
    t0 = time.perf_counter()
    for _ in range(steps):
        o.step()
    dt = time.perf_counter() - t0
    o.close()
    out = {"value": zones_per_step * steps / dt, "unit": "zone-updates/s", "cores": cores,
           "kind": "port",
           "sample": f"{steps} steps of the same workload after 1 untimed step, "
                     f"{dt:.1f} s, OpenMP over {cores} threads"}
    cal = calibration(key)
    if cal:
        # the flang-built reference takes `ratio` times the restatement's time on the same steps:
        # its own rate on this host would be about value / ratio
        out["calibration"] = cal
        out["reference_equivalent_value"] = out["value"] / cal["all_cores_ratio"]
    return out


def profiled_traffic(workload, kernel, post_tw=None):
    """HBM bytes per launch of `kernel` from the committed rocprofv3 counter summary of the disk
    workload (FETCH_SIZE/WRITE_SIZE runs scaled by the copy calibration), or None.  The post
    kernel's traffic is the one of the row segment width the run was tuned to."""
    if workload != "disk":
        return None
    path = os.path.join(ROOT, "profiles", "disk_1024x2048_n16_hbm_traffic.md")
    if not os.path.exists(path):
        return None
    vals = {}
    for line in open(path):
        parts = line.split()
        if len(parts) == 3 and parts[0] == "traffic_bytes":
            vals[parts[1]] = int(parts[2])
    if kernel == "hydro" and "hydro_gas" in vals and "hydro_dust" in vals:
        return vals["hydro_gas"] + vals["hydro_dust"]
    if kernel == "post":
        return vals.get(f"post_tw{post_tw}")
    return vals.get(kernel)


def per_bin_l1(a, ref):
    """Relative L1 error of each bin's density field over the cells: sum_c |a_k - ref_k| / sum_c ref_k,
    three significant digits."""
    import numpy as np
    e = np.abs(a - ref).sum(axis=1) / np.maximum(ref.sum(axis=1), 1e-300)
    return [float(f"{x:.3g}") for x in e]


def growth_block(steps_probe_dt=2e-8, surrogate_cells=16384, seed=99):
    """Growth timing and error per mode (see module docstring)."""
    import numpy as np
    from shark_amd import growth as G
    from shark_amd import surrogate as S
    from shark_amd.train_surrogate import tables_sdust

    nb = 64
    p = G.disk_params(nbins=nb, max_substeps=1000000)
    sd = tables_sdust(p)
    rhod, sigma, cs, r = G.disk_state(p, nr=512, nphi=512, bins=sd)
    n = sigma.size
    g = G.Growth(p, n)
    modes = {}
    ref = None
    try:
        for mode in (G.FP64, G.MFMA_FP32, G.MFMA_BF16, G.MFMA_BF16_PLAIN):
            g.upload(rhod, sigma, cs, r)
            g.run(mode, steps_probe_dt)
            out, nsub = g.download()
            ms = g.kernel_ms()
            if mode == G.FP64:
                ref = out
            err = S.mass_l1(out, ref)
            modes[G.MODE_NAMES[mode]] = {
                "ms": ms, "substeps_max": int(nsub.max()), "substeps_mean": float(nsub.mean()),
                "ns_per_cell_substep": ms * 1e6 / float(nsub.sum()),
                "l1_vs_fp64_mean": float(err.mean()), "l1_vs_fp64_max": float(err.max()),
                "l1_per_bin": per_bin_l1(out, ref),
                "mass_residual_max": float(np.abs(out.sum(0) / rhod.sum(0) - 1).max())}
    finally:
        g.close()
    base = modes["fp64"]["ms"]
    for v in modes.values():
        v["speedup_vs_fp64"] = base / v["ms"]
    out = {"workload": "disk_512x512_b64_mrn", "cells": n, "dt": steps_probe_dt, "modes": modes}

    # mode 0's two launch shapes (one lane per cell, one wave per cell) on the first m cells of
    # the same disk: their times, and that they give the same bits
    shapes = []
    for m in (2048, 16384, 65536):
        res = {}
        for wc in (0, 1 << 40):
            g = G.Growth(p, m)
            try:
                g.set_wave_cells(wc)
                g.upload(np.ascontiguousarray(rhod[:, :m]), sigma[:m].copy(), cs[:m].copy(), r[:m].copy())
                g.run(G.FP64, steps_probe_dt)
                res[wc] = g.download() + (g.kernel_ms(),)
            finally:
                g.close()
        lane, wave = res[0], res[1 << 40]
        shapes.append({"cells": m, "lane_per_cell_ms": lane[2], "wave_per_cell_ms": wave[2],
                       "bit_identical": bool(np.array_equal(lane[0], wave[0]) and np.array_equal(lane[1], wave[1]))})
    out["fp64_launch_shapes"] = shapes

    wpath = os.path.join(ROOT, "shark_amd", "data", "surrogate_disk64.npz")
    if os.path.exists(wpath):
        L = S.Learned(wpath)
        meta = L.meta
        nphi = max(1, surrogate_cells // 1024)
        _, _, st = S.sample_states(p, sd, 1024, nphi, seed=seed, age_max=meta["age_max"],
                                   dt_lo=meta["dt_range"][0], dt_hi=meta["dt_range"][1], gen_mode=G.FP64)
        ns = st["sigma"].size
        g = G.Growth(p, ns)
        try:
            L.attach(g)
            g.set_cell_dt(st["dt"])
            g.upload(st["rhod"], st["sigma"], st["cs"], st["r"])
            g.run(G.SURROGATE, 0.0)   # untimed first call (code objects, first-touch)
            g.upload(st["rhod"], st["sigma"], st["cs"], st["r"])
            g.run(G.SURROGATE, 0.0)
            so, _ = g.download()
            ms_s = g.kernel_ms()
            nfb, raw = g.surrogate_stats()
            g.upload(st["rhod"], st["sigma"], st["cs"], st["r"])
            g.run(G.FP64, 0.0)
            fo, fsub = g.download()
            ms_f = g.kernel_ms()
        finally:
            g.close()
            L.close()
        err = S.mass_l1(so, fo)
        out["surrogate"] = {
            "cells": ns, "dt_range": meta["dt_range"], "age_max": meta["age_max"],
            "train_samples": meta["samples"], "ms": ms_s, "fp64_ms": ms_f,
            "speedup_vs_fp64": ms_f / ms_s, "fp64_substeps_mean": float(fsub.mean()),
            "l1_vs_fp64_mean": float(err.mean()), "l1_vs_fp64_p99": float(np.quantile(err, 0.99)),
            "no_change_l1_mean": float(S.mass_l1(st["rhod"], fo).mean()),
            "l1_per_bin": per_bin_l1(so, fo),
            "networks": 2 if L.net2 is not None else 1,
            "activations": "bf16 hi+lo (first network), bf16 (residual network)",
            "fallback_cells": int(nfb), "fallback_fraction": nfb / ns,
            "fallback_threshold": S.resid_max(meta), "raw_mass_residual_mean": raw,
            "mass_residual_max": float(np.abs(so.sum(0) / st["rhod"].sum(0) - 1).max())}
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--workload", choices=sorted(WORKLOADS), default="disk")
    ap.add_argument("--cpu-steps", type=int, default=5)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--graph", action="store_true", help="(default) time the hipGraph replay")
    ap.add_argument("--events", action="store_true", help="time per-kernel event-bracketed launches "
                    "instead of the hipGraph replay")
    ap.add_argument("--copy-probe", action="store_true")
    ap.add_argument("--no-growth", action="store_true", help="skip the growth timing block")
    ap.add_argument("--growth-only", action="store_true",
                    help="print only the growth timing block (for profiling its kernels)")
    ap.add_argument("--no-tune", action="store_true", help="keep the default launch shapes")
    ap.add_argument("--no-workloads", action="store_true",
                    help="skip the shearing-box and learned-growth sub-blocks of the default line")
    args = ap.parse_args()

    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU: this build runs the step on a single MI355X; "
              "multi-GPU (radial slab decomposition) is not built.", file=sys.stderr)
        sys.exit(2)
    if args.copy_probe:
        copy_probe()
        return
    if args.growth_only:
        print(json.dumps({"growth": growth_block()}))
        return

    from shark_amd import setups

    attainable = hbm_attainable()
    res = time_workload(args.workload, args.steps, args.warmup, tune=not args.no_tune,
                        events=args.events, attainable=attainable)
    cfg, geom, state = res.pop("_problem")
    learned = args.workload == "c5"
    out = {
        "metric": METRIC,
        "value": res["value"],
        "unit": "zone-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": res["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
    }
    out.update({k: v for k, v in res.items() if k not in ("value", "ms_per_step")})
    out["hbm_attainable"] = attainable
    if learned:
        out["dtype"] = "f64 step; learned growth bf16 weights x bf16 hi+lo activations -> f32"
        # the oracle's FP64 coagulation needs up to ~1700 sub-steps per cell at this dt: a CPU
        # sample of the coupled step would not finish in a bounded time
        out["cpu_baseline"] = None
        out["cpu_baseline_note"] = "not run for the coupled learned-growth workload"
    elif not args.no_cpu_baseline:
        zones_per_step = res["zones_per_step"]
        out["cpu_baseline"] = cpu_baseline(cfg, geom, state, zones_per_step, args.cpu_steps,
                                           key=args.workload)
        out["speedup_vs_cpu"] = out["value"] / out["cpu_baseline"]["value"]
    del cfg, geom, state
    if args.workload == "disk" and not args.no_workloads:
        # the other workloads of the metric, each timed the same way (the 2D shearing box and the
        # 64-bin disk with the learned growth step inside every step)
        out["workloads"] = {}
        for key in ("si", "c5"):
            r = time_workload(key, min(args.steps, 50), min(args.warmup, 5), tune=not args.no_tune,
                              attainable=attainable)
            r.pop("_problem")
            out["workloads"][WORKLOADS[key][0]] = r
    if not args.no_growth and not learned:
        out["growth"] = growth_block()
    print(json.dumps(out))


def hbm_attainable(nbytes=1 << 30, reps=20):
    """On-box HBM bandwidth a plain streaming kernel attains (shark_hbm_probe: write, copy and read
    of 1 GiB FP64 buffers, 16-byte accesses, HIP events): the figure the kernels' traffic is also
    rated against.  The best of the three is the attainable figure."""
    from shark_amd import gpu
    res = gpu.hbm_probe(nbytes, reps)
    res["GBs"] = max(res["write_GBs"], res["copy_GBs"], res["read_GBs"])
    res["source"] = "shark_hbm_probe, 1 GiB buffers, 20 reps, HIP events"
    return res


def time_workload(key, steps, warmup, tune=True, events=False, attainable=None):
    """Time `steps` steps of a workload after `warmup` (graph replay unless a growth step is inside
    the step or events=True), then per-kernel times from event-bracketed launches; returns the
    measured fields of the JSON line."""
    from shark_amd import gpu, setups
    from shark_amd.cfg import algorithmic_bytes

    name, make = WORKLOADS[key]
    learned = key == "c5"
    made = make(setups)
    cfg, geom, state = made[:3]
    cells = cfg.nx * (cfg.ny if cfg.ny > 1 else 1)
    fluids = 1 + cfg.ndust
    zones_per_step = cells * fluids
    g = gpu.Step(cfg, geom, state)
    tn = g.autotune(reps=5) if tune else None
    gr = net = None
    if learned:
        from shark_amd import growth as G
        from shark_amd import surrogate as S
        net = S.Learned()
        gr = G.Growth(made[3], (cfg.nx + 4) * (cfg.ny + 4))
        net.attach(gr)
        g.attach_growth(gr, G.SURROGATE)
    g.run(warmup)

    # the timed steps replay the step's hipGraph (the production path); per-kernel times come
    # from a second, event-bracketed pass. A growth step inside the step is launched directly.
    use_graph = not learned and not events
    g.set_timing(not use_graph)
    nfb_steps = []
    t0 = time.perf_counter()
    if learned:
        # one step per call (each already synchronises the host once): the fallback count of
        # every timed step is read between them
        for _ in range(steps):
            g.step()
            nfb_steps.append(int(gr.surrogate_stats()[0]))
        n = steps
    else:
        n, _ = g.run(steps)
    wall = time.perf_counter() - t0
    assert n == steps, (n, steps)
    if use_graph:
        g.set_timing(True)
        g.run(min(steps, 50))
    kt = {k: g.kernel_time(k)[0] for k in ("bc", "hydro", "post", "dt") + (("growth",) if learned else ())}
    g.set_timing(False)
    learned_info = None
    if learned:
        nfb, raw = gr.surrogate_stats()
        learned_info = {"growth_ms_per_step": kt["growth"], "fallback_cells_last_step": int(nfb),
                        "fallback_fraction_last_step": nfb / cells,
                        "fallback_cells_timed_steps": int(sum(nfb_steps)),
                        "fallback_fraction_timed_steps": sum(nfb_steps) / (cells * steps),
                        "steps_before_timing": warmup, "raw_mass_residual_mean": raw,
                        "networks": 2 if net.net2 is not None else 1}

    nv = 5 + 4 * cfg.ndust
    bytes_hydro = cells * 8 * (2 * nv + 1)            # read u_prim + cs, write u_prim
    bytes_post = cells * 16 * nv if cfg.fargo else 0     # FARGO's second full pass
    if learned:
        # the post pass runs in two parts around the growth call: each reads and writes u_prim,
        # and the dust densities before growth are written by the first and read by the second
        bytes_post = cells * 8 * (4 * nv + 2 * cfg.ndust)
    step_alg = algorithmic_bytes(cfg)
    ms_step = wall / steps * 1e3
    att = attainable["GBs"] if attainable else None
    if not cfg.fargo:
        # without FARGO the algorithmic traffic is one pass; rate the whole step against it
        dom, dom_ms, dom_bytes = "step", ms_step, step_alg
    else:
        dom = max(("hydro", "post"), key=lambda k: kt[k])
        dom_ms = kt[dom]
        dom_bytes = bytes_hydro if dom == "hydro" else bytes_post
    achieved = dom_bytes / (dom_ms * 1e-3) / 1e9
    roof = {"bound": "hbm", "kernel": dom, "achieved": achieved, "peak": HBM_PEAK_GBS,
            "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS,
            "frac_of_attainable": achieved / att if att else None,
            "traffic": profiled_traffic(key, dom, g.get_tune()["post_tw"]),
            "traffic_source": "profiles/disk_1024x2048_n16_hbm_traffic.md" if key == "disk" else None,
            "algorithmic_bytes": dom_bytes, "avg_ms": dom_ms}
    ncell_all = (cfg.nx + 4) * (cfg.ny + 4)
    mlp = None
    if learned:
        mlp = mlp_roofline(ncell_all)
        if kt["growth"] > dom_ms:
            # the learned step dominates: rate its network products against the dense BF16 peak
            # (2 flops per multiply-add over every layer of each network, every cell incl. ghosts)
            flops = mlp["networks"] * 2.0 * ncell_all * sum(a * b for a, b in zip(MLP_DIMS[:-1], MLP_DIMS[1:]))
            tf = flops / (kt["growth"] * 1e-3) / 1e12
            roof = {"bound": "mfma", "kernel": "growth", "achieved": tf, "peak": 2500.0,
                    "unit": "TFLOP/s", "frac": tf / 2500.0, "traffic": None,
                    "algorithmic_flops": flops, "avg_ms": kt["growth"]}
    step_gbs = step_alg / (wall / steps) / 1e9
    out = {
        "value": zones_per_step * steps / wall,
        "ms_per_step": ms_step,
        "zones_per_step": zones_per_step,
        "config": {"workload": name, "nx": cfg.nx, "ny": cfg.ny, "ndust": cfg.ndust,
                   "fluids": fluids, "geom": "polar" if cfg.geom == 2 else "cartesian",
                   "fargo": bool(cfg.fargo), "drag": bool(cfg.drag),
                   "solver": GAS_SOLVERS[cfg.solver], "solverdust": DUST_SOLVERS[cfg.solverdust],
                   "slope_type": cfg.slope_type},
        "roofline": roof,
        "kernels_ms": kt,
        "tune": tn,
        "timed_path": "hipGraph replay, one graph per step" if use_graph else "direct launches",
        "step_roofline": {"algorithmic_bytes": step_alg, "achieved_GBs": step_gbs,
                          "frac": step_gbs / HBM_PEAK_GBS,
                          "frac_of_attainable": step_gbs / att if att else None},
    }
    if learned:
        out["learned_growth"] = learned_info
        out["mlp_infer_roofline"] = mlp
    g.close()
    if gr is not None:
        gr.close()
        net.close()
    out["_problem"] = (cfg, geom, state)
    return out


def mlp_roofline(ncells, seed=5):
    """k_mlp_infer (the fused network of the learned step) alone over `ncells` feature rows, one
    launch per network, as the learned step runs it (BF16 hi + lo activations: two products per
    weight fragment), against the dense BF16 MFMA peak; the single-product variant beside it."""
    import numpy as np
    from shark_amd import surrogate as S
    L = S.Learned(max_batch=ncells)
    nets = [n for n in (L.net, L.net2) if n is not None]
    try:
        x = np.random.default_rng(seed).standard_normal((ncells, MLP_DIMS[0])).astype(np.float32)
        ms = 0.0
        for i, n in enumerate(nets):       # the residual network runs the single-product kernel
            n.forward(x, hilo=i == 0)
            ms += n.kernel_ms()
        L.net.forward(x)
        ms_single = L.net.kernel_ms()
    finally:
        L.close()
    flops1 = 2.0 * ncells * sum(a * b for a, b in zip(MLP_DIMS[:-1], MLP_DIMS[1:]))
    flops = flops1 * len(nets)
    tf = flops / (ms * 1e-3) / 1e12
    issued = flops1 * (len(nets) + 1)      # two products per fragment in the first network
    return {"kernel": "k_mlp_infer", "rows": ncells, "networks": len(nets),
            "activations": "bf16 hi+lo (first network), bf16 (residual network)",
            "ms": ms, "algorithmic_flops": flops, "achieved_TFLOPs": tf, "peak_TFLOPs": 2500.0,
            "frac": tf / 2500.0, "mfma_flops_issued": issued,
            "frac_issued": issued / (ms * 1e-3) / 1e12 / 2500.0,
            "single_product_variant": {"ms_per_network": ms_single,
                                       "achieved_TFLOPs": flops1 / (ms_single * 1e-3) / 1e12}}


if __name__ == "__main__":
    main()
