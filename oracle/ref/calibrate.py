#!/usr/bin/env python3
"""Calibrate the CPU restatement (oracle/liboracle.so) against the flang-built reference step.

Development container only (needs /root/reference and flang).  For each configuration the
reference is built with oracle/ref/build_ref.sh and run by its harness (which times the step loop
alone, `loop_seconds` in meta.txt) with OMP_NUM_THREADS=1 and with every core; the oracle then
steps the reference's own initial state the same number of steps, timed the same way, in a child
process with the same thread count.  Writes oracle/calibration.json:
  {config: {threads: {"reference_s", "oracle_s", "ratio"}}}  with ratio = reference_s / oracle_s,
so a CPU figure of the oracle reads as that many times the reference's speed.

The reference's OpenMP loops have two data races (godunov.f90:43, dust_growth.f90:63); with
several threads its results are not used here, only its time.
"""
import json
import os
import subprocess
import sys
import tempfile

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.abspath(os.path.join(HERE, "..", ".."))
sys.path[:0] = [os.path.join(REPO, "tests", "golden"), os.path.join(REPO, "tests"), os.path.join(REPO, "oracle"), REPO]

import make_ref_fixtures as M  # noqa: E402

# name: (setup, nx, ny, ndust, geom, gridspace, solver, solverdust, namelist, nsteps)
CONFIGS = {
    "dustywave_512_n1": ("DUSTYWAVE", 512, 1, 1, 0, 0, 1, 0,
                         M.WAVE.format(cfl=0.5, tend="1d10", slope=1, vx0="1.0d-4", drho="0.0d0", K="10d0"), 1000),
    "si_256x256_n8": ("SI", 256, 256, 8, 0, 0, 1, 0,
                      M.test_nml("PSI/psi.nml", freq_out="100000", tend="1d10"), 20),
    "disk_256x512_n16": ("disk_polar", 256, 512, 16, 2, 0, 1, 0,
                         M.DISK.format(slope=1, r_relax="0.3d0", visc=".false.", alpha="1d-2"), 10),
}


def run_reference(spec, threads, td):
    """The reference harness (built by make_ref_fixtures.run_case) on spec for spec's step count."""
    setup, nx, ny, nd, geom, gs, solver, sdust, nml, nsteps = spec
    bname = f"{setup}_{nx}x{ny}_n{nd}_g{geom}{gs}_s{solver}{sdust}"
    exe = os.path.join(REPO, "oracle", "_ref", bname, "shark_ref")
    open(os.path.join(td, "run.nml"), "w").write(nml)
    subprocess.run([exe, os.path.join(td, "run.nml"), str(nsteps), td], check=True, cwd=td,
                   env=dict(os.environ, OMP_NUM_THREADS=str(threads)), stdout=subprocess.DEVNULL)
    return M._meta(os.path.join(td, "meta.txt"))


ORACLE_CHILD = r"""
import json, sys, time
sys.path[:0] = {paths!r}
import pyoracle, refcases
cfg, geom, state, final, dts, meta = refcases.load_ref({case!r})
o = pyoracle.Oracle(cfg, geom, state)
t0 = time.perf_counter()
o.run({nsteps})
print(json.dumps({{"oracle_s": time.perf_counter() - t0}}))
"""


def main():
    ncores = os.cpu_count() or 1
    out = {}
    for name, spec in CONFIGS.items():
        nsteps = spec[9]
        out[name] = {"nsteps": nsteps}
        # the reference's initial state and grid, as a fixture file the oracle loads
        case = f"zz_calib_{name}"
        npz = os.path.join(M.HERE, f"ref_{case}.npz")
        try:
            M.run_case(case, spec[:9] + (0,), force_build=True)
            for threads in sorted({1, ncores}):
                with tempfile.TemporaryDirectory() as td:
                    meta = run_reference(spec, threads, td)
                code = ORACLE_CHILD.format(paths=[os.path.join(REPO, "tests"), os.path.join(REPO, "oracle"), REPO],
                                           case=case, nsteps=nsteps)
                r = subprocess.run([sys.executable, "-c", code], check=True, capture_output=True, text=True,
                                   env=dict(os.environ, OMP_NUM_THREADS=str(threads)))
                osec = json.loads(r.stdout.strip().splitlines()[-1])["oracle_s"]
                out[name][str(threads)] = {"reference_s": meta["loop_seconds"], "oracle_s": osec,
                                           "ratio": meta["loop_seconds"] / osec}
                print(name, threads, out[name][str(threads)], flush=True)
        finally:
            if os.path.exists(npz):
                os.remove(npz)
    out["host"] = {"cores": ncores, "cpu": _cpu_model()}
    with open(os.path.join(os.path.dirname(HERE), "calibration.json"), "w") as f:
        json.dump(out, f, indent=1)


def _cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return None


if __name__ == "__main__":
    main()
