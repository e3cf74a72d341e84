#!/usr/bin/env python3
"""Generate tests/golden/ref_output_<case>.npz: the files the reference's own `output(iout)`
(src/commons/write_output.f90:2-300) writes for the final state of a fixture case.

Runs in the development container only (reference sources + flang): builds the case with
oracle/ref/build_ref.sh, runs the harness (oracle/ref/ref_harness.f90) with an output number, and
stores every file of output_NNNNN/ as bytes, with the run's time and units.  The final state is
the one of tests/golden/ref_<case>.npz (same namelist and step count; checked here).
"""
import json
import os
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)
from make_ref_fixtures import BUILD, CASES, REPO, _meta  # noqa: E402

OUT_CASES = {"disk_2d": 7, "si_2d": 12, "dustywave_dhll": 3, "sod_hll_s3": 4, "kh_n1_hllc": 9,
             "cdisk_visc_n1": 5, "disk_2d_logr": 6}


def run(name, iout):
    setup, nx, ny, nd, geom, gs, solver, sdust, nml, nsteps = CASES[name]
    subprocess.run([BUILD, setup, str(nx), str(ny), str(nd), str(geom), str(gs), str(solver), str(sdust)],
                   check=True, stdout=subprocess.DEVNULL, stderr=subprocess.DEVNULL)
    exe = os.path.join(REPO, "oracle", "_ref", f"{setup}_{nx}x{ny}_n{nd}_g{geom}{gs}_s{solver}{sdust}", "shark_ref")
    with tempfile.TemporaryDirectory() as td:
        open(os.path.join(td, "run.nml"), "w").write(nml)
        env = dict(os.environ, OMP_NUM_THREADS="1")
        subprocess.run([exe, "run.nml", str(nsteps), td, str(iout)], check=True, cwd=td, env=env,
                       stdout=subprocess.DEVNULL)
        m = _meta(os.path.join(td, "meta.txt"))
        final = np.fromfile(os.path.join(td, "final.bin"))
        ref = np.load(os.path.join(HERE, f"ref_{name}.npz"))
        assert np.array_equal(final, ref["final"].reshape(-1)), name
        odir = os.path.join(td, f"output_{iout:05d}")
        files = {f: np.frombuffer(open(os.path.join(odir, f), "rb").read(), dtype=np.uint8)
                 for f in sorted(os.listdir(odir))}
    meta = dict(case=name, iout=iout, time=m["time"], gridspace=gs,
                units={k: m[k] for k in ("unit_t", "unit_d", "unit_l", "unit_v", "unit_p", "unit_dcol")},
                files=sorted(files))
    np.savez_compressed(os.path.join(HERE, f"ref_output_{name}.npz"), meta=json.dumps(meta),
                        **{"file_" + k.replace(".", "_"): v for k, v in files.items()})
    return meta


def main():
    names = sys.argv[1:] or list(OUT_CASES)
    for name in names:
        iout = OUT_CASES[name]
        m = run(name, iout)
        print(name, m["files"])


if __name__ == "__main__":
    main()
