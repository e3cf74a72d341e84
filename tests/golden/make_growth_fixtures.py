#!/usr/bin/env python3
"""Generate tests/golden/ref_growth_<case>.npz from the reference's own dust_growth.

Builds the reference Collapse_1D setup (NX=8, NDUST=16, GEOM=1) with the growth harness
oracle/ref/ref_growth_harness.f90, which calls read_params, setup, ctoprim, compute_tstop
and dust_growth (src/dust/dust_growth.f90:3) once over dt.  Stores the inputs the oracle
needs and the dust densities after the call.  The "charge" case does the same for
analytical_charge with oracle/ref/ref_charge_harness.f90 (ref_charge.npz).
Development container only.
"""
import json
import os
import re
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.abspath(os.path.join(HERE, "..", ".."))
BUILD = os.path.join(REPO, "oracle", "ref", "build_ref.sh")
NAME = "growth_Collapse_1D_8_n16"
EXE = os.path.join(REPO, "oracle", "_ref", NAME, "shark_ref")

BASE = """&GRID_PARAMS
CFL        = 0.5
tend       = 1
gamma      = 1.66667d0
slope_type = 1
rin        = 1d-1
/
&SETUP_PARAMS
M_cloud     = 1.0d0
/
&DUST_PARAMS
drag                = .true.
dust_distribution   = 'mrn'
smin     = 1d-6
scutmin  = 1d-6
smax     = 1d-1
scut     = 1d-3
rhograin = 2.3d0
dust2gas = 0.01d0
growth              = .true.
fragmentation       = .true.
eps_threshold       = 1d-15
eps_threshold_frag  = 1d-7
dust_ratio_min      = 1d-30
CFL_growth          = 0.5d0
drift_in_growth     = .true.
turb_in_growth      = .true.
brownian_in_growth  = .true.
ambipolar_in_growth = .true.
slope_mono          = 3.5d0
gamma_grains        = 190.0d0
estar_grains        = 3.7e10
alpha_turb          = 1.0d0
kernel_type         = 0
/
&GRAVITY_PARAMS
/
&IONISATION_PARAMS
charging=.false.
/
"""

CASES = {
    "all_terms": ({}, 1e-4),
    "no_fragmentation": ({"fragmentation": ".false."}, 1e-3),
    "turb_brownian_clustered": ({"drift_in_growth": ".false.", "clustered_fraction": "2.0d0"}, 3e-4),
}
# kernel_type 1 and 2 are not run here: at the densities of this setup (up to 1e21 times the
# cloud density) their rates need more sub-steps than the reference finishes in an hour; the
# oracle's constant and additive kernels are pinned by the analytic checks instead.


def _meta(path):
    m = {}
    for line in open(path):
        k, v = line.split(None, 1)
        v = v.strip()
        if v in ("T", "F"):
            m[k] = v == "T"
        else:
            try:
                m[k] = int(v)
            except ValueError:
                m[k] = float(v.replace("D", "E"))
    return m


def run_case(name, flags, dt):
    if not os.path.exists(EXE):
        env = dict(os.environ, REF_NAME=NAME, REF_HARNESS="ref_growth_harness.f90")
        subprocess.run([BUILD, "Collapse_1D", "8", "1", "16", "1", "0", "1", "0"], check=True,
                       env=env, stdout=subprocess.DEVNULL, stderr=subprocess.DEVNULL)
    nml = BASE
    for k, v in flags.items():
        if re.search(r"(?m)^%s\s*=" % k, nml):
            nml = re.sub(r"(?m)^%s\s*=.*$" % k, "%s = %s" % (k, v), nml)
        else:
            nml = nml.replace("kernel_type", "%s = %s\nkernel_type" % (k, v), 1)
    with tempfile.TemporaryDirectory() as td:
        open(os.path.join(td, "g.nml"), "w").write(nml)
        subprocess.run([EXE, "g.nml", repr(dt), "."], cwd=td, check=True,
                       env=dict(os.environ, OMP_NUM_THREADS="1"), stdout=subprocess.DEVNULL)
        m = _meta(os.path.join(td, "meta.txt"))
        nc, nv, nd = m["ncells"], m["nvar"], m["ndust"]
        ic = np.fromfile(os.path.join(td, "ic.bin")).reshape(nv, nc)
        fin = np.fromfile(os.path.join(td, "final.bin")).reshape(nv, nc)
        aux = np.fromfile(os.path.join(td, "aux.bin"))
    o = [0]

    def take(n, shape=None):
        a = aux[o[0]:o[0] + n]
        o[0] += n
        return a.reshape(shape) if shape else a

    arrs = dict(q_rho=take(nc), tstop=take(nc * nd, (nd, nc)), sdust=take(nc * nd, (nd, nc)),
                mdust=take(nc * nd, (nd, nc)), cs=take(nc), active=take(nc), aplus=take(nd),
                aminus=take(nd), mminus=take(nd), q=take(nc * nv, (nv, nc)))
    assert o[0] == aux.size
    arrs["u_rho"] = ic[0]
    arrs["rhod_ic"] = ic[5:5 + nd]
    arrs["rhod_final"] = fin[5:5 + nd]
    m.update(case=name, namelist=nml, T0_cloud=10.0)
    np.savez_compressed(os.path.join(HERE, f"ref_growth_{name}.npz"), meta=json.dumps(m), **arrs)
    return m


CHARGE_NAME = "charge_Collapse_1D_8_n16"
CHARGE_EXE = os.path.join(REPO, "oracle", "_ref", CHARGE_NAME, "shark_ref")


def run_charge():
    """ref_charge.npz: the reference's analytical_charge (src/chemistry/ionisation.f90:809-922)
    on the Collapse_1D cells rescaled by oracle/ref/ref_charge_harness.f90 (growth and
    fragmentation off, so the setup calls no charge routine itself)."""
    if not os.path.exists(CHARGE_EXE):
        env = dict(os.environ, REF_NAME=CHARGE_NAME, REF_HARNESS="ref_charge_harness.f90")
        subprocess.run([BUILD, "Collapse_1D", "8", "1", "16", "1", "0", "1", "0"], check=True,
                       env=env, stdout=subprocess.DEVNULL, stderr=subprocess.DEVNULL)
    nml = BASE.replace("growth              = .true.", "growth              = .false.")
    nml = nml.replace("fragmentation       = .true.", "fragmentation       = .false.")
    with tempfile.TemporaryDirectory() as td:
        open(os.path.join(td, "c.nml"), "w").write(nml)
        subprocess.run([CHARGE_EXE, "c.nml", "."], cwd=td, check=True,
                       env=dict(os.environ, OMP_NUM_THREADS="1"), stdout=subprocess.DEVNULL)
        m = _meta(os.path.join(td, "meta.txt"))
        aux = np.fromfile(os.path.join(td, "aux.bin"))
    nc, nd = m["ncells"], m["ndust"]
    o = [0]

    def take(n, shape=None):
        a = aux[o[0]:o[0] + n]
        o[0] += n
        return a.reshape(shape) if shape else a

    arrs = dict(q_rho=take(nc), rhod=take(nc * nd, (nd, nc)), mdust=take(nc * nd, (nd, nc)),
                sdust=take(nc * nd, (nd, nc)), zd=take(nc * nd, (nd, nc)), ni=take(nc), ne=take(nc))
    assert o[0] == aux.size
    m.update(namelist=nml)
    np.savez_compressed(os.path.join(HERE, "ref_charge.npz"), meta=json.dumps(m), **arrs)
    return m


MARCHAND_NAME = "marchand_Collapse_1D_8_n16"
MARCHAND_EXE = os.path.join(REPO, "oracle", "_ref", MARCHAND_NAME, "shark_ref")
MARCHAND_OUT = ("zd", "ni", "ne", "psi", "eta_o", "eta_h", "eta_a", "sigma_o", "sigma_p", "sigma_h",
                "gamma_d")


def run_marchand():
    """ref_marchand.npz: the reference's Marchand charge with resistivities
    (src/chemistry/ionisation.f90:4-214) on the rescaled Collapse_1D cells of
    oracle/ref/ref_marchand_harness.f90, called twice (cold start, then from psi_old)."""
    if not os.path.exists(MARCHAND_EXE):
        env = dict(os.environ, REF_NAME=MARCHAND_NAME, REF_HARNESS="ref_marchand_harness.f90")
        subprocess.run([BUILD, "Collapse_1D", "8", "1", "16", "1", "0", "1", "0"], check=True,
                       env=env, stdout=subprocess.DEVNULL, stderr=subprocess.DEVNULL)
    nml = BASE.replace("growth              = .true.", "growth              = .false.")
    nml = nml.replace("fragmentation       = .true.", "fragmentation       = .false.")
    with tempfile.TemporaryDirectory() as td:
        open(os.path.join(td, "c.nml"), "w").write(nml)
        subprocess.run([MARCHAND_EXE, "c.nml", "."], cwd=td, check=True,
                       env=dict(os.environ, OMP_NUM_THREADS="1"), stdout=subprocess.DEVNULL)
        m = _meta(os.path.join(td, "meta.txt"))
        aux = np.fromfile(os.path.join(td, "aux.bin"))
    nc, nd = m["ncells"], m["ndust"]
    o = [0]

    def take(n, shape=None):
        a = aux[o[0]:o[0] + n]
        o[0] += n
        return a.reshape(shape) if shape else a

    arrs = dict(u_rho=take(nc), rhod=take(nc * nd, (nd, nc)), T=take(nc), cs=take(nc),
                sdust=take(nc * nd, (nd, nc)), mdust=take(nc * nd, (nd, nc)))
    for call in (1, 2):
        for k in MARCHAND_OUT:
            shape = (nd, nc) if k in ("zd", "gamma_d") else None
            arrs[f"{k}_{call}"] = take(nc * nd if shape else nc, shape)
    assert o[0] == aux.size
    m.update(namelist=nml)
    np.savez_compressed(os.path.join(HERE, "ref_marchand.npz"), meta=json.dumps(m), **arrs)
    return m


if __name__ == "__main__":
    names = sys.argv[1:] or list(CASES) + ["charge", "marchand"]
    for n in names:
        if n == "charge":
            m = run_charge()
            print("charge cells", m["ncells"], "bins", m["ndust"])
        elif n == "marchand":
            m = run_marchand()
            print("marchand cells", m["ncells"], "bins", m["ndust"])
            continue
        f, dt = CASES[n]
        m = run_case(n, f, dt)
        print(n, "dt", dt, "cells", m["ncells"], "bins", m["ndust"])
