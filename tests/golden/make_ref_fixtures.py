#!/usr/bin/env python3
"""Generate reference fixtures tests/golden/ref_<case>.npz from the flang-built reference.

Runs in the development container only (needs /root/reference and flang).  For each case
it builds the reference Fortran step with oracle/ref/build_ref.sh (the reference's own
sources, our harness main oracle/ref/ref_harness.f90), runs it single-threaded on the
namelist below for a fixed number of steps, and stores:
  ic (nvar,ncells)      u_prim after the reference setup
  final (nvar,ncells)   u_prim after `nsteps` calls of solve (src/solver/solve.f90:10)
  dt (nsteps,)          dt of each step
  cs, dx, surf, vol, radii_c, position, active, active_pred, [sdust], [condinit]
  meta                  JSON: macro set, namelist scalars, sizes
The fixtures are data (inputs and expected outputs); no reference source is stored.
"""
import json
import os
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.abspath(os.path.join(HERE, "..", ".."))
BUILD = os.path.join(REPO, "oracle", "ref", "build_ref.sh")
REF = os.environ.get("SHARK_REFERENCE", "/root/reference")

WAVE = """&GRID_PARAMS
freq_out=100000
CFL = {cfl}
tend = {tend}
gamma = 1.6666667d0
slope_type = {slope}
/
&SETUP_PARAMS
box_l = 1.0d0
cs_0 = 1.0d0
rho_0 = 1.0d0
kx_wave = 1.0d0
vx_0 = {vx0}
delta_rho = {drho}
delta_rho_d = 0.0d0
K(1) = {K}
dust2gas_ratio(1) = 1.0d0
/
&DUST_PARAMS
drag=.true.
/
&GRAVITY_PARAMS
/
&IONISATION_PARAMS
charging=.false.
/
"""

SI = """&GRID_PARAMS
freq_out=100000
CFL=0.5
tend=1.0d10
gamma=1.666667d0
force_kick=.true.
slope_type={slope}
/
&SETUP_PARAMS
box_l       = 2d0
box_l_y     = 2d0
rho_init    = 1.0d0
omega_shear = 1.0d0
hoverr      = 0.1
Stokes_species(1) = 0.1d0
Stokes_species(2) = 1.0d0
dust2gas_species(1) = 0.5d0
dust2gas_species(2) = 1.0d0
/
&DUST_PARAMS
dust_back_reaction = .true.
drag               = .true.
/
&GRAVITY_PARAMS
/
&IONISATION_PARAMS
charging=.false.
/
"""

DISK = """&GRID_PARAMS
freq_out=100000
CFL  = 0.5
tend = 1.0d10
gamma=1.666667d0
force_kick =.true.
slope_type = {slope}
viscosity = {visc}
/
&SETUP_PARAMS
alpha_visc   = {alpha}
box_l        = 2.1d0
R_out_disk   = 1d30
smooth_r     = 0.4d0
sigma_R0     = 1700.0d0
hoverr       = 0.05d0
r_relax      = {r_relax}
/
&DUST_PARAMS
scut     = 1.0d0
dust2gas = 0.01d0
dust_back_reaction=.true.
drag=.true.
/
&GRAVITY_PARAMS
/
&IONISATION_PARAMS
charging=.false.
/
"""

# disk_polar with a planet on a circular orbit inside the disk (setup.f90:323-378)
DISK_PLANET = DISK.replace("r_relax      = {r_relax}\n",
                           "r_relax      = {r_relax}\nMplanet      = {mplanet}\nRplanet      = 1.3d0\n"
                           "a_smooth_pl  = 0.05d0\n")

# disk_polar with the disk's own potential (setup.f90:300-322)
DISK_SG = DISK.replace("&GRAVITY_PARAMS\n/", "&GRAVITY_PARAMS\nself_gravity = .true.\n/")

SOD = """&GRID_PARAMS
freq_out=100000
CFL = {cfl}
tend = {tend}
gamma = 1.4d0
slope_type = {slope}
/
&SETUP_PARAMS
direction_shock = {dir}
/
&DUST_PARAMS
/
&GRAVITY_PARAMS
/
&IONISATION_PARAMS
charging=.false.
/
"""

# Kelvin-Helmholtz (setups/KH): adiabatic gas with dust, drag and drag heating; the parameters of
# the reference's test/KH/kh.nml on a 1 au box, tend in Myr
KH = """&GRID_PARAMS
freq_out=100000
CFL = {cfl}
tend = 1d3
gamma = 1.666667d0
slope_type = {slope}
force_kick = .false.
/
&SETUP_PARAMS
box_l   = 4.8481e-6
box_l_y = 4.8481e-6
rho_dense    = 1.0d-18
rho_diffuse  = {rho_diffuse}
T_dense      = 10.0d0
T_diffuse    = 10.0d0
Mach         = 0.25d0
vy0          = 1d-1
kx           = 2.0d0
/
&DUST_PARAMS
dust_distribution='mrn'
smin     = 5d-7
scutmin  = 1d-7
scut     = {scut}
smax     = 0.10d0
rhograin = 2.3d0
dust2gas = 0.01d0
mrn      = 3.5d0
dust_back_reaction = .true.
growth             = .false.
/
&GRAVITY_PARAMS
/
&IONISATION_PARAMS
charging=.false.
/
"""

# Cartesian disk (setups/disk): softened point-mass gravity, zero-gradient boundaries, buffer
# zones reset every step, alpha viscosity; box and disk in au, Mstar in solar masses
CDISK = """&GRID_PARAMS
freq_out=100000
CFL = {cfl}
tend = 1d10
slope_type = {slope}
force_kick = .true.
viscosity = {visc}
/
&SETUP_PARAMS
box_l = 400.0d0
sigma_R0 = 1700.0d0
R0_disk = 1.0d0
Mstar = 1.0d0
HoverR = 0.05d0
smooth_r = {smooth}
alpha_visc = {alpha}
Mplanet = {mplanet}
frac_R_disk = 0.1d0
buffer_in = 0.1d-2
buffer_out = 0.25d0
/
&DUST_PARAMS
scut = 1.0d-4
dust2gas = 0.01d0
rhograin = 2.3d0
/
&GRAVITY_PARAMS
/
&IONISATION_PARAMS
charging=.false.
/
"""

# Collapse_1D: 1D spherical log grid from rin (au) to the cloud radius, barotropic EOS, 1D
# self-gravity (built with GRAVITY=1); the reference's test-suite collapse run
COLLAPSE = """&GRID_PARAMS
freq_out = 100000
CFL = {cfl}
tend = 1
gamma = 1.66667d0
slope_type = {slope}
rin = 1d-1
force_kick = .false.
/
&SETUP_PARAMS
M_cloud = 1
more_outputs = .false.
/
&DUST_PARAMS
{dust}
/
&GRAVITY_PARAMS
/
&IONISATION_PARAMS
charging=.false.
/
"""

# Sedov blast (setups/sedov): adiabatic gas at rest, the pressure raised by P_cen within r_expl of
# the box centre, periodic boundaries
SEDOV = """&GRID_PARAMS
freq_out=100000
CFL = {cfl}
tend = 1d10
gamma = 1.4d0
slope_type = {slope}
/
&SETUP_PARAMS
box_l = 1.0d0
rho_ext = 1.0d0
P_ext = 1d-5
P_cen = 1.0d0
/
&DUST_PARAMS
/
&GRAVITY_PARAMS
/
&IONISATION_PARAMS
charging=.false.
/
"""


def suite_nml(rel, **subst):
    """A namelist of the reference's own test suite (test/test-suite/<rel>), with the run length
    and output cadence replaced so that a fixed number of steps is taken."""
    return test_nml(os.path.join("test-suite", rel), **subst)


def test_nml(rel, **subst):
    """A namelist of the reference's test directory (test/<rel>), run length and output cadence
    replaced as in suite_nml."""
    text = open(os.path.join(REF, "test", rel)).read()
    out = []
    for line in text.splitlines():
        key = line.split("=")[0].strip().lower()
        if key in subst:
            line = f"{line.split('=')[0]}= {subst[key]}"
        out.append(line)
    return "\n".join(out) + "\n"


# name: (setup, nx, ny, ndust, geom, gridspace, solver, solverdust, namelist, nsteps[, extra -D])
CASES = {
    "dustywave_k50": ("DUSTYWAVE", 128, 1, 1, 0, 0, 1, 0,
                      WAVE.format(cfl=0.8, tend=4.5, slope=2, vx0="1.0d-4", drho="1.0d-4", K="50d0"), -1),
    "dustywave_c1": ("DUSTYWAVE", 512, 1, 1, 0, 0, 1, 0,
                     WAVE.format(cfl=0.5, tend="1d10", slope=1, vx0="1.0d-4", drho="0.0d0", K="10d0"), 1000),
    "dustywave_hllc_s3": ("DUSTYWAVE", 64, 1, 1, 0, 0, 2, 0,
                          WAVE.format(cfl=0.5, tend="1d10", slope=3, vx0="3.0d-1", drho="1.0d-1", K="100d0"), 60),
    "dustywave_llf_dllf": ("DUSTYWAVE", 64, 1, 1, 0, 0, 0, 1,
                           WAVE.format(cfl=0.5, tend="1d10", slope=2, vx0="3.0d-1", drho="1.0d-1", K="5d0"), 60),
    "dustywave_dhll": ("DUSTYWAVE", 64, 1, 1, 0, 0, 1, 2,
                       WAVE.format(cfl=0.5, tend="1d10", slope=1, vx0="3.0d-1", drho="1.0d-1", K="5d0"), 60),
    "dustywave_dhllc": ("DUSTYWAVE", 64, 1, 1, 0, 0, 1, 3,
                        WAVE.format(cfl=0.5, tend="1d10", slope=3, vx0="3.0d-1", drho="1.0d-1", K="5d0"), 60),
    "si_2d": ("SI", 24, 24, 2, 0, 0, 1, 0, SI.format(slope=1), 25),
    "si_2d_s3_hllc": ("SI", 24, 24, 2, 0, 0, 2, 0, SI.format(slope=3), 25),
    # explicit drag in the force instead of the implicit drag step (setup.f90:355-362)
    "si_2d_explicit_drag": ("SI", 24, 24, 2, 0, 0, 1, 0,
                            SI.format(slope=1).replace("drag               = .true.", "drag               = .false."), 25),
    "si_2d_explicit_drag_s2": ("SI", 24, 24, 2, 0, 0, 2, 1,
                               SI.format(slope=2).replace("drag               = .true.", "drag               = .false."), 25),
    "disk_2d": ("disk_polar", 24, 32, 2, 2, 0, 1, 0, DISK.format(slope=1, r_relax="0.3d0", visc=".false.", alpha="1d-2"), 25),
    "disk_2d_relax_s2": ("disk_polar", 24, 32, 2, 2, 0, 1, 0, DISK.format(slope=2, r_relax="3.0d0", visc=".false.", alpha="1d-2"), 25),
    "sedov_hllc_s3": ("sedov", 32, 32, 0, 0, 0, 2, 0, SEDOV.format(cfl=0.5, slope=3), 40),
    "sedov_llf_s1": ("sedov", 40, 40, 0, 0, 0, 0, 0, SEDOV.format(cfl=0.8, slope=1), 40),
    "suite_disk_dust": ("disk_polar", 24, 32, 4, 2, 0, 1, 0,
                        suite_nml("DISK_DUST/disk_dust.nml", freq_out="100000", tend="1d10"), 30),
    "suite_disk_eq_accrete": ("disk_polar", 24, 32, 4, 2, 0, 1, 0,
                              suite_nml("DISK_EQ/disk_accrete.nml", freq_out="100000", tend="1d10"), 30),
    "suite_s71_accrete": ("disk_polar", 32, 24, 2, 2, 0, 1, 0,
                          suite_nml("DISK_S7-1/disk_accrete.nml", freq_out="100000", tend="1d10"), 30),
    "test_polar_disk": ("disk_polar", 24, 32, 2, 2, 0, 1, 0,
                        test_nml("polar_disk/disk.nml", freq_out="100000", tend="1d10"), 30),
    "test_polar_disk_s3": ("disk_polar", 32, 24, 1, 2, 0, 2, 0,
                           test_nml("polar_disk/disk_test.nml", freq_out="100000", tend="1d10"), 30),
    "test_cartesian_disk": ("disk", 32, 32, 2, 0, 0, 1, 0,
                            test_nml("CARTESIAN_disk/disk.nml", freq_out="100000", tend="1d10"), 30),
    "test_si": ("SI", 32, 32, 2, 0, 0, 1, 0, test_nml("SI/si.nml", freq_out="100000", tend="1d10"), 30),
    "test_kh": ("KH", 32, 32, 4, 0, 0, 2, 0, test_nml("KH/kh.nml", freq_out="100000", tend="1d10"), 30),
    "test_sod": ("SOD", 200, 1, 0, 0, 0, 1, 0, test_nml("Sod/sod.nml", freq_out="100000"), -1),
    "test_sod2d_x_up": ("SOD", 64, 8, 0, 0, 0, 2, 0,
                        test_nml("Sod2d/sod_x_up.nml", freq_out="100000", tend="1d10"), 40),
    "test_sod2d_y": ("SOD", 8, 64, 0, 0, 0, 1, 0, test_nml("Sod2d/sod_y.nml", freq_out="100000", tend="1d10"), 40),
    "test_dustywave": ("DUSTYWAVE", 128, 1, 1, 0, 0, 1, 0,
                       test_nml("DUSTYWAVE/dustywave.nml", freq_out="100000"), -1),
    "test_psi": ("SI", 32, 32, 8, 0, 0, 1, 0, test_nml("PSI/psi.nml", freq_out="100000", tend="1d10"), 30),
    "suite_si_ba": ("SI", 24, 24, 1, 0, 0, 1, 0, suite_nml("SI_BA/si.nml", freq_out="100000", tend="1d10"), 30),
    "disk_2d_logr": ("disk_polar", 24, 32, 2, 2, 1, 1, 0,
                     DISK.format(slope=3, r_relax="0.3d0", visc=".true.", alpha="1d-2"), 25),
    "disk_2d_visc": ("disk_polar", 24, 32, 2, 2, 0, 1, 0,
                     DISK.format(slope=1, r_relax="0.3d0", visc=".true.", alpha="5d-2"), 25),
    "disk_2d_visc_gas": ("disk_polar", 32, 24, 0, 2, 0, 2, 0,
                         DISK.format(slope=3, r_relax="0.3d0", visc=".true.", alpha="1d-1"), 25),
    "sod_hll_s3": ("SOD", 128, 1, 0, 0, 0, 1, 0, SOD.format(cfl=0.8, tend=0.245, slope=3, dir=0), -1),
    "sod_hllc_s1": ("SOD", 128, 1, 0, 0, 0, 2, 0, SOD.format(cfl=0.8, tend=0.245, slope=1, dir=0), -1),
    "sod_llf_s2": ("SOD", 96, 1, 0, 0, 0, 0, 0, SOD.format(cfl=0.5, tend=0.2, slope=2, dir=0), -1),
    "sod_2d_y_hll": ("SOD", 16, 64, 0, 0, 0, 1, 0, SOD.format(cfl=0.5, tend="1d10", slope=1, dir=1), 40),
    "sod_2d_x_hllc": ("SOD", 48, 8, 0, 0, 0, 2, 0, SOD.format(cfl=0.5, tend="1d10", slope=3, dir=0), 30),
    "kh_n1_hllc": ("KH", 32, 32, 1, 0, 0, 2, 0,
                   KH.format(cfl=0.5, slope=3, rho_diffuse="0.5d-18", scut="250d-7"), 40),
    "kh_n4_hll": ("KH", 24, 32, 4, 0, 0, 1, 2,
                  KH.format(cfl=0.4, slope=1, rho_diffuse="1.0d-18", scut="250d-7"), 40),
    "kh_n4_llf_dhllc": ("KH", 24, 32, 4, 0, 0, 0, 3,
                        KH.format(cfl=0.5, slope=2, rho_diffuse="0.7d-18", scut="250d-7"), 30),
    "cdisk_visc_n1": ("disk", 32, 32, 1, 0, 0, 1, 0,
                      CDISK.format(cfl=0.5, slope=1, visc=".true.", smooth="0.5d0", alpha="1d-2", mplanet="0.0d0"), 30),
    "cdisk_n2_s3": ("disk", 40, 32, 2, 0, 0, 2, 2,
                    CDISK.format(cfl=0.4, slope=3, visc=".false.", smooth="2.0d0", alpha="1d-2", mplanet="0.0d0"), 30),
    "cdisk_visc_gas": ("disk", 32, 24, 0, 0, 0, 1, 0,
                       CDISK.format(cfl=0.5, slope=2, visc=".true.", smooth="0.5d0", alpha="5d-2", mplanet="0.0d0"), 30),
    # species counts without an exact kernel instantiation (they run in the capacity kernels)
    "disk_2d_n5": ("disk_polar", 24, 32, 5, 2, 0, 1, 0,
                   DISK.format(slope=1, r_relax="0.3d0", visc=".false.", alpha="1d-2"), 25),
    "disk_2d_n12": ("disk_polar", 20, 24, 12, 2, 0, 1, 0,
                    DISK.format(slope=3, r_relax="0.3d0", visc=".false.", alpha="1d-2"), 25),
    "disk_2d_n20": ("disk_polar", 16, 24, 20, 2, 0, 1, 0,
                    DISK.format(slope=2, r_relax="0.3d0", visc=".false.", alpha="1d-2"), 20),
    # 100 steps: the shearing box, the polar disk and the PSI namelist's 8 species
    "si_2d_100": ("SI", 32, 32, 2, 0, 0, 1, 0, SI.format(slope=1), 100),
    "disk_2d_planet": ("disk_polar", 28, 40, 2, 2, 0, 1, 0,
                       DISK_PLANET.format(slope=1, r_relax="0.3d0", visc=".false.", alpha="1d-2", mplanet="3d-3"), 40),
    "disk_2d_planet_gas": ("disk_polar", 28, 40, 0, 2, 0, 2, 0,
                           DISK_PLANET.format(slope=3, r_relax="0.3d0", visc=".false.", alpha="1d-2", mplanet="1d-2"), 40),
    "disk_2d_selfgrav": ("disk_polar", 16, 24, 2, 2, 0, 1, 0,
                         DISK_SG.format(slope=1, r_relax="0.3d0", visc=".false.", alpha="1d-2"), 30),
    "disk_2d_selfgrav_gas": ("disk_polar", 20, 28, 0, 2, 0, 2, 0,
                             DISK_SG.format(slope=2, r_relax="0.3d0", visc=".false.", alpha="1d-2"), 30),
    "disk_2d_100": ("disk_polar", 32, 48, 4, 2, 0, 1, 0,
                    DISK.format(slope=1, r_relax="0.3d0", visc=".false.", alpha="1d-2"), 100),
    "test_psi_100": ("SI", 32, 32, 8, 0, 0, 1, 0, test_nml("PSI/psi.nml", freq_out="100000", tend="1d10"), 100),
    "cdisk_planet_n1": ("disk", 32, 32, 1, 0, 0, 1, 1,
                        CDISK.format(cfl=0.5, slope=1, visc=".true.", smooth="1.0d0", alpha="1d-2",
                                     mplanet="3.0d-3"), 40),
}

# FARGO runs: the reference's fargo_scheme as written is undefined (`one` undeclared and never
# assigned, fargo_scheme.f90:57,69; `eps` implicit REAL(4), :47), so these cases build it with
# `one` = 1 and double-precision implicit reals (REF_FARGO_FLAGS, applied to that file alone).
# "shifts": one signed integer per column (nx+4 of them); the harness then sets each column's
# fargo_velocity before every step so that the ring shifts by exactly that many cells (eps = 0;
# the velocities used are stored as fargo_v, one row per step).  Without "shifts" the setup's own
# velocity -Omega r is used; the reference then leaves the slopes of the last two active columns
# and of the last two rows unset (:20-31), so those fixtures are compared on the cells whose
# remap reads only set slopes.
FARGO_FLAGS = "-Done=1.0d0 -fdefault-real-8 -fdefault-double-8"
DISK_FARGO = DISK.replace("force_kick =.true.\n", "force_kick =.true.\nfargo = .true.\n")
assert DISK_FARGO != DISK


def _shifts(nx, ny, seed):
    rng = np.random.default_rng(seed)
    m = [0, 0]
    for _ in range(nx):
        mag = int(rng.choice([0, 1, 2, 4, 8, 16, ny // 2, ny - 1]))
        m.append(mag * int(rng.choice([-1, 1])))
    return m + [0, 0]


FARGO_CASES = {
    "fargo_int_disk": ("disk_polar", 24, 32, 2, 2, 0, 1, 0,
                       DISK_FARGO.format(slope=1, r_relax="0.3d0", visc=".false.", alpha="1d-2"), 6,
                       dict(shifts=_shifts(24, 32, 1))),
    "fargo_int_disk_s3": ("disk_polar", 20, 48, 4, 2, 0, 2, 0,
                          DISK_FARGO.format(slope=3, r_relax="0.3d0", visc=".false.", alpha="1d-2"), 6,
                          dict(shifts=_shifts(20, 48, 2))),
    "fargo_frac_disk": ("disk_polar", 24, 32, 2, 2, 0, 1, 0,
                        DISK_FARGO.format(slope=1, r_relax="0.3d0", visc=".false.", alpha="1d-2"), 1,
                        dict()),
    "fargo_frac_disk_s2": ("disk_polar", 28, 40, 4, 2, 0, 1, 0,
                           DISK_FARGO.format(slope=2, r_relax="0.3d0", visc=".false.", alpha="1d-2"), 1,
                           dict()),
    "fargo_frac_disk_s3": ("disk_polar", 24, 36, 3, 2, 0, 2, 0,
                           DISK_FARGO.format(slope=3, r_relax="0.3d0", visc=".false.", alpha="1d-2"), 1,
                           dict()),
}

# Collapse_1D (built with GRAVITY=1): its step turns the first active cell into NaN on the first
# step (gridinit_sphere1D, grid.f90:335-337, leaves radii_c = 0 in the inner ghost cells, so the
# spherical predictor source -rho u / radii_c, godunov.f90:207, is 0/0 there and the inner-face
# flux carries it in), and the NaN reaches every cell through the enclosed mass of the next step.
# The one-step runs are fixtures for every other active cell and for dt.
COLLAPSE_DUST = """drag = .true.
dust_distribution = 'mrn'
smin = 5d-7
scutmin = 1d-7
smax = 250d-7
scut = 1.0d0
rhograin = 2.3d0
dust2gas = 0.01d0
mrn = 3.5d0
growth = .false.
dust_ratio_min = 1d-30"""
COLLAPSE_CASES = {
    "collapse_gas_1": ("Collapse_1D", 128, 1, 0, 1, 0, 1, 0, COLLAPSE.format(cfl=0.8, slope=1, dust=""), 1,
                       "-DGRAVITY=1"),
    "collapse_dust_1": ("Collapse_1D", 96, 1, 4, 1, 0, 1, 0,
                        COLLAPSE.format(cfl=0.5, slope=1, dust=COLLAPSE_DUST), 1, "-DGRAVITY=1"),
}
DEFECTIVE = {"collapse_gas": COLLAPSE_CASES["collapse_gas_1"]}


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


def run_case(name, spec, force_build=False, fargo=None):
    setup, nx, ny, nd, geom, gs, solver, sdust, nml, nsteps = spec[:10]
    extra = list(spec[10:])
    bname = f"{setup}_{nx}x{ny}_n{nd}_g{geom}{gs}_s{solver}{sdust}" + ("_grav" if extra else "")
    benv = dict(os.environ)
    if fargo is not None:
        bname += "_fargofix"
        benv["REF_FARGO_FLAGS"] = FARGO_FLAGS
    benv["REF_NAME"] = bname
    exe = os.path.join(REPO, "oracle", "_ref", bname, "shark_ref")
    if force_build or not os.path.exists(exe):
        subprocess.run([BUILD, setup, str(nx), str(ny), str(nd), str(geom), str(gs), str(solver),
                        str(sdust)] + extra, check=True, stdout=subprocess.DEVNULL,
                       stderr=subprocess.DEVNULL, env=benv)
    fargo_v = None
    with tempfile.TemporaryDirectory() as td:
        nmlp = os.path.join(td, "run.nml")
        open(nmlp, "w").write(nml)
        env = dict(os.environ, OMP_NUM_THREADS="1")
        if fargo and "shifts" in fargo:
            sp = os.path.join(td, "shifts.txt")
            open(sp, "w").write(" ".join(str(x) for x in fargo["shifts"]) + "\n")
            env["REF_FARGO_SHIFTS"] = sp
        subprocess.run([exe, nmlp, str(nsteps), td], check=True, cwd=td, env=env,
                       stdout=subprocess.DEVNULL)
        if fargo and "shifts" in fargo:
            fargo_v = np.fromfile(os.path.join(td, "fargo_v.bin"))
        m = _meta(os.path.join(td, "meta.txt"))
        nc, nv, ndim = m["ncells"], m["nvar"], m["ndim"]
        ic = np.fromfile(os.path.join(td, "ic.bin")).reshape(nv, nc)
        final = np.fromfile(os.path.join(td, "final.bin")).reshape(nv, nc)
        dts = np.fromfile(os.path.join(td, "dt.bin"))
        aux = np.fromfile(os.path.join(td, "aux.bin"))
        evp = os.path.join(td, "eta_visc.bin")
        eta_visc = np.fromfile(evp) if os.path.exists(evp) else None
        dvp = os.path.join(td, "dvol.bin")
        dvol = np.fromfile(dvp) if os.path.exists(dvp) else None
    off = 0

    def take(n, shape=None):
        nonlocal off
        a = aux[off:off + n]
        off += n
        return a.reshape(shape) if shape else a

    arrs = dict(ic=ic, final=final, dt=dts)
    if eta_visc is not None:
        arrs["eta_visc"] = eta_visc
    if dvol is not None:
        arrs["dvol"] = dvol
    arrs["cs"] = take(nc)
    arrs["dx"] = take(nc * ndim, (ndim, nc))
    arrs["surf"] = take(nc * ndim, (ndim, nc))
    arrs["vol"] = take(nc)
    arrs["radii_c"] = take(nc)
    arrs["position"] = take(nc * ndim, (ndim, nc))
    arrs["active"] = take(nc)
    arrs["active_pred"] = take(nc)
    if nd > 0:
        arrs["sdust"] = take(nc * nd, (nd, nc))
    if m.get("fargo"):
        arrs["fargo_velocity"] = take(nc)
    if setup == "disk_polar":
        arrs["condinit"] = take(nc * nv, (nv, nc))
    assert off == aux.size, (name, off, aux.size)
    if fargo_v is not None:
        arrs["fargo_v"] = fargo_v.reshape(nsteps, nc)
        arrs["shifts"] = np.array(fargo["shifts"])
    m.update(setup=setup, geom=geom, gridspace=gs, solver=solver, solverdust=sdust,
             namelist=nml, case=name)
    np.savez_compressed(os.path.join(HERE, f"ref_{name}.npz"), meta=json.dumps(m), **arrs)
    return m


def check_collapse():
    """Runs the Collapse_1D case for one step into a scratch directory and reports the NaN cells."""
    name = "collapse_gas"
    spec = DEFECTIVE[name]
    setup, nx, ny, nd, geom, gs, solver, sdust, nml, nsteps = spec[:10]
    bname = f"{setup}_{nx}x{ny}_n{nd}_g{geom}{gs}_s{solver}{sdust}_grav"
    exe = os.path.join(REPO, "oracle", "_ref", bname, "shark_ref")
    if not os.path.exists(exe):
        subprocess.run([BUILD, setup, str(nx), str(ny), str(nd), str(geom), str(gs), str(solver), str(sdust)]
                       + list(spec[10:]), check=True, stdout=subprocess.DEVNULL, stderr=subprocess.DEVNULL,
                       env=dict(os.environ, REF_NAME=bname))
    with tempfile.TemporaryDirectory() as td:
        open(os.path.join(td, "run.nml"), "w").write(nml)
        subprocess.run([exe, os.path.join(td, "run.nml"), "1", td], check=True, cwd=td,
                       env=dict(os.environ, OMP_NUM_THREADS="1"), stdout=subprocess.DEVNULL)
        m = _meta(os.path.join(td, "meta.txt"))
        u = np.fromfile(os.path.join(td, "final.bin")).reshape(m["nvar"], m["ncells"])
        aux = np.fromfile(os.path.join(td, "aux.bin"))
        radii_c = aux[4 * m["ncells"]:5 * m["ncells"]]
    bad = np.flatnonzero(~np.isfinite(u).all(axis=0))
    print("cells with NaN after one step (1-based):", (bad + 1).tolist(), "radii_c of cells 1-3:", radii_c[:3])


def main(argv):
    if argv[1:] == ["--check-collapse"]:
        check_collapse()
        return
    names = argv[1:] or list(CASES) + list(FARGO_CASES) + list(COLLAPSE_CASES)
    for n in names:
        if n in FARGO_CASES:
            m = run_case(n, FARGO_CASES[n][:10], fargo=FARGO_CASES[n][10])
        elif n in COLLAPSE_CASES:
            m = run_case(n, COLLAPSE_CASES[n])
        else:
            m = run_case(n, CASES[n])
        print(n, "steps", m["nsteps"], "ncells", m["ncells"], "nvar", m["nvar"])


if __name__ == "__main__":
    main(sys.argv)
