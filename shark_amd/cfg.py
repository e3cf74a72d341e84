"""ctypes mirror of ``shark_cfg`` (include/shark_gpu.h) and the derived sizes.

The struct carries the reference's compile-time macro set (bin/Makefile:5-36) and the
namelist scalars the step reads (&GRID_PARAMS, src/commons/read_params.f90:10;
&DUST_PARAMS, src/dust/dust_commons.f90; &SETUP_PARAMS of each setup).
"""
import ctypes as C
from dataclasses import dataclass

SHARK_CFG_VERSION = 2
SHARK_MAX_DUST = 64

GEOM_CARTESIAN, GEOM_SPHERE, GEOM_POLAR = 0, 1, 2
SOLVER_LLF, SOLVER_HLL, SOLVER_HLLC = 0, 1, 2
DUST_HUANG_BAI, DUST_LLF, DUST_HLL, DUST_HLLC = 0, 1, 2, 3
BC_PERIODIC, BC_SHEAR_SI, BC_DISK, BC_SOD, BC_ZERO_GRADIENT, BC_COLLAPSE = 0, 1, 2, 3, 4, 5
TSTOP_DUSTYWAVE, TSTOP_SI, TSTOP_EPSTEIN_DISK, TSTOP_EPSTEIN, TSTOP_EPSTEIN_CDISK, TSTOP_COLLAPSE = 0, 1, 2, 3, 4, 5
FORCE_NONE, FORCE_SI, FORCE_DISK, FORCE_CDISK = 0, 1, 2, 3
INLOOP_NONE, INLOOP_SI, INLOOP_DISK, INLOOP_CDISK = 0, 1, 2, 3


class ShARKCfg(C.Structure):
    _fields_ = [
        ("version", C.c_int32), ("nx", C.c_int32), ("ny", C.c_int32), ("ndust", C.c_int32),
        ("nghost", C.c_int32), ("geom", C.c_int32), ("solver", C.c_int32),
        ("solverdust", C.c_int32), ("slope_type", C.c_int32), ("iso_cs", C.c_int32),
        ("force_kick", C.c_int32), ("fargo", C.c_int32), ("drag", C.c_int32),
        ("dust_back_reaction", C.c_int32), ("growth", C.c_int32), ("viscosity", C.c_int32),
        ("bc_kind", C.c_int32), ("tstop_kind", C.c_int32), ("force_kind", C.c_int32),
        ("inloop_kind", C.c_int32), ("disk_bc_mode", C.c_int32), ("shock_dir", C.c_int32),
        ("gamma", C.c_double), ("cfl", C.c_double), ("smallr", C.c_double),
        ("smallc", C.c_double), ("smallp", C.c_double), ("dust_ratio_min", C.c_double),
        ("K", C.c_double * SHARK_MAX_DUST),
        ("q_shear", C.c_double), ("omega_shear", C.c_double), ("box_l", C.c_double),
        ("rad0", C.c_double), ("eta_stream", C.c_double), ("rho_init", C.c_double),
        ("hoverr", C.c_double), ("rhograin", C.c_double),
        ("mstar", C.c_double), ("r_relax", C.c_double), ("n_rel", C.c_double),
        ("smooth_r", C.c_double), ("rho_l", C.c_double), ("rho_r", C.c_double),
        ("p_l", C.c_double), ("p_r", C.c_double),
        ("box_l_y", C.c_double), ("mplanet", C.c_double), ("rplanet", C.c_double),
        ("sigma_r0", C.c_double), ("r0_disk", C.c_double), ("frac_r_disk", C.c_double),
        ("buffer_in", C.c_double), ("buffer_out", C.c_double), ("alpha_visc", C.c_double),
        ("dust2gas", C.c_double), ("reserved", C.c_double * 2),
        ("accretion", C.c_int32), ("gravity", C.c_int32), ("m_acc", C.c_double),
        ("phi_mom", C.c_double), ("l_out", C.c_double), ("eos_baro", C.c_int32),
        ("self_gravity", C.c_int32), ("l_soft", C.c_double), ("m_cent", C.c_double),
        ("t0_cloud", C.c_double), ("mu_gas", C.c_double), ("unit_v", C.c_double),
        ("unit_nh", C.c_double), ("a_smooth_pl", C.c_double),
    ]


def make_cfg(**kw) -> ShARKCfg:
    """Build a cfg with the reference's defaults (hydro_commons.f90:98-109,
    parameters.f90:47-54, dust_commons.f90:85-108) overridden by keyword arguments."""
    c = ShARKCfg()
    c.version = SHARK_CFG_VERSION
    c.nghost = 2
    c.ny = 1
    c.geom = GEOM_CARTESIAN
    c.solver = SOLVER_HLL
    c.solverdust = DUST_HUANG_BAI
    c.slope_type = 1
    c.iso_cs = 1
    c.drag = 1
    c.dust_back_reaction = 1
    c.gamma = 1.66667
    c.cfl = 0.5
    c.phi_mom = 2.0
    c.l_out = 1.0
    c.smallr = 1e-27
    c.smallc = 1e-27
    c.smallp = 1e-27 * 1e-27 ** 2
    c.dust_ratio_min = 1e-11
    for k, v in kw.items():
        if k == "K":
            for i, kk in enumerate(v):
                c.K[i] = kk
        else:
            if not hasattr(c, k):
                raise KeyError(k)
            setattr(c, k, v)
    return c


@dataclass
class Sizes:
    nx: int
    ny: int
    ndust: int

    @property
    def ndim(self):
        return 2 if self.ny > 1 else 1

    @property
    def nx_max(self):
        return self.nx + 4

    @property
    def ny_max(self):
        return self.ny + 4 if self.ny > 1 else 1

    @property
    def ncells(self):
        return self.nx_max * self.ny_max

    @property
    def nvar(self):
        return 5 + 4 * self.ndust


def sizes_of(cfg: ShARKCfg) -> Sizes:
    return Sizes(cfg.nx, cfg.ny, cfg.ndust)


def algorithmic_bytes(cfg: ShARKCfg) -> int:
    """HBM bytes per step a perfect implementation must move: one read and one write of
    u_prim plus one read of cs per active cell, and a second read+write pass for FARGO."""
    s = sizes_of(cfg)
    per_cell = 8 * (2 * s.nvar + 1) + (16 * s.nvar if cfg.fargo else 0)
    return per_cell * s.nx * (s.ny if s.ny > 1 else 1)
