"""Load the reference fixtures (tests/golden/ref_*.npz) as (cfg, geometry, state)."""
import json
import os
import re

import numpy as np

from shark_amd import cfg as C

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
REF_CASES = sorted(f[4:-4] for f in os.listdir(GOLDEN)
                   if f.startswith("ref_") and f.endswith(".npz")
                   and not f.startswith(("ref_growth_", "ref_charge", "ref_output_", "ref_marchand",
                                         "ref_fargo_", "ref_collapse_")))
# Collapse_1D fixtures (make_ref_fixtures.py COLLAPSE_CASES): one reference step, compared on the
# active cells the reference leaves finite (tests/test_collapse_oracle.py)
COLLAPSE_REF_CASES = sorted(f[4:-4] for f in os.listdir(GOLDEN)
                            if f.startswith("ref_collapse_") and f.endswith(".npz"))
# FARGO fixtures (make_ref_fixtures.py FARGO_CASES): integer shifts set before every step, or the
# setup's own shift velocity for one step, compared where the reference's remap is defined
FARGO_REF_CASES = sorted(f[4:-4] for f in os.listdir(GOLDEN)
                         if f.startswith("ref_fargo_") and f.endswith(".npz"))


def _K(nml):
    return [float(v.replace("d", "e").replace("D", "e"))
            for v in re.findall(r"K\(\d+\)\s*=\s*([0-9.eEdD+-]+)", nml)]


def load_ref(name):
    z = np.load(os.path.join(GOLDEN, f"ref_{name}.npz"))
    m = json.loads(str(z["meta"]))
    kw = dict(nx=m["nx"], ny=m["ny"], ndust=m["ndust"], geom=m["geom"], solver=m["solver"],
              solverdust=m["solverdust"], slope_type=m["slope_type"], iso_cs=m["iso_cs"],
              force_kick=int(m["force_kick"]), fargo=int(m["fargo"]), drag=int(m.get("drag", 1)),
              dust_back_reaction=int(m.get("dust_back_reaction", 1)), gamma=m["gamma"],
              cfl=m["CFL"], smallr=m["smallr"], smallc=m["smallc"], smallp=m["smallp"],
              dust_ratio_min=m.get("dust_ratio_min", 1e-11))
    setup = m["setup"]
    if setup == "DUSTYWAVE":
        kw.update(bc_kind=C.BC_PERIODIC, tstop_kind=C.TSTOP_DUSTYWAVE, force_kind=C.FORCE_NONE,
                  inloop_kind=C.INLOOP_NONE, K=_K(m["namelist"]))
    elif setup == "SI":
        kw.update(bc_kind=C.BC_SHEAR_SI, tstop_kind=C.TSTOP_SI, force_kind=C.FORCE_SI,
                  inloop_kind=C.INLOOP_SI, q_shear=m["q_shear"], omega_shear=m["omega_shear"],
                  box_l=m["box_l"], rad0=m["rad0"], eta_stream=m["eta_stream"],
                  rho_init=m["rho_init"], hoverr=m["hoverr"], rhograin=m["rhograin"] / m["unit_d"])
    elif setup == "disk_polar":
        kw.update(bc_kind=C.BC_DISK, tstop_kind=C.TSTOP_EPSTEIN_DISK, force_kind=C.FORCE_DISK,
                  inloop_kind=C.INLOOP_DISK, mstar=m["mstar"], r_relax=m["r_relax"],
                  n_rel=m["n_rel"], smooth_r=m["smooth_r"], hoverr=m["hoverr"],
                  rhograin=m.get("rhograin", 0.0) / m["unit_d"], disk_bc_mode=0,
                  viscosity=int(m.get("viscosity", False)), alpha_visc=m.get("alpha_visc", 0.0),
                  box_l=m.get("box_l", 0.0), accretion=int(m.get("accretion", False)),
                  m_acc=m.get("M_acc", 0.0) / (365.25 * 24.0 * 3600.0) * m["unit_t"],
                  phi_mom=m.get("phi_mom", 2.0), l_out=m.get("L_out", 1.0), dust2gas=m.get("dust2gas", 0.0),
                  mplanet=m.get("mplanet", 0.0), rplanet=m.get("rplanet", 0.0),
                  a_smooth_pl=m.get("a_smooth_pl", 0.0), self_gravity=int(m.get("self_gravity", False)))
    elif setup == "SOD":
        kw.update(bc_kind=C.BC_SOD, tstop_kind=C.TSTOP_DUSTYWAVE, force_kind=C.FORCE_NONE,
                  inloop_kind=C.INLOOP_NONE, rho_l=m["rho_l"], rho_r=m["rho_r"], p_l=m["p_l"],
                  p_r=m["p_r"], shock_dir=m["shock_dir"])
    elif setup == "disk":
        kw.update(bc_kind=C.BC_ZERO_GRADIENT, tstop_kind=C.TSTOP_EPSTEIN_CDISK, force_kind=C.FORCE_CDISK,
                  inloop_kind=C.INLOOP_CDISK, viscosity=int(m["viscosity"]), box_l=m["box_l"],
                  box_l_y=m["box_l_y"], sigma_r0=m["sigma_R0"], r0_disk=m["R0_disk"], mstar=m["mstar"],
                  mplanet=m["mplanet"], rplanet=m["rplanet"], frac_r_disk=m["frac_R_disk"],
                  buffer_in=m["buffer_in"], buffer_out=m["buffer_out"], hoverr=m["hoverr"],
                  smooth_r=m["smooth_r"], alpha_visc=m["alpha_visc"], dust2gas=m.get("dust2gas", 0.0),
                  rhograin=m.get("rhograin", 0.0) / m["unit_d"])
    elif setup == "sedov":
        kw.update(bc_kind=C.BC_PERIODIC, tstop_kind=C.TSTOP_DUSTYWAVE, force_kind=C.FORCE_NONE,
                  inloop_kind=C.INLOOP_NONE)
    elif setup == "KH":
        kw.update(bc_kind=C.BC_PERIODIC, tstop_kind=C.TSTOP_EPSTEIN, force_kind=C.FORCE_NONE,
                  inloop_kind=C.INLOOP_NONE, rhograin=m["rhograin"] / m["unit_d"])
    elif setup == "Collapse_1D":
        kw.update(bc_kind=C.BC_COLLAPSE, tstop_kind=C.TSTOP_COLLAPSE, force_kind=C.FORCE_NONE,
                  inloop_kind=C.INLOOP_NONE, gravity=1, eos_baro=int(m["non_standard_eos"] == 1),
                  l_soft=m["l_soft"] / m["unit_l"], m_cent=m["m_cent"], t0_cloud=m["t0_cloud"],
                  mu_gas=m["mu_gas"], unit_v=m["unit_v"], unit_nh=m["unit_nh"],
                  rhograin=m.get("rhograin", 0.0) / m["unit_d"])
    cfg = C.make_cfg(**kw)
    geom = dict(dx=z["dx"], surf=z["surf"], vol=z["vol"], radii_c=z["radii_c"], position=z["position"])
    if "dvol" in z:
        geom["dvol"] = z["dvol"]
    state = dict(u=z["ic"], cs=z["cs"], sdust=z["sdust"] if "sdust" in z else None,
                 condinit=z["condinit"] if "condinit" in z else None,
                 fargo_velocity=z["fargo_velocity"] if "fargo_velocity" in z else None)
    return cfg, geom, state, z["final"], z["dt"], m


def active_mask(cfg):
    s = C.sizes_of(cfg)
    if s.ndim == 1:
        a = np.zeros(s.nx_max, bool)
        a[2:2 + s.nx] = True
        return a
    a = np.zeros((s.ny_max, s.nx_max), bool)
    a[2:2 + s.ny, 2:2 + s.nx] = True
    return a.reshape(-1)


def rel_l1(a, b, mask=None):
    """Per-field relative L1 over active cells: sum|a-b| / max(sum|b|, tiny); absolute L1
    for fields that are identically zero in the reference."""
    if mask is not None:
        a = a[:, mask]
        b = b[:, mask]
    num = np.abs(a - b).sum(axis=1)
    den = np.abs(b).sum(axis=1)
    return np.where(den > 0, num / np.where(den > 0, den, 1.0), num)


def run_fargo_fixture(name, make, get_u):
    """Run a FARGO fixture through a stepper: make(cfg, geom, state) builds it (oracle or GPU),
    get_u(stepper) returns its u_prim.  Integer-shift fixtures restart from the previous state
    with that step's fargo_velocity before every step (the reference harness sets it per step).
    Returns (dts, u, final, mask): mask selects the cells to compare (all cells for integer
    shifts; for the one-step fixtures with the setup's velocity, the active cells whose remap
    reads only the slopes the reference sets, fargo_scheme.f90:20-31)."""
    cfg, geom, state, final, dts, meta = load_ref(name)
    z = np.load(os.path.join(GOLDEN, f"ref_{name}.npz"))
    st = dict(state)
    got = []
    if "fargo_v" in z:
        for s in range(len(dts)):
            st["fargo_velocity"] = z["fargo_v"][s]
            x = make(cfg, geom, st)
            got.append(x.step())
            st["u"] = get_u(x)
        mask = np.ones(final.shape[1], bool)
    else:
        x = make(cfg, geom, st)
        got = [x.step() for _ in range(len(dts))]
        st["u"] = get_u(x)
        mask = fargo_defined_cells(cfg, geom, state["fargo_velocity"], dts[0])
    return np.array(got), st["u"], final, mask


def fargo_defined_cells(cfg, geom, vk, dt):
    """Active cells of columns 3..nx whose two source rows (fargo_scheme.f90:46-67, one wrap) lie
    in rows 3..ny, where the reference has set the slopes."""
    nxm, nym = cfg.nx + 4, cfg.ny + 4
    mask = np.zeros((nym, nxm), bool)
    dx2 = geom["dx"][1].reshape(nym, nxm)
    rc = geom["radii_c"].reshape(nym, nxm)
    v = np.asarray(vk).reshape(nym, nxm)
    ny = cfg.ny
    for iy in range(3, ny + 3):
        for ix in range(3, cfg.nx + 1):
            c = (iy - 1, ix - 1)
            dy = float(dx2[c]) * float(rc[c])
            nshft = int(abs(float(v[c]) * dt) / dy)
            if v[c] > 0.0:
                j = iy - nshft - 1
                if j < 3:
                    j += ny
                jp = j + 1
                if jp < 3:
                    jp += ny
            else:
                j = iy + nshft
                if j > ny + 2:
                    j -= ny
                jp = j + 1
                if jp > ny + 2:
                    jp -= ny
            mask[c] = 3 <= j <= ny and 3 <= jp <= ny
    return mask.reshape(-1)
