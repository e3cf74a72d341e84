"""Initial states for the workloads, built the way the reference's setups build them.

Each function returns ``(cfg, geom, state)``: the ``shark_cfg`` struct, the grid arrays of
gridinit (src/grid/grid.f90:404-496) and the state after setup (u_prim after primtoc,
src/solver/solve.f90:265-302, plus cs, sdust, uprim_condinit, fargo_velocity).

* ``dustywave``  — src/setups/DUSTYWAVE/setup.f90:1-79 (1D periodic sound wave + drag)
* ``si_box``     — src/setups/SI/setup.f90:1-163: multi-species NSH equilibrium in a
                   shearing box, here with seeded Gaussian velocity noise on every fluid
* ``disk``       — src/setups/disk_polar/setup.f90:1-84 on the linear polar grid
                   (gridinit_disk_log with GRIDSPACE=0), with per-species grain sizes and
                   seeded density noise
Noise comes from numpy's Philox generator keyed by the seed, so every run of the same
workload starts from the same state.
"""
import math

import numpy as np

from . import cfg as C

PI = 3.141592653589793238462643383279  # src/commons/phys_const.f90:14
AU = 1.5e13
MSUN = 2e33
GRAV = 6.67e-8


def _f32(x):
    """A Fortran default-REAL literal or float() conversion promoted to double."""
    return float(np.float32(x))


# --------------------------------------------------------------------------- grids
def grid_cartesian_1d(nx, box_l):
    """gridinit, GEOM=0, ndim=1 (grid.f90:408-414): positions accumulate dx over active cells."""
    nxm = nx + 4
    dx = box_l / _f32(nx)
    g = dict(dx=np.full((1, nxm), dx), surf=np.ones((1, nxm)), vol=np.full(nxm, dx),
             radii_c=np.ones(nxm), position=np.zeros((1, nxm)))
    pos = 0.0
    for ix in range(1, nxm + 1):
        if 3 <= ix <= nx + 2:
            pos = pos + dx
            g["position"][0, ix - 1] = pos
    return g


def _ix_iy(nx, ny):
    nxm, nym = nx + 4, ny + 4
    ix = np.tile(np.arange(1, nxm + 1, dtype=np.float64), nym)
    iy = np.repeat(np.arange(1, nym + 1, dtype=np.float64), nxm)
    return ix, iy


def grid_cartesian_2d(nx, ny, lx, ly):
    """gridinit, GEOM=0, ndim=2 (grid.f90:416-428)."""
    ix, iy = _ix_iy(nx, ny)
    d1 = lx / _f32(nx)
    d2 = ly / _f32(ny)
    nc = ix.size
    dx = np.stack([np.full(nc, d1), np.full(nc, d2)])
    pos = np.stack([(((ix - 3) + 0.5) * lx) / _f32(nx), (((iy - 3) + 0.5) * ly) / _f32(ny)])
    return dict(dx=dx, surf=dx.copy(), vol=np.full(nc, d1 * d2), radii_c=np.ones(nc), position=pos)


def grid_polar(nx, ny, box_l, inner_r):
    """gridinit_disk_log, GEOM=2, GRIDSPACE=0 (grid.f90:480-495)."""
    ix, iy = _ix_iy(nx, ny)
    nc = ix.size
    d1 = box_l / _f32(nx)
    d2 = (2.0 * PI) / _f32(ny)
    rc = inner_r + (((ix - 3) + 0.5) * box_l) / _f32(nx)
    phi = (((iy - 3) + 0.5) * (2.0 * PI)) / _f32(ny)
    dx = np.stack([np.full(nc, d1), np.full(nc, d2)])
    surf = np.stack([(rc - 0.5 * d1) * d2, np.full(nc, d1)])
    vol = (rc * d1) * d2
    pos = np.stack([rc * np.cos(phi), rc * np.sin(phi)])
    return dict(dx=dx, surf=surf, vol=vol, radii_c=rc, position=pos)


def primtoc(q, nd, gamma):
    """primtoc, src/solver/solve.f90:265-302."""
    u = np.empty_like(q)
    u[0] = q[0]
    u[1] = q[0] * q[1]
    u[2] = q[0] * q[2]
    u[3] = q[0] * q[3]
    u[4] = q[4] / (gamma - 1.0) + 0.5 * q[0] * q[1] ** 2 + 0.5 * q[0] * q[2] ** 2 + 0.5 * q[0] * q[3] ** 2
    for k in range(nd):
        r, x, y, z = 5 + k, 5 + nd + k, 5 + 2 * nd + k, 5 + 3 * nd + k
        u[r] = q[r]
        u[x] = q[r] * q[x]
        u[y] = q[r] * q[y]
        u[z] = q[r] * q[z]
    return u


# --------------------------------------------------------------------------- setups
def sod(nx=1024, ny=1, shock_dir=0, cfl=0.8, slope_type=3, solver=C.SOLVER_HLL, gamma=1.4,
        rho_l=1.0, rho_r=0.125, p_l=1.0, p_r=0.1, box_l=1.0, box_l_y=1.0):
    """Sod shock tube (SOD setup.f90:1-43, boundaries.f90:10-80): adiabatic gas (iso_cs < 1), left
    state where position(:, shock_dir+1) <= 1/2, right state beyond, at rest; the ghost layers
    across the shock keep the two states.  Defaults: the reference's test-suite tube
    (test/test-suite/Sod-tube/sod.nml, 1024 cells)."""
    cfg = C.make_cfg(nx=nx, ny=ny, ndust=0, iso_cs=-1, cfl=cfl, slope_type=slope_type, solver=solver,
                     gamma=gamma, bc_kind=C.BC_SOD, tstop_kind=C.TSTOP_DUSTYWAVE, force_kind=C.FORCE_NONE,
                     inloop_kind=C.INLOOP_NONE, shock_dir=shock_dir, rho_l=rho_l, rho_r=rho_r, p_l=p_l,
                     p_r=p_r)
    g = grid_cartesian_1d(nx, box_l) if ny == 1 else grid_cartesian_2d(nx, ny, box_l, box_l_y)
    pos = g["position"][shock_dir]
    right = pos > 0.5
    nc = pos.size
    q = np.zeros((5, nc))
    q[0] = np.where(right, rho_r, rho_l)
    q[4] = np.where(right, p_r, p_l)
    u = primtoc(q, 0, gamma)
    return cfg, g, dict(u=u, cs=np.zeros(nc), sdust=None, condinit=None, fargo_velocity=None)


# src/commons/phys_const.f90:8-12
MH, KB, GRAV, PC = 1.67e-24, 1.38e-16, 6.67e-8, 3e18


def _powi(a, b):
    """a**b for an integer b, by binary powering as flang lowers it."""
    recip = b < 0
    b = abs(b)
    r = 1.0
    while True:
        if b & 1:
            r *= a
        b //= 2
        if b == 0:
            break
        a *= a
    return 1.0 / r if recip else r


def units_kh(mu_gas=2.31):
    """The KH setup's units module (src/setups/KH/units.f90)."""
    unit_l = PC
    unit_d = mu_gas * MH
    unit_t = 1.0 / math.sqrt(GRAV * unit_d)
    unit_v = unit_l / unit_t
    return dict(unit_l=unit_l, unit_d=unit_d, unit_t=unit_t, unit_v=unit_v,
                unit_p=unit_d * (unit_v * unit_v), unit_m=_powi(unit_l, 3) * unit_d,
                unit_dcol=unit_d * unit_l)


def mrn_sizes_and_ratios(ndust, units, smin=1e-7, smax=1e-2, scutmin=5e-7, scut=250e-7, mrn=3.5,
                         rhograin=2.7, dust2gas=0.01, ice_mantle=8.7e-7, dust_ratio_min=1e-11):
    """Grain sizes (code units) and dust-to-gas ratios of distribution_dust(.true.) with the MRN
    law (src/dust/dust_init.f90:11-103, mrn_distri :164-197; kernel_type 0): logarithmic bins
    from smin to smax, the power law normalised over the bins between scutmin and scut, other
    bins at dust_ratio_min, each size sqrt(a- a+) plus the ice mantle."""
    ul, um = units["unit_l"], units["unit_m"]
    f43 = _f32(4.0 / 3.0)
    zeta = math.pow(smax / smin, 1.0 / ndust)
    eta = zeta * zeta * zeta
    m_min = (((f43 * PI) * rhograin) * (smin * smin * smin)) / um
    am = [(smin * _powi(zeta, k - 1)) / ul for k in range(1, ndust + 1)]
    ap = [(smin * _powi(zeta, k)) / ul for k in range(1, ndust + 1)]
    massmin = m_min * _powi(eta, 0)
    mono1 = (((f43 * PI) * rhograin) * (scutmin * scutmin * scutmin)) / um
    mono2 = (((f43 * PI) * rhograin) * (scut * scut * scut)) / um
    icutmin = max(1, math.floor(math.log(mono1 / massmin) / math.log(eta) + 1))
    icutmax = min(math.floor(math.log(mono2 / massmin) / math.log(eta) + 1), ndust)
    e = 4.0 - mrn
    eps = [0.0] * ndust
    norm = 0.0
    for k in range(icutmin, icutmax + 1):
        t = math.pow(ap[k - 1], e) - math.pow(am[k - 1], e)
        norm = norm + t
        eps[k - 1] = t
    eps = [max((dust2gas * x) / norm, dust_ratio_min) for x in eps]
    sizes = [math.sqrt(am[k] * ap[k]) + ice_mantle / ul for k in range(ndust)]
    return sizes, eps


def kh(nx=256, ny=256, ndust=1, cfl=0.5, slope_type=3, solver=C.SOLVER_HLLC, solverdust=0,
       gamma=1.666667, box_l=4.8481e-6, box_l_y=4.8481e-6, rho_dense=1e-18, rho_diffuse=1e-18,
       T_dense=10.0, mach=0.25, vy0=0.1, kx=2.0, smin=5e-7, scutmin=1e-7, scut=250e-7, smax=0.1,
       rhograin=2.3, dust2gas=0.01, mrn=3.5, mu_gas=2.31):
    """Kelvin-Helmholtz shear layer with dust (src/setups/KH/setup.f90:1-52): adiabatic gas at
    Mach/2 of the T_dense sound speed, streaming in +x outside |y| < box_l_y/4 and in -x inside,
    with a vy0 cos(2 kx pi x / box_l) seed; dust at rest, one species at dust-to-gas 0.01 and size
    scut, or the MRN distribution; Epstein stopping times with the step's sound speed
    (:218-236); periodic boundaries; drag with heating of the gas.  Defaults: the reference's
    test/KH/kh.nml (1 au box, cgs inputs, units of src/setups/KH/units.f90)."""
    u = units_kh(mu_gas)
    smallr = 1e-27 / u["unit_d"]
    smallc = 1e-27 / u["unit_v"]
    cfg = C.make_cfg(nx=nx, ny=ny, ndust=ndust, iso_cs=-1, cfl=cfl, slope_type=slope_type,
                     solver=solver, solverdust=solverdust, gamma=gamma, bc_kind=C.BC_PERIODIC,
                     tstop_kind=C.TSTOP_EPSTEIN, force_kind=C.FORCE_NONE, inloop_kind=C.INLOOP_NONE,
                     rhograin=rhograin / u["unit_d"], smallr=smallr, smallc=smallc,
                     smallp=smallr * (smallc * smallc))
    g = grid_cartesian_2d(nx, ny, box_l, box_l_y)
    nc = g["vol"].size
    cs = math.sqrt(((gamma * KB) * T_dense) / (mu_gas * MH))
    nv = 5 + 4 * ndust
    q = np.zeros((nv, nc))
    acos1 = math.acos(-1.0)
    for c in range(nc):
        xx = g["position"][0, c] - 0.5 * box_l
        yy = g["position"][1, c] - 0.5 * box_l_y
        rho = rho_dense / u["unit_d"]
        vx = ((cs * mach) / u["unit_v"]) * 0.5
        vy = (((vy0 * cs) * mach) * math.cos((((xx * 2.0) * kx) * acos1) / box_l)) / u["unit_v"]
        if abs(yy / box_l_y) < 0.25:
            rho = rho_diffuse / u["unit_d"]
            vx = -vx
        q[0, c], q[1, c], q[2, c] = rho, vx, vy
        q[4, c] = (rho_dense * (cs * cs)) / u["unit_p"]
    sd = np.zeros((ndust, nc))
    if ndust == 1:
        sizes, eps = [scut / u["unit_l"]], [_f32(0.01)]
    elif ndust > 1:
        sizes, eps = mrn_sizes_and_ratios(ndust, u, smin=smin, smax=smax, scutmin=scutmin, scut=scut,
                                          mrn=mrn, rhograin=rhograin, dust2gas=dust2gas)
    for k in range(ndust):
        q[5 + k] = eps[k] * q[0]
        sd[k] = sizes[k]
    return cfg, g, dict(u=primtoc(q, ndust, gamma), cs=np.zeros(nc), sdust=sd if ndust else None,
                        condinit=None, fargo_velocity=None)


AU, MSUN = 1.5e13, 2e33   # src/commons/phys_const.f90:11,13


AU, MSUN = 1.5e13, 2e33   # phys_const.f90:11, 13


def barotrop(nH, t0_cloud=10.0, unit_nh=1.0):
    """barotrop (src/setups/Collapse_1D/eos.f90:1-11); the first exponent is the default-real
    literal 0.8 (single precision)."""
    x = nH * unit_nh
    return ((t0_cloud * math.sqrt(1.0 + math.pow(x / 1e11, float(np.float32(0.8))))) *
            math.pow(1.0 + x / 1e16, -0.3)) * math.pow(1.0 + x / 1e21, 0.56667)


def cs_eos(T, gamma, unit_v, mu_gas=2.31):
    """cs_eos (eos.f90:13-20)."""
    return math.sqrt(((gamma * KB) * T) / (mu_gas * MH)) / unit_v


def grid_sphere_1d(nx, rmax, rin, unit_l):
    """gridinit_sphere1D (src/grid/grid.f90:304-357): the faces radii(i) = inner_r zeta^(i-fa) of
    the active cells (integer powers as flang lowers them), the first cell spanning [0, inner_r],
    from rin (au) to rmax (cm).  Ghost cells as the reference leaves them: radii_c 0 in the inner
    two and (0 + r_last)/2, 0 in the outer two; dx of the inner ghosts that of the first cell, 0
    in the outer; vol, dvol and surf 0 (no flux through the centre or the outer edge).
    dvol = (radii^3 - radii_c^3)/3 is the part of a cell between its centre and its outer face."""
    nxm = nx + 4
    fa, la = 3, nx + 2
    zeta = math.pow(rmax / (rin * AU), 1.0 / (nx - 1))
    inner = (rin * AU) / unit_l
    radii = np.zeros(nxm)
    for i in range(fa, la + 1):
        radii[i - 1] = inner * _powi(zeta, i - fa)
    dx = np.zeros(nxm)
    for i in range(fa, la + 1):
        dx[i - 1] = radii[i - 1] - radii[i - 2]
    dx[:fa] = dx[fa - 1]
    rc = np.zeros(nxm)
    for i in range(fa, nxm + 1):
        rc[i - 1] = (radii[i - 1] + radii[i - 2]) / 2.0
    vol, dvol, surf = np.zeros(nxm), np.zeros(nxm), np.zeros(nxm)
    for i in range(fa, la + 1):
        vol[i - 1] = (_powi(radii[i - 1], 3) - _powi(radii[i - 2], 3)) / 3.0
        dvol[i - 1] = (_powi(radii[i - 1], 3) - _powi(rc[i - 1], 3)) / 3.0
        surf[i - 1] = radii[i - 2] * radii[i - 2]
    return dict(dx=dx.reshape(1, nxm), surf=surf.reshape(1, nxm), vol=vol, radii_c=rc,
                position=rc.reshape(1, nxm).copy(), dvol=dvol)


def collapse(nx=1024, ndust=0, cfl=0.5, slope_type=1, solver=C.SOLVER_HLL, solverdust=0, gamma=1.66667,
             m_cloud=1.0, alpha_cloud=0.25, t0_cloud=10.0, rin=0.1, mu_gas=2.31, rhograin=2.3,
             dust2gas=0.01, smin=5e-7, smax=250e-7, scutmin=1e-7, scut=1.0, mrn=3.5, dust_ratio_min=1e-30):
    """Collapse_1D (src/setups/Collapse_1D/setup.f90:1-60, units.f90, eos.f90): a uniform cloud of
    M_cloud solar masses at rest on the spherical 1D grid, radius 2/5 G M/(kB T0) mu mH alpha_cloud,
    barotropic EOS (P = rho cs_eos(barotrop(rho))^2, iso_cs < 1), 1D self-gravity, the collapse
    boundaries; dust from the MRN distribution at dust2gas.  Defaults: the reference's test-suite
    collapse runs (test/test-suite/collapse, dustycollapse)."""
    u = units_kh(mu_gas)
    unit_nh = u["unit_d"] / (mu_gas * MH)
    mcloud = m_cloud * MSUN
    r_cloud = 2.0 / 5.0 * GRAV * mcloud / (KB * t0_cloud) * mu_gas * MH * alpha_cloud
    rho_cloud = mcloud / ((4.0 / 3.0 * PI) * _powi(r_cloud, 3)) / u["unit_d"]
    g = grid_sphere_1d(nx, r_cloud, rin, u["unit_l"])
    nc = nx + 4
    nv = 5 + 4 * ndust
    cs = cs_eos(barotrop(rho_cloud, t0_cloud, unit_nh), gamma, u["unit_v"], mu_gas)
    q = np.zeros((nv, nc))
    q[0] = rho_cloud
    q[4] = rho_cloud * (cs * cs)
    sdust = None
    if ndust:
        sizes, eps = mrn_sizes_and_ratios(ndust, u, smin=smin, smax=smax, scutmin=scutmin, scut=scut,
                                          mrn=mrn, rhograin=rhograin, dust2gas=dust2gas,
                                          dust_ratio_min=dust_ratio_min)
        for k in range(ndust):
            q[5 + k] = eps[k] * q[0]
        sdust = np.repeat(np.asarray(sizes)[:, None], nc, axis=1)
    cfg = C.make_cfg(nx=nx, ny=1, ndust=ndust, geom=C.GEOM_SPHERE, iso_cs=-1, cfl=cfl,
                     slope_type=slope_type, solver=solver, solverdust=solverdust, gamma=gamma,
                     smallr=1e-27 / u["unit_d"], smallc=1e-27 / u["unit_v"],
                     smallp=(1e-27 / u["unit_d"]) * ((1e-27 / u["unit_v"]) * (1e-27 / u["unit_v"])),
                     dust_ratio_min=dust_ratio_min, bc_kind=C.BC_COLLAPSE, tstop_kind=C.TSTOP_COLLAPSE,
                     force_kind=C.FORCE_NONE, inloop_kind=C.INLOOP_NONE, gravity=1, eos_baro=1,
                     l_soft=0.0, m_cent=0.0, t0_cloud=t0_cloud, mu_gas=mu_gas, unit_v=u["unit_v"],
                     unit_nh=unit_nh, rhograin=rhograin / u["unit_d"])
    return cfg, g, dict(u=primtoc(q, ndust, gamma), cs=np.zeros(nc), sdust=sdust, condinit=None,
                        fargo_velocity=None)


def units_cdisk():
    """The cartesian disk's units module (src/setups/disk/units.f90)."""
    unit_l = AU
    unit_d = MSUN / _powi(AU, 3)
    unit_t = 1.0 / math.sqrt(GRAV * unit_d)
    unit_v = unit_l / unit_t
    return dict(unit_l=unit_l, unit_d=unit_d, unit_dcol=MSUN / _powi(AU, 2), unit_t=unit_t, unit_m=MSUN,
                unit_v=unit_v, unit_p=unit_v * unit_v)


def cdisk(nx=256, ny=256, ndust=1, cfl=0.5, slope_type=1, solver=C.SOLVER_HLL, solverdust=0, box_l=400.0,
          sigma_r0=1700.0, r0_disk=1.0, mstar=1.0, mplanet=0.0, rplanet=20.0, hoverr=0.05, smooth_r=0.5,
          alpha_visc=1e-2, frac_r_disk=0.1, buffer_in=0.1e-2, buffer_out=0.25, viscosity=1, smin=1e-7,
          smax=1e-2, scutmin=5e-7, scut=1e-4, rhograin=2.3, dust2gas=0.01, gamma=1.66667):
    """Cartesian disk (src/setups/disk/setup.f90:1-70): a softened point mass (plus an optional
    planet on a circular orbit) at the centre of a square box, surface density
    sigma_R0 (R0/r) out to frac_R_disk box_l and a floor of 1/100 of its edge value beyond,
    Keplerian rotation with the softened radius r = sqrt(x^2 + y^2 + smooth_r^2), isothermal
    cs = HoverR sqrt(Mstar / r), dust at dust2gas with the gas velocity and the MRN bin sizes;
    zero-gradient boundaries, the inner and outer buffers reset to this state every step, alpha
    viscosity eta = alpha_visc H cs.  Inputs in the namelist's units (au, g/cm^2, solar
    masses), converted as the setup does; defaults: the setup module's, except HoverR taken as
    the double 0.05."""
    u = units_cdisk()
    sigma_r0 = sigma_r0 / u["unit_dcol"]
    r0_disk = r0_disk * AU / u["unit_l"]
    rplanet = rplanet * AU / u["unit_l"]
    smooth_r = smooth_r * AU / u["unit_l"]
    mstar = mstar * MSUN / u["unit_m"]
    mplanet = mplanet * MSUN / u["unit_m"]
    box_l = box_l * AU / u["unit_l"]
    box_l_y = box_l
    r_out = frac_r_disk * box_l
    cfg = C.make_cfg(nx=nx, ny=ny, ndust=ndust, iso_cs=1, cfl=cfl, slope_type=slope_type, solver=solver,
                     solverdust=solverdust, gamma=gamma, force_kick=1, viscosity=int(viscosity),
                     bc_kind=C.BC_ZERO_GRADIENT, tstop_kind=C.TSTOP_EPSTEIN_CDISK, force_kind=C.FORCE_CDISK,
                     inloop_kind=C.INLOOP_CDISK, box_l=box_l, box_l_y=box_l_y, sigma_r0=sigma_r0,
                     r0_disk=r0_disk, mstar=mstar, mplanet=mplanet, rplanet=rplanet, hoverr=hoverr,
                     smooth_r=smooth_r, alpha_visc=alpha_visc, frac_r_disk=frac_r_disk, buffer_in=buffer_in,
                     buffer_out=buffer_out, dust2gas=dust2gas, rhograin=rhograin / u["unit_d"],
                     smallr=1e-27 / u["unit_d"], smallc=1e-27 / u["unit_v"],
                     smallp=(1e-27 / u["unit_d"]) * ((1e-27 / u["unit_v"]) * (1e-27 / u["unit_v"])))
    g = grid_cartesian_2d(nx, ny, box_l, box_l_y)
    nc = g["vol"].size
    nv = 5 + 4 * ndust
    q = np.zeros((nv, nc))
    cs = np.zeros(nc)
    eta = np.zeros(nc)
    for c in range(nc):
        xx = g["position"][0, c] - 0.5 * box_l
        yy = g["position"][1, c] - 0.5 * box_l_y
        r_cyl = math.sqrt(xx * xx + yy * yy)
        rr = math.sqrt((xx * xx + yy * yy) + smooth_r * smooth_r)
        x_soft, y_soft = xx * rr / r_cyl, yy * rr / r_cyl
        H = hoverr * rr
        cs0 = hoverr * math.sqrt(mstar / rr)
        om = math.sqrt(mstar / _powi(rr, 3))
        rho = sigma_r0 * (r0_disk / rr)
        if rr > r_out:
            rho = sigma_r0 * (r0_disk / r_out) / 100.0
        q[0, c], q[1, c], q[2, c] = rho, -(om * y_soft), om * x_soft
        q[4, c] = rho * (cs0 * cs0)
        cs[c] = cs0
        for k in range(ndust):
            q[5 + k, c] = dust2gas * rho
            q[5 + ndust + k, c] = -(om * y_soft)
            q[5 + 2 * ndust + k, c] = om * x_soft
        eta[c] = (alpha_visc * H) * cs0
    sd = None
    if ndust:
        sizes, _ = mrn_sizes_and_ratios(ndust, u, smin=smin, smax=smax, scutmin=scutmin, scut=scut,
                                        rhograin=rhograin, dust2gas=dust2gas)
        sd = np.repeat(np.asarray(sizes)[:, None], nc, axis=1)
    return cfg, g, dict(u=primtoc(q, ndust, gamma), cs=cs, sdust=sd, condinit=None, fargo_velocity=None,
                        eta_visc=eta)


def sedov(n=256, cfl=0.5, slope_type=3, solver=C.SOLVER_HLLC, gamma=1.4, box_l=1.0, rho_ext=1.0,
          p_ext=1e-5, p_cen=1.0, r_expl=None):
    """Sedov blast (src/setups/sedov/setup.f90:1-34): adiabatic gas at rest in a square periodic
    box, pressure P_ext + P_cen within r_expl of the centre; only the active cells are set (the
    ghosts are filled by the first step's boundaries).  r_expl defaults to the setup module's
    default-real 0.1 (single precision)."""
    r = _f32(0.1) if r_expl is None else r_expl
    cfg = C.make_cfg(nx=n, ny=n, ndust=0, iso_cs=-1, cfl=cfl, slope_type=slope_type, solver=solver,
                     gamma=gamma, bc_kind=C.BC_PERIODIC, tstop_kind=C.TSTOP_DUSTYWAVE, force_kind=C.FORCE_NONE,
                     inloop_kind=C.INLOOP_NONE)
    g = grid_cartesian_2d(n, n, box_l, box_l)
    nm = n + 4
    q = np.zeros((5, nm * nm))
    ix, iy = _ix_iy(n, n)
    act = (ix >= 3) & (ix <= n + 2) & (iy >= 3) & (iy <= n + 2)
    xn = np.abs(g["position"][0] - 0.5 * box_l)
    yn = np.abs(g["position"][1] - 0.5 * box_l)
    q[0, act] = rho_ext
    q[4, act] = p_ext
    q[4, act & (xn * xn + yn * yn < r * r)] = p_ext + p_cen
    return cfg, g, dict(u=primtoc(q, 0, gamma), cs=np.zeros(nm * nm), sdust=None, condinit=None,
                        fargo_velocity=None)


def dustywave(nx=512, K=10.0, dust2gas=1.0, vx0=1e-4, drho=0.0, cs0=1.0, rho0=1.0, box_l=1.0,
              kx=1.0, cfl=0.5, slope_type=1, solver=C.SOLVER_HLL, solverdust=C.DUST_HUANG_BAI,
              gamma=1.6666667):
    """1D dusty wave (DUSTYWAVE setup.f90:12-78); defaults are the 512-cell, K=10 case."""
    cfg = C.make_cfg(nx=nx, ny=1, ndust=1, cfl=cfl, slope_type=slope_type, solver=solver,
                     solverdust=solverdust, gamma=gamma, bc_kind=C.BC_PERIODIC,
                     tstop_kind=C.TSTOP_DUSTYWAVE, force_kind=C.FORCE_NONE,
                     inloop_kind=C.INLOOP_NONE, K=[K])
    g = grid_cartesian_1d(nx, box_l)
    x = g["position"][0]
    s = np.sin((((2.0 * PI) * x) / box_l) * kx)
    nv = 9
    q = np.zeros((nv, nx + 4))
    q[0] = rho0 * (1.0 + drho * s)
    q[1] = vx0 * s
    cs = np.full(nx + 4, cs0)
    q[4] = cs ** 2 * q[0]
    q[5] = dust2gas * q[0]
    q[6] = vx0 * s
    u = primtoc(q, 1, gamma)
    return cfg, g, dict(u=u, cs=cs, sdust=None, condinit=None, fargo_velocity=None)


def si_box(n=1024, nd=8, st_min=1e-3, st_max=1.0, eps_tot=0.01, hoverr=0.1, noise=1e-3,
           box_l=1.0, seed=20261202, cfl=0.5, slope_type=1, gamma=1.666667):
    """Shearing box (SI setup.f90:20-163): NSH equilibrium of gas + nd species with Stokes
    numbers log-spaced in [st_min, st_max], dust-to-gas eps_tot split as an MRN a^-3.5 size
    distribution (bin mass ~ a^0.5), pressure-gradient drift eta = 0.05 H/R (0.05 cs), and
    Gaussian velocity noise of `noise` cs on every velocity component of every fluid."""
    rad0, om, qsh, rho_init = 1.0, 1.0, 1.5, 1.0
    eta = _f32(0.05) * hoverr                    # setup.f90:24 (0.05 is a REAL(4) literal)
    L = box_l * eta * rad0
    vk = rad0 * om
    cs0 = hoverr * vk
    theta_dust = 1e-11
    rhograin = rho_init / theta_dust
    st = np.logspace(math.log10(st_min), math.log10(st_max), nd) if nd > 1 else np.array([st_min])
    w = st ** 0.5
    eps = eps_tot * w / w.sum()
    A = 0.0
    B = 1.0
    for k in range(nd):
        A = A + eps[k] * st[k] / (1.0 + st[k] ** 2)
        B = B + eps[k] / (1.0 + st[k] ** 2)
    vx_n = 2.0 * eta * vk * A / (A + B)
    vy_n = -eta * vk * B / (A + B)
    ux = (vx_n + 2.0 * st * vy_n) / (1.0 + st ** 2)
    uy = (vy_n - 0.5 * st * vx_n) / (1.0 + st ** 2)
    cfg = C.make_cfg(nx=n, ny=n, ndust=nd, cfl=cfl, slope_type=slope_type, gamma=gamma,
                     force_kick=1, bc_kind=C.BC_SHEAR_SI, tstop_kind=C.TSTOP_SI,
                     force_kind=C.FORCE_SI, inloop_kind=C.INLOOP_SI, q_shear=qsh,
                     omega_shear=om, box_l=L, rad0=rad0, eta_stream=eta, rho_init=rho_init,
                     hoverr=hoverr, rhograin=rhograin)
    g = grid_cartesian_2d(n, n, L, L)
    nc = g["vol"].size
    xx = g["position"][0] - 0.5 * L
    rng = np.random.Generator(np.random.Philox(seed))
    nv = 5 + 4 * nd
    q = np.zeros((nv, nc))
    q[0] = rho_init
    q[1] = vx_n + noise * cs0 * rng.standard_normal(nc)
    q[2] = noise * cs0 * rng.standard_normal(nc)
    q[3] = qsh * om * xx + vy_n + noise * cs0 * rng.standard_normal(nc)
    q[4] = q[0] * cs0 ** 2
    sd = np.zeros((nd, nc))
    for k in range(nd):
        q[5 + k] = eps[k] * rho_init
        q[5 + nd + k] = ux[k] + noise * cs0 * rng.standard_normal(nc)
        q[5 + 2 * nd + k] = noise * cs0 * rng.standard_normal(nc)
        q[5 + 3 * nd + k] = qsh * om * xx + uy[k] + noise * cs0 * rng.standard_normal(nc)
        sd[k] = st[k] * rho_init * cs0 / rhograin / om
    u = primtoc(q, nd, gamma)
    return cfg, g, dict(u=u, cs=np.full(nc, cs0), sdust=sd, condinit=None, fargo_velocity=None)


def disk(nr=1024, nphi=2048, nd=16, st_min=1e-4, st_max=1.0, eps_tot=0.01, hoverr=0.05,
         r_in=0.4, r_out=2.5, sigma_r0=1700.0, noise=1e-3, fargo=True, frame="reference",
         seed=20261202, cfl=0.5, slope_type=1, gamma=1.666667, r_relax=0.3, n_rel=8.0,
         sizes=None, weights=None):
    """Polar disk (disk_polar setup.f90:12-84, linear radial grid): Sigma ~ r^-1 (R_out_disk
    -> infinity), H/R = hoverr, Keplerian rotation about Mstar = 1, nd species whose Epstein
    Stokes number at r = 1 is log-spaced in [st_min, st_max], dust-to-gas eps_tot split equally,
    Gaussian density noise of relative amplitude `noise` on the gas (dust follows).

    frame="reference": velocities as the reference setup stores them with fargo on — gas with
    the full azimuthal velocity, dust with Omega*r + vfarg, FARGO shift velocity
    vfarg = -Omega*r (setup.f90:35-37,45,64).
    frame="residual": every fluid stores its velocity relative to Omega*r and the FARGO shift is
    +Omega*r.
    r_relax <= r_in keeps the in-loop relaxation of setup.f90:228 off (its default 3.0 would
    relax every cell toward the initial state).
    sizes / weights: explicit grain sizes (code units) and dust-mass fractions per species (the
    growth bins' sizes with an MRN split, for runs with dust growth) instead of the Stokes grid
    and the equal split."""
    unit_l = AU
    unit_dcol = MSUN / AU ** 2
    unit_d = MSUN / AU ** 3
    mstar = 1.0
    box_l = r_out - r_in
    g = grid_polar(nr, nphi, box_l, r_in)
    rr = g["radii_c"]
    nc = rr.size
    H = hoverr * rr
    om = np.sqrt(mstar / rr ** 3)
    cs0 = om * H
    sig0 = sigma_r0 / unit_dcol
    rng = np.random.Generator(np.random.Philox(seed))
    rho = sig0 * (1.0 / rr) * (1.0 + noise * rng.standard_normal(nc))
    vkep = om * rr
    vgas = om * rr * np.sqrt(1.0 - hoverr ** 2 * 2.0)
    rhograin = 2.7 / unit_d
    # grain size giving Stokes number st at r=1 (Epstein law, setup.f90:60 / :436)
    st = np.logspace(math.log10(st_min), math.log10(st_max), nd) if nd > 1 else np.array([st_min])
    rho1 = sig0 / (math.sqrt(2.0 * PI) * hoverr)
    if sizes is None:
        sizes = st / (math.sqrt(PI * gamma / 8.0) * rhograin / (rho1 * hoverr))
    eta_drift = np.sqrt(1.0 - hoverr ** 2 * 2.0) - 1.0
    nv = 5 + 4 * nd
    q = np.zeros((nv, nc))
    q[0] = rho
    if frame == "reference":
        vfarg = -vkep
        q[2] = vgas
    elif frame == "residual":
        vfarg = vkep
        q[2] = vgas - vkep
    else:
        raise ValueError(frame)
    q[4] = rho * cs0 ** 2
    sd = np.zeros((nd, nc))
    for k in range(nd):
        ts = math.sqrt(PI * gamma / 8.0) * rhograin * sizes[k] / (rho / (math.sqrt(2.0 * PI) * H) * cs0)
        stk = ts * om
        q[5 + k] = eps_tot / nd * rho if weights is None else (eps_tot * weights[k]) * rho
        q[5 + nd + k] = eta_drift * om * rr / (stk + 1.0 / stk)
        q[5 + 2 * nd + k] = (vkep + vfarg) if frame == "reference" else 0.0
        sd[k] = sizes[k]
    u = primtoc(q, nd, gamma)
    cfg = C.make_cfg(nx=nr, ny=nphi, ndust=nd, geom=C.GEOM_POLAR, cfl=cfl, slope_type=slope_type,
                     gamma=gamma, force_kick=1, fargo=1 if fargo else 0, bc_kind=C.BC_DISK,
                     tstop_kind=C.TSTOP_EPSTEIN_DISK, force_kind=C.FORCE_DISK,
                     inloop_kind=C.INLOOP_DISK, mstar=mstar, r_relax=r_relax, n_rel=n_rel,
                     smooth_r=r_in, hoverr=hoverr, rhograin=rhograin, disk_bc_mode=0)
    fv = vfarg if fargo else None
    return cfg, g, dict(u=u, cs=cs0.copy(), sdust=sd, condinit=u.copy(), fargo_velocity=fv)


def disk_growth(nr=1024, nphi=2048, nbins=64, **kw):
    """The polar disk with nbins dust species that are the growth bins (1e-4..10 cm, MRN mass
    split), for steps with dust growth; returns (cfg, geom, state, growth params)."""
    from . import growth as G
    p = G.disk_params(nbins=nbins, max_substeps=1000000)
    zeta = (p.smax / p.smin) ** (1.0 / nbins)
    k = np.arange(nbins)
    sz = np.sqrt((p.smin * zeta ** k / p.unit_l) * (p.smin * zeta ** (k + 1) / p.unit_l))
    w = sz ** 0.5
    cfg, geom, state = disk(nr=nr, nphi=nphi, nd=nbins, sizes=sz, weights=w / w.sum(), **kw)
    return cfg, geom, state, p


WORKLOADS = {
    "dustywave_512_n1": lambda: dustywave(),
    "si_1024x1024_n8": lambda: si_box(),
    "disk_1024x2048_n16_fargo": lambda: disk(),
}
