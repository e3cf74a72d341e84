// step_kernels.hip — the per-step kernels of SHARK's Godunov step on gfx950.
//
// One step (src/solver/solve.f90:10-181) runs as four launches on one stream:
//   k_bc       ghost fill of the state (apply_boundaries); k_bc_zg for the zero-gradient
//              boundaries, whose source cells are themselves ghosts
//   k_hydro    per fluid: ctoprim + force + kick + predictor + Riemann fluxes + update +
//              sources (polar, drag heating, viscosity), fused; a wavefront owns 60 columns x
//              a strip of rows and marches the strip row by row with a 3-row register window;
//              x neighbours are exchanged with cross-lane shuffles.  qm/qp/flux/dq never reach
//              HBM.
//   k_post2    per row segment, all fluids staged in LDS: FARGO remap, implicit drag + floor,
//              in-loop hook, and the next step's Courant candidate (dt fused into the
//              epilogue of the previous step; the min over cells is an integer atomic min
//              on the ordered bit pattern, exact and order independent)
//   k_finalize time += dt; dt <- min candidate
// All arithmetic keeps the reference's evaluation order; see physics.h.
#include <cfloat>

#include "physics.h"
#include "step_kernels.h"

namespace shark {

#define DEV __device__ __forceinline__

DEV int cell_of(const StepArgs& a, int ix, int iy) {
  return (a.ndim == 1) ? (ix - 1) : (ix - 1) + (iy - 1) * a.nxm;
}
DEV bool is_active(const StepArgs& a, int ix, int iy) {
  return ix >= 3 && ix <= a.nx + 2 && (a.ndim == 1 || (iy >= 3 && iy <= a.ny + 2));
}

// External accelerations of update_force_setup for an active cell (SI: setup.f90:346-353,
// disk_polar: setup.f90:366-377).  qa/qb: gas (vx, vz) or dust (vdx, vdz) velocities.
// next: 0 for the step being taken, 1 for the next step (its Courant candidate), which selects
// the planet's position in time.  CD: the cartesian-disk force is compiled in (the hydro
// kernels of other grids leave it out, so it costs them no registers)
template <bool CD = true>
DEV void cell_force_setup(const StepArgs& a, int ix, int iy, int next, double qx, double qz, bool gas,
                          double& f1, double& f2, double& f3) {
  f1 = 0.0; f2 = 0.0; f3 = 0.0;
  if (CD && a.force_kind == 3) {
    // cartesian disk (disk/setup.f90:274-301): softened star, plus a planet on a circular orbit
    // (its terms per step from k_planet); rr**3 with the integer literal 3 is rr*(rr*rr)
    const double xx = a.g.pos1[ix - 1] - kHalf * a.box_l, yy = a.g.pos2[iy - 1] - kHalf * a.box_l_y;
    const double rr = sqrt((xx * xx + yy * yy) + a.smooth_r * a.smooth_r);
    const double r3 = rr * (rr * rr);
    f1 = -(a.mstar * xx / r3);
    f2 = -(a.mstar * yy / r3);
    if (a.mplanet > 0.0) {
      f1 = f1 - a.sc->planet[2 * next];
      f2 = f2 - a.sc->planet[2 * next + 1];
    }
    return;
  }
  if (a.force_kind == 1) {
    const double Om = a.omega_shear;
    double t = ((2.0 * a.q_shear) * (Om * Om)) * (a.g.pos1[ix - 1] - kHalf * a.box_l);
    if (gas) f1 = (t - (2.0 * Om) * qz) - ((2.0 * a.rad0) * Om) * a.eta_stream;
    else f1 = t - (2.0 * Om) * qz;
    f3 = (2.0 * Om) * qx;
  } else if (a.force_kind == 2) {
    const double r = a.g.rc[ix - 1];
    f1 = -(a.mstar / (r * r));
    if (a.pgrad) {
      // planet (disk_polar/setup.f90:323-378): the gradient of its potential, per step from
      // k_planet_grad
      const size_t cell = (size_t)(ix - 1) + (size_t)(iy - 1) * a.nxm;
      f1 = f1 - a.pgrad[(size_t)(2 * next) * a.nc + cell];
      f2 = f2 - a.pgrad[(size_t)(2 * next + 1) * a.nc + cell];
    }
  }
}

// The force of fluid k (-1: gas) in an active cell: with explicit drag, the step's forces from
// k_explicit_force (setup forces plus the drag terms, all fluids of the cell coupled); else the
// setup's forces of this cell and fluid alone.  EXPL compiles the explicit-drag lookup in: the
// shearing box's (2D cartesian) hydro kernels and the standalone Courant kernel (the post pass's candidate is
// replaced by the latter whenever explicit drag is on), so the other kernels keep their registers.
template <bool CD = true, bool EXPL = true>
DEV void cell_force(const StepArgs& a, int ix, int iy, int next, double qx, double qz, bool gas, int k,
                    double& f1, double& f2, double& f3) {
  if (EXPL && a.fexp) {
    const size_t cell = (size_t)(ix - 1) + (size_t)(iy - 1) * a.nxm, fl = (size_t)(k + 1);
    f1 = a.fexp[(3 * fl) * a.nc + cell];
    f2 = a.fexp[(3 * fl + 1) * a.nc + cell];
    f3 = a.fexp[(3 * fl + 2) * a.nc + cell];
    return;
  }
  cell_force_setup<CD>(a, ix, iy, next, qx, qz, gas, f1, f2, f3);
}

// ------------------------------------------------------------------ ghost fill
// apply_boundaries: periodic (src/grid/boundary_types.f90:149-186), SI shearing box
// (src/setups/SI/boundaries.f90:10-85), disk (src/setups/disk_polar/boundaries.f90:10-62).
// The reference fills y-ghost rows of active columns first, then x-ghost columns of every
// row; each thread here resolves its ghost cell through those (at most two) copies.
__global__ void k_bc(StepArgs a) {
  if (!a.sc->active) return;
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  const int n1 = (a.ndim == 2) ? 4 * a.nx : 0;
  const int n2 = 4 * a.nym;
  if (t >= n1 + n2) return;
  const size_t nc = a.nc;
  auto ysrc = [&](int iy) { return iy <= 2 ? a.ny + iy : iy - a.ny; };
  auto is_grow = [&](int iy) { return a.ndim == 2 && (iy <= 2 || iy >= a.ny + 3); };
  // Sod tube (SOD/boundaries.f90:10-80): the ghost layers across the shock hold the initial
  // left / right state at rest with E = P/(gamma-1)
  auto sod_state = [&](int dst, bool left) {
    a.U[0 * nc + dst] = left ? a.rho_l : a.rho_r;
    a.U[4 * nc + dst] = (left ? a.p_l : a.p_r) / (a.gamma - 1.0);
    a.U[1 * nc + dst] = 0.0;
    a.U[2 * nc + dst] = 0.0;
    a.U[3 * nc + dst] = 0.0;
  };
  const bool sod = a.bc_kind == 3;
  if (t < n1) {
    const int r = t / a.nx, ix = 3 + t % a.nx;
    const int iy = (r == 0) ? 1 : (r == 1) ? 2 : (r == 2) ? a.nym : a.nym - 1;
    const int dst = cell_of(a, ix, iy), src = cell_of(a, ix, ysrc(iy));
    if (sod && a.shock_dir == 1) {
      sod_state(dst, iy <= 2);
      return;
    }
    for (int v = 0; v < a.nvar; ++v) a.U[v * nc + dst] = a.U[v * nc + src];
    return;
  }
  const int t2 = t - n1;
  const int c = t2 / a.nym, iy = 1 + t2 % a.nym;
  const int ix = (c == 0) ? 1 : (c == 1) ? 2 : (c == 2) ? a.nxm : a.nxm - 1;
  const bool left = ix <= 2;
  int sx, sy = iy;
  if (a.bc_kind == 2 && a.disk_bc_mode == 0) {
    // boundaries.f90:45-46: first_active/last_active used as flat cell indices
    const int flat = left ? 3 : a.nx + 2;
    if (a.ndim == 2) { sx = (flat - 1) % a.nxm + 1; sy = (flat - 1) / a.nxm + 1; }
    else { sx = flat; sy = 1; }
  } else if (a.bc_kind == 2) {
    sx = left ? 3 : a.nx + 2;
  } else {
    sx = left ? a.nx + ix : ix - a.nx;
  }
  if (sod && a.shock_dir == 0) {
    sod_state(cell_of(a, ix, iy), left);
    return;
  }
  if (sod && is_grow(sy)) {   // the periodic source is a y-ghost holding the fixed state
    sod_state(cell_of(a, ix, iy), sy <= 2);
    return;
  }
  if (is_grow(sy)) sy = ysrc(sy);  // the source itself is a y-ghost filled first
  const int dst = cell_of(a, ix, iy), src = cell_of(a, sx, sy);
  for (int v = 0; v < a.nvar; ++v) a.U[v * nc + dst] = a.U[v * nc + src];
  if (a.bc_kind == 1) {
    const double s = left ? -1.0 : 1.0;
    // boundaries.f90:66-77
    const double sh = ((a.U[0 * nc + src] * a.q_shear) * a.omega_shear) * a.box_l;
    a.U[3 * nc + dst] = left ? a.U[3 * nc + src] - sh : a.U[3 * nc + src] + sh;
    for (int k = 0; k < a.nd; ++k) {
      const int ir = 5 + k, iz = 5 + 3 * a.nd + k;
      const double shd = ((a.U[ir * nc + src] * a.q_shear) * a.omega_shear) * a.box_l;
      a.U[iz * nc + dst] = left ? a.U[iz * nc + src] - shd : a.U[iz * nc + src] + shd;
    }
    (void)s;
  } else if (a.bc_kind == 2) {
    // boundaries.f90:51-57 inflow clamps
    double& vx = a.U[1 * nc + dst];
    vx = left ? fmaxf_(vx, 0.0) : fminf_(vx, 0.0);
    for (int k = 0; k < a.nd; ++k) {
      double& vd = a.U[(5 + a.nd + k) * nc + dst];
      vd = left ? fmaxf_(vd, 0.0) : fminf_(vd, 0.0);
    }
  }
}

// ------------------------------------------------------------------ Collapse_1D
constexpr double kKB = 1.38e-16, kMH = 1.67e-24;   // phys_const.f90:8-9

// barotrop (src/setups/Collapse_1D/eos.f90:1-11; its first exponent is the default-real literal
// 0.8, single precision widened) and cs_eos (:13-20): the barotropic sound speed of density rho
DEV double cs_baro(const StepArgs& a, double rho) {
  const double e08 = (double)0.8f;
  const double x = rho * a.unit_nh;
  const double T = ((a.t0_cloud * sqrt(1.0 + pow(x / 1e11, e08))) * pow(1.0 + x / 1e16, -0.3)) *
                   pow(1.0 + x / 1e21, 0.56667);
  return sqrt(((a.gamma * kKB) * T) / (a.mu_gas * kMH)) / a.unit_v;
}

// boundary_collapse_1D (src/grid/boundary_types.f90:4-44), 1D: ghosts copy the first / last
// active cell; nothing flows through the centre (the inner ghosts and the first active cell at
// rest) or out of the outer edge (the outer ghosts and the last active cell keep only inward
// velocities).  One thread per variable plane.
__global__ void k_bc_collapse(StepArgs a) {
  if (!a.sc->active) return;
  const size_t nc = a.nc;
  const int fa = 2, la = a.nx + 1;   // 0-based first / last active cell
  for (int v = threadIdx.x; v < a.nvar; v += blockDim.x) {
    const bool mom = v == 1 || (v >= 5 + a.nd && v < 5 + 2 * a.nd);   // vx of the gas and dust
    double* U = a.U + (size_t)v * nc;
    const double l = U[fa], r = U[la];
    const double lv = mom ? 0.0 : l, rv = mom ? fminf_(r, 0.0) : r;
    U[0] = lv;
    U[1] = lv;
    U[nc - 2] = rv;
    U[nc - 1] = rv;
    if (mom) {
      U[fa] = 0.0;
      U[la] = rv;
    }
  }
}

// mtot (src/gravity/gravity_utils.f90:11-51, NY = 1): the mass within each cell centre,
// Mc = M_cent + 4 pi sum_{j<=i} rho_j vol_j - 4 pi rho_i dvol_i over gas and dust, from the
// start-of-step densities.  The running sum is one chain in the reference's order (one thread
// over each block-wide tile of per-cell terms); the ghost cells hold 0, the inner one next to the
// grid M_cent.
__global__ __launch_bounds__(1024) void k_mtot(StepArgs a) {
  __shared__ double mint[1024], dmint[1024];
  __shared__ double carry;
  const double fourpi = 4.0 * kPi;
  const size_t nc = a.nc;
  for (int c = threadIdx.x; c < a.nc; c += blockDim.x) a.mc[c] = 0.0;
  if (threadIdx.x == 0) carry = a.m_cent;
  __syncthreads();
  if (threadIdx.x == 0) a.mc[1] = a.m_cent;
  for (int base = 2; base <= a.nx + 1; base += 1024) {
    const int c = base + (int)threadIdx.x;
    if (c <= a.nx + 1) {
      const double rho = fmaxf_(a.U[c], a.smallr), vol = a.g.vol[c], dv = a.dvol[c];
      double mi = (rho * vol) * fourpi, dm = (rho * dv) * fourpi;
      for (int k = 0; k < a.nd; ++k) {
        const double rd = a.U[(5 + k) * nc + c];
        mi = mi + (rd * vol) * fourpi;
        dm = dm + (rd * dv) * fourpi;
      }
      mint[threadIdx.x] = mi;
      dmint[threadIdx.x] = dm;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
      const int n = (a.nx + 2 - base) < 1024 ? (a.nx + 2 - base) : 1024;
      double m = carry;
      for (int i = 0; i < n; ++i) {
        m = m + mint[i];
        mint[i] = m;
      }
      carry = m;
    }
    __syncthreads();
    if (c <= a.nx + 1) a.mc[c] = mint[threadIdx.x] - dmint[threadIdx.x];
    __syncthreads();
  }
}

// ------------------------------------------------------------------ hydro
// Per-lane state of one loaded row of one fluid.
struct RowF {
  double u[5];   // conservative (gas: rho, mx, my, mz, E; dust: rhod, mdx, mdy, mdz)
  double q[5];   // primitive rho, vx, vy, vz, and P for adiabatic gas (ctoprim, solve.f90:211-232)
  double cs;
};
struct FaceSet {
  double m[5], p[5];  // qm (state at the right face) and qp (left face) of rho, vx, vy, vz [, P]
};

DEV double shfl_up1(double x) { return __shfl_up(x, 1); }
DEV double shfl_dn1(double x) { return __shfl_down(x, 1); }

template <int ND, bool CD = true>
DEV double tstop_of(const StepArgs& a, int k, int cell, int ix, double rho_q, double rhod_q, double cs_q);

// GEOM: 0 cartesian, 2 polar, 3 cartesian with the cartesian disk's point-mass force and
// viscosity compiled in, 4 polar with viscosity compiled in.  SLOPE: limiter (1..3); SOLV: compile-time Riemann solver (gas 0..2,
// dust 0..3) or -1 to read it from the arguments.  ADI: adiabatic gas (P from the total energy, its own slopes,
// Hancock source and face states, sound speeds from the face states; iso_cs < 1).
template <int NDIM, int GEOM, bool GAS, int SLOPE, int SOLV, bool ADI>
__device__ void hydro_fluid(const StepArgs& a, int fluid, int g, int strip, int lane) {
  constexpr bool POLAR = GEOM == 2 || GEOM == 4;
  constexpr int NU = GAS ? 5 : 4;
  constexpr int NQ = ADI ? 5 : 4;   // primitives carried through slopes and faces
  const size_t nc = a.nc;
  int p[5];
  if (GAS) { p[0] = 0; p[1] = 1; p[2] = 2; p[3] = 3; p[4] = 4; }
  else {
    const int k = fluid - 1;
    p[0] = 5 + k; p[1] = 5 + a.nd + k; p[2] = 5 + 2 * a.nd + k; p[3] = 5 + 3 * a.nd + k; p[4] = p[3];
  }
  // sound speed: the gas, and the HLLC dust solver; known at compile time when SOLV is
  const bool need_cs = GAS || (SOLV < 0 ? a.solverdust == 3 : SOLV == 3);
  const int ix = 1 + 60 * g + lane;
  const int ixc = ix < a.nxm ? ix : a.nxm;
  const int ixl = ixc > 1 ? ixc - 1 : 1;
  const int ixr = ixc < a.nxm ? ixc + 1 : a.nxm;
  const bool own_col = lane >= 2 && lane <= 61 && ix <= a.nx + 2;
  const double dt = a.sc->dt;
  constexpr int st = SLOPE;
  const double dx1 = a.g.dx1[ixc - 1], dx1m = a.g.dx1[ixl - 1];
  const double rp = (POLAR) ? a.g.rc[ixc - 1] : 1.0;
  const double dx2 = (NDIM == 2) ? a.g.dx2[ixc - 1] : 0.0;
  const double surf1 = a.g.surf1[ixc - 1], surf1r = a.g.surf1[ixr - 1];
  const double surf2 = (NDIM == 2) ? a.g.surf2[ixc - 1] : 0.0;
  const double vol = a.g.vol[ixc - 1];
  const bool col_active = ix >= 3 && ix <= a.nx + 2;
  // correctly rounded reciprocals of the per-column divisors (exdiv)
  const double dsx = dx1 + dx1m, ysx = 1.0 / dsx;
  const double dsy = dx2 + dx2, ysy = (NDIM == 2) ? 1.0 / dsy : 0.0;
  const double yrp = 1.0 / rp, yvol = 1.0 / vol;
  // spherical 1D grid (GEOM 1): the cell-centre radius, and the gravity of the enclosed mass
  // Mc / (r^2 + l_soft^2), 0 where both vanish (the inner ghost cells: radii_c = 0, Mc = M_cent)
  constexpr bool SPH = GEOM == 1;
  const double rsph = SPH ? a.g.rc[ixc - 1] : 0.0;
  double gacc = 0.0;
  if (SPH && a.gravity) {
    const double d = rsph * rsph + a.l_soft * a.l_soft;
    gacc = d != 0.0 ? a.mc[ixc - 1] / d : 0.0;
  }

  int y0 = 1, y1 = 1;
  if (NDIM == 2) {
    const int sh = GAS ? a.strip_g : a.strip_h;
    y0 = 3 + strip * sh;
    y1 = y0 + sh - 1;
    if (y1 > a.ny + 2) y1 = a.ny + 2;
  }

  // raw loads of a row are issued one iteration ahead of their use (prefetch)
  auto fetch_row = [&](int j, RowF& r) {
    const int cell = (NDIM == 1) ? ixc - 1 : (ixc - 1) + (j - 1) * a.nxm;
#pragma unroll
    for (int c = 0; c < NU; ++c) r.u[c] = a.U[p[c] * nc + cell];
    r.cs = (need_cs && !ADI) ? a.cs[cell] : 0.0;
  };
  auto convert_row = [&](RowF& r) {
    const double yr = 1.0 / r.u[0];
    r.q[0] = GAS ? fmaxf_(r.u[0], a.smallr) : r.u[0];
    r.q[1] = exdiv(r.u[1], r.u[0], yr);
    r.q[2] = exdiv(r.u[2], r.u[0], yr);
    r.q[3] = exdiv(r.u[3], r.u[0], yr);
    if constexpr (ADI) {
      // solve.f90:214-220: P = max((gamma-1)(E - ekin), smallp), cs = sqrt(gamma P / rho)
      const double ekin = kHalf * r.u[0] * (r.q[1] * r.q[1]) + kHalf * r.u[0] * (r.q[2] * r.q[2]) +
                          kHalf * r.u[0] * (r.q[3] * r.q[3]);
      r.q[4] = fmaxf_((a.gamma - 1.0) * (r.u[4] - ekin - 0.0), a.smallp);
      r.cs = sqrt(a.gamma * r.q[4] / r.q[0]);
      if (a.eos_baro) {   // non_standard_eos = 1 (solve.f90:225-226)
        r.cs = cs_baro(a, r.q[0]);
        r.q[4] = r.u[0] * (r.cs * r.cs);
      }
    }
  };

  // face states of a row (predictor, godunov.f90:89-611); dqy = 0 in 1D
  auto faces = [&](const RowF& r, const double dqx[5], const double dqy[5], int iy, FaceSet& fx,
                   FaceSet& fy) {
    double f1 = 0.0, f2 = 0.0, f3 = 0.0;
    if (a.force_kick && col_active && (NDIM == 1 || (iy >= 3 && iy <= a.ny + 2)))
      cell_force<GEOM == 3, GEOM == 0 && NDIM == 2>(a, ix, iy, 0, r.q[1], r.q[3], GAS, fluid - 1, f1, f2, f3);
    const double r_rho = r.q[0];
    double u = r.q[1], v = r.q[2], w = r.q[3];
    const double drx = dqx[0], dux = dqx[1], dvx = dqx[2], dwx = dqx[3];
    const double dry = dqy[0], duy = dqy[1], dvy = dqy[2], dwy = dqy[3];
    double sr0, su0, sv0, sw0, sp0 = 0.0, P = 0.0, dpx = 0.0, dpy = 0.0;
    if (a.force_kick) {
      u = u + (f1 * kHalf) * dt;
      v = v + (f2 * kHalf) * dt;
      w = w + (f3 * kHalf) * dt;
    }
    if (GAS) {
      if constexpr (ADI) {   // godunov.f90:123-124, 131-133
        P = r.q[4];
        dpx = dqx[4];
        dpy = (NDIM == 2) ? dqy[4] : 0.0;
        if (POLAR) sp0 = ((-(u * dpx)) - v * dpy - (dux + dvy) * a.gamma * P) - exdiv(a.gamma * P * u, rp, yrp);
        else sp0 = (-(u * dpx)) - v * dpy - (dux + dvy) * a.gamma * P;
      } else {               // :125-128
        const double cs2 = r.cs * r.cs;
        dpx = dqx[0] * cs2;
        dpy = (NDIM == 2) ? dqy[0] * cs2 : 0.0;
      }
      if (POLAR) {
        sr0 = ((-(u * drx)) - v * dry - (dux + dvy) * r_rho) - exdiv(r_rho * u, rp, yrp);
        su0 = ((-(u * dux)) - v * duy - dpx / r_rho) + exdiv(v * v, rp, yrp);
        sv0 = ((-(u * dvx)) - v * dvy - dpy / r_rho) - exdiv(u * v, rp, yrp);
        sw0 = (-(u * dwx)) - v * dwy;
      } else {
        sr0 = (-(u * drx)) - v * dry - (dux + dvy) * r_rho;
        if (SPH && rsph != 0.0) sr0 = sr0 - (r_rho * u) / rsph;   // godunov.f90:205-207
        su0 = (-(u * dux)) - v * duy - dpx / r_rho;
        sv0 = (-(u * dvx)) - v * dvy - dpy / r_rho;
        sw0 = (-(u * dwx)) - v * dwy;
      }
      if (SPH && a.gravity) su0 = su0 + (-gacc);   // :243-246
    } else {
      sr0 = (-(u * drx)) - v * dry - (dux + dvy) * r_rho;
      su0 = (-(u * dux)) - v * duy;
      sv0 = (-(u * dvx)) - v * dvy;
      sw0 = (-(u * dwx)) - v * dwy;
      if (POLAR) {
        sr0 = sr0 - exdiv(r_rho * u, rp, yrp);
        su0 = su0 + exdiv(v * v, rp, yrp);
        sv0 = sv0 - exdiv(u * v, rp, yrp);
      }
      if (SPH && rsph != 0.0) sr0 = sr0 - (r_rho * u) / rsph;   // :386-389
      if (SPH && a.gravity) su0 = su0 + (-gacc);               // :442-445
    }
    const double hdt = kHalf * dt;
    const double b0 = r_rho + hdt * sr0, b1 = u + hdt * su0, b2 = v + hdt * sv0, b3 = w + hdt * sw0;
    fx.m[0] = b0 + kHalf * drx * dx1; fx.p[0] = b0 - kHalf * drx * dx1;
    fx.m[1] = b1 + kHalf * dux * dx1; fx.p[1] = b1 - kHalf * dux * dx1;
    fx.m[2] = b2 + kHalf * dvx * dx1; fx.p[2] = b2 - kHalf * dvx * dx1;
    fx.m[3] = b3 + kHalf * dwx * dx1; fx.p[3] = b3 - kHalf * dwx * dx1;
    if constexpr (ADI) {   // :261, :266
      const double b4 = P + hdt * sp0;
      fx.m[4] = fmaxf_(b4 + kHalf * dpx * dx1, a.smallp);
      fx.p[4] = fmaxf_(b4 - kHalf * dpx * dx1, a.smallp);
      if (a.eos_baro) {   // :612-617
        const double cm = cs_baro(a, fx.m[0]), cp = cs_baro(a, fx.p[0]);
        fx.m[4] = (cm * cm) * fx.m[0];
        fx.p[4] = (cp * cp) * fx.p[0];
      }
    }
    if (NDIM == 2) {
      const double dyl = rp * dx2;
      fy.m[0] = b0 + kHalf * dry * dyl; fy.p[0] = b0 - kHalf * dry * dyl;
      fy.m[1] = b1 + kHalf * duy * dyl; fy.p[1] = b1 - kHalf * duy * dyl;
      fy.m[2] = b2 + kHalf * dvy * dyl; fy.p[2] = b2 - kHalf * dvy * dyl;
      fy.m[3] = b3 + kHalf * dwy * dyl; fy.p[3] = b3 - kHalf * dwy * dyl;
      if constexpr (ADI) {   // :275, :280
        const double b4 = P + hdt * sp0;
        fy.m[4] = fmaxf_(b4 + kHalf * dpy * dyl, a.smallp);
        fy.p[4] = fmaxf_(b4 - kHalf * dpy * dyl, a.smallp);
        if (a.eos_baro) {
          const double cm = cs_baro(a, fy.m[0]), cp = cs_baro(a, fy.p[0]);
          fy.m[4] = (cm * cm) * fy.m[0];
          fy.p[4] = (cp * cp) * fy.p[0];
        }
      }
    }
  };

  // Riemann flux between left state L(q) and right state R(q); dir 0: x (normal vx),
  // dir 1: y (normal vy). Returns conservative components in plane order.
  auto riemann = [&](const double L[5], double csl, const double R[5], double csr, int dir,
                     double F[5]) {
    if constexpr (ADI) {   // sound speeds of the face states, godunov.f90:675-680, 702-710
      if (a.eos_baro) {
        csl = cs_baro(a, L[0]);
        csr = cs_baro(a, R[0]);
      } else {
        csl = sqrt(a.gamma * L[4] / L[0]);
        csr = sqrt(a.gamma * R[4] / R[0]);
      }
    }
    Face fl, fr;
    fl.r = L[0]; fr.r = R[0];
    fl.un = dir == 0 ? L[1] : L[2]; fl.ut = dir == 0 ? L[2] : L[1]; fl.w = L[3];
    fr.un = dir == 0 ? R[1] : R[2]; fr.ut = dir == 0 ? R[2] : R[1]; fr.w = R[3];
    Flux5 f;
    if (GAS && ADI) {
      fl.p = L[4];
      fr.p = R[4];
      f = gas_flux(SOLV >= 0 ? SOLV : a.solver, a.gamma, fl, fr, csl, csr);
    } else if (GAS) {
      fl.p = (csl * csl) * L[0];   // isothermal face pressure, godunov.f90:606-611
      fr.p = (csr * csr) * R[0];
      f = gas_flux(SOLV >= 0 ? SOLV : a.solver, a.gamma, fl, fr, csl, csr);
    } else {
      fl.p = fr.p = 0.0;
      f = dust_flux(SOLV >= 0 ? SOLV : a.solverdust, fl, fr, csl, csr);
    }
    F[0] = f.r;
    F[1] = dir == 0 ? f.un : f.ut;
    F[2] = dir == 0 ? f.ut : f.un;
    F[3] = f.w;
    F[4] = f.e;
  };

  // ---- march
  RowF R0, R1, R2;
  FaceSet fx1, fy1;
  double sLy_prev[5] = {0, 0, 0, 0, 0};
  double fym_prev[5] = {0, 0, 0, 0, 0};
  double cs_prev = 0.0;
  double Fx_own_prev[5] = {0, 0, 0, 0, 0}, Fx_rt_prev[5] = {0, 0, 0, 0, 0};
  double Fy_bot[5] = {0, 0, 0, 0, 0};
  // viscosity (cartesian disk): velocities of the row below the one being finalized
  constexpr bool VISC = GAS && (GEOM == 3 || GEOM == 4) && NDIM == 2 && !ADI;
  const bool visc = VISC && a.viscosity;
  double vbx = 0.0, vby = 0.0;

  const int jstart = (NDIM == 2) ? y0 - 2 : 1;
  const int jend = (NDIM == 2) ? y1 + 2 : 1;
  // PF rows are in flight ahead of the row being used
  constexpr int PF = GAS ? 2 : 3;
  RowF pend[PF];
#pragma unroll
  for (int i = 0; i < PF; ++i)
    if (jstart + i <= jend) fetch_row(jstart + i, pend[i]);
  for (int j = jstart; j <= jend; ++j) {
    if (GEOM == 3 && visc && j >= jstart + 3) { vbx = R0.q[1]; vby = R0.q[2]; }   // row j-3
    R0 = R1;
    R1 = R2;
    R2 = pend[0];
#pragma unroll
    for (int i = 0; i + 1 < PF; ++i) pend[i] = pend[i + 1];
    if (j + PF <= jend) fetch_row(j + PF, pend[PF - 1]);
    convert_row(R2);
    // in 1D the single row is row "R1" of the loop body below
    const RowF& Rm = (NDIM == 2) ? R1 : R2;
    const int m = (NDIM == 2) ? j - 1 : 1;   // row whose face states are built now
    double sLy_new[5] = {0, 0, 0, 0, 0};
    if (NDIM == 2 && j >= y0 - 1) {
#pragma unroll
      for (int c = 0; c < NQ; ++c) sLy_new[c] = exdiv(2.0 * (R2.q[c] - R1.q[c]), dsy, ysy);
    }
    const bool have_faces = (NDIM == 1) || (j >= y0);
    double Fx_own[5] = {0, 0, 0, 0, 0}, Fx_rt[5] = {0, 0, 0, 0, 0};
    double Fy_top[5] = {0, 0, 0, 0, 0};
    if (have_faces) {
      double dqx[5] = {0, 0, 0, 0, 0}, dqy[5] = {0, 0, 0, 0, 0};
#pragma unroll
      for (int c = 0; c < NQ; ++c) {
        const double ql = shfl_up1(Rm.q[c]);
        const double sL = exdiv(2.0 * (Rm.q[c] - ql), dsx, ysx);
        const double sR = shfl_dn1(sL);
        dqx[c] = slope_limit(st, sL, sR);
        dqy[c] = (NDIM == 2) ? exdiv(slope_limit(st, sLy_prev[c], sLy_new[c]), rp, yrp) : 0.0;
      }
      faces(Rm, dqx, dqy, m, fx1, fy1);
      // x fluxes of row m, needed when m is an owned row
      if (NDIM == 1 || (m >= y0 && m <= y1)) {
        double L[5] = {0, 0, 0, 0, 0};
#pragma unroll
        for (int c = 0; c < NQ; ++c) L[c] = shfl_up1(fx1.m[c]);
        const double csl = shfl_up1(Rm.cs);
        const int fcell = (NDIM == 1) ? ixc - 1 : (ixc - 1) + (m - 1) * a.nxm;
        if (!GAS && a.fcs) {   // adiabatic gas: the gas pass's face sound speeds
          riemann(L, a.fcs[fcell], fx1.p, a.fcs[nc + fcell], 0, Fx_own);
        } else {
          riemann(L, csl, fx1.p, Rm.cs, 0, Fx_own);
        }
        if constexpr (GAS && ADI) {
          // the face on the left of every owned column and of the ghost column nx+3
          if (a.fcs && (own_col || ix == a.nx + 3)) {
            a.fcs[fcell] = a.eos_baro ? cs_baro(a, L[0]) : sqrt(a.gamma * L[4] / L[0]);
            a.fcs[nc + fcell] = a.eos_baro ? cs_baro(a, fx1.p[0]) : sqrt(a.gamma * fx1.p[4] / fx1.p[0]);
          }
        }
#pragma unroll
        for (int c = 0; c < 5; ++c) Fx_rt[c] = shfl_dn1(Fx_own[c]);
      }
      // y flux through the face between rows m-1 and m
      if (NDIM == 2 && m - 1 >= y0 - 1 && m - 1 <= y1) {
        const int fcell = (ixc - 1) + (m - 1) * a.nxm;
        if (!GAS && a.fcs) {
          riemann(fym_prev, a.fcs[2 * nc + fcell], fy1.p, a.fcs[3 * nc + fcell], 1, Fy_top);
        } else {
          riemann(fym_prev, cs_prev, fy1.p, Rm.cs, 1, Fy_top);
        }
        if constexpr (GAS && ADI) {
          if (a.fcs && own_col) {
            a.fcs[2 * nc + fcell] = a.eos_baro ? cs_baro(a, fym_prev[0]) : sqrt(a.gamma * fym_prev[4] / fym_prev[0]);
            a.fcs[3 * nc + fcell] = a.eos_baro ? cs_baro(a, fy1.p[0]) : sqrt(a.gamma * fy1.p[4] / fy1.p[0]);
          }
        }
      }
    }
    // finalize row m-1 (2D) or the single row (1D)
    const int fr = (NDIM == 2) ? m - 1 : 1;
    const RowF& Rf = (NDIM == 2) ? R0 : R2;
    const double* Fxo = (NDIM == 2) ? Fx_own_prev : Fx_own;
    const double* Fxr = (NDIM == 2) ? Fx_rt_prev : Fx_rt;
    // x neighbours of the row's velocities for the viscous Laplacian (all lanes shuffle)
    double vlx = 0.0, vrx = 0.0, vly = 0.0, vry = 0.0;
    if (visc) {
      vlx = shfl_up1(Rf.q[1]); vrx = shfl_dn1(Rf.q[1]);
      vly = shfl_up1(Rf.q[2]); vry = shfl_dn1(Rf.q[2]);
    }
    if (own_col && ((NDIM == 1) || (have_faces && fr >= y0 && fr <= y1))) {
      const int cell = (NDIM == 1) ? ixc - 1 : (ixc - 1) + (fr - 1) * a.nxm;
      double uk[5];
#pragma unroll
      for (int c = 0; c < NU; ++c) uk[c] = Rf.u[c];
      if (a.force_kick) {  // kick(1.0d0), force_kicks.f90:12-39
        double f1, f2, f3;
        cell_force<GEOM == 3, GEOM == 0 && NDIM == 2>(a, ix, fr, 0, Rf.q[1], Rf.q[3], GAS, fluid - 1, f1, f2, f3);
        const double rho = Rf.u[0];
        if (GAS) {
          const double eold = ((kHalf * (uk[1] * uk[1])) / rho + (kHalf * (uk[2] * uk[2])) / rho) +
                              (kHalf * (uk[3] * uk[3])) / rho;
          uk[1] = uk[1] + ((rho * 1.0) * dt) * f1;
          uk[2] = uk[2] + ((rho * 1.0) * dt) * f2;
          uk[3] = uk[3] + ((rho * 1.0) * dt) * f3;
          const double enew = ((kHalf * (uk[1] * uk[1])) / rho + (kHalf * (uk[2] * uk[2])) / rho) +
                              (kHalf * (uk[3] * uk[3])) / rho;
          uk[4] = uk[4] + (enew - eold);
        } else {
          uk[1] = uk[1] + ((rho * 1.0) * dt) * f1;
          uk[2] = uk[2] + ((rho * 1.0) * dt) * f2;
          uk[3] = uk[3] + ((rho * 1.0) * dt) * f3;
        }
      }
      // sources (source_terms.f90:37-47), from the start-of-step primitives
      double S[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
      if (POLAR) {
        const double r = rp;
        if (GAS) {
          const double P = ADI ? Rf.q[4] : Rf.u[0] * (Rf.cs * Rf.cs);
          S[1] = 0.0 + exdiv(dt * (Rf.q[0] * (Rf.q[2] * Rf.q[2]) + P), r, yrp);
          S[2] = 0.0 - dt * exdiv(Rf.q[0] * Rf.q[2] * Rf.q[1], r, yrp);
        } else {
          S[1] = 0.0 + dt * exdiv(Rf.q[0] * (Rf.q[2] * Rf.q[2]), r, yrp);
          S[2] = 0.0 - dt * exdiv(Rf.q[0] * Rf.q[2] * Rf.q[1], r, yrp);
        }
      }
      if constexpr (SPH) {
        // source_terms.f90:33-35 (the pressure term of the spherical momentum equation) and the
        // gravity of the enclosed mass :79-85, from the start-of-step primitives
        if (GAS) S[1] = S[1] - dt * (-(((2.0 * Rf.q[0]) * (Rf.cs * Rf.cs)) / rsph));
        if (a.gravity) {
          const double d = rsph * rsph + a.l_soft * a.l_soft;
          S[1] = S[1] - dt * ((Rf.q[0] * a.mc[cell]) / d);
        }
      }
      if constexpr (GAS && ADI) {
        // drag heating of adiabatic gas (source_terms.f90:88-95), species in order, from the
        // start-of-step primitives and stopping times
        if (a.nd > 0 && a.back_reaction) {
          const double vx = Rf.q[1], vy = Rf.q[2], vz = Rf.q[3];
          const int nd = a.nd;
          double s4 = 0.0;
          for (int k = 0; k < nd; ++k) {
            const double rd = a.U[(5 + k) * nc + cell];
            const double vdx = a.U[(5 + nd + k) * nc + cell] / rd;
            const double vdy = a.U[(5 + 2 * nd + k) * nc + cell] / rd;
            const double vdz = a.U[(5 + 3 * nd + k) * nc + cell] / rd;
            const double ts = tstop_of<1>(a, k, cell, ixc, Rf.q[0], rd, Rf.cs);
            s4 = ((s4 + (dt * (vx * (vx - vdx))) / ts) + (dt * (vy * (vy - vdy))) / ts) +
                 (dt * (vz * (vz - vdz))) / ts;
          }
          S[4] = s4;
        }
      }
      if constexpr (VISC && GEOM == 4) {
        // viscosity (source_terms.f90:111-116, GEOM 2): radial Laplacians from the start-of-step
        // primitives, eta_visc = alpha_visc H cs with H = HoverR r (disk_polar/setup.f90:29-30, 67),
        // added to the polar sources
        if (visc) {
          const double r = rp;
          const double eta = (a.alpha_visc * (a.hoverr * r)) * Rf.cs;
          const double vx = Rf.q[1], vy = Rf.q[2];
          const double su = vrx + vlx, sv = vry + vly;
          const double d1 = dx1 * dx1;
          const double lxu = (su - 2.0 * vx) / d1 + su / r / dx1 / 2.0 - vx / r / r;
          const double lxv = (sv - 2.0 * vy) / d1 - vy / (r * r) + sv / r / dx1 / 2.0;
          const double k = Rf.q[0] * dt * eta;
          S[1] = S[1] + k * lxu;
          S[2] = S[2] + k * lxv;
        }
      }
      if constexpr (VISC && GEOM == 3) {
        // viscosity (source_terms.f90:99-110, GEOM 0) from the start-of-step primitives, with
        // eta_visc = alpha_visc H cs (disk/setup.f90:60); rows fr-1 and fr+1 are vb* and R1
        if (visc) {
          const double xx = a.g.pos1[ixc - 1] - kHalf * a.box_l, yy = a.g.pos2[fr - 1] - kHalf * a.box_l_y;
          const double rr = sqrt((xx * xx + yy * yy) + a.smooth_r * a.smooth_r);
          const double eta = (a.alpha_visc * (a.hoverr * rr)) * Rf.cs;
          const double vx = Rf.q[1], vy = Rf.q[2];
          const double d1 = dx1 * dx1, d2 = dx2 * dx2;
          const double lxu = ((vrx + vlx) - 2.0 * vx) / d1;
          const double lyu = ((R1.q[1] + vbx) - 2.0 * vx) / d2;
          const double lxv = ((vry + vly) - 2.0 * vy) / d1;
          const double lyv = ((R1.q[2] + vby) - 2.0 * vy) / d2;
          const double k = Rf.q[0] * dt * eta;
          S[1] = (0.0 + k * lxu) + k * lyu;
          S[2] = 0.0 + k * (lxv + lyv);
        }
      }
#pragma unroll
      for (int c = 0; c < NU; ++c) {
        double t = uk[c] + exdiv(Fxo[c] * surf1 - Fxr[c] * surf1r, vol, yvol) * dt;
        if (NDIM == 2) t = t + exdiv(Fy_bot[c] * surf2 - Fy_top[c] * surf2, vol, yvol) * dt;
        t = t + S[c];
        a.B[p[c] * nc + cell] = t;
      }
    }
    // rotate
    if (NDIM == 2) {
#pragma unroll
      for (int c = 0; c < NQ; ++c) { sLy_prev[c] = sLy_new[c]; fym_prev[c] = fy1.m[c]; }
      cs_prev = R1.cs;
#pragma unroll
      for (int c = 0; c < 5; ++c) { Fx_own_prev[c] = Fx_own[c]; Fx_rt_prev[c] = Fx_rt[c]; Fy_bot[c] = Fy_top[c]; }
    }
  }
}

// the gas kernel is allowed two waves per SIMD (a few registers spill; measured faster on the
// 1024 x 2048 disk), the dust kernel keeps its full register budget
template <int NDIM, int GEOM, bool GAS, int SLOPE, int SOLV, bool ADI>
__global__ __launch_bounds__(256, (GAS ? 2 : 1)) void k_hydro(StepArgs a) {
  if (!a.sc->active) return;
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  // XCD-aware order: each of the 8 XCDs takes a contiguous range of work items, so the halo rows
  // shared by neighbouring strips of a column group are read through one L2
  unsigned b = blockIdx.x;
  const unsigned nb = gridDim.x;
  if (a.xcd_order && (nb & 7u) == 0) b = (b & 7u) * (nb >> 3) + (b >> 3);
  const long item = (long)b * 4 + wave;
  const long G = a.ngroups, S = GAS ? a.nstrips_g : a.nstrips, F = GAS ? 1 : a.nd;
  if (item >= G * S * F) return;
  const int g = (int)(item % G);
  const long t = item / G;
  const int strip = (int)(t % S);
  const int fluid = GAS ? 0 : 1 + (int)(t / S);
  hydro_fluid<NDIM, GEOM, GAS, SLOPE, SOLV, ADI>(a, fluid, g, strip, lane);
}

// ------------------------------------------------------------------ post
// Courant candidate of one active cell (courant.f90:68-158), from its conservative state.
// Sound speed of a cell from its conservative gas state: the setup's cs when isothermal, else
// ctoprim's sqrt(gamma P / max(rho, smallr)) with P = max((gamma-1)(E - ekin), smallp)
// (solve.f90:214-221), from the velocities u/rho.
DEV double cs_cell(const StepArgs& a, double rho, double E, double vx, double vy, double vz, int cell) {
  if (a.eos_baro) return cs_baro(a, fmaxf_(rho, a.smallr));
  if (a.iso) return a.cs[cell];
  const double ekin = kHalf * rho * (vx * vx) + kHalf * rho * (vy * vy) + kHalf * rho * (vz * vz);
  const double P = fmaxf_((a.gamma - 1.0) * (E - ekin - 0.0), a.smallp);
  return sqrt(a.gamma * P / fmaxf_(rho, a.smallr));
}

// ctoprim's sound speed of a cell's start-of-step state (a.U)
DEV double start_cs(const StepArgs& a, int cell) {
  if (a.eos_baro) return cs_baro(a, fmaxf_(a.U[cell], a.smallr));
  if (a.iso) return a.cs[cell];
  const size_t nc = a.nc;
  const double rho = a.U[cell];
  return cs_cell(a, rho, a.U[4 * nc + cell], a.U[nc + cell] / rho, a.U[2 * nc + cell] / rho,
                 a.U[3 * nc + cell] / rho, cell);
}

// next: 1 when the candidate is for the step after the one being taken (cell_force)
// ND: species capacity of the arrays; nd (<= ND): species in the run
template <int ND>
DEV double courant_cell(const StepArgs& a, const double* un, int ix, int iy, int next, int cell, int nd) {
  double dxx;
  if (a.ndim == 1) dxx = a.g.dx1[ix - 1];
  else if (a.geom == 2) dxx = fminf_(a.g.dx1[ix - 1], a.g.rc[ix - 1] * a.g.dx2[ix - 1]);
  else dxx = fminf_(a.g.dx1[ix - 1], a.g.dx2[ix - 1]);
  const double rho = un[0];
  const double vx = un[1] / rho, vy = un[2] / rho, vz = un[3] / rho;
  const double cs = cs_cell(a, rho, un[4], vx, vy, vz, cell);
  double vv = (fabs(vx) + fabs(vy)) + fabs(vz);
  double vmax = cs + vv;
  double vdx[ND > 0 ? ND : 1], vdz[ND > 0 ? ND : 1];
#pragma unroll
  for (int k = 0; k < ND; ++k) {
    if (k >= nd) break;
    const double rd = un[5 + k];
    vdx[k] = un[5 + nd + k] / rd;
    const double vdy = un[5 + 2 * nd + k] / rd;
    vdz[k] = un[5 + 3 * nd + k] / rd;
    vv = (fabs(vdx[k]) + fabs(vdy)) + fabs(vdz[k]);
    vmax = fmaxf_(vmax, 0.0 + vv);
  }
  double dt = (a.cfl * dxx) / fabs(vmax);
  if (a.gravity) {   // courant.f90:62-63
    const double r = a.g.rc[ix - 1];
    dt = fminf_(dt, (a.cfl * dxx) / sqrt(a.mc[cell] / sqrt(r * r + a.l_soft * a.l_soft)));
  }
  if (a.force_kick) {
    double f1, f2, f3;
    cell_force(a, ix, iy, next, vx, vz, true, -1, f1, f2, f3);
    double fm = sqrt((f1 * f1 + f2 * f2) + f3 * f3);
#pragma unroll
    for (int k = 0; k < ND; ++k) {
      if (k >= nd) break;
      cell_force(a, ix, iy, next, vdx[k], vdz[k], false, k, f1, f2, f3);
      fm = fmaxf_(fm, sqrt((f1 * f1 + f2 * f2) + f3 * f3));
    }
    if (vv != 0.0) {
      const double fratio = fmaxf_((fm * dxx) / (vv * vv), 1e-3);
      dt = fminf_(dt, (((a.cfl * dxx) / vv) * (sqrt(1.0 + (2.0 * a.cfl) * fratio) - 1.0)) / fratio);
    }
  }
  return dt;
}

// The inflow value of the polar disk's accretion boundary (disk_polar/boundaries.f90:63-94) for
// plane v of the last active column in row iy, as k_bc_accrete writes it; false where the row
// takes no inflow (phi > phi_mom pi) or the plane is not rewritten (vz, E and the dust vz).
DEV bool accretion_inflow(const StepArgs& a, int v, int iy, double& out) {
  const double phi = (((iy - 3) + kHalf) * (2.0 * kPi)) / (double)(float)a.ny;
  if (!(phi <= a.phi_mom * kPi)) return false;
  const bool gas = v < 5;
  const int f = gas ? v : (v - 5) / (a.nd > 0 ? a.nd : 1);   // dust: density, x, y, z plane
  if (f > 2) return false;
  const double vrout = a.cs[cell_of(a, a.nx + 2, iy)];
  const double sigma_out = a.m_acc / vrout / a.phi_mom / kPi / a.box_l;
  const double rs = gas ? sigma_out : a.dust2gas * sigma_out;
  if (f == 0) out = rs;
  else if (f == 1) out = -(rs * vrout);
  else out = rs * a.l_out * sqrt(a.mstar / 100.0) * 100.0 / a.box_l;
  return true;
}

DEV unsigned long long wave_min_u64(unsigned long long v) {
#pragma unroll
  for (int o = 32; o >= 1; o >>= 1) {
    unsigned long long w = __shfl_xor(v, o);
    v = w < v ? w : v;
  }
  return v;
}

// The block's minimum goes to the device-wide running minimum.  Every block's atomic lands on the
// same address, where they serialise (24% of the shearing-box post pass when each block issued
// one), so a block first reads the stored minimum and skips the atomic when it cannot lower it.
// The stored value only decreases within a launch, so any value read is >= the final minimum and
// the skip never drops the smallest candidate: the result is the same bit pattern.  cur0, a value of
// the stored minimum read when the block started (with that read's latency hidden behind the
// block's work), settles most blocks without waiting on a second read.
DEV void block_atomic_min(DevScalars* sc, double d, unsigned long long cur0 = ~0ull) {
  __shared__ unsigned long long red[4];
  unsigned long long v = wave_min_u64(dbits(d));
  const int wave = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0) red[wave] = v;
  __syncthreads();
  if (threadIdx.x == 0) {
    unsigned long long m = red[0];
    for (int w = 1; w < (int)(blockDim.x >> 6); ++w) m = red[w] < m ? red[w] : m;
    if (m < cur0) {
      const unsigned long long cur = __hip_atomic_load(&sc->dtbits, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      if (m < cur) atomicMin(&sc->dtbits, m);
    }
  }
}

// Stopping time of species k from the start-of-step primitives; cs_q is that state's sound speed
// (read only by the uniform Epstein law, KH setup.f90:238 / SOD setup.f90:182).  CD: the
// cartesian disk's law is compiled in (left out of the FARGO kernels, which never need it)
template <int ND, bool CD>
DEV double tstop_of(const StepArgs& a, int k, int cell, int ix, double rho_q, double rhod_q, double cs_q) {
  if (a.tstop_kind == 0) return ((rhod_q * rho_q) / a.K[k]) / rho_q;  // DUSTYWAVE setup.f90:217
  const double s = a.sdust_uniform ? a.sdust_k[k] : a.sdust[(size_t)k * a.nc + cell];
  if (a.tstop_kind == 1)  // SI setup.f90:392
    return ((a.rhograin * s) / a.rho_init) / ((a.rad0 * a.omega_shear) * a.hoverr);
  // uniform Epstein (kind 3) and Collapse_1D's (kind 5, src/setups/Collapse_1D/setup.f90:310-320,
  // with the barotropic cs_q)
  if (a.tstop_kind == 3 || a.tstop_kind == 5) return ((sqrt((kPi * a.gamma) / 8.0) * a.rhograin) * s) / (rho_q * cs_q);
  if (CD && a.tstop_kind == 4) {   // cartesian disk, disk/setup.f90:333-341
    const int iy = cell / a.nxm + 1;
    const double xx = a.g.pos1[ix - 1] - kHalf * a.box_l, yy = a.g.pos2[iy - 1] - kHalf * a.box_l_y;
    const double H = a.hoverr * sqrt((xx * xx + yy * yy) + a.smooth_r * a.smooth_r);
    return ((sqrt((kPi * a.gamma) / 8.0) * a.rhograin) * s) / ((rho_q / sqrt(2.0 * kPi * H)) * a.cs[cell]);
  }
  const double H = a.hoverr * a.g.rc[ix - 1];  // disk_polar setup.f90:436
  return (sqrt((kPi * a.gamma) / 8.0) * a.rhograin) * s /
         ((rho_q / (sqrt(2.0 * kPi) * H)) * a.cs[cell]);
}

// k_post2: the post pass with all fluids of a row segment of TW cells staged in LDS and the
// species spread over the NS = 256/TW thread slots of the block.  The drag sums run in
// species order in one thread per (cell, component), so results equal the reference bit for
// bit; loads, divisions and the Courant work are shared by all slots.
// PART 0: the whole post pass.  With a growth step between drag and the in-loop hook
// (solve.f90:110-152) the pass is split: PART 1 remaps and drags (writes U, no Courant candidate),
// PART 2 reloads U after growth, applies the in-loop hook and computes the Courant candidate.
template <int ND, bool FARGO, int NDIM, int TW, int PART, bool CAP>
__global__ __launch_bounds__(256, (TW == 32 ? 6 : 3)) void k_post2(StepArgs a) {
  if (!a.sc->active) return;
  const unsigned long long cur0 =
      (PART != 1 && threadIdx.x == 0) ? __hip_atomic_load(&a.sc->dtbits, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) : ~0ull;
  constexpr int NV = 5 + 4 * ND;
  // CAP: ND is a capacity and the run's species count nd (<= ND) is read at run time; otherwise
  // nd is the compile-time ND
  const int nd = CAP ? a.nd : ND;
  const int nv = 5 + 4 * nd;
  constexpr int NS = 256 / TW;
  constexpr int NDL = ND > 0 ? ND : 1;
  // the Courant scratch of phase 5 reuses the drag arrays, which are dead by then, to keep the
  // block's LDS small enough for several blocks per CU
  __shared__ double un[NV][TW];
  __shared__ double akS[NDL > 2 * NS ? NDL : 2 * NS][TW];   // alpha_k; phase 5: vmax, |F| per slot
  __shared__ double gS[NDL][TW];   // f_k = alpha_k / (1 + alpha_k)
  __shared__ double sumS[4][TW];   // drag sums; phase 5: vv of the last species
  double (*vmaxS)[TW] = akS;
  double (*fmS)[TW] = akS + NS;
  double* vvlast = sumS[0];
  const int c = threadIdx.x % TW, s = threadIdx.x / TW;
  // XCD-aware block order: the hardware deals consecutive blocks to the 8 XCDs in turn; the
  // swizzle gives each XCD a contiguous band of rows, so the FARGO source rows a block reads were
  // loaded into the same XCD's L2 by its neighbours in the band
  int bx = blockIdx.x, by = blockIdx.y;
  if (NDIM == 2 && a.xcd_order) {
    const unsigned nbx = gridDim.x, nb = nbx * gridDim.y;
    if ((nb & 7u) == 0) {
      const unsigned b = blockIdx.x + blockIdx.y * nbx;
      const unsigned l = (b & 7u) * (nb >> 3) + (b >> 3);
      bx = (int)(l % nbx);
      by = (int)(l / nbx);
    }
  }
  const int ix = 3 + bx * TW + c;
  const int iy = (NDIM == 2) ? 3 + by : 1;
  const bool valid = ix <= a.nx + 2;
  const int ixs = valid ? ix : a.nx + 2;  // clamped column for loads of idle lanes
  const size_t nc = a.nc;
  const int cell = cell_of(a, ixs, iy);
  const double dt = a.sc->dt;
  // the start-of-step density and sound speed the stopping times read, and the isothermal sound
  // speed of the Courant candidate (a field no kernel writes), loaded ahead of phase 1's loads so
  // that their latency overlaps those
  const bool drag_pre = nd > 0 && a.drag && PART != 2;
  const double u0_pre = drag_pre ? a.U[cell] : 0.0;
  const bool cs_iso = a.iso && !a.eos_baro && PART != 1;
  const double cs_pre = ((drag_pre && a.tstop_kind == 2) || (cs_iso && s == 0)) ? a.cs[cell] : 0.0;

  // ---- phase 1: remap (FARGO) or copy every variable of the segment into LDS
  if (PART == 2) {
#pragma unroll
    for (int v = s; v < nv; v += NS) un[v][c] = a.U[v * nc + cell];
    __syncthreads();   // the planes below were loaded by other slots
    // growth changed the dust densities only; each species keeps its velocity:
    // momentum = (momentum / density before growth) * density after growth
    for (int k = s; k < nd; k += NS) {
      const double r0 = a.Rold[(size_t)k * nc + cell], r1 = un[5 + k][c];
      un[5 + nd + k][c] = (un[5 + nd + k][c] / r0) * r1;
      un[5 + 2 * nd + k][c] = (un[5 + 2 * nd + k][c] / r0) * r1;
      un[5 + 3 * nd + k][c] = (un[5 + 3 * nd + k][c] / r0) * r1;
    }
    __syncthreads();
  } else if (FARGO) {
    const double rp = (a.geom == 2) ? a.g.rc[ixs - 1] : 1.0;
    const double vk = a.fargo_vk[ixs - 1];
    const double dy = a.g.dx2[ixs - 1] * rp;
    const double sh = fabs(vk * dt);
    const int nshft = (int)(sh / dy);
    const double eps = fmod(sh, dy) / dy;
    const int ny = a.ny;
    int j0 = (vk > 0.0) ? iy - nshft - 1 : iy + nshft;
    j0 = ((j0 - 3) % ny + ny) % ny + 3;
    const int j1 = (j0 - 3 + 1) % ny + 3;
    const int jm = (j0 - 3 + ny - 1) % ny + 3;
    const int jp = (j1 - 3 + 1) % ny + 3;
    const int cm = cell_of(a, ixs, jm), c0 = cell_of(a, ixs, j0), c1 = cell_of(a, ixs, j1),
              cp = cell_of(a, ixs, jp);
    const double h2 = (kHalf - eps) * (kHalf - eps);
    const double w0 = (vk > 0.0) ? eps : (1.0 - eps), w1 = (vk > 0.0) ? (1.0 - eps) : eps;
    // the slot's variables are remapped CH at a time: the 4·CH loads of a chunk are issued
    // before any is used, and the small chunk keeps the kernel at 68 VGPRs, so six blocks fit a
    // CU (the LDS limit) where issuing every load at once needed 96 VGPRs and allowed five
    constexpr int NPS = (NV + NS - 1) / NS;
    constexpr int CH = 3;
#pragma unroll
    for (int i0 = 0; i0 < NPS; i0 += CH) {
      double um[CH], u0[CH], u1[CH], up[CH];
#pragma unroll
      for (int j = 0; j < CH; ++j) {
        const int v = s + (i0 + j) * NS;
        if (i0 + j < NPS && v < nv) {
          const double* Bv = a.B + v * nc;
          um[j] = Bv[cm]; u0[j] = Bv[c0]; u1[j] = Bv[c1]; up[j] = Bv[cp];
        }
      }
#pragma unroll
      for (int j = 0; j < CH; ++j) {
        const int v = s + (i0 + j) * NS;
        if (i0 + j < NPS && v < nv) {
          const double d0 = slope_limit(a.slope, u0[j] - um[j], u1[j] - u0[j]);
          const double d1 = slope_limit(a.slope, u1[j] - u0[j], up[j] - u1[j]);
          double x = u0[j] * w0 + u1[j] * w1;
          x = x + kHalf * d0 * (0.25 - h2) + kHalf * d1 * (h2 - 0.25);
          un[v][c] = x;
        }
      }
    }
    if (a.accretion && ix == a.nx + 2) {
      // fargo_scheme applies the boundaries again after the remap (fargo_scheme.f90:80), and
      // with accretion they reset the last active column to the inflow state (as k_bc_accrete)
      for (int v = s; v < nv; v += NS) {
        double x;
        if (accretion_inflow(a, v, iy, x)) un[v][c] = x;
      }
    }
  } else {
#pragma unroll
    for (int v = s; v < nv; v += NS) un[v][c] = a.B[v * nc + cell];
  }
  const bool do_drag = nd > 0 && a.drag && PART != 2;
  if (do_drag) {
    // stopping times of the step (compute_tstop on the start-of-step primitives); alpha_k and
    // f_k = alpha_k / (1 + alpha_k) are shared by the drag sums and the momentum update, f_k
    // computed once per species here rather than in each of the four sums and the update
    const double rho_q = fmaxf_(u0_pre, a.smallr);
    if (a.tstop_kind == 2) {
      // Epstein (disk_polar setup.f90:436): the gas factor is the same for every species
      const double H = a.hoverr * a.g.rc[ixs - 1];
      const double den = (rho_q / (a.s2pi * H)) * cs_pre;
      const double yden = 1.0 / den;
      for (int k = s; k < nd; k += NS) {
        const double sk = a.sdust_uniform ? a.sdust_k[k] : a.sdust[(size_t)k * a.nc + cell];
        const double ak = (1.0 * dt) / exdiv(a.eps_k1 * sk, den, yden);
        akS[k][c] = ak;
        gS[k][c] = exdiv(ak, 1.0 + ak, 1.0 / (1.0 + ak));
      }
    } else {
      const double cs_q = (a.tstop_kind == 3 || a.tstop_kind == 5) ? start_cs(a, cell) : 0.0;
      for (int k = s; k < nd; k += NS) {
        const double rq = (a.tstop_kind == 0) ? a.U[(5 + k) * nc + cell] : 0.0;
        const double ak = (1.0 * dt) / tstop_of<ND, !FARGO>(a, k, cell, ixs, rho_q, rq, cs_q);
        akS[k][c] = ak;
        gS[k][c] = exdiv(ak, 1.0 + ak, 1.0 / (1.0 + ak));
      }
    }
  }
  __syncthreads();
  // ---- phase 2: drag sums in species order (dust_drag.f90:24-38)
  if (do_drag) {
    if (s < 4) {
      const int base = (s == 3) ? 5 : 5 + (s + 1) * nd;   // rhod, vdx, vdy, vdz planes
      double acc = un[s == 3 ? 0 : s + 1][c];
#pragma unroll 4
      for (int k = 0; k < nd; ++k) {
        acc = acc + gS[k][c] * un[base + k][c];
      }
      sumS[s][c] = acc;
    }
    __syncthreads();
    // ---- phase 3: momentum update and floor (dust_drag.f90:40-73)
    const double pnx = sumS[0][c], pny = sumS[1][c], pnz = sumS[2][c], rhon = sumS[3][c];
    const double yrho = 1.0 / rhon;
    for (int k = s; k < nd; k += NS) {
      const double ak = akS[k][c], ap1 = 1.0 + ak, y1 = 1.0 / ap1;
      const double f = gS[k][c];
      const double rd = un[5 + k][c];
      un[5 + nd + k][c] = exdiv(un[5 + nd + k][c], ap1, y1) + exdiv(f * pnx, rhon, yrho) * rd;
      un[5 + 2 * nd + k][c] = exdiv(un[5 + 2 * nd + k][c], ap1, y1) + exdiv(f * pny, rhon, yrho) * rd;
      un[5 + 3 * nd + k][c] = exdiv(un[5 + 3 * nd + k][c], ap1, y1) + exdiv(f * pnz, rhon, yrho) * rd;
      un[5 + k][c] = fmaxf_(un[0][c] * a.dust_ratio_min, rd);
    }
    if (s == NS - 1 && a.back_reaction) {
      const double rho = un[0][c];
      un[1][c] = exdiv(pnx, rhon, yrho) * rho;
      un[2][c] = exdiv(pny, rhon, yrho) * rho;
      un[3][c] = exdiv(pnz, rhon, yrho) * rho;
    }
    __syncthreads();
  }
  // ---- phase 4: in-loop hook, per fluid
  if (PART == 1) {
    // growth runs next; the hook and the Courant candidate follow it in PART 2
  } else if (a.inloop_kind == 2) {
    const double r = a.g.rc[ixs - 1];
    if (r < a.r_relax) {
      const double t_rel = a.n_rel * 2.0 * kPi / (sqrt(a.mstar / (r * r * r))) -
                           a.n_rel * 2.0 * kPi / (sqrt(a.mstar / (a.smooth_r * a.smooth_r * a.smooth_r)));
      const double e = exp(-dt / t_rel);
      const double* ci = a.condinit;
      for (int f = s; f <= nd; f += NS) {
        const int ir = f == 0 ? 0 : 5 + f - 1;
        const int iu = f == 0 ? 1 : 5 + nd + f - 1;
        const int iv = f == 0 ? 2 : 5 + 2 * nd + f - 1;
        const double rho_i = ci[ir * nc + cell], vx_i = ci[iu * nc + cell] / rho_i,
                     vy_i = ci[iv * nc + cell] / rho_i;
        const double rho_o = un[ir][c], vx_o = un[iu][c] / rho_o, vy_o = un[iv][c] / rho_o;
        const double rho_n = (rho_i) * (1.0 - e) + rho_o * e;
        const double vx_n = (vx_i) * (1.0 - e) + vx_o * e;
        const double vy_n = (vy_i) * (1.0 - e) + vy_o * e;
        un[ir][c] = rho_n; un[iu][c] = rho_n * vx_n; un[iv][c] = rho_n * vy_n;
      }
    }
    __syncthreads();
    if (ix == 3)
      for (int v = s; v < nv; v += NS) un[v][c] = a.condinit[v * nc + cell];
    __syncthreads();
  } else if (!FARGO && a.inloop_kind == 3) {
    // cartesian disk buffers (disk/setup.f90:203-262): reset to the initial disk (rate = 0)
    const double xx = a.g.pos1[ixs - 1] - kHalf * a.box_l, yy = a.g.pos2[iy - 1] - kHalf * a.box_l_y;
    const double rr = sqrt((xx * xx + yy * yy) + a.smooth_r * a.smooth_r);
    if (rr < a.buffer_in * a.box_l || rr > a.buffer_out * a.box_l) {
      const double r_cyl = sqrt(xx * xx + yy * yy);
      const double x_soft = xx * rr / r_cyl, y_soft = yy * rr / r_cyl;
      const double Omega = sqrt(a.mstar / (rr * (rr * rr)));
      const double cs0 = (a.hoverr * rr) * Omega;
      const double R_out_disk = a.frac_r_disk * a.box_l, rate = 0.0;
      double rho_ana = a.sigma_r0 * (a.r0_disk / rr);
      if (rr > R_out_disk) rho_ana = a.sigma_r0 * (a.r0_disk / R_out_disk) / 100.0;
      const double v_ana = -(Omega * rho_ana);
      for (int f = s; f <= nd; f += NS) {
        if (f == 0) {
          un[0][c] = rho_ana + (un[0][c] - rho_ana) * rate;
          un[1][c] = -(y_soft * v_ana) + (un[1][c] + yy * v_ana) * rate;
          un[2][c] = x_soft * v_ana + (un[2][c] - xx * v_ana) * rate;
          un[4][c] = (rho_ana + (un[0][c] - rho_ana) * rate) * (cs0 * cs0);
        } else {
          const int k = f - 1;
          const double dust_to = a.dust2gas;
          un[5 + k][c] = dust_to * rho_ana + (un[5 + k][c] - rho_ana * dust_to) * rate;
          un[5 + nd + k][c] = -(yy * v_ana * dust_to) + (un[5 + nd + k][c] + dust_to * yy * v_ana) * rate;
          un[5 + 2 * nd + k][c] = xx * v_ana * dust_to + (un[5 + 2 * nd + k][c] - dust_to * xx * v_ana) * rate;
        }
      }
    }
    __syncthreads();
  } else if (a.inloop_kind == 1 && !a.force_kick) {
    const double Om = a.omega_shear, O = Om * dt;
    const double x0 = a.g.pos1[ixs - 1] - kHalf * a.box_l;
    for (int f = s; f <= nd; f += NS) {
      const int ir = f == 0 ? 0 : 5 + f - 1;
      const int ixv = f == 0 ? 1 : 5 + nd + f - 1;
      const int iz = f == 0 ? 3 : 5 + 3 * nd + f - 1;
      const double AA = (f == 0) ? (2.0 * a.q_shear) * Om * x0 - (2.0 * a.rad0) * a.eta_stream
                                 : (2.0 * a.q_shear) * Om * x0;
      const double u = un[iz][c] / un[ir][c], v = un[ixv][c] / un[ir][c], d = un[ir][c];
      un[iz][c] = d * (u * (1. - O * O) + 2. * v * O + AA * (O * O)) / (1. + O * O);
      un[ixv][c] = d * (v * (1. - O * O) - 2. * u * O + AA * O) / (1. + O * O);
    }
    __syncthreads();
  }
  // ---- write the final state
  if (valid)
    for (int v = s; v < nv; v += NS) a.U[v * nc + cell] = un[v][c];
  if (PART == 1) {
    if (valid)
      for (int k = s; k < nd; k += NS) a.Rold[(size_t)k * nc + cell] = un[5 + k][c];
    return;
  }
  // ---- phase 5: next step's Courant candidate (courant.f90:68-158), species split over slots
  {
    double f1, f2, f3;
    double vmax = 0.0, fm = 0.0;
    bool any = false;
    if (s == 0) {
      const double rho = un[0][c], yr = 1.0 / rho;
      const double vx = exdiv(un[1][c], rho, yr), vy = exdiv(un[2][c], rho, yr),
                   vz = exdiv(un[3][c], rho, yr);
      const double vv = (fabs(vx) + fabs(vy)) + fabs(vz);
      vmax = (cs_iso ? cs_pre : cs_cell(a, rho, un[4][c], vx, vy, vz, cell)) + vv;
      if (nd == 0) vvlast[c] = vv;
      if (a.force_kick) {
        cell_force<!FARGO, false>(a, ixs, iy, 1, vx, vz, true, -1, f1, f2, f3);
        fm = sqrt((f1 * f1 + f2 * f2) + f3 * f3);
      }
      any = true;
    }
    // the disk's dust acceleration does not depend on the species' velocity: one evaluation
    double fk_disk = 0.0;
    if (a.force_kick && a.force_kind == 2 && s < nd) {
      cell_force<!FARGO, false>(a, ixs, iy, 1, 0.0, 0.0, false, 0, f1, f2, f3);
      fk_disk = sqrt((f1 * f1 + f2 * f2) + f3 * f3);
    }
    for (int k = s; k < nd; k += NS) {
      const double rd = un[5 + k][c], yd = 1.0 / rd;
      const double vdx = exdiv(un[5 + nd + k][c], rd, yd), vdy = exdiv(un[5 + 2 * nd + k][c], rd, yd),
                   vdz = exdiv(un[5 + 3 * nd + k][c], rd, yd);
      const double vv = (fabs(vdx) + fabs(vdy)) + fabs(vdz);
      vmax = any ? fmaxf_(vmax, 0.0 + vv) : 0.0 + vv;
      if (a.force_kick) {
        double fk = fk_disk;
        if (a.force_kind != 2) {
          cell_force<!FARGO, false>(a, ixs, iy, 1, vdx, vdz, false, k, f1, f2, f3);
          fk = sqrt((f1 * f1 + f2 * f2) + f3 * f3);
        }
        fm = any ? fmaxf_(fm, fk) : fk;
      }
      any = true;
      if (k == nd - 1) vvlast[c] = vv;
    }
    vmaxS[s][c] = any ? vmax : 0.0;
    fmS[s][c] = any ? fm : 0.0;
  }
  __syncthreads();
  double dtc = 2e44;
  if (s == 0 && valid) {
    double vmax = vmaxS[0][c], fm = fmS[0][c];
    for (int w = 1; w < NS; ++w) {
      vmax = fmaxf_(vmax, vmaxS[w][c]);
      fm = fmaxf_(fm, fmS[w][c]);
    }
    double dxx;
    if (a.ndim == 1) dxx = a.g.dx1[ix - 1];
    else if (a.geom == 2) dxx = fminf_(a.g.dx1[ix - 1], a.g.rc[ix - 1] * a.g.dx2[ix - 1]);
    else dxx = fminf_(a.g.dx1[ix - 1], a.g.dx2[ix - 1]);
    dtc = (a.cfl * dxx) / fabs(vmax);
    const double vv = vvlast[c];
    if (a.force_kick && vv != 0.0) {
      const double fratio = fmaxf_((fm * dxx) / (vv * vv), 1e-3);
      dtc = fminf_(dtc, (((a.cfl * dxx) / vv) * (sqrt(1.0 + (2.0 * a.cfl) * fratio) - 1.0)) / fratio);
    }
  }
  block_atomic_min(a.sc, dtc, cur0);
}

template <int ND, int NDIM, bool CAP>
__global__ __launch_bounds__(256) void k_courant(StepArgs a) {
  constexpr int NV = 5 + 4 * ND;
  const int nd = CAP ? a.nd : ND;   // CAP: ND is a capacity, the run has a.nd species
  int ix, iy;
  if (NDIM == 2) {
    ix = 3 + blockIdx.x * 64 + (threadIdx.x & 63);
    iy = 3 + blockIdx.y * 4 + (threadIdx.x >> 6);
  } else {
    ix = 3 + blockIdx.x * 256 + threadIdx.x;
    iy = 1;
  }
  double dtc = 2e44;
  if (ix <= a.nx + 2 && (NDIM == 1 || iy <= a.ny + 2)) {
    const int cell = cell_of(a, ix, iy);
    double un[NV];
#pragma unroll
    for (int v = 0; v < NV; ++v)
      if (v < 5 + 4 * nd) un[v] = a.U[(size_t)v * a.nc + cell];
    dtc = courant_cell<ND>(a, un, ix, iy, 0, cell, nd);
  }
  block_atomic_min(a.sc, dtc);
}

__global__ void k_set_dt(DevScalars* sc) {
  sc->dt = __longlong_as_double((long long)sc->dtbits);
  sc->dtbits = dbits(2e44);
}

__global__ void k_finalize(DevScalars* sc) {
  if (!sc->active) return;
  sc->time = sc->time + sc->dt;
  sc->dt_last = sc->dt;
  sc->nstep = sc->nstep + 1;
  sc->dt = __longlong_as_double((long long)sc->dtbits);
  sc->dtbits = dbits(2e44);
  sc->active = sc->time < sc->tend ? 1 : 0;
}

__global__ void k_arm(DevScalars* sc, double tend) {
  sc->tend = tend;
  sc->active = sc->time < tend ? 1 : 0;
}

// ctoprim over every cell (solve.f90:206-248, isothermal)
__global__ void k_ctoprim(StepArgs a, double* q) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= a.nc) return;
  const size_t nc = a.nc;
  const double rho = a.U[c];
  q[c] = fmaxf_(rho, a.smallr);
  q[nc + c] = a.U[nc + c] / rho;
  q[2 * nc + c] = a.U[2 * nc + c] / rho;
  q[3 * nc + c] = a.U[3 * nc + c] / rho;
  if (a.iso) {
    q[4 * nc + c] = rho * (a.cs[c] * a.cs[c]);
  } else {
    const double vx = a.U[nc + c] / rho, vy = a.U[2 * nc + c] / rho, vz = a.U[3 * nc + c] / rho;
    const double ekin = kHalf * rho * (vx * vx) + kHalf * rho * (vy * vy) + kHalf * rho * (vz * vz);
    q[4 * nc + c] = fmaxf_((a.gamma - 1.0) * (a.U[4 * nc + c] - ekin - 0.0), a.smallp);
  }
  for (int k = 0; k < a.nd; ++k) {
    const double rd = a.U[(5 + k) * nc + c];
    q[(5 + k) * nc + c] = rd;
    q[(5 + a.nd + k) * nc + c] = a.U[(5 + a.nd + k) * nc + c] / rd;
    q[(5 + 2 * a.nd + k) * nc + c] = a.U[(5 + 2 * a.nd + k) * nc + c] / rd;
    q[(5 + 3 * a.nd + k) * nc + c] = a.U[(5 + 3 * a.nd + k) * nc + c] / rd;
  }
}

// ------------------------------------------------------------------ launchers
// zero_gradients (src/grid/boundary_types.f90:48-100).  In 2D its source cells are
// first_active_y / last_active_y and first_active / last_active taken as flat cell indices
// (:70-71, :88-89): cells 3, ny+2 and nx+2 of the first ghost row, themselves written by the
// loops.  In the reference's loop order every ghost cell receives cell 3's value (cell 3 copies
// onto itself, cell nx+2 receives it in the first loop before the second loop reads it), except
// the upper ghost rows of the first loop, which receive cell ny+2's value as it was before the
// first loop reached it and cell 3's value after.  One block stages both source values in LDS
// before any ghost cell is written.
__global__ __launch_bounds__(1024) void k_bc_zg(StepArgs a) {
  if (!a.sc->active) return;
  const size_t nc = a.nc;
  const int nv = a.nvar;
  if (a.ndim == 1) {
    for (int v = threadIdx.x; v < nv; v += blockDim.x) {
      const double l = a.U[v * nc + 2], r = a.U[v * nc + a.nx + 1];
      a.U[v * nc + 0] = l;
      a.U[v * nc + 1] = l;
      a.U[v * nc + a.nxm - 1] = r;
      a.U[v * nc + a.nxm - 2] = r;
    }
    return;
  }
  __shared__ double s3[5 + 4 * 64], fo[5 + 4 * 64];
  const int f = a.ny + 1;   // flat last_active_y, 0-based
  for (int v = threadIdx.x; v < nv; v += blockDim.x) {
    s3[v] = a.U[v * nc + 2];
    fo[v] = a.U[v * nc + f];
  }
  __syncthreads();
  const int fx = f % a.nxm + 1, fy = f / a.nxm + 1;
  const bool f_written = fy <= 2 && fx >= 3 && fx <= a.nx + 2;   // a target of the first loop
  const int nA = 2 * a.nx, nB = 2 * a.nym;
  for (int t = threadIdx.x; t < nA + nB; t += blockDim.x) {
    if (t < nA) {
      const int ix = 3 + t / 2, iy = 1 + t % 2;
      const int bl = cell_of(a, ix, iy), br = cell_of(a, ix, a.nym + 1 - iy);
      const bool late = f_written && (ix > fx || (ix == fx && iy >= fy));
      for (int v = 0; v < nv; ++v) {
        a.U[v * nc + bl] = s3[v];
        a.U[v * nc + br] = late ? s3[v] : fo[v];
      }
    } else {
      const int u = t - nA, ix = 1 + u / a.nym, iy = 1 + u % a.nym;
      const int bl = cell_of(a, ix, iy), br = cell_of(a, a.nxm + 1 - ix, iy);
      for (int v = 0; v < nv; ++v) {
        a.U[v * nc + bl] = s3[v];
        a.U[v * nc + br] = s3[v];
      }
    }
  }
}

// The planet at the step's time and at time + dt (the time loop's time = time + dt,
// shark.f90:107), once per step: theta_p = 2 pi mod(t / T_p, 1) (cartesian disk
// disk/setup.f90:292-301, polar disk disk_polar/setup.f90:331).
__global__ void k_planet(StepArgs a) {
  DevScalars* sc = a.sc;
  const double period = 2.0 * kPi / sqrt(a.mstar / (a.rplanet * (a.rplanet * a.rplanet)));
  for (int w = 0; w < 2; ++w) {
    const double t = w == 0 ? sc->time : sc->time + sc->dt;
    const double theta_p = 2.0 * kPi * fmod(t / period, 1.0);
    const double xx_p = cos(theta_p) * a.rplanet, yy_p = sin(theta_p) * a.rplanet;
    if (a.force_kind == 2) {   // polar disk: the planet's position (its potential per cell)
      sc->planet[2 * w] = xx_p;
      sc->planet[2 * w + 1] = yy_p;
      continue;
    }
    const double rr_p = sqrt((xx_p * xx_p + yy_p * yy_p) + a.smooth_r * a.smooth_r);
    const double rp3 = rr_p * (rr_p * rr_p);
    sc->planet[2 * w] = a.mplanet * xx_p / rp3;
    sc->planet[2 * w + 1] = a.mplanet * yy_p / rp3;
  }
}

// The polar disk's planet (disk_polar/setup.f90:323-358) at the step's time and at time + dt:
// its softened potential in the four neighbours of every active cell, from the grid's position
// and the planet's position (k_planet), and the centred gradient of it, written to
// pgrad[(2 w + component) * nc + cell].  The potential's ghost values follow
// apply_boundaries_phi (boundaries.f90:98-140): y-ghost rows of active columns are periodic
// copies; every x-ghost cell takes the potential of cell number first_active / last_active, i.e.
// cells (3,1) and (nx+2,1) of the first ghost row, which hold rows ny+1 of columns 3 and nx+2.
// With self-gravity and no planet the potential is the one k_selfgrav left in phi_sg, the same
// for both times (the state of the step's start; the next step's dt is taken again from its own).
__global__ void k_planet_grad(StepArgs a) {
  if (!a.sc->active) return;
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= a.nx * a.ny) return;
  const int ix = 3 + t % a.nx, iy = 3 + t / a.nx;
  const size_t cell = (size_t)(ix - 1) + (size_t)(iy - 1) * a.nxm;
  const double r = a.g.rc[ix - 1];
  const double sm2 = a.a_smooth_pl * a.a_smooth_pl;
  const bool planet = a.mplanet > 0.0;
  for (int w = 0; w < 2; ++w) {
    const double cx = a.sc->planet[2 * w], cy = a.sc->planet[2 * w + 1];
    auto pot = [&](int jx, int jy) {
      if (jx < 3) { jx = 3; jy = a.ny + 1; }
      else if (jx > a.nx + 2) { jx = a.nx + 2; jy = a.ny + 1; }
      else if (jy < 3) jy += a.ny;
      else if (jy > a.ny + 2) jy -= a.ny;
      const size_t c = (size_t)(jx - 1) + (size_t)(jy - 1) * a.nxm;
      if (!planet) return a.phi_sg[c];
      const double ex = a.posxy[c] - cx, ey = a.posxy[a.nc + c] - cy;
      return -(a.mplanet / sqrt((ex * ex + ey * ey) + sm2));
    };
    const double d1 = a.g.dx1[ix - 1], d2 = a.g.dx2[ix - 1];
    a.pgrad[(size_t)(2 * w) * a.nc + cell] =
        (pot(ix + 1, iy) - pot(ix - 1, iy)) / ((kHalf * a.g.dx1[ix - 2] + d1) + kHalf * a.g.dx1[ix]);
    a.pgrad[(size_t)(2 * w + 1) * a.nc + cell] =
        ((pot(ix, iy + 1) - pot(ix, iy - 1)) / ((kHalf * d2 + d2) + kHalf * d2)) / r;
  }
}

// The disk's own potential (disk_polar/setup.f90:300-322) from the state at the step's start:
// for every active cell a direct sum over the active cells, ix outer and iy inner as the
// reference sums them, with its terms: the summing cell's own density (q(i,irho)), the source's
// dx(:,1) dx(:,2), radii and the azimuth difference's cosine from the host's table.  A wave's
// lanes hold consecutive rows of one column, so the sources' values are uniform and the table's
// rows are read coalesced.
__global__ __launch_bounds__(256) void k_selfgrav(StepArgs a) {
  if (!a.sc->active) return;
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= a.nx * a.ny) return;
  const int iy = 3 + t % a.ny, ix = 3 + t / a.ny;
  const size_t cell = (size_t)(ix - 1) + (size_t)(iy - 1) * a.nxm;
  const double rr = a.g.rc[ix - 1], rho = fmaxf_(a.U[cell], a.smallr);
  const double base = a.a_smooth_pl * a.a_smooth_pl + rr * rr, two_rr = 2.0 * rr;
  const double* __restrict__ ct = a.sgcos + (iy - 3);
  double acc = 0.0;
  for (int sx = 3; sx <= a.nx + 2; ++sx) {
    const double rp = a.g.rc[sx - 1], w = rho * a.g.dx1[sx - 1] * a.g.dx2[sx - 1];
    const double b2 = base + rp * rp, k2 = two_rr * rp;
    for (int sy = 0; sy < a.ny; ++sy) acc = acc - w / sqrt(b2 - k2 * ct[(size_t)sy * a.ny]);
  }
  a.phi_sg[cell] = acc;
}

// update_force_setup with explicit drag (SI setup.f90:340-365, drag = .false.): per active cell,
// from ctoprim's primitives of the step's start (solve.f90:29), the setup's force of every fluid,
// then for each species in order its drag term subtracted from the gas force and added to its own
// force.  The stopping time is the shearing box's, constant in time, so the reference reading the
// previous step's value (update_force_setup runs before compute_tstop) changes nothing.
__global__ void k_explicit_force(StepArgs a) {
  if (!a.sc->active) return;
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= a.nx * a.ny) return;
  const int ix = 3 + t % a.nx, iy = 3 + t / a.nx;
  const size_t nc = a.nc, cell = (size_t)(ix - 1) + (size_t)(iy - 1) * a.nxm;
  const int nd = a.nd;
  const double rho = a.U[cell], rho_q = fmaxf_(rho, a.smallr);
  const double vx = a.U[nc + cell] / rho, vy = a.U[2 * nc + cell] / rho, vz = a.U[3 * nc + cell] / rho;
  double f1, f2, f3;
  cell_force_setup(a, ix, iy, 0, vx, vz, true, f1, f2, f3);
  for (int k = 0; k < nd; ++k) {
    const double rd = a.U[(5 + k) * nc + cell];
    const double vdx = a.U[(5 + nd + k) * nc + cell] / rd, vdy = a.U[(5 + 2 * nd + k) * nc + cell] / rd,
                 vdz = a.U[(5 + 3 * nd + k) * nc + cell] / rd;
    double g1, g2, g3;
    cell_force_setup(a, ix, iy, 0, vdx, vdz, false, g1, g2, g3);
    const double ts = tstop_of<1, false>(a, k, (int)cell, ix, rho_q, rd, a.cs[cell]);
    const double eps = rd / rho_q;
    f1 = f1 - (eps * (vx - vdx)) / ts;
    f2 = f2 - (eps * (vy - vdy)) / ts;
    f3 = f3 - (eps * (vz - vdz)) / ts;
    const size_t fl = (size_t)(k + 1);
    a.fexp[(3 * fl) * nc + cell] = g1 + (vx - vdx) / ts;
    a.fexp[(3 * fl + 1) * nc + cell] = g2 + (vy - vdy) / ts;
    a.fexp[(3 * fl + 2) * nc + cell] = g3 + (vz - vdz) / ts;
  }
  a.fexp[cell] = f1;
  a.fexp[nc + cell] = f2;
  a.fexp[2 * nc + cell] = f3;
}

static void launch_planet(const StepArgs& a, hipStream_t s) {
  if ((a.force_kind == 3 || a.force_kind == 2) && a.mplanet > 0.0)
    hipLaunchKernelGGL(k_planet, dim3(1), dim3(1), 0, s, a);
  if (a.pgrad && a.mplanet > 0.0)
    hipLaunchKernelGGL(k_planet_grad, dim3((a.nx * a.ny + 255) / 256), dim3(256), 0, s, a);
}

// self-gravity without a planet: the potential of this state and its gradient (with a planet,
// the reference's planet potential replaces the disk's own, so k_planet_grad alone applies)
static void launch_selfgrav(const StepArgs& a, hipStream_t s) {
  if (!a.self_gravity || a.mplanet > 0.0) return;
  hipLaunchKernelGGL(k_selfgrav, dim3((a.nx * a.ny + 255) / 256), dim3(256), 0, s, a);
  hipLaunchKernelGGL(k_planet_grad, dim3((a.nx * a.ny + 255) / 256), dim3(256), 0, s, a);
}

// Accretion through the outer boundary of the polar disk (disk_polar/boundaries.f90:63-94),
// after the ghost fill: the two outer ghost columns and the last active column of every row with
// phi <= phi_mom pi take the inflow state of the accretion rate at the last active cell's cs.
// X is the state written (U at the start of the step); ncol columns from the right edge are
// written, skipping the first col0 (FARGO's own boundary call before its remap,
// fargo_scheme.f90:15, writes only the last active column of B: col0 2, ncol 1).
__global__ void k_bc_accrete(StepArgs a, double* X, int col0, int ncol) {
  if (!a.sc->active) return;
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= ncol * a.ny) return;
  const int ix = a.nxm - col0 - t / a.ny, iy = 3 + t % a.ny;
  const double phi = (((iy - 3) + kHalf) * (2.0 * kPi)) / (double)(float)a.ny;
  if (!(phi <= a.phi_mom * kPi)) return;
  const size_t nc = a.nc;
  const int br = cell_of(a, ix, iy);
  const double vrout = a.cs[cell_of(a, a.nx + 2, iy)];
  const double sigma_out = a.m_acc / vrout / a.phi_mom / kPi / a.box_l;
  const double sq = sqrt(a.mstar / 100.0);
  X[br] = sigma_out;
  X[nc + br] = -(sigma_out * vrout);
  X[2 * nc + br] = sigma_out * a.l_out * sq * 100.0 / a.box_l;
  const int nd = a.nd;
  for (int k = 0; k < nd; ++k) {
    X[(5 + k) * nc + br] = a.dust2gas * sigma_out;
    X[(5 + nd + k) * nc + br] = -(a.dust2gas * sigma_out * vrout);
    X[(5 + 2 * nd + k) * nc + br] = a.dust2gas * sigma_out * a.l_out * sq * 100.0 / a.box_l;
  }
}

static void launch_courant_now(const StepArgs& a, hipStream_t s);

void launch_bc(const StepArgs& a, hipStream_t s) {
  launch_planet(a, s);
  if (a.bc_kind == 5) {
    hipLaunchKernelGGL(k_bc_collapse, dim3(1), dim3(256), 0, s, a);
    // the boundary rewrites the first and last active cells' velocities, so this step's dt is
    // taken here from the filled state (with the enclosed mass of this state), as courant
    // follows apply_boundaries and ctoprim in solve (solve.f90:29-56)
    launch_courant_now(a, s);
    return;
  }
  if (a.bc_kind == 4) {
    hipLaunchKernelGGL(k_bc_zg, dim3(1), dim3(1024), 0, s, a);
    return;
  }
  const int n = (a.ndim == 2 ? 4 * a.nx : 0) + 4 * a.nym;
  hipLaunchKernelGGL(k_bc, dim3((n + 255) / 256), dim3(256), 0, s, a);
  const bool sg = a.self_gravity && !(a.mplanet > 0.0);
  if (a.accretion && a.bc_kind == 2)
    hipLaunchKernelGGL(k_bc_accrete, dim3((3 * a.ny + 255) / 256), dim3(256), 0, s, a, a.U, 0, 3);
  if (sg) launch_selfgrav(a, s);
  if (a.fexp) hipLaunchKernelGGL(k_explicit_force, dim3((a.nx * a.ny + 255) / 256), dim3(256), 0, s, a);
  if ((a.accretion && a.bc_kind == 2) || sg || a.fexp) {
    // the inflow rewrites the last active column, and the disk's potential and the explicit drag
    // are this state's, so this step's dt is taken here, from the filled state, as courant does
    // after apply_boundaries and update_force_setup (solve.f90:29-74); the candidate the previous
    // post pass left is replaced
    launch_courant_now(a, s);
  }
}

template <int NDIM, int GEOM, bool GAS, int SLOPE>
static void hydro_launch1(const StepArgs& a, hipStream_t s) {
  const long items = (long)a.ngroups * (GAS ? a.nstrips_g : a.nstrips) * (GAS ? 1 : a.nd);
  const dim3 grid((unsigned)((items + 3) / 4));
  const bool fast = GAS ? a.solver == 1 : a.solverdust == 0;   // HLL gas / Huang-Bai dust
  if constexpr (GAS && (GEOM == 0 || GEOM == 1)) {
    if (!a.iso) {   // adiabatic gas (cartesian and spherical 1D grids: shark_gpu_create rejects the others)
      if (fast) hipLaunchKernelGGL((k_hydro<NDIM, GEOM, true, SLOPE, 1, true>), grid, dim3(256), 0, s, a);
      else hipLaunchKernelGGL((k_hydro<NDIM, GEOM, true, SLOPE, -1, true>), grid, dim3(256), 0, s, a);
      return;
    }
  }
  if (fast) hipLaunchKernelGGL((k_hydro<NDIM, GEOM, GAS, SLOPE, GAS ? 1 : 0, false>), grid, dim3(256), 0, s, a);
  else hipLaunchKernelGGL((k_hydro<NDIM, GEOM, GAS, SLOPE, -1, false>), grid, dim3(256), 0, s, a);
}

template <int NDIM, int GEOM, bool GAS>
static void hydro_launch_slope(const StepArgs& a, hipStream_t s) {
  switch (a.slope) {
    case 1: hydro_launch1<NDIM, GEOM, GAS, 1>(a, s); break;
    case 2: hydro_launch1<NDIM, GEOM, GAS, 2>(a, s); break;
    default: hydro_launch1<NDIM, GEOM, GAS, 3>(a, s); break;
  }
}

template <bool GAS>
static void hydro_launch_geom(const StepArgs& a, hipStream_t s) {
  if (a.ndim == 1 && a.geom == 1) hydro_launch_slope<1, 1, GAS>(a, s);   // spherical (Collapse_1D)
  else if (a.ndim == 1) hydro_launch_slope<1, 0, GAS>(a, s);
  else if (a.geom == 2 && a.viscosity) hydro_launch_slope<2, 4, GAS>(a, s);   // viscous polar disk
  else if (a.geom == 2) hydro_launch_slope<2, 2, GAS>(a, s);
  else if (a.force_kind == 3 || a.viscosity) hydro_launch_slope<2, 3, GAS>(a, s);   // cartesian disk
  else hydro_launch_slope<2, 0, GAS>(a, s);
}

// gas and dust fluids run as two launches so each gets its own register allocation
void launch_hydro(const StepArgs& a, hipStream_t s) {
  hydro_launch_geom<true>(a, s);
  if (a.nd > 0) hydro_launch_geom<false>(a, s);
}
void launch_hydro_fluid(const StepArgs& a, bool gas, hipStream_t s) {
  if (gas) hydro_launch_geom<true>(a, s);
  else if (a.nd > 0) hydro_launch_geom<false>(a, s);
}

// CAP: the kernels take the species count at run time, with ND as the capacity (numbers of dust
// species without an exact instantiation run in the next larger one)
template <int ND, bool CAP, int PART>
static void post_nd_part(const StepArgs& a, hipStream_t s);
template <int ND, bool CAP>
static void post_nd(const StepArgs& a, hipStream_t s) { post_nd_part<ND, CAP, 0>(a, s); }
template <int ND, bool CAP>
static void post1_nd(const StepArgs& a, hipStream_t s) { post_nd_part<ND, CAP, 1>(a, s); }
template <int ND, bool CAP>
static void post2_nd(const StepArgs& a, hipStream_t s) { post_nd_part<ND, CAP, 2>(a, s); }
// Row-segment widths of the post pass.  Up to 16 species: 16 or 32 cells (32 cells with 8 slots
// take 27 KB of LDS and <= 96 VGPRs at 16 species, so five blocks share a CU); more species (up to
// 261 variables): 8 or 16 cells, which keep the segment's LDS under 64 KB.
template <int ND>
struct PostTW {
  static constexpr int A = ND > 16 ? 8 : 16;
  static constexpr int B = ND > 16 ? 16 : 32;
};
bool post_tw_ok(int nd, int tw) {
  return nd > 16 ? (tw == 8 || tw == 16) : (tw == 16 || tw == 32);
}
int post_tw_default(int nd) { return nd > 16 ? 16 : 32; }
int post_tw_candidates(int nd, int* out) {
  out[0] = nd > 16 ? 8 : 16;
  out[1] = nd > 16 ? 16 : 32;
  return 2;
}
template <int ND, bool CAP, int PART, int TW>
static void post_nd_tw(const StepArgs& a, hipStream_t s);
template <int ND, bool CAP, int PART>
static void post_nd_part(const StepArgs& a, hipStream_t s) {
  if (a.post_tw == PostTW<ND>::A) post_nd_tw<ND, CAP, PART, PostTW<ND>::A>(a, s);
  else post_nd_tw<ND, CAP, PART, PostTW<ND>::B>(a, s);
}
template <int ND, bool CAP, int PART, int TW>
static void post_nd_tw(const StepArgs& a, hipStream_t s) {
  if (a.ndim == 1) {
    dim3 grid((a.nx + TW - 1) / TW);
    hipLaunchKernelGGL((k_post2<ND, false, 1, TW, PART, CAP>), grid, dim3(256), 0, s, a);
  } else {
    dim3 grid((a.nx + TW - 1) / TW, a.ny);
    if (a.fargo && PART != 2) {
      if (a.accretion && a.bc_kind == 2)
        hipLaunchKernelGGL(k_bc_accrete, dim3((a.ny + 255) / 256), dim3(256), 0, s, a, a.B, 2, 1);
      hipLaunchKernelGGL((k_post2<ND, true, 2, TW, PART, CAP>), grid, dim3(256), 0, s, a);
    }
    else hipLaunchKernelGGL((k_post2<ND, false, 2, TW, PART, CAP>), grid, dim3(256), 0, s, a);
  }
}

template <int ND, bool CAP>
static void courant_nd(const StepArgs& a, hipStream_t s) {
  if (a.ndim == 1) hipLaunchKernelGGL((k_courant<ND, 1, CAP>), dim3((a.nx + 255) / 256), dim3(256), 0, s, a);
  else hipLaunchKernelGGL((k_courant<ND, 2, CAP>), dim3((a.nx + 63) / 64, (a.ny + 3) / 4), dim3(256), 0, s, a);
}

// Exact instantiations for the species counts of the reference's test and benchmark runs; any other
// count up to 64 runs in the capacity instantiation of the next size up.
#define SHARK_ND_SWITCH(FN)                                   \
  switch (a.nd) {                                             \
    case 0: FN<0, false>(a, s); break;                        \
    case 1: FN<1, false>(a, s); break;                        \
    case 2: FN<2, false>(a, s); break;                        \
    case 3: FN<3, false>(a, s); break;                        \
    case 4: FN<4, false>(a, s); break;                        \
    case 8: FN<8, false>(a, s); break;                        \
    case 16: FN<16, false>(a, s); break;                      \
    case 64: FN<64, false>(a, s); break;                      \
    default:                                                  \
      if (a.nd <= 8) FN<8, true>(a, s);                       \
      else if (a.nd <= 16) FN<16, true>(a, s);                \
      else if (a.nd <= 32) FN<32, true>(a, s);                \
      else FN<64, true>(a, s);                                \
      break;                                                  \
  }

void launch_post(const StepArgs& a, hipStream_t s) { SHARK_ND_SWITCH(post_nd) }
void launch_post_part(const StepArgs& a, int part, hipStream_t s) {
  if (part == 1) { SHARK_ND_SWITCH(post1_nd) }
  else { SHARK_ND_SWITCH(post2_nd) }
}
static void launch_courant_now(const StepArgs& a, hipStream_t s) {
  if (a.gravity) hipLaunchKernelGGL(k_mtot, dim3(1), dim3(1024), 0, s, a);
  SHARK_ND_SWITCH(courant_nd)
  hipLaunchKernelGGL(k_set_dt, dim3(1), dim3(1), 0, s, a.sc);
}

void launch_courant_init(const StepArgs& a, hipStream_t s) {
  launch_planet(a, s);
  launch_selfgrav(a, s);
  if (a.fexp) hipLaunchKernelGGL(k_explicit_force, dim3((a.nx * a.ny + 255) / 256), dim3(256), 0, s, a);
  launch_courant_now(a, s);
}
void launch_set_dt(const StepArgs& a, hipStream_t s) {
  hipLaunchKernelGGL(k_set_dt, dim3(1), dim3(1), 0, s, a.sc);
}
void launch_finalize(const StepArgs& a, hipStream_t s) {
  hipLaunchKernelGGL(k_finalize, dim3(1), dim3(1), 0, s, a.sc);
}
void launch_arm(const StepArgs& a, double tend, hipStream_t s) {
  hipLaunchKernelGGL(k_arm, dim3(1), dim3(1), 0, s, a.sc, tend);
}
void launch_ctoprim(const StepArgs& a, double* q, hipStream_t s) {
  hipLaunchKernelGGL(k_ctoprim, dim3((a.nc + 255) / 256), dim3(256), 0, s, a, q);
}

bool nd_supported(int nd) { return nd >= 0 && nd <= 64; }

}  // namespace shark
