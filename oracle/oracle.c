/*
 * oracle.c — CPU restatement of SHARK's Godunov step (TEST INFRASTRUCTURE ONLY).
 *
 * This file is the checker for libshark_gpu.so.  Only tests/, __graft_entry__.smoke()
 * and bench.py's cpu_baseline leg may load it; the product never links or calls it.
 *
 * It restates, loop by loop and in the reference's own expression order (Fortran
 * left-to-right evaluation, `x**2` as x*x, no FMA contraction: build with
 * -ffp-contract=off), the routines `solve` calls (src/solver/solve.f90:10-181):
 *   apply_boundaries (3 setups), ctoprim, update_force_setup, distribution_dust(.false.),
 *   compute_tstop, courant, kick, predictor + slope_limit, add_delta_u + Riemann solvers,
 *   source_terms, fargo_scheme, dust_drag, setup_inloop.
 * Layout is the reference's: u(ncells,nvar) column-major, 0-based cell index
 * c = (ix-1) + (iy-1)*nx_max for 1-based (ix,iy).
 *
 * Deliberate, documented departures from the reference (its FARGO is undefined as
 * written, src/solver/fargo_scheme.f90): `one` = 1.0d0 (undeclared there, :57,69), `eps`
 * kept in FP64 (implicit REAL(4) there, :47), slopes taken for every active cell with
 * periodic neighbours (:20-31 leave edge slopes unset), shifts wrapped modulo ny.
 * Every other routine is restated without change, including the reference's own quirks
 * (courant uses the last dust species' |v| in the force constraint, courant.f90:114-116;
 * kick updates E even for isothermal gas, force_kicks.f90:18; disk ghost cells copy flat
 * cell indices first_active/last_active, disk_polar/boundaries.f90:45-46).
 *
 * Parity status of each function is pinned in tests/test_oracle_*.py against the
 * flang-built reference (oracle/_ref) fixtures and closed-form checks.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <stdio.h>
#include "oracle.h"
#include "growth_oracle.h"

#define HALF 0.5
static const double PI = 3.141592653589793238462643383279;  /* phys_const.f90:14 */

struct oracle {
  shark_cfg cfg;
  int nx, ny, nd, ndim, nxm, nym, nc, nvar;
  int fa, la, fay, lay;              /* first/last active (1-based) */
  int irho, ivx, ivy, ivz, iP;       /* 0-based variable planes */
  double *u, *q, *qm, *qp, *flux, *dq, *cs, *force, *force_dust, *tstop, *sdust, *eps_dust;
  double *dx, *surf, *vol, *radii_c, *position, *condinit, *fargo_v, *S_U, *uold, *duprim;
  double *dvol, *mc;                 /* 1D self-gravity: inner half-cell volume, enclosed mass */
  double *phi_sg;                    /* disk_polar planet: its potential per cell */
  unsigned char *active, *active_pred;
  double dt, time;
  int growth_on;                     /* dust_growth after the drag (solve.f90:112), disk env */
  shark_growth_params gp;
  growth_tables gt;
};

static inline int cidx(const oracle* o, int ix, int iy) {
  /* icell (src/commons/allocate_init.f90:284-296); 1D ignores iy */
  return (o->ndim == 1) ? (ix - 1) : (ix - 1) + (iy - 1) * o->nxm;
}
static inline int ixx(const oracle* o, int c) { return (o->ndim == 1) ? c + 1 : c % o->nxm + 1; }
static inline int iyy(const oracle* o, int c) { return c / o->nxm + 1; }
#define U(v, c) o->u[(size_t)(v) * o->nc + (c)]
#define Q(v, c) o->q[(size_t)(v) * o->nc + (c)]
#define QM(v, c, d) o->qm[((size_t)(d) * o->nvar + (v)) * o->nc + (c)]
#define QP(v, c, d) o->qp[((size_t)(d) * o->nvar + (v)) * o->nc + (c)]
#define DQ(v, c, d) o->dq[((size_t)(d) * o->nvar + (v)) * o->nc + (c)]
#define FL(v, c, d) o->flux[((size_t)(d) * o->nvar + (v)) * o->nc + (c)]
#define DX(c, d) o->dx[(size_t)(d) * o->nc + (c)]
#define SURF(c, d) o->surf[(size_t)(d) * o->nc + (c)]
#define POS(c, d) o->position[(size_t)(d) * o->nc + (c)]
#define FORCE(c, k) o->force[(size_t)(k) * o->nc + (c)]
#define FDUST(c, k, d) o->force_dust[((size_t)(d) * 3 + (k)) * o->nc + (c)]
#define TSTOP(c, d) o->tstop[(size_t)(d) * o->nc + (c)]
#define SDUST(c, d) o->sdust[(size_t)(d) * o->nc + (c)]
#define IRHOD(k) (o->iP + 1 + (k))
#define IVDX(k) (o->iP + 1 + o->nd + (k))
#define IVDY(k) (o->iP + 1 + 2 * o->nd + (k))
#define IVDZ(k) (o->iP + 1 + 3 * o->nd + (k))

static double fmax_f(double a, double b) { return a > b ? a : b; }   /* Fortran MAX */
static double fmin_f(double a, double b) { return a < b ? a : b; }   /* Fortran MIN */

oracle* oracle_create(const shark_cfg* cfg) {
  oracle* o = (oracle*)calloc(1, sizeof(oracle));
  o->cfg = *cfg;
  o->nx = cfg->nx; o->ny = cfg->ny; o->nd = cfg->ndust;
  o->ndim = (cfg->ny > 1) ? 2 : 1;
  o->nxm = cfg->nx + 4;
  o->nym = (o->ndim == 2) ? cfg->ny + 4 : 1;
  o->nc = o->nxm * o->nym;
  o->nvar = 5 + 4 * o->nd;                       /* parameters.f90:43 */
  o->fa = 3; o->la = cfg->nx + 2;                /* parameters.f90:29-36 */
  o->fay = (o->ndim == 2) ? 3 : 1; o->lay = (o->ndim == 2) ? cfg->ny + 2 : 1;
  o->irho = 0; o->ivx = 1; o->ivy = 2; o->ivz = 3; o->iP = 4;  /* allocate_init.f90:225-234 */
  size_t nc = o->nc, nv = o->nvar, nd = o->nd > 0 ? o->nd : 1;
  o->u = calloc(nc * nv, 8); o->q = calloc(nc * nv, 8);
  o->qm = calloc(nc * nv * o->ndim, 8); o->qp = calloc(nc * nv * o->ndim, 8);
  o->flux = calloc(nc * nv * o->ndim, 8); o->dq = calloc(nc * nv * o->ndim, 8);
  o->cs = calloc(nc, 8); o->force = calloc(nc * 3, 8); o->force_dust = calloc(nc * 3 * nd, 8);
  o->tstop = calloc(nc * nd, 8); o->sdust = calloc(nc * nd, 8); o->eps_dust = calloc(nc * nd, 8);
  o->dx = calloc(nc * o->ndim, 8); o->surf = calloc(nc * o->ndim, 8); o->vol = calloc(nc, 8);
  o->radii_c = calloc(nc, 8); o->position = calloc(nc * o->ndim, 8);
  o->condinit = calloc(nc * nv, 8); o->fargo_v = calloc(nc, 8); o->S_U = calloc(nc * nv, 8);
  o->uold = calloc(nc * nv, 8); o->duprim = calloc(nc * nv, 8);
  o->dvol = calloc(nc, 8); o->mc = calloc(nc, 8); o->phi_sg = calloc(nc, 8);
  o->active = calloc(nc, 1); o->active_pred = calloc(nc, 1);
  /* get_active_cells (src/commons/allocate_init.f90:339-381) */
  for (int iy = 1; iy <= o->nym; ++iy)
    for (int ix = 1; ix <= o->nxm; ++ix) {
      int c = cidx(o, ix, iy);
      if (o->ndim == 1) {
        o->active[c] = (ix >= o->fa && ix <= o->la);
        o->active_pred[c] = (ix >= 2 && ix <= o->nxm - 1);
      } else {
        o->active[c] = (ix >= o->fa && ix <= o->la && iy >= o->fay && iy <= o->lay);
        o->active_pred[c] = (ix >= 2 && ix <= o->nxm - 1 && iy >= 2 && iy <= o->nym - 1);
      }
    }
  return o;
}

void oracle_destroy(oracle* o) {
  if (!o) return;
  double* ps[] = {o->u, o->q, o->qm, o->qp, o->flux, o->dq, o->cs, o->force, o->force_dust,
                  o->tstop, o->sdust, o->eps_dust, o->dx, o->surf, o->vol, o->radii_c,
                  o->position, o->condinit, o->fargo_v, o->S_U, o->uold, o->duprim, o->dvol, o->mc,
                  o->phi_sg};
  for (size_t i = 0; i < sizeof(ps) / sizeof(ps[0]); ++i) free(ps[i]);
  free(o->active); free(o->active_pred); free(o);
}

int oracle_ncells(const oracle* o) { return o->nc; }
int oracle_nvar(const oracle* o) { return o->nvar; }

void oracle_set_geometry(oracle* o, const double* dx, const double* surf, const double* vol,
                         const double* radii_c, const double* position) {
  size_t nc = o->nc;
  memcpy(o->dx, dx, nc * o->ndim * 8); memcpy(o->surf, surf, nc * o->ndim * 8);
  memcpy(o->vol, vol, nc * 8); memcpy(o->radii_c, radii_c, nc * 8);
  if (position) memcpy(o->position, position, nc * o->ndim * 8);
}

/* dvol of gridinit_sphere1D (grid.f90:343-346): (radii^3 - radii_c^3)/3, the part of each cell
 * between its centre and its outer face (the enclosed mass of mtot subtracts it) */
void oracle_set_dvol(oracle* o, const double* dvol) { memcpy(o->dvol, dvol, (size_t)o->nc * 8); }

static void ctoprim(oracle* o);
static void distribution_dust_refresh(oracle* o);
static void compute_tstop(oracle* o);

void oracle_upload(oracle* o, const double* u, const double* cs, const double* sdust,
                   const double* condinit, const double* fargo_v) {
  size_t nc = o->nc;
  memcpy(o->u, u, nc * o->nvar * 8); memcpy(o->cs, cs, nc * 8);
  if (sdust && o->nd) memcpy(o->sdust, sdust, nc * o->nd * 8);
  if (condinit) memcpy(o->condinit, condinit, nc * o->nvar * 8);
  if (fargo_v) memcpy(o->fargo_v, fargo_v, nc * 8);
  o->time = 0.0;
  if (!o->cfg.drag && o->cfg.force_kick && o->nd > 0) {
    /* explicit drag reads tstop in update_force_setup, which runs before the step's
     * compute_tstop (solve.f90:29-35): the first step reads the stopping times the setup
     * computed at initialisation (SI setup.f90:160, disk_polar :77) */
    ctoprim(o);
    distribution_dust_refresh(o);
    compute_tstop(o);
  }
}

void oracle_download(const oracle* o, double* u, double* q) {
  if (u) memcpy(u, o->u, (size_t)o->nc * o->nvar * 8);
  if (q) memcpy(q, o->q, (size_t)o->nc * o->nvar * 8);
}
void oracle_get_aux(const oracle* o, double* tstop, double* force, double* force_dust,
                    double* eps_dust) {
  size_t nc = o->nc, nd = o->nd;
  if (tstop && nd) memcpy(tstop, o->tstop, nc * nd * 8);
  if (force) memcpy(force, o->force, nc * 3 * 8);
  if (force_dust && nd) memcpy(force_dust, o->force_dust, nc * 3 * nd * 8);
  if (eps_dust && nd) memcpy(eps_dust, o->eps_dust, nc * nd * 8);
}
double oracle_time(const oracle* o) { return o->time; }
double oracle_dt(const oracle* o) { return o->dt; }

/* ---------------------------------------------------------------- boundaries */
/* periodic_boundaries, src/grid/boundary_types.f90:149-186 */
static void bc_periodic(oracle* o) {
  for (int ix = o->fa; ix <= o->la; ++ix)
    for (int iy = 1; iy <= 2; ++iy) {
      int bl = cidx(o, ix, iy), br = cidx(o, ix, o->nym + 1 - iy);
      int al = cidx(o, ix, o->lay - 2 + iy), ar = cidx(o, ix, o->fay + 2 - iy);
      for (int v = 0; v < o->nvar; ++v) { U(v, bl) = U(v, al); U(v, br) = U(v, ar); }
    }
  for (int ix = 1; ix <= 2; ++ix)
    for (int iy = 1; iy <= o->nym; ++iy) {
      int bl = cidx(o, ix, iy), br = cidx(o, o->nxm + 1 - ix, iy);
      int al = cidx(o, o->la - 2 + ix, iy), ar = cidx(o, o->fa + 2 - ix, iy);
      for (int v = 0; v < o->nvar; ++v) { U(v, bl) = U(v, al); U(v, br) = U(v, ar); }
    }
}

/* SI shearing-box boundaries, src/setups/SI/boundaries.f90:10-85 */
static void bc_shear_si(oracle* o) {
  const shark_cfg* p = &o->cfg;
  for (int ix = o->fa; ix <= o->la; ++ix)
    for (int iy = 1; iy <= 2; ++iy) {
      int bl = cidx(o, ix, iy), br = cidx(o, ix, o->nym + 1 - iy);
      int al = cidx(o, ix, o->lay - 2 + iy), ar = cidx(o, ix, o->fay + 2 - iy);
      for (int v = 0; v < o->nvar; ++v) { U(v, bl) = U(v, al); U(v, br) = U(v, ar); }
      /* :32-49 compute mom_new for corner cells but assign nothing */
    }
  for (int ix = 1; ix <= 2; ++ix)
    for (int iy = 1; iy <= o->nym; ++iy) {
      int bl = cidx(o, ix, iy), br = cidx(o, o->nxm + 1 - ix, iy);
      int al = cidx(o, o->la - 2 + ix, iy), ar = cidx(o, o->fa + 2 - ix, iy);
      for (int v = 0; v < o->nvar; ++v) { U(v, bl) = U(v, al); U(v, br) = U(v, ar); }
      U(o->ivz, bl) = U(o->ivz, al) - ((U(o->irho, al) * p->q_shear) * p->omega_shear) * p->box_l;
      U(o->ivz, br) = U(o->ivz, ar) + ((U(o->irho, ar) * p->q_shear) * p->omega_shear) * p->box_l;
      for (int k = 0; k < o->nd; ++k) {
        U(IVDZ(k), bl) = U(IVDZ(k), al) - ((U(IRHOD(k), al) * p->q_shear) * p->omega_shear) * p->box_l;
        U(IVDZ(k), br) = U(IVDZ(k), ar) + ((U(IRHOD(k), ar) * p->q_shear) * p->omega_shear) * p->box_l;
      }
    }
}

/* disk boundaries, src/setups/disk_polar/boundaries.f90:10-94, with the accretion inflow of
 * :63-94 through the two outer ghost columns and the last active column */
static void bc_disk(oracle* o) {
  for (int ix = o->fa; ix <= o->la; ++ix)
    for (int iy = 1; iy <= 2; ++iy) {
      int bl = cidx(o, ix, iy), br = cidx(o, ix, o->nym + 1 - iy);
      int al = cidx(o, ix, o->lay - 2 + iy), ar = cidx(o, ix, o->fay + 2 - iy);
      for (int v = 0; v < o->nvar; ++v) { U(v, bl) = U(v, al); U(v, br) = U(v, ar); }
    }
  for (int ix = 1; ix <= 2; ++ix)
    for (int iy = 1; iy <= o->nym; ++iy) {
      int bl = cidx(o, ix, iy), br = cidx(o, o->nxm + 1 - ix, iy);
      int al, ar;
      if (o->cfg.disk_bc_mode == 0) {
        /* :45-46 use first_active/last_active as flat cell indices */
        al = o->fa - 1; ar = o->la - 1;
      } else {
        al = cidx(o, o->fa, iy); ar = cidx(o, o->la, iy);
      }
      for (int v = 0; v < o->nvar; ++v) { U(v, bl) = U(v, al); U(v, br) = U(v, ar); }
      U(o->ivx, br) = fmin_f(U(o->ivx, br), 0.0);
      U(o->ivx, bl) = fmax_f(U(o->ivx, bl), 0.0);
      for (int k = 0; k < o->nd; ++k) {
        U(IVDX(k), br) = fmin_f(U(IVDX(k), br), 0.0);
        U(IVDX(k), bl) = fmax_f(U(IVDX(k), bl), 0.0);
      }
    }
  if (!o->cfg.accretion) return;
  const shark_cfg* p = &o->cfg;
  for (int ix = 1; ix <= 3; ++ix)
    for (int iy = o->fay; iy <= o->lay; ++iy) {
      int br = cidx(o, o->nxm + 1 - ix, iy);
      double phi = (((iy - o->fay) + HALF) * (2.0 * PI)) / (double)(float)p->ny;
      if (!(phi <= p->phi_mom * PI)) continue;
      double vrout = o->cs[cidx(o, o->la, iy)];
      double sigma_out = p->m_acc / vrout / p->phi_mom / PI / p->box_l;
      double vfac = sigma_out * p->l_out * sqrt(p->mstar / 100.0) * 100.0 / p->box_l;
      U(o->irho, br) = sigma_out;
      U(o->ivx, br) = -(sigma_out * vrout);
      U(o->ivy, br) = vfac;
      for (int k = 0; k < o->nd; ++k) {
        U(IRHOD(k), br) = p->dust2gas * sigma_out;
        U(IVDX(k), br) = -(p->dust2gas * sigma_out * vrout);
        U(IVDY(k), br) = p->dust2gas * sigma_out * p->l_out * sqrt(p->mstar / 100.0) * 100.0 / p->box_l;
      }
    }
}

/* Sod shock tube, src/setups/SOD/boundaries.f90:10-80: the ghost layers across the shock
 * direction hold the initial left/right states (E = P/(gamma-1), at rest), the other direction is
 * periodic; y layers first (2D, active columns), then x layers over every row */
static void sod_state(oracle* o, int c, double rho, double P) {
  U(o->irho, c) = rho;
  U(o->iP, c) = P / (o->cfg.gamma - 1.0);
  U(o->ivx, c) = 0.0;
  U(o->ivy, c) = 0.0;
  U(o->ivz, c) = 0.0;
}

static void bc_sod(oracle* o) {
  const shark_cfg* p = &o->cfg;
  if (o->ndim == 2)
    for (int ix = o->fa; ix <= o->la; ++ix)
      for (int iy = 1; iy <= 2; ++iy) {
        int bl = cidx(o, ix, iy), br = cidx(o, ix, o->nym + 1 - iy);
        int al = cidx(o, ix, o->lay - 2 + iy), ar = cidx(o, ix, o->fay + 2 - iy);
        if (p->shock_dir == 1) {
          sod_state(o, bl, p->rho_l, p->p_l);
          sod_state(o, br, p->rho_r, p->p_r);
        } else {
          for (int v = 0; v < o->nvar; ++v) { U(v, bl) = U(v, al); U(v, br) = U(v, ar); }
        }
      }
  for (int ix = 1; ix <= 2; ++ix)
    for (int iy = 1; iy <= o->nym; ++iy) {
      int bl = cidx(o, ix, iy), br = cidx(o, o->nxm + 1 - ix, iy);
      int al = cidx(o, o->la - 2 + ix, iy), ar = cidx(o, o->fa + 2 - ix, iy);
      if (p->shock_dir == 0) {
        sod_state(o, bl, p->rho_l, p->p_l);
        sod_state(o, br, p->rho_r, p->p_r);
      } else {
        for (int v = 0; v < o->nvar; ++v) { U(v, bl) = U(v, al); U(v, br) = U(v, ar); }
      }
    }
}

/* zero_gradients, src/grid/boundary_types.f90:48-100 (cartesian disk).  In 2D the source cells
 * are first_active_y / last_active_y (:70-71) and first_active / last_active (:88-89) taken as
 * flat cell indices, i.e. cells of the first ghost row; kept as written, in the loop order */
static void bc_zero_gradient(oracle* o) {
  if (o->ndim == 1) {
    for (int ix = 1; ix <= 2; ++ix)
      for (int v = 0; v < o->nvar; ++v) {
        U(v, ix - 1) = U(v, o->fa - 1);
        U(v, o->nxm - ix) = U(v, o->la - 1);
      }
    return;
  }
  for (int ix = o->fa; ix <= o->la; ++ix)
    for (int iy = 1; iy <= 2; ++iy) {
      int bl = cidx(o, ix, iy), br = cidx(o, ix, o->nym + 1 - iy);
      for (int v = 0; v < o->nvar; ++v) { U(v, bl) = U(v, o->fay - 1); U(v, br) = U(v, o->lay - 1); }
    }
  for (int ix = 1; ix <= 2; ++ix)
    for (int iy = 1; iy <= o->nym; ++iy) {
      int bl = cidx(o, ix, iy), br = cidx(o, o->nxm + 1 - ix, iy);
      for (int v = 0; v < o->nvar; ++v) { U(v, bl) = U(v, o->fa - 1); U(v, br) = U(v, o->la - 1); }
    }
}

/* boundary_collapse_1D, src/grid/boundary_types.f90:4-44: zero-gradient ghosts, no flow through
 * the centre (the inner ghosts and the first active cell at rest), outflow-free outer edge (the
 * outer ghosts and the last active cell keep only inward velocities); first_active/last_active
 * are 1D cell indices */
static void bc_collapse(oracle* o) {
  const int al = o->fa - 1, ar = o->la - 1;
  for (int ix = 1; ix <= 2; ++ix) {
    const int bl = ix - 1, br = o->nxm - ix;
    for (int v = 0; v < o->nvar; ++v) { U(v, bl) = U(v, al); U(v, br) = U(v, ar); }
    U(o->ivx, bl) = 0.0;
    U(o->ivx, al) = 0.0;
    U(o->ivx, br) = fmin_f(U(o->ivx, br), 0.0);
    U(o->ivx, ar) = fmin_f(U(o->ivx, ar), 0.0);
    for (int k = 0; k < o->nd; ++k) {
      U(IVDX(k), bl) = 0.0;
      U(IVDX(k), al) = 0.0;
      U(IVDX(k), br) = fmin_f(U(IVDX(k), br), 0.0);
      U(IVDX(k), ar) = fmin_f(U(IVDX(k), ar), 0.0);
    }
  }
}

static void apply_boundaries(oracle* o) {
  switch (o->cfg.bc_kind) {
    case SHARK_BC_COLLAPSE: bc_collapse(o); break;
    case SHARK_BC_ZERO_GRADIENT: bc_zero_gradient(o); break;
    case SHARK_BC_PERIODIC: bc_periodic(o); break;
    case SHARK_BC_SHEAR_SI: bc_shear_si(o); break;
    case SHARK_BC_DISK: bc_disk(o); break;
    case SHARK_BC_SOD: bc_sod(o); break;
  }
}

/* ---------------------------------------------------------------- barotropic EOS, gravity */
#define KB_CGS 1.38e-16   /* phys_const.f90:9 */
#define MH_CGS 1.67e-24   /* phys_const.f90:8 */
/* barotrop, src/setups/Collapse_1D/eos.f90:1-11; its first exponent is the default-real
 * literal 0.8 (single precision, widened) */
double oracle_barotrop(const shark_cfg* p, double nH) {
  const double e08 = (double)0.8f;
  return ((p->t0_cloud * sqrt(1.0 + pow((nH * p->unit_nh) / 1e11, e08))) *
          pow(1.0 + (nH * p->unit_nh) / 1e16, -0.3)) *
         pow(1.0 + (nH * p->unit_nh) / 1e21, 0.56667);
}
/* cs_eos, eos.f90:13-20 */
double oracle_cs_eos(const shark_cfg* p, double T) {
  return sqrt(((p->gamma * KB_CGS) * T) / (p->mu_gas * MH_CGS)) / p->unit_v;
}
static double cs_baro(const shark_cfg* p, double rho) { return oracle_cs_eos(p, oracle_barotrop(p, rho)); }

/* r^2 + l_soft^2 of the gravity terms, and the enclosed mass over it: zero where both vanish
 * (the inner ghost cells, radii_c = 0 and Mc = M_cent = 0, where the reference divides 0 by 0) */
static double grav_acc(const oracle* o, int c) {
  const double r = o->radii_c[c], ls = o->cfg.l_soft;
  const double d = r * r + ls * ls;
  return d != 0.0 ? o->mc[c] / d : 0.0;
}

/* mtot, src/gravity/gravity_utils.f90:11-51 (NY = 1): enclosed mass at each cell centre,
 * Mc(i) = M_cent + 4 pi sum_{j<=i} rho_j vol_j - 4 pi rho_i dvol_i (gas and dust), from the
 * primitive densities; ghost cells other than the inner one next to the grid hold 0 (unset in
 * the reference) */
static double dmint_of(const oracle* o, int c) {
  const double fourpi = 4.0 * PI;
  double dmint = (Q(o->irho, c) * o->dvol[c]) * fourpi;
  for (int k = 0; k < o->nd; ++k) dmint = dmint + (Q(IRHOD(k), c) * o->dvol[c]) * fourpi;
  return dmint;
}

static void mtot(oracle* o) {
  const double fourpi = 4.0 * PI;
  memset(o->mc, 0, (size_t)o->nc * 8);
  o->mc[o->fa - 2] = o->cfg.m_cent;
  for (int c = o->fa - 1; c <= o->la - 1; ++c) {
    double mint = (Q(o->irho, c) * o->vol[c]) * fourpi;
    for (int k = 0; k < o->nd; ++k) mint = mint + (Q(IRHOD(k), c) * o->vol[c]) * fourpi;
    o->mc[c] = o->mc[c - 1] + mint;
  }
  for (int c = o->fa - 1; c <= o->la - 1; ++c) o->mc[c] = o->mc[c] - dmint_of(o, c);
}

/* ---------------------------------------------------------------- ctoprim */
/* src/solver/solve.f90:195-253: isothermal (:224-226, P = rho cs^2 with the setup's cs) or
 * adiabatic (:218-222, P from the total energy, cs = sqrt(gamma P / rho) written back) */
static void ctoprim(oracle* o) {
  const int iso = o->cfg.iso_cs == 1;
  const double gamma = o->cfg.gamma;
  #pragma omp parallel for schedule(static)
  for (int c = 0; c < o->nc; ++c) {
    double rho = U(o->irho, c);
    Q(o->irho, c) = fmax_f(rho, o->cfg.smallr);
    Q(o->ivx, c) = U(o->ivx, c) / rho;
    Q(o->ivy, c) = U(o->ivy, c) / rho;
    Q(o->ivz, c) = U(o->ivz, c) / rho;
    if (iso) {
      Q(o->iP, c) = rho * (o->cs[c] * o->cs[c]);
    } else {
      const double vx = U(o->ivx, c) / rho, vy = U(o->ivy, c) / rho, vz = U(o->ivz, c) / rho;
      const double ekin = HALF * rho * (vx * vx) + HALF * rho * (vy * vy) + HALF * rho * (vz * vz);
      const double P = fmax_f((gamma - 1.0) * (U(o->iP, c) - ekin - 0.0), o->cfg.smallp);
      Q(o->iP, c) = P;
      o->cs[c] = sqrt(gamma * P / Q(o->irho, c));
    }
    if (o->cfg.eos_baro) {   /* non_standard_eos = 1 (:225-226) */
      o->cs[c] = cs_baro(&o->cfg, Q(o->irho, c));
      Q(o->iP, c) = rho * (o->cs[c] * o->cs[c]);
    }
    for (int k = 0; k < o->nd; ++k) {
      double rd = U(IRHOD(k), c);
      Q(IRHOD(k), c) = rd;
      Q(IVDX(k), c) = U(IVDX(k), c) / rd;
      Q(IVDY(k), c) = U(IVDY(k), c) / rd;
      Q(IVDZ(k), c) = U(IVDZ(k), c) / rd;
    }
  }
  if (o->cfg.gravity) mtot(o);   /* :249-251 */
}

/* ---------------------------------------------------------------- setup hooks */
/* x**3 for the integer literal 3, as flang lowers it */
static double cube_i(double x) { return x * (x * x); }

/* cartesian disk: offsets from the box centre and the softened radius (disk/setup.f90:266-272) */
static void cdisk_r(const oracle* o, int c, double* xx, double* yy, double* rr) {
  const shark_cfg* p = &o->cfg;
  *xx = POS(c, 0) - HALF * p->box_l;
  *yy = POS(c, 1) - HALF * p->box_l_y;
  *rr = sqrt((*xx * *xx + *yy * *yy) + p->smooth_r * p->smooth_r);
}

/* disk_polar potential (setup.f90:300-340): the disk's own (self_gravity) and/or the softened
 * potential of a planet on a circular orbit in every active cell, at the step's time; x = radii_c cos(phi), y = radii_c sin(phi) are the
 * grid's position (grid.f90 evaluates the same products).  Then apply_boundaries_phi
 * (boundaries.f90:98-140): y-ghost rows of the active columns are periodic copies, and every
 * x-ghost cell takes phi_sg(first_active) / phi_sg(last_active), the reference using those column
 * numbers as cell numbers: cells (3,1) and (nx+2,1) of the first ghost row, which the y copy set
 * to rows ny+1 of columns 3 and nx+2. */
static void planet_potential(oracle* o) {
  const shark_cfg* p = &o->cfg;
  if (p->self_gravity) {
    /* the disk's own potential (setup.f90:300-322): for every active cell a direct sum over
     * the active cells, ix outer and iy inner, with the reference's terms: the density is the
     * summing cell's own (q(i,irho)), the azimuths are float(iy) * dx(:,2) */
    #pragma omp parallel for schedule(static)
    for (int c = 0; c < o->nc; ++c) {
      if (!o->active[c]) continue;
      const double rr = o->radii_c[c], rho = Q(o->irho, c);
      const double phi_loc = (double)iyy(o, c) * DX(c, 1);
      double acc = 0.0;
      for (int ix = o->fa; ix <= o->la; ++ix)
        for (int iy = o->fay; iy <= o->lay; ++iy) {
          const int s = cidx(o, ix, iy);
          const double rr_p = o->radii_c[s], phi_p = (double)iy * DX(s, 1);
          acc = acc - ((rho * DX(s, 0)) * DX(s, 1)) /
                          sqrt(((p->a_smooth_pl * p->a_smooth_pl + rr * rr) + rr_p * rr_p) -
                               ((2.0 * rr) * rr_p) * cos(phi_p - phi_loc));
        }
      o->phi_sg[c] = acc;
    }
  }
  if (p->mplanet > 0.0) {
    /* the planet's potential replaces the self-gravity one (:323-340 assign, not add) */
    const double theta_p = 2.0 * PI * fmod(o->time / (2.0 * PI / sqrt(p->mstar / cube_i(p->rplanet))), 1.0);
    const double cx = cos(theta_p) * p->rplanet, cy = sin(theta_p) * p->rplanet;
    for (int c = 0; c < o->nc; ++c) {
      if (!o->active[c]) continue;
      const double ex = POS(c, 0) - cx, ey = POS(c, 1) - cy;
      const double rr_p = sqrt((ex * ex + ey * ey) + p->a_smooth_pl * p->a_smooth_pl);
      o->phi_sg[c] = -(p->mplanet / rr_p);
    }
  }
  for (int ix = o->fa; ix <= o->la; ++ix)
    for (int g = 1; g <= 2; ++g) {
      o->phi_sg[cidx(o, ix, g)] = o->phi_sg[cidx(o, ix, o->lay - 2 + g)];
      o->phi_sg[cidx(o, ix, o->nym + 1 - g)] = o->phi_sg[cidx(o, ix, o->fay + 2 - g)];
    }
  for (int g = 1; g <= 2; ++g)
    for (int iy = 1; iy <= o->nym; ++iy) {
      o->phi_sg[cidx(o, g, iy)] = o->phi_sg[o->fa - 1];
      o->phi_sg[cidx(o, o->nxm + 1 - g, iy)] = o->phi_sg[o->la - 1];
    }
}

/* update_force_setup: SI src/setups/SI/setup.f90:335-368; disk :360-403 */
static void update_force(oracle* o) {
  const shark_cfg* p = &o->cfg;
  const int planet = p->force_kind == SHARK_FORCE_DISK && (p->mplanet > 0.0 || p->self_gravity);
  if (planet) planet_potential(o);
  #pragma omp parallel for schedule(static)
  for (int c = 0; c < o->nc; ++c) {
    if (!o->active[c]) continue;
    if (p->force_kind == SHARK_FORCE_SI) {
      double Om = p->omega_shear;
      double a = ((2.0 * p->q_shear) * (Om * Om)) * (POS(c, 0) - HALF * p->box_l);
      FORCE(c, 0) = (a - (2.0 * Om) * Q(o->ivz, c)) - ((2.0 * p->rad0) * Om) * p->eta_stream;
      FORCE(c, 1) = 0.0;
      FORCE(c, 2) = (2.0 * Om) * Q(o->ivx, c);
      for (int k = 0; k < o->nd; ++k) {
        FDUST(c, 0, k) = a - (2.0 * Om) * Q(IVDZ(k), c);
        FDUST(c, 1, k) = 0.0;
        FDUST(c, 2, k) = (2.0 * Om) * Q(IVDX(k), c);
      }
    } else if (p->force_kind == SHARK_FORCE_CDISK) {
      /* disk/setup.f90:274-301: softened star, plus a planet on a circular orbit */
      double xx, yy, rr;
      cdisk_r(o, c, &xx, &yy, &rr);
      double f1 = -(p->mstar * xx / cube_i(rr)), f2 = -(p->mstar * yy / cube_i(rr));
      double g1 = f1, g2 = f2;
      if (p->mplanet > 0.0) {
        double theta_p = 2.0 * PI * fmod(o->time / (2.0 * PI / sqrt(p->mstar / cube_i(p->rplanet))), 1.0);
        double xx_p = cos(theta_p) * p->rplanet, yy_p = sin(theta_p) * p->rplanet;
        double rr_p = sqrt((xx_p * xx_p + yy_p * yy_p) + p->smooth_r * p->smooth_r);
        f1 = f1 - p->mplanet * xx_p / cube_i(rr_p);
        f2 = f2 - p->mplanet * yy_p / cube_i(rr_p);
        g1 = g1 - p->mplanet * xx_p / cube_i(rr_p);
        g2 = g2 - p->mplanet * yy_p / cube_i(rr_p);
      }
      FORCE(c, 0) = f1; FORCE(c, 1) = f2; FORCE(c, 2) = 0.0;
      for (int k = 0; k < o->nd; ++k) {
        FDUST(c, 0, k) = g1; FDUST(c, 1, k) = g2; FDUST(c, 2, k) = 0.0;
      }
    } else if (p->force_kind == SHARK_FORCE_DISK) {
      double r = o->radii_c[c];
      FORCE(c, 0) = -(p->mstar / (r * r));
      FORCE(c, 1) = 0.0; FORCE(c, 2) = 0.0;
      for (int k = 0; k < o->nd; ++k) {
        FDUST(c, 0, k) = -(p->mstar / (r * r));
        FDUST(c, 1, k) = 0.0; FDUST(c, 2, k) = 0.0;
      }
      if (planet) {
        /* centred gradient of the potential (setup.f90:343-356), subtracted from every fluid's
         * force (:369-382) */
        const int ix = ixx(o, c), iy = iyy(o, c);
        const int xm = cidx(o, ix - 1, iy), xp = cidx(o, ix + 1, iy), ym = cidx(o, ix, iy - 1), yp = cidx(o, ix, iy + 1);
        const double g1 = (o->phi_sg[xp] - o->phi_sg[xm]) / ((HALF * DX(xm, 0) + DX(c, 0)) + HALF * DX(xp, 0));
        const double g2 = ((o->phi_sg[yp] - o->phi_sg[ym]) / ((HALF * DX(ym, 1) + DX(c, 1)) + HALF * DX(yp, 1))) / r;
        FORCE(c, 0) = FORCE(c, 0) - g1;
        FORCE(c, 1) = FORCE(c, 1) - g2;
        for (int k = 0; k < o->nd; ++k) {
          FDUST(c, 0, k) = FDUST(c, 0, k) - g1;
          FDUST(c, 1, k) = FDUST(c, 1, k) - g2;
        }
      }
    }
    if (!p->drag && p->force_kind != SHARK_FORCE_NONE) {
      /* explicit drag in the force when the implicit solve is off (SI :355-362, disk :386-395) */
      for (int k = 0; k < o->nd; ++k) {
        double ts = TSTOP(c, k), eps = Q(IRHOD(k), c) / Q(o->irho, c);
        FORCE(c, 0) = FORCE(c, 0) - (eps * (Q(o->ivx, c) - Q(IVDX(k), c))) / ts;
        FORCE(c, 1) = FORCE(c, 1) - (eps * (Q(o->ivy, c) - Q(IVDY(k), c))) / ts;
        FORCE(c, 2) = FORCE(c, 2) - (eps * (Q(o->ivz, c) - Q(IVDZ(k), c))) / ts;
        FDUST(c, 0, k) = FDUST(c, 0, k) + (Q(o->ivx, c) - Q(IVDX(k), c)) / ts;
        FDUST(c, 1, k) = FDUST(c, 1, k) + (Q(o->ivy, c) - Q(IVDY(k), c)) / ts;
        FDUST(c, 2, k) = FDUST(c, 2, k) + (Q(o->ivz, c) - Q(IVDZ(k), c)) / ts;
      }
    }
  }
}

/* distribution_dust(.false.), src/dust/dust_init.f90:76-82 */
static void distribution_dust_refresh(oracle* o) {
  for (int k = 0; k < o->nd; ++k)
    for (int c = 0; c < o->nc; ++c)
      o->eps_dust[(size_t)k * o->nc + c] = Q(IRHOD(k), c) / Q(o->irho, c);
}

/* compute_tstop: DUSTYWAVE setup.f90:212-221, SI :389-397, disk_polar :425-451, uniform Epstein
 * (KH setup.f90:218-236, SOD :162-190) with the cs of this step's ctoprim */
static void compute_tstop(oracle* o) {
  const shark_cfg* p = &o->cfg;
  #pragma omp parallel for schedule(static)
  for (int c = 0; c < o->nc; ++c) {
    if (!o->active[c]) continue;
    for (int k = 0; k < o->nd; ++k) {
      double ts = 0.0;
      if (p->tstop_kind == SHARK_TSTOP_DUSTYWAVE) {
        ts = ((Q(IRHOD(k), c) * Q(o->irho, c)) / p->K[k]) / Q(o->irho, c);
      } else if (p->tstop_kind == SHARK_TSTOP_SI) {
        ts = ((p->rhograin * SDUST(c, k)) / p->rho_init) / ((p->rad0 * p->omega_shear) * p->hoverr);
      } else if (p->tstop_kind == SHARK_TSTOP_EPSTEIN) {
        ts = (sqrt((PI * p->gamma) / 8.0) * p->rhograin) * SDUST(c, k) / (Q(o->irho, c) * o->cs[c]);
      } else if (p->tstop_kind == SHARK_TSTOP_COLLAPSE) {
        /* src/setups/Collapse_1D/setup.f90:310-320 */
        ts = ((sqrt((PI * p->gamma) / 8.0) * p->rhograin) * SDUST(c, k)) /
             (Q(o->irho, c) * cs_baro(p, Q(o->irho, c)));
      } else if (p->tstop_kind == SHARK_TSTOP_EPSTEIN_CDISK) {
        double xx, yy, rr;
        cdisk_r(o, c, &xx, &yy, &rr);
        double H = p->hoverr * rr;
        ts = (sqrt((PI * p->gamma) / 8.0) * p->rhograin) * SDUST(c, k) /
             ((Q(o->irho, c) / sqrt(2.0 * PI * H)) * o->cs[c]);
      } else {
        double H = p->hoverr * o->radii_c[c];
        ts = (sqrt((PI * p->gamma) / 8.0) * p->rhograin) * SDUST(c, k) /
             ((Q(o->irho, c) / (sqrt(2.0 * PI) * H)) * o->cs[c]);
      }
      TSTOP(c, k) = ts;
    }
  }
}

/* ---------------------------------------------------------------- courant */
/* src/solver/courant.f90:11-124 (isothermal, MHD=0, GRAVITY=0) */
static void courant(oracle* o) {
  const shark_cfg* p = &o->cfg;
  double dt = 2e44;
  for (int c = 0; c < o->nc; ++c) {
    if (!o->active[c]) continue;
    double dxx;
    if (o->ndim == 1) dxx = DX(c, 0);
    else if (p->geom == SHARK_GEOM_POLAR) dxx = fmin_f(DX(c, 0), o->radii_c[c] * DX(c, 1));
    else dxx = fmin_f(DX(c, 0), DX(c, 1));
    double vv = (fabs(Q(o->ivx, c)) + fabs(Q(o->ivy, c))) + fabs(Q(o->ivz, c));
    double vmax = o->cs[c] + vv;
    for (int k = 0; k < o->nd; ++k) {
      vv = (fabs(Q(IVDX(k), c)) + fabs(Q(IVDY(k), c))) + fabs(Q(IVDZ(k), c));
      vmax = fmax_f(vmax, 0.0 + vv);
    }
    dt = fmin_f(dt, (p->cfl * dxx) / fabs(vmax));
    if (p->gravity) {   /* :62-63 */
      const double r = o->radii_c[c], ls = p->l_soft;
      dt = fmin_f(dt, (p->cfl * dxx) / sqrt(o->mc[c] / sqrt(r * r + ls * ls)));
    }
    if (p->force_kick) {
      double fm = sqrt((FORCE(c, 0) * FORCE(c, 0) + FORCE(c, 1) * FORCE(c, 1)) + FORCE(c, 2) * FORCE(c, 2));
      for (int k = 0; k < o->nd; ++k)
        fm = fmax_f(fm, sqrt((FDUST(c, 0, k) * FDUST(c, 0, k) + FDUST(c, 1, k) * FDUST(c, 1, k)) +
                             FDUST(c, 2, k) * FDUST(c, 2, k)));
      if (vv != 0.0) {
        double fratio = fmax_f((fm * dxx) / (vv * vv), 1e-3);
        dt = fmin_f(dt, (((p->cfl * dxx) / vv) * (sqrt(1.0 + (2.0 * p->cfl) * fratio) - 1.0)) / fratio);
      }
    }
  }
  o->dt = dt;
}

/* ---------------------------------------------------------------- kick */
/* kick(coeffdt), src/solver/force_kicks.f90:2-41 */
static void kick(oracle* o, double coeffdt) {
  const double dt = o->dt;
  #pragma omp parallel for schedule(static)
  for (int c = 0; c < o->nc; ++c) {
    if (!o->active[c]) continue;
    double rho = U(o->irho, c);
    /* (iso_cs.ne.1 .or. non_standard_eos.ne.1) is true for isothermal gas (:18) */
    double eold = ((HALF * (U(o->ivx, c) * U(o->ivx, c))) / rho + (HALF * (U(o->ivy, c) * U(o->ivy, c))) / rho) +
                  (HALF * (U(o->ivz, c) * U(o->ivz, c))) / rho;
    U(o->ivx, c) = U(o->ivx, c) + ((rho * coeffdt) * dt) * FORCE(c, 0);
    U(o->ivy, c) = U(o->ivy, c) + ((rho * coeffdt) * dt) * FORCE(c, 1);
    U(o->ivz, c) = U(o->ivz, c) + ((rho * coeffdt) * dt) * FORCE(c, 2);
    double enew = ((HALF * (U(o->ivx, c) * U(o->ivx, c))) / rho + (HALF * (U(o->ivy, c) * U(o->ivy, c))) / rho) +
                  (HALF * (U(o->ivz, c) * U(o->ivz, c))) / rho;
    U(o->iP, c) = U(o->iP, c) + (enew - eold);
    for (int k = 0; k < o->nd; ++k) {
      double rd = U(IRHOD(k), c);
      U(IVDX(k), c) = U(IVDX(k), c) + ((rd * coeffdt) * dt) * FDUST(c, 0, k);
      U(IVDY(k), c) = U(IVDY(k), c) + ((rd * coeffdt) * dt) * FDUST(c, 1, k);
      U(IVDZ(k), c) = U(IVDZ(k), c) + ((rd * coeffdt) * dt) * FDUST(c, 2, k);
    }
  }
}

/* ---------------------------------------------------------------- predictor */
/* slope_limit, src/solver/slope_limiters.f90:5-37 */
double oracle_slope_limit(int type, double sl, double sr) {
  double s;
  if (type == 1) {
    s = sl;
    if (fabs(sr) < fabs(sl)) s = sr;
    if (sr * sl < 0.0) s = 0.0;
  } else if (type == 2) {
    s = ((2.0 * sl) * sr) / (sl + sr);
    if (sr * sl <= 0.0) s = 0.0;
  } else {
    double dcen = HALF * (sl + sr);
    double dsgn = copysign(1.0, dcen);
    double slop = fmin_f(1.5 * fabs(sl), 1.5 * fabs(sr));
    double dlim = slop;
    if ((sl * sr) <= 0.0) dlim = 0.0;
    s = dsgn * fmin_f(dlim, fabs(dcen));
  }
  return s;
}

/* predictor, src/solver/godunov.f90:7-629 (isothermal, MHD=0) */
static void predictor(oracle* o) {
  const shark_cfg* p = &o->cfg;
  const double dt = o->dt, gamma = p->gamma;
  const int st = p->slope_type;
  memset(o->dq, 0, (size_t)o->nc * o->nvar * o->ndim * 8);
  memset(o->qp, 0, (size_t)o->nc * o->nvar * o->ndim * 8);
  memset(o->qm, 0, (size_t)o->nc * o->nvar * o->ndim * 8);
  #pragma omp parallel for schedule(static)
  for (int c = 0; c < o->nc; ++c) {
    if (!o->active_pred[c]) continue;
    int ix = ixx(o, c), iy = iyy(o, c);
    double rp = (p->geom == SHARK_GEOM_POLAR) ? o->radii_c[c] : 1.0;
    for (int v = 0; v < o->nvar; ++v) {
      int il = cidx(o, ix - 1, iy), ir = cidx(o, ix + 1, iy);
      DQ(v, c, 0) = oracle_slope_limit(st, (2.0 * (Q(v, c) - Q(v, il))) / (DX(c, 0) + DX(il, 0)),
                                       (2.0 * (Q(v, ir) - Q(v, c))) / (DX(ir, 0) + DX(c, 0)));
      if (o->ndim == 2) {
        il = cidx(o, ix, iy - 1); ir = cidx(o, ix, iy + 1);
        DQ(v, c, 1) = oracle_slope_limit(st, (2.0 * (Q(v, c) - Q(v, il))) / (DX(c, 1) + DX(il, 1)),
                                         (2.0 * (Q(v, ir) - Q(v, c))) / (DX(ir, 1) + DX(c, 1))) / rp;
      }
    }
    double dry = 0, dpy = 0, duy = 0, dvy = 0, dwy = 0;
    double r_rho = Q(o->irho, c), u = Q(o->ivx, c), dux = DQ(o->ivx, c, 0), drx = DQ(o->irho, c, 0);
    double w = Q(o->ivz, c), dwx = DQ(o->ivz, c, 0), v = Q(o->ivy, c), dvx = DQ(o->ivy, c, 0);
    const int iso = p->iso_cs == 1;
    double cs2 = o->cs[c] * o->cs[c];
    double P = Q(o->iP, c), dpx = DQ(o->iP, c, 0);   /* :123-124 */
    if (iso) {                                     /* :125-128 */
      P = Q(o->irho, c) * cs2;
      dpx = DQ(o->irho, c, 0) * cs2;
    }
    if (o->ndim == 2) {
      duy = DQ(o->ivx, c, 1); dry = DQ(o->irho, c, 1); dvy = DQ(o->ivy, c, 1);
      dwy = DQ(o->ivz, c, 1); dpy = DQ(o->iP, c, 1);
      if (iso) dpy = DQ(o->irho, c, 1) * cs2;
    }
    if (p->force_kick) {
      u = u + (FORCE(c, 0) * HALF) * dt;
      v = v + (FORCE(c, 1) * HALF) * dt;
      w = w + (FORCE(c, 2) * HALF) * dt;
    }
    double sr0, su0, sv0, sw0, sp0;
    if (p->geom == SHARK_GEOM_POLAR) {          /* :214-220 (overwrites the :196 block) */
      sr0 = ((-(u * drx)) - v * dry - (dux + dvy) * r_rho) - (r_rho * u) / rp;
      sp0 = ((-(u * dpx)) - v * dpy - (dux + dvy) * gamma * P) - gamma * P * u / rp;
      su0 = ((-(u * dux)) - v * duy - dpx / r_rho) + (v * v) / rp;
      sv0 = ((-(u * dvx)) - v * dvy - dpy / r_rho) - (u * v) / rp;
      sw0 = (-(u * dwx)) - v * dwy;
    } else {                                     /* :196-202; GEOM 1 :205-211 */
      sr0 = (-(u * drx)) - v * dry - (dux + dvy) * r_rho;
      if (p->geom == SHARK_GEOM_SPHERE && o->radii_c[c] != 0.0) sr0 = sr0 - (r_rho * u) / o->radii_c[c];
      sp0 = (-(u * dpx)) - v * dpy - (dux + dvy) * gamma * P;
      su0 = (-(u * dux)) - v * duy - dpx / r_rho;
      sv0 = (-(u * dvx)) - v * dvy - dpy / r_rho;
      sw0 = (-(u * dwx)) - v * dwy;
    }
    if (p->gravity) su0 = su0 + (-grav_acc(o, c));   /* :243-246 */
    double dxl = DX(c, 0);                       /* :258-268 */
    QM(o->irho, c, 0) = r_rho + HALF * dt * sr0 + HALF * drx * dxl;
    QM(o->ivx, c, 0) = u + HALF * dt * su0 + HALF * dux * dxl;
    QM(o->ivy, c, 0) = v + HALF * dt * sv0 + HALF * dvx * dxl;
    QM(o->ivz, c, 0) = w + HALF * dt * sw0 + HALF * dwx * dxl;
    QP(o->irho, c, 0) = r_rho + HALF * dt * sr0 - HALF * drx * dxl;
    QP(o->ivx, c, 0) = u + HALF * dt * su0 - HALF * dux * dxl;
    QP(o->ivy, c, 0) = v + HALF * dt * sv0 - HALF * dvx * dxl;
    QP(o->ivz, c, 0) = w + HALF * dt * sw0 - HALF * dwx * dxl;
    if (!iso) {                                  /* :261, :266 */
      QM(o->iP, c, 0) = fmax_f(P + HALF * dt * sp0 + HALF * dpx * dxl, p->smallp);
      QP(o->iP, c, 0) = fmax_f(P + HALF * dt * sp0 - HALF * dpx * dxl, p->smallp);
    }
    if (o->ndim == 2) {                          /* :271-282 */
      double dyl = rp * DX(c, 1);
      QM(o->irho, c, 1) = r_rho + HALF * dt * sr0 + HALF * dry * dyl;
      QM(o->ivx, c, 1) = u + HALF * dt * su0 + HALF * duy * dyl;
      QM(o->ivy, c, 1) = v + HALF * dt * sv0 + HALF * dvy * dyl;
      QM(o->ivz, c, 1) = w + HALF * dt * sw0 + HALF * dwy * dyl;
      QP(o->irho, c, 1) = r_rho + HALF * dt * sr0 - HALF * dry * dyl;
      QP(o->ivx, c, 1) = u + HALF * dt * su0 - HALF * duy * dyl;
      QP(o->ivy, c, 1) = v + HALF * dt * sv0 - HALF * dvy * dyl;
      QP(o->ivz, c, 1) = w + HALF * dt * sw0 - HALF * dwy * dyl;
      if (!iso) {                                /* :275, :280 */
        QM(o->iP, c, 1) = fmax_f(P + HALF * dt * sp0 + HALF * dpy * dyl, p->smallp);
        QP(o->iP, c, 1) = fmax_f(P + HALF * dt * sp0 - HALF * dpy * dyl, p->smallp);
      }
    }
    /* dust, :359-510 */
    for (int k = 0; k < o->nd; ++k) {
      int ir_ = IRHOD(k), iu = IVDX(k), iv = IVDY(k), iw = IVDZ(k);
      r_rho = Q(ir_, c); u = Q(iu, c); dux = DQ(iu, c, 0); drx = DQ(ir_, c, 0);
      v = Q(iv, c); dvx = DQ(iv, c, 0); w = Q(iw, c); dwx = DQ(iw, c, 0);
      if (o->ndim == 2) { dry = DQ(ir_, c, 1); duy = DQ(iu, c, 1); dvy = DQ(iv, c, 1); dwy = DQ(iw, c, 1); }
      if (p->force_kick) {
        u = u + (FDUST(c, 0, k) * HALF) * dt;
        v = v + (FDUST(c, 1, k) * HALF) * dt;
        w = w + (FDUST(c, 2, k) * HALF) * dt;
      }
      sr0 = (-(u * drx)) - v * dry - (dux + dvy) * r_rho;
      su0 = (-(u * dux)) - v * duy;
      sv0 = (-(u * dvx)) - v * dvy;
      sw0 = (-(u * dwx)) - v * dwy;
      if (p->geom == SHARK_GEOM_POLAR) {
        sr0 = sr0 - (r_rho * u) / rp;
        su0 = su0 + (v * v) / rp;
        sv0 = sv0 - (u * v) / rp;
      }
      if (p->geom == SHARK_GEOM_SPHERE && o->radii_c[c] != 0.0) sr0 = sr0 - (r_rho * u) / o->radii_c[c];  /* :386-389 */
      if (p->gravity) su0 = su0 + (-grav_acc(o, c));   /* :442-445 */
      QM(ir_, c, 0) = r_rho + HALF * dt * sr0 + HALF * drx * dxl;
      QM(iu, c, 0) = u + HALF * dt * su0 + HALF * dux * dxl;
      QM(iv, c, 0) = v + HALF * dt * sv0 + HALF * dvx * dxl;
      QM(iw, c, 0) = w + HALF * dt * sw0 + HALF * dwx * dxl;
      QP(ir_, c, 0) = r_rho + HALF * dt * sr0 - HALF * drx * dxl;
      QP(iu, c, 0) = u + HALF * dt * su0 - HALF * dux * dxl;
      QP(iv, c, 0) = v + HALF * dt * sv0 - HALF * dvx * dxl;
      QP(iw, c, 0) = w + HALF * dt * sw0 - HALF * dwx * dxl;
      if (o->ndim == 2) {
        double dyl = rp * DX(c, 1);
        QM(ir_, c, 1) = r_rho + HALF * dt * sr0 + HALF * dry * dyl;
        QM(iu, c, 1) = u + HALF * dt * su0 + HALF * duy * dyl;
        QM(iv, c, 1) = v + HALF * dt * sv0 + HALF * dvy * dyl;
        QM(iw, c, 1) = w + HALF * dt * sw0 + HALF * dwy * dyl;
        QP(ir_, c, 1) = r_rho + HALF * dt * sr0 - HALF * dry * dyl;
        QP(iu, c, 1) = u + HALF * dt * su0 - HALF * duy * dyl;
        QP(iv, c, 1) = v + HALF * dt * sv0 - HALF * dvy * dyl;
        QP(iw, c, 1) = w + HALF * dt * sw0 - HALF * dwy * dyl;
      }
    }
    /* isothermal face pressure, :606-611 */
    if (iso)
      for (int d = 0; d < o->ndim; ++d) {
        QP(o->iP, c, d) = cs2 * QP(o->irho, c, d);
        QM(o->iP, c, d) = cs2 * QM(o->irho, c, d);
      }
    if (p->eos_baro)                             /* :612-617 */
      for (int d = 0; d < o->ndim; ++d) {
        const double cp = cs_baro(p, QP(o->irho, c, d)), cm = cs_baro(p, QM(o->irho, c, d));
        QP(o->iP, c, d) = (cp * cp) * QP(o->irho, c, d);
        QM(o->iP, c, d) = (cm * cm) * QM(o->irho, c, d);
      }
  }
}

/* ---------------------------------------------------------------- Riemann solvers */
/* gas: solver_llf :4-155, solver_hll :157-306, solver_hllc :308-500 of src/solver/solvers.f90.
 * ql/qr: face states (nvar), vn/vt: normal/transverse velocity index (allocate_init.f90:249-252) */
void oracle_gas_flux(int solver, double gamma, const double* ql, const double* qr, double* flx,
                     double csl, double csr, int vn, int vt) {
  const int irho = 0, ivz = 3, iP = 4;
  double rho_r = qr[irho], rho_l = ql[irho], u_r = qr[vn], u_l = ql[vn], v_r = qr[vt], v_l = ql[vt];
  double w_r = qr[ivz], w_l = ql[ivz], P_r = qr[iP], P_l = ql[iP];
  double mu_r = rho_r * u_r, mu_l = rho_l * u_l;
  double mv_r = rho_r * v_r, mv_l = rho_l * v_l, mw_r = rho_r * w_r, mw_l = rho_l * w_l;
  if (solver == SHARK_SOLVER_HLLC) {
    double E_r = P_r / (gamma - 1.0) + HALF * rho_r * (u_r * u_r) + HALF * rho_r * (v_r * v_r) + HALF * rho_r * (w_r * w_r);
    double E_l = P_l / (gamma - 1.0) + HALF * rho_l * (u_l * u_l) + HALF * rho_l * (v_l * v_l) + HALF * rho_l * (w_l * w_l);
    double S_l = fmin_f(fmin_f(u_l, u_r) - fmax_f(csl, csr), 0.0);
    double S_r = fmax_f(fmax_f(u_l, u_r) + fmax_f(csl, csr), 0.0);
    double hl = rho_l * (u_l - S_l), hr = rho_r * (S_r - u_r);
    double ustar = (hr * u_r + hl * u_l + (P_l - P_r)) / (hr + hl);
    double Pstar = (hr * P_l + hl * P_r + hl * hr * (u_l - u_r)) / (hr + hl);
    double rsl = rho_l * (S_l - u_l) / (S_l - ustar);
    double esl = ((S_l - u_l) * E_l - P_l * u_l + Pstar * ustar) / (S_l - ustar);
    double rsr = rho_r * (S_r - u_r) / (S_r - ustar);
    double esr = ((S_r - u_r) * E_r - P_r * u_r + Pstar * ustar) / (S_r - ustar);
    double r_o, u_o, P_o, e_o;
    if (S_l > 0.0) { r_o = rho_l; u_o = u_l; P_o = P_l; e_o = E_l; }
    else if (ustar > 0.0) { r_o = rsl; u_o = ustar; P_o = Pstar; e_o = esl; }
    else if (S_r > 0.0) { r_o = rsr; u_o = ustar; P_o = Pstar; e_o = esr; }
    else { r_o = rho_r; u_o = u_r; P_o = P_r; e_o = E_r; }
    flx[irho] = r_o * u_o;
    flx[vn] = r_o * u_o * u_o + P_o;
    flx[iP] = (e_o + P_o) * u_o;
    if (ustar > 0.0) { flx[vt] = r_o * u_o * v_l; flx[ivz] = r_o * u_o * w_l; }
    else { flx[vt] = r_o * u_o * v_r; flx[ivz] = r_o * u_o * w_r; }
    return;
  }
  double E_r = P_r / (gamma - 1.0) + HALF * rho_r * (u_r * u_r);
  double E_l = P_l / (gamma - 1.0) + HALF * rho_l * (u_l * u_l);
  E_r = E_r + HALF * rho_r * (v_r * v_r);
  E_l = E_l + HALF * rho_l * (v_l * v_l);
  E_r = E_r + HALF * rho_r * (w_r * w_r);
  E_l = E_l + HALF * rho_l * (w_l * w_l);
  double fr_r = rho_r * u_r, fr_l = rho_l * u_l;
  double fu_r = rho_r * (u_r * u_r) + P_r, fu_l = rho_l * (u_l * u_l) + P_l;
  double fv_r = rho_r * u_r * v_r, fv_l = rho_l * u_l * v_l;
  double fw_r = rho_r * u_r * w_r, fw_l = rho_l * u_l * w_l;
  double fP_r = (E_r + P_r) * u_r, fP_l = (E_l + P_l) * u_l;
  if (solver == SHARK_SOLVER_LLF) {
    double lam = fmax_f(fabs(u_l) + csl, fabs(u_r) + csr);
    flx[irho] = HALF * (fr_l + fr_r) - HALF * lam * (rho_r - rho_l);
    flx[vn] = HALF * (fu_l + fu_r) - HALF * lam * (mu_r - mu_l);
    flx[vt] = HALF * (fv_l + fv_r) - HALF * lam * (mv_r - mv_l);
    flx[ivz] = HALF * (fw_l + fw_r) - HALF * lam * (mw_r - mw_l);
    flx[iP] = HALF * (fP_l + fP_r) - HALF * lam * (E_r - E_l);
    return;
  }
  double S_l = fmin_f(fmin_f(u_l, u_r) - fmax_f(csl, csr), 0.0);
  double S_r = fmax_f(fmax_f(u_l, u_r) + fmax_f(csl, csr), 0.0);
  flx[irho] = (S_r * fr_l - S_l * fr_r + S_r * S_l * (rho_r - rho_l)) / (S_r - S_l);
  flx[vn] = (S_r * fu_l - S_l * fu_r + S_r * S_l * (mu_r - mu_l)) / (S_r - S_l);
  flx[vt] = (S_r * fv_l - S_l * fv_r + S_r * S_l * (mv_r - mv_l)) / (S_r - S_l);
  flx[ivz] = (S_r * fw_l - S_l * fw_r + S_r * S_l * (mw_r - mw_l)) / (S_r - S_l);
  flx[iP] = (S_r * fP_l - S_l * fP_r + S_r * S_l * (E_r - E_l)) / (S_r - S_l);
}

/* dust: solver_dust_Huang_Bai :506-698, solver_dust_llf :840-1001, solver_dust_hll :1010-1194,
 * solver_hllc_dust :704-836.  One species; returns (rho, mom_n, mom_t, mom_w) fluxes. */
void oracle_dust_flux(int solver, double gamma, const double ql[4], const double qr[4],
                      double f[4], double csl, double csr) {
  double rho_r = qr[0], rho_l = ql[0], u_r = qr[1], u_l = ql[1], v_r = qr[2], v_l = ql[2];
  double w_r = qr[3], w_l = ql[3];
  double mu_r = rho_r * u_r, mu_l = rho_l * u_l, mv_r = rho_r * v_r, mv_l = rho_l * v_l;
  double mw_r = rho_r * w_r, mw_l = rho_l * w_l;
  double fr_r = rho_r * u_r, fr_l = rho_l * u_l;
  double fu_r = rho_r * (u_r * u_r), fu_l = rho_l * (u_l * u_l);
  double fv_r = rho_r * u_r * v_r, fv_l = rho_l * u_l * v_l;
  double fw_r = rho_r * u_r * w_r, fw_l = rho_l * u_l * w_l;
  f[0] = f[1] = f[2] = f[3] = 0.0;
  if (solver == SHARK_DUST_HUANG_BAI) {
    if (u_r > 0.0 && u_l > 0.0) { f[0] = fr_l; f[1] = fu_l; f[2] = fv_l; f[3] = fw_l; }
    else if (u_l < 0.0 && u_r < 0.0) { f[0] = fr_r; f[1] = fu_r; f[2] = fv_r; f[3] = fw_r; }
    else if (u_l < 0.0 && u_r > 0.0) { f[0] = f[1] = f[2] = f[3] = 0.0; }
    else if (u_l > 0.0 && u_r < 0.0) { f[0] = fr_l + fr_r; f[1] = fu_l + fu_r; f[2] = fv_l + fv_r; f[3] = fw_l + fw_r; }
  } else if (solver == SHARK_DUST_LLF) {
    double lam = fmax_f(fabs(u_l), fabs(u_r));
    f[0] = HALF * (fr_l + fr_r) - HALF * lam * (rho_r - rho_l);
    f[1] = HALF * (fu_l + fu_r) - HALF * lam * (mu_r - mu_l);
    f[2] = HALF * (fv_l + fv_r) - HALF * lam * (mv_r - mv_l);
    f[3] = HALF * (fw_l + fw_r) - HALF * lam * (mw_r - mw_l);
  } else if (solver == SHARK_DUST_HLL) {
    double S_r = fmax_f(fmax_f(u_l, u_r), 0.0), S_l = fmin_f(fmin_f(u_l, u_r), 0.0);
    if (u_l != u_r) {
      f[0] = (S_r * fr_l - S_l * fr_r + S_r * S_l * (rho_r - rho_l)) / (S_r - S_l);
      f[1] = (S_r * fu_l - S_l * fu_r + S_r * S_l * (mu_r - mu_l)) / (S_r - S_l);
      f[2] = (S_r * fv_l - S_l * fv_r + S_r * S_l * (mv_r - mv_l)) / (S_r - S_l);
      f[3] = (S_r * fw_l - S_l * fw_r + S_r * S_l * (mw_r - mw_l)) / (S_r - S_l);
    }
    if (u_l == u_r) {
      double lam = fmax_f(fabs(u_l), fabs(u_r));
      f[0] = HALF * (fr_l + fr_r) - HALF * lam * (rho_r - rho_l);
      f[1] = HALF * (fu_l + fu_r) - HALF * lam * (mu_r - mu_l);
      f[2] = HALF * (fv_l + fv_r) - HALF * lam * (mv_r - mv_l);
      f[3] = HALF * (fw_l + fw_r) - HALF * lam * (mw_r - mw_l);
    }
  } else { /* HLLC with P=0 */
    double P_r = 0.0, P_l = 0.0;
    double E_r = P_r / (gamma - 1.0) + HALF * rho_r * (u_r * u_r) + HALF * rho_r * (v_r * v_r) + HALF * rho_r * (w_r * w_r);
    double E_l = P_l / (gamma - 1.0) + HALF * rho_l * (u_l * u_l) + HALF * rho_l * (v_l * v_l) + HALF * rho_l * (w_l * w_l);
    double S_l = fmin_f(fmin_f(u_l, u_r) - fmax_f(csl, csr), 0.0);
    double S_r = fmax_f(fmax_f(u_l, u_r) + fmax_f(csl, csr), 0.0);
    double hl = rho_l * (u_l - S_l), hr = rho_r * (S_r - u_r);
    double ustar = (hr * u_r + hl * u_l + (P_l - P_r)) / (hr + hl);
    double Pstar = (hr * P_l + hl * P_r + hl * hr * (u_l - u_r)) / (hr + hl);
    double rsl = rho_l * (S_l - u_l) / (S_l - ustar);
    double rsr = rho_r * (S_r - u_r) / (S_r - ustar);
    (void)E_r; (void)E_l;
    double r_o, u_o, P_o;
    if (S_l > 0.0) { r_o = rho_l; u_o = u_l; P_o = P_l; }
    else if (ustar > 0.0) { r_o = rsl; u_o = ustar; P_o = Pstar; }
    else if (S_r > 0.0) { r_o = rsr; u_o = ustar; P_o = Pstar; }
    else { r_o = rho_r; u_o = u_r; P_o = P_r; }
    f[0] = r_o * u_o;
    f[1] = r_o * u_o * u_o + P_o;
    if (ustar > 0.0) { f[2] = r_o * u_o * v_l; f[3] = r_o * u_o * w_l; }
    else { f[2] = r_o * u_o * v_r; f[3] = r_o * u_o * w_r; }
  }
}

/* add_delta_u, src/solver/godunov.f90:633-760 */
static void add_delta_u(oracle* o) {
  const shark_cfg* p = &o->cfg;
  const double dt = o->dt;
  const int nv = o->nvar;
  memset(o->flux, 0, (size_t)o->nc * nv * o->ndim * 8);
  #pragma omp parallel for schedule(static)
  for (int c = 0; c < o->nc; ++c) {
    if (!o->active_pred[c]) continue;
    int ix = ixx(o, c), iy = iyy(o, c);
    double ql[5 + 4 * SHARK_MAX_DUST], qr[5 + 4 * SHARK_MAX_DUST], flx[5 + 4 * SHARK_MAX_DUST];
    for (int d = 0; d < o->ndim; ++d) {
      int il = (d == 0) ? cidx(o, ix - 1, iy) : cidx(o, ix, iy - 1);
      for (int v = 0; v < nv; ++v) { ql[v] = QM(v, il, d); qr[v] = QP(v, c, d); flx[v] = 0.0; }
      double csl = o->cs[il], csr = o->cs[c];
      if (p->iso_cs != 1 && p->eos_baro) {       /* :675-677, :702-706 */
        csl = cs_baro(p, ql[o->irho]);
        csr = cs_baro(p, qr[o->irho]);
      } else if (p->iso_cs != 1) {               /* :679-680, :709-710 */
        csl = sqrt(p->gamma * ql[o->iP] / ql[o->irho]);
        csr = sqrt(p->gamma * qr[o->iP] / qr[o->irho]);
      }
      int vn = (d == 0) ? o->ivx : o->ivy, vt = (d == 0) ? o->ivy : o->ivx;
      oracle_gas_flux(p->solver, p->gamma, ql, qr, flx, csl, csr, vn, vt);
      for (int k = 0; k < o->nd; ++k) {
        int dn = (d == 0) ? IVDX(k) : IVDY(k), dt_ = (d == 0) ? IVDY(k) : IVDX(k);
        double l4[4] = {ql[IRHOD(k)], ql[dn], ql[dt_], ql[IVDZ(k)]};
        double r4[4] = {qr[IRHOD(k)], qr[dn], qr[dt_], qr[IVDZ(k)]};
        double f4[4];
        oracle_dust_flux(p->solverdust, p->gamma, l4, r4, f4, csl, csr);
        flx[IRHOD(k)] = f4[0]; flx[dn] = f4[1]; flx[dt_] = f4[2]; flx[IVDZ(k)] = f4[3];
      }
      for (int v = 0; v < nv; ++v) FL(v, c, d) = flx[v];
    }
  }
  #pragma omp parallel for schedule(static)
  for (int c = 0; c < o->nc; ++c) {
    if (!o->active[c]) continue;
    int ix = ixx(o, c), iy = iyy(o, c);
    int il = cidx(o, ix + 1, iy);
    if (o->ndim == 1) {
      for (int v = 0; v < nv; ++v)
        U(v, c) = U(v, c) + (FL(v, c, 0) * SURF(c, 0) - FL(v, il, 0) * SURF(il, 0)) / o->vol[c] * dt;
    } else {
      int ily = cidx(o, ix, iy + 1);
      for (int v = 0; v < nv; ++v)
        U(v, c) = U(v, c) + (FL(v, c, 0) * SURF(c, 0) - FL(v, il, 0) * SURF(il, 0)) / o->vol[c] * dt +
                  (FL(v, c, 1) * SURF(c, 1) - FL(v, ily, 1) * SURF(ily, 1)) / o->vol[c] * dt;
    }
  }
}

/* ---------------------------------------------------------------- sources */
/* Source_terms, src/solver/source_terms.f90:1-224 (GEOM 0/2, no viscosity): polar terms and, for
 * adiabatic gas with dust back-reaction, the drag heating of :88-95 */
static void source_terms(oracle* o) {
  const shark_cfg* p = &o->cfg;
  const double dt = o->dt;
  const size_t n = (size_t)o->nc * o->nvar;
  memset(o->S_U, 0, n * 8);
  if (p->geom == SHARK_GEOM_POLAR) {
    #pragma omp parallel for schedule(static)
    for (int c = 0; c < o->nc; ++c) {
      if (!o->active[c]) continue;
      double r = o->radii_c[c];
      double* S = o->S_U;
      size_t nc = o->nc;
      S[o->ivx * nc + c] = S[o->ivx * nc + c] + dt * (Q(o->irho, c) * (Q(o->ivy, c) * Q(o->ivy, c)) + Q(o->iP, c)) / r;
      S[o->ivy * nc + c] = S[o->ivy * nc + c] - dt * (Q(o->irho, c) * Q(o->ivy, c) * Q(o->ivx, c) / r);
      for (int k = 0; k < o->nd; ++k) {
        S[IVDX(k) * nc + c] = S[IVDX(k) * nc + c] + dt * (Q(IRHOD(k), c) * (Q(IVDY(k), c) * Q(IVDY(k), c)) / r);
        S[IVDY(k) * nc + c] = S[IVDY(k) * nc + c] - dt * (Q(IRHOD(k), c) * Q(IVDY(k), c) * Q(IVDX(k), c) / r);
      }
    }
  }
  if (p->geom == SHARK_GEOM_SPHERE || p->gravity) {
    /* GEOM 1 :33-35 (the pressure term of the spherical momentum equation) and gravity
     * :79-85, in the reference's order within the active-cell loop */
    #pragma omp parallel for schedule(static)
    for (int c = 0; c < o->nc; ++c) {
      if (!o->active[c]) continue;
      double* S = o->S_U;
      size_t nc = o->nc;
      const double r = o->radii_c[c];
      if (p->geom == SHARK_GEOM_SPHERE)
        S[o->ivx * nc + c] = S[o->ivx * nc + c] - dt * (-(((2.0 * Q(o->irho, c)) * (o->cs[c] * o->cs[c])) / r));
      if (p->gravity) {
        const double d = r * r + p->l_soft * p->l_soft;
        S[o->ivx * nc + c] = S[o->ivx * nc + c] - dt * ((Q(o->irho, c) * o->mc[c]) / d);
        for (int k = 0; k < o->nd; ++k)
          S[IVDX(k) * nc + c] = S[IVDX(k) * nc + c] - dt * ((Q(IRHOD(k), c) * o->mc[c]) / d);
      }
    }
  }
  if (o->nd > 0 && p->dust_back_reaction && p->iso_cs < 1) {
    #pragma omp parallel for schedule(static)
    for (int c = 0; c < o->nc; ++c) {
      if (!o->active[c]) continue;
      double* S = o->S_U + (size_t)o->iP * o->nc + c;
      const double vx = Q(o->ivx, c), vy = Q(o->ivy, c), vz = Q(o->ivz, c);
      for (int k = 0; k < o->nd; ++k) {
        const double ts = TSTOP(c, k);
        *S = ((*S + (dt * (vx * (vx - Q(IVDX(k), c)))) / ts) + (dt * (vy * (vy - Q(IVDY(k), c)))) / ts) +
             (dt * (vz * (vz - Q(IVDZ(k), c)))) / ts;
      }
    }
  }
  if (p->viscosity && p->geom == SHARK_GEOM_CARTESIAN && o->ndim == 2) {
    /* :99-110, GEOM 0: eta_visc = alpha_visc H cs of the cartesian disk (disk/setup.f90:60) */
    #pragma omp parallel for schedule(static)
    for (int c = 0; c < o->nc; ++c) {
      if (!o->active[c]) continue;
      int ix = ixx(o, c), iy = iyy(o, c);
      int xp = cidx(o, ix + 1, iy), xm = cidx(o, ix - 1, iy), yp = cidx(o, ix, iy + 1), ym = cidx(o, ix, iy - 1);
      double xx, yy, rr;
      cdisk_r(o, c, &xx, &yy, &rr);
      double eta = (p->alpha_visc * (p->hoverr * rr)) * o->cs[c];
      double d1 = DX(c, 0), d2 = DX(c, 1);
      double* S = o->S_U;
      size_t nc = o->nc;
      double lxu = ((Q(o->ivx, xp) + Q(o->ivx, xm)) - 2.0 * Q(o->ivx, c)) / (d1 * d1);
      S[o->ivx * nc + c] = S[o->ivx * nc + c] + Q(o->irho, c) * dt * eta * lxu;
      double lyu = ((Q(o->ivx, yp) + Q(o->ivx, ym)) - 2.0 * Q(o->ivx, c)) / (d2 * d2);
      S[o->ivx * nc + c] = S[o->ivx * nc + c] + Q(o->irho, c) * dt * eta * lyu;
      double lxv = ((Q(o->ivy, xp) + Q(o->ivy, xm)) - 2.0 * Q(o->ivy, c)) / (d1 * d1);
      double lyv = ((Q(o->ivy, yp) + Q(o->ivy, ym)) - 2.0 * Q(o->ivy, c)) / (d2 * d2);
      S[o->ivy * nc + c] = S[o->ivy * nc + c] + Q(o->irho, c) * dt * eta * (lxv + lyv);
    }
  }
  if (p->viscosity && p->geom == SHARK_GEOM_POLAR) {
    /* :111-116, GEOM 2: radial Laplacians only; eta_visc = alpha_visc H cs with H = HoverR r
     * (disk_polar/setup.f90:29-30, 67) */
    #pragma omp parallel for schedule(static)
    for (int c = 0; c < o->nc; ++c) {
      if (!o->active[c]) continue;
      int ix = ixx(o, c), iy = iyy(o, c);
      int xp = cidx(o, ix + 1, iy), xm = cidx(o, ix - 1, iy);
      double r = o->radii_c[c], d1 = DX(c, 0);
      double eta = (p->alpha_visc * (p->hoverr * r)) * o->cs[c];
      double* S = o->S_U;
      size_t nc = o->nc;
      double su = Q(o->ivx, xp) + Q(o->ivx, xm), sv = Q(o->ivy, xp) + Q(o->ivy, xm);
      double lxu = (su - 2.0 * Q(o->ivx, c)) / (d1 * d1) + su / r / d1 / 2.0 - Q(o->ivx, c) / r / r;
      S[o->ivx * nc + c] = S[o->ivx * nc + c] + Q(o->irho, c) * dt * eta * lxu;
      double lxv = (sv - 2.0 * Q(o->ivy, c)) / (d1 * d1) - Q(o->ivy, c) / (r * r) + sv / r / d1 / 2.0;
      S[o->ivy * nc + c] = S[o->ivy * nc + c] + Q(o->irho, c) * dt * eta * lxv;
    }
  }
  /* :221 u_prim = u_prim + S_U over the whole array */
  #pragma omp parallel for schedule(static)
  for (size_t i = 0; i < n; ++i) o->u[i] = o->u[i] + o->S_U[i];
}

/* ---------------------------------------------------------------- FARGO */
/* The remap of fargo_scheme (src/solver/fargo_scheme.f90:17-77), with `one`=1.0, FP64 eps,
 * slopes for every active cell from its y neighbours, and shifts wrapped modulo ny.  uold is the
 * state after apply_boundaries (its y-ghost rows hold the periodic copies the slopes read);
 * the active cells of columns fa..la are written to u.  A cell's dy is dx(:,2)*radii_c on polar
 * grids (radii == NULL: cartesian, radius 1), its shift velocity vk_cell (fargo_velocity). */
static void fargo_remap_core(int nxm, int nc, int ny, int nvar, int fa, int la, int st, double dt,
                             const double* dx2, const double* radii, const double* vk_cell,
                             const double* uold, double* duprim, double* u) {
  const double forth = 0.25;
  const int fay = 3, lay = ny + 2;
  /* :20-31 slopes (every active cell) */
  #pragma omp parallel for schedule(static)
  for (int ix = fa; ix <= la; ++ix)
    for (int iy = fay; iy <= lay; ++iy) {
      int c = (ix - 1) + (iy - 1) * nxm, il = c - nxm, ir = c + nxm;
      for (int v = 0; v < nvar; ++v)
        duprim[(size_t)v * nc + c] =
            oracle_slope_limit(st, uold[(size_t)v * nc + c] - uold[(size_t)v * nc + il],
                               uold[(size_t)v * nc + ir] - uold[(size_t)v * nc + c]);
    }
  /* :34-77 remap */
  #pragma omp parallel for schedule(static)
  for (int ix = fa; ix <= la; ++ix)
    for (int iy = fay; iy <= lay; ++iy) {
      int c = (ix - 1) + (iy - 1) * nxm;
      double rp = radii ? radii[c] : 1.0;
      double vk = vk_cell[c];
      double dy = dx2[c] * rp;
      double sh = fabs(vk * dt);
      int nshft = (int)(sh / dy);
      double eps = fmod(sh, dy) / dy;
      int j0, j1;
      /* the active rows are 3..ny+2; shifts wrap modulo ny */
      if (vk > 0.0) j0 = iy - nshft - 1; else j0 = iy + nshft;
      j0 = ((j0 - 3) % ny + ny) % ny + 3;
      j1 = (j0 - 3 + 1) % ny + 3;
      int a = (ix - 1) + (j0 - 1) * nxm, b = (ix - 1) + (j1 - 1) * nxm;
      double h2 = (HALF - eps) * (HALF - eps);
      for (int v = 0; v < nvar; ++v) {
        size_t pv = (size_t)v * nc;
        double un;
        if (vk > 0.0) un = uold[pv + a] * eps + uold[pv + b] * (1.0 - eps);
        else un = uold[pv + a] * (1.0 - eps) + uold[pv + b] * eps;
        un = un + HALF * duprim[pv + a] * (forth - h2) + HALF * duprim[pv + b] * (h2 - forth);
        u[pv + c] = un;
      }
    }
}

/* fargo_scheme, src/solver/fargo_scheme.f90:2-81: boundaries, remap, boundaries */
static void fargo_scheme(oracle* o) {
  apply_boundaries(o);
  memcpy(o->uold, o->u, (size_t)o->nc * o->nvar * 8);
  fargo_remap_core(o->nxm, o->nc, o->ny, o->nvar, o->fa, o->la, o->cfg.slope_type, o->dt,
                   o->dx + o->nc, (o->cfg.geom == SHARK_GEOM_POLAR) ? o->radii_c : NULL, o->fargo_v,
                   o->uold, o->duprim, o->u);
  apply_boundaries(o);
}

/* The remap alone (the code fargo_scheme runs) on a 2D state of nvar planes of (nx+4) x (ny+4)
 * cells: uin's y-ghost rows are taken as periodic copies of its active rows (as the disk's
 * apply_boundaries leaves them), every cell of uout outside the active block is a copy of uin. */
void oracle_fargo_remap(int nx, int ny, int nvar, int slope, double dt, const double* dx2,
                        const double* radii_c, const double* fargo_v, const double* uin, double* uout) {
  const int nxm = nx + 4, nym = ny + 4, nc = nxm * nym;
  const size_t n = (size_t)nc * nvar;
  double* uold = malloc(n * 8);
  double* du = calloc(n, 8);
  memcpy(uold, uin, n * 8);
  for (int v = 0; v < nvar; ++v)
    for (int ix = 1; ix <= nxm; ++ix)
      for (int g = 1; g <= 2; ++g) {
        double* p = uold + (size_t)v * nc + (ix - 1);
        p[(size_t)(g - 1) * nxm] = p[(size_t)(ny + g - 1) * nxm];        /* rows 1,2 <- ny+1,ny+2 */
        p[(size_t)(ny + 1 + g) * nxm] = p[(size_t)(1 + g) * nxm];        /* rows ny+3,ny+4 <- 3,4 */
      }
  memcpy(uout, uin, n * 8);
  fargo_remap_core(nxm, nc, ny, nvar, 3, nx + 2, slope, dt, dx2, radii_c, fargo_v, uold, du, uout);
  free(uold);
  free(du);
}

/* ---------------------------------------------------------------- drag */
/* dust_drag(coeffdt), src/dust/dust_drag.f90:2-75 */
static void dust_drag(oracle* o, double coeffdt) {
  const double dt = o->dt;
  const int nd = o->nd;
  #pragma omp parallel for schedule(static)
  for (int c = 0; c < o->nc; ++c) {
    if (!o->active[c]) continue;
    double ak[SHARK_MAX_DUST];
    double rhon = U(o->irho, c), pnx = U(o->ivx, c), pny = U(o->ivy, c), pnz = U(o->ivz, c);
    for (int k = 0; k < nd; ++k) {
      ak[k] = coeffdt * dt / TSTOP(c, k);
      pnx = pnx + ak[k] / (1.0 + ak[k]) * U(IVDX(k), c);
      pny = pny + ak[k] / (1.0 + ak[k]) * U(IVDY(k), c);
      pnz = pnz + ak[k] / (1.0 + ak[k]) * U(IVDZ(k), c);
      rhon = rhon + ak[k] / (1.0 + ak[k]) * U(IRHOD(k), c);
    }
    for (int k = 0; k < nd; ++k) {
      U(IVDX(k), c) = U(IVDX(k), c) / (1.0 + ak[k]) + (ak[k] / (1.0 + ak[k])) * pnx / rhon * U(IRHOD(k), c);
      U(IVDY(k), c) = U(IVDY(k), c) / (1.0 + ak[k]) + (ak[k] / (1.0 + ak[k])) * pny / rhon * U(IRHOD(k), c);
      U(IVDZ(k), c) = U(IVDZ(k), c) / (1.0 + ak[k]) + (ak[k] / (1.0 + ak[k])) * pnz / rhon * U(IRHOD(k), c);
    }
    if (o->cfg.dust_back_reaction) {
      U(o->ivx, c) = pnx / rhon * U(o->irho, c);
      U(o->ivy, c) = pny / rhon * U(o->irho, c);
      U(o->ivz, c) = pnz / rhon * U(o->irho, c);
    }
  }
  #pragma omp parallel for schedule(static)
  for (int c = 0; c < o->nc; ++c) {
    if (!o->active[c]) continue;
    for (int k = 0; k < nd; ++k)
      U(IRHOD(k), c) = fmax_f(U(o->irho, c) * o->cfg.dust_ratio_min, U(IRHOD(k), c));
  }
}

/* ---------------------------------------------------------------- setup_inloop */
/* SI: src/setups/SI/setup.f90:299-332; disk: src/setups/disk_polar/setup.f90:206-284 */
static void setup_inloop(oracle* o) {
  const shark_cfg* p = &o->cfg;
  const double dt = o->dt;
  if (p->inloop_kind == SHARK_INLOOP_SI) {
    if (p->force_kick) return;
    for (int c = 0; c < o->nc; ++c) {
      if (!o->active[c]) continue;
      double Om = p->omega_shear;
      double u = U(o->ivz, c) / U(o->irho, c), v = U(o->ivx, c) / U(o->irho, c), d = U(o->irho, c);
      double Ohmdt = Om * dt;
      double AA = (2.0 * p->q_shear) * Om * (POS(c, 0) - HALF * p->box_l) - (2.0 * p->rad0) * p->eta_stream;
      U(o->ivz, c) = d * (u * (1. - Ohmdt * Ohmdt) + 2. * v * Ohmdt + AA * (Ohmdt * Ohmdt)) / (1. + Ohmdt * Ohmdt);
      U(o->ivx, c) = d * (v * (1. - Ohmdt * Ohmdt) - 2. * u * Ohmdt + AA * Ohmdt) / (1. + Ohmdt * Ohmdt);
      AA = (2.0 * p->q_shear) * Om * (POS(c, 0) - HALF * p->box_l);
      for (int k = 0; k < o->nd; ++k) {
        u = U(IVDZ(k), c) / U(IRHOD(k), c); v = U(IVDX(k), c) / U(IRHOD(k), c); d = U(IRHOD(k), c);
        U(IVDZ(k), c) = d * (u * (1. - Ohmdt * Ohmdt) + 2. * v * Ohmdt + AA * (Ohmdt * Ohmdt)) / (1. + Ohmdt * Ohmdt);
        U(IVDX(k), c) = d * (v * (1. - Ohmdt * Ohmdt) - 2. * u * Ohmdt + AA * Ohmdt) / (1. + Ohmdt * Ohmdt);
      }
    }
  } else if (p->inloop_kind == SHARK_INLOOP_CDISK) {
    /* disk/setup.f90:203-262: cells of the inner and outer buffers reset to the initial disk
     * (rate = 0), every expression kept as written */
    const double rate = 0.0, R_out_disk = p->frac_r_disk * p->box_l;
    #pragma omp parallel for schedule(static)
    for (int c = 0; c < o->nc; ++c) {
      if (!o->active[c]) continue;
      double xx, yy, rr;
      cdisk_r(o, c, &xx, &yy, &rr);
      double r_cyl = sqrt(xx * xx + yy * yy);
      double x_soft = xx * rr / r_cyl, y_soft = yy * rr / r_cyl;
      if (!(rr < p->buffer_in * p->box_l || rr > p->buffer_out * p->box_l)) continue;
      double H = p->hoverr * rr;
      double Omega = sqrt(p->mstar / cube_i(rr));
      double cs0 = H * Omega;
      double rho_ana = p->sigma_r0 * (p->r0_disk / rr);
      if (rr > R_out_disk) rho_ana = p->sigma_r0 * (p->r0_disk / R_out_disk) / 100.0;
      double v_ana = -(Omega * rho_ana);
      U(o->irho, c) = rho_ana + (U(o->irho, c) - rho_ana) * rate;
      U(o->ivx, c) = -(y_soft * v_ana) + (U(o->ivx, c) + yy * v_ana) * rate;
      U(o->ivy, c) = x_soft * v_ana + (U(o->ivy, c) - xx * v_ana) * rate;
      U(o->iP, c) = (rho_ana + (U(o->irho, c) - rho_ana) * rate) * (cs0 * cs0);
      for (int k = 0; k < o->nd; ++k) {
        double dust_to = p->dust2gas;
        U(IRHOD(k), c) = dust_to * rho_ana + (U(IRHOD(k), c) - rho_ana * dust_to) * rate;
        U(IVDX(k), c) = -(yy * v_ana * dust_to) + (U(IVDX(k), c) + dust_to * yy * v_ana) * rate;
        U(IVDY(k), c) = xx * v_ana * dust_to + (U(IVDY(k), c) - dust_to * xx * v_ana) * rate;
      }
    }
  } else if (p->inloop_kind == SHARK_INLOOP_DISK) {
    const double* ci = o->condinit;
    size_t nc = o->nc;
    #pragma omp parallel for schedule(static)
    for (int c = 0; c < o->nc; ++c) {
      if (!o->active[c]) continue;
      double r = o->radii_c[c];
      if (!(r < p->r_relax)) continue;
      double t_rel = p->n_rel * 2.0 * PI / (sqrt(p->mstar / (r * r * r))) -
                     p->n_rel * 2.0 * PI / (sqrt(p->mstar / (p->smooth_r * p->smooth_r * p->smooth_r)));
      double e = exp(-dt / t_rel);
      double rho_i = ci[o->irho * nc + c], vx_i = ci[o->ivx * nc + c] / rho_i, vy_i = ci[o->ivy * nc + c] / rho_i;
      double rho_o = U(o->irho, c), vx_o = U(o->ivx, c) / rho_o, vy_o = U(o->ivy, c) / rho_o;
      double rho_n = (rho_i) * (1.0 - e) + rho_o * e;
      double vx_n = (vx_i) * (1.0 - e) + vx_o * e;
      double vy_n = (vy_i) * (1.0 - e) + vy_o * e;
      U(o->irho, c) = rho_n; U(o->ivx, c) = rho_n * vx_n; U(o->ivy, c) = rho_n * vy_n;
      for (int k = 0; k < o->nd; ++k) {
        rho_i = ci[IRHOD(k) * nc + c]; vx_i = ci[IVDX(k) * nc + c] / rho_i; vy_i = ci[IVDY(k) * nc + c] / rho_i;
        rho_o = U(IRHOD(k), c); vx_o = U(IVDX(k), c) / rho_o; vy_o = U(IVDY(k), c) / rho_o;
        rho_n = (rho_i) * (1.0 - e) + rho_o * e;
        vx_n = (vx_i) * (1.0 - e) + vx_o * e;
        vy_n = (vy_i) * (1.0 - e) + vy_o * e;
        U(IRHOD(k), c) = rho_n; U(IVDX(k), c) = rho_n * vx_n; U(IVDY(k), c) = rho_n * vy_n;
      }
    }
    for (int ix = 1; ix <= o->fa; ++ix)
      for (int iy = o->fay; iy <= o->lay; ++iy)
        for (int v = 0; v < o->nvar; ++v)
          U(v, cidx(o, ix, iy)) = ci[(size_t)v * nc + cidx(o, o->fa, iy)];
  }
}

/* ---------------------------------------------------------------- growth */
/* dust_growth over the active cells (solve.f90:112) in the disk environment of growth_oracle.c:
 * surface density u_prim(:,irho), the step's cs and cell radius, the step's dt.  The reference
 * changes only the dust densities (dust_growth.f90:243-245), so each species' velocity m/rho_d
 * jumps by the density ratio and the next Courant step collapses (1.8e-2 -> 4e-10 on a 32x48
 * disk); here every species keeps its velocity, its momentum scaled with its density. */
static void growth_disk_step(oracle* o) {
  const int nb = o->nd;
  int64_t na = 0;
  for (int c = 0; c < o->nc; ++c) na += o->active[c] ? 1 : 0;
  double* sg = (double*)malloc(sizeof(double) * na);
  double* cs = (double*)malloc(sizeof(double) * na);
  double* r = (double*)malloc(sizeof(double) * na);
  double* rd = (double*)malloc(sizeof(double) * na * nb);
  int* map = (int*)malloc(sizeof(int) * na);
  int64_t n = 0;
  for (int c = 0; c < o->nc; ++c)
    if (o->active[c]) {
      map[n] = c;
      sg[n] = U(0, c);
      cs[n] = o->cs[c];
      r[n] = o->radii_c[c];
      ++n;
    }
  for (int k = 0; k < nb; ++k)
    for (int64_t i = 0; i < na; ++i) rd[k * na + i] = U(IRHOD(k), map[i]);
  growth_oracle_disk(&o->gp, &o->gt, na, sg, cs, r, o->dt, rd, NULL, NULL);
  /* each species keeps its velocity: momentum = (momentum / old density) * new density */
  for (int k = 0; k < nb; ++k)
    for (int64_t i = 0; i < na; ++i) {
      const int c = map[i];
      const double r0 = U(IRHOD(k), c), r1 = rd[k * na + i];
      U(IVDX(k), c) = (U(IVDX(k), c) / r0) * r1;
      U(IVDY(k), c) = (U(IVDY(k), c) / r0) * r1;
      U(IVDZ(k), c) = (U(IVDZ(k), c) / r0) * r1;
      U(IRHOD(k), c) = r1;
    }
  free(sg); free(cs); free(r); free(rd); free(map);
}

int oracle_set_growth(oracle* o, const shark_growth_params* gp) {
  if (!gp) { o->growth_on = 0; return 0; }
  if (gp->nbins != o->nd) return 1;
  o->gp = *gp;
  growth_oracle_tables(gp, &o->gt);
  o->growth_on = 1;
  return 0;
}

/* ---------------------------------------------------------------- solve */
/* solve, src/solver/solve.f90:10-181 (no charging, MHD, TURB) */
double oracle_step(oracle* o) {
  const shark_cfg* p = &o->cfg;
  apply_boundaries(o);
  ctoprim(o);
  if (p->force_kick) update_force(o);
  if (o->nd > 0) {
    distribution_dust_refresh(o);
    compute_tstop(o);
  }
  courant(o);
  if (p->force_kick) kick(o, 1.0);
  predictor(o);
  add_delta_u(o);
  source_terms(o);
  if (p->fargo) fargo_scheme(o);
  if (o->nd > 0 && p->drag) dust_drag(o, 1.0);
  if (o->growth_on) growth_disk_step(o);
  setup_inloop(o);
  o->time = o->time + o->dt;
  return o->dt;
}

int64_t oracle_run(oracle* o, int64_t nsteps, double tend) {
  int64_t n = 0;
  while (n < nsteps && o->time < tend) { oracle_step(o); ++n; }
  return n;
}

/* primitive variables of the current state (ctoprim), for download comparisons */
void oracle_primitives(oracle* o, double* q) {
  ctoprim(o);
  if (q) memcpy(q, o->q, (size_t)o->nc * o->nvar * 8);
}
