// shark_gpu.hip — C-ABI of libshark_gpu.so (include/shark_gpu.h): device state, upload and
// download, and the step driver (one hipGraph per step, replayed; or per-kernel launches
// bracketed by HIP events when timing is on).
#include <hip/hip_runtime.h>

#include <errno.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/shark_gpu.h"
#include "growth.h"
#include "../../include/shark_growth.h"
#include "../../include/shark_output.h"
#include "output.h"
#include "physics.h"
#include "step_kernels.h"

using shark::DevScalars;
using shark::StepArgs;

namespace shark {
bool nd_supported(int nd);
}

static thread_local char g_err[512] = "";

static int fail(int code, const char* fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof(g_err), fmt, ap);
  va_end(ap);
  return code;
}

namespace shark {
// error reporting shared with the growth entry points (same last-error buffer)
int fail_msg(int code, const char* msg) {
  snprintf(g_err, sizeof(g_err), "%s", msg);
  return code;
}
}  // namespace shark

#define HIPCHK(x)                                                                   \
  do {                                                                              \
    hipError_t e_ = (x);                                                            \
    if (e_ != hipSuccess)                                                           \
      return fail(SHARK_EHIP, "%s failed: %s (%s:%d)", #x, hipGetErrorString(e_),   \
                  __FILE__, __LINE__);                                              \
  } while (0)

enum { KT_BC = 0, KT_HYDRO, KT_POST, KT_DT, KT_GROWTH, KT_N };
static const char* kKernelNames[KT_N] = {"bc", "hydro", "post", "dt", "growth"};

struct shark_ctx {
  shark_cfg cfg;
  StepArgs a;
  hipStream_t stream = nullptr;
  double* dU = nullptr;
  double* dB = nullptr;
  double* dcs = nullptr;
  double* dgeom = nullptr;   // 7 per-column arrays
  double* dsdust = nullptr;
  double* dcond = nullptr;
  double* dfargo = nullptr;
  DevScalars* dsc = nullptr;
  DevScalars* hsc = nullptr;  // pinned mirror
  bool geom_set = false, uploaded = false, dvol_set = false;
  hipGraphExec_t gexec = nullptr;
  bool timing = false;
  double kt_ms[KT_N] = {0, 0, 0, 0, 0};
  int64_t kt_n[KT_N] = {0, 0, 0, 0, 0};
  std::vector<double> hpos, hrc;    // position(ncells,ndim) and radii_c(ncells) for outputs
  shark_growth* growth = nullptr;   // attached growth state (dust_growth inside each step)
  double* dRold = nullptr;          // dust densities before growth, for the velocity-keeping momenta
  double* dfcs = nullptr;           // gas face sound speeds for the HLLC dust solver (adiabatic gas)
  double* ddvol = nullptr;          // spherical grid: the cells' inner volumes (shark_gpu_set_dvol)
  double* dmc = nullptr;            // 1D self-gravity: enclosed mass per cell
  double* dposxy = nullptr;         // polar-disk planet: the grid's position (2, ncells), then the
                                    // gradient of its potential (4, ncells)
  double* dphisg = nullptr;         // polar-disk self-gravity: potential (ncells), then the cosine
                                    // table (ny, ny)
  double* dfexp = nullptr;          // explicit drag: every fluid's force of the step
  int growth_mode = 0;
};

static int nvar_of(const shark_cfg* c) { return 5 + 4 * c->ndust; }

extern "C" const char* shark_gpu_last_error(void) { return g_err; }
extern "C" int shark_gpu_version(void) { return SHARK_CFG_VERSION; }

extern "C" int64_t shark_gpu_algorithmic_bytes(const shark_cfg* c) {
  const int64_t nv = nvar_of(c);
  const int64_t cells = (int64_t)c->nx * (c->ny > 1 ? c->ny : 1);
  return cells * (8 * (2 * nv + 1) + (c->fargo ? 16 * nv : 0));
}

static int validate(const shark_cfg* c) {
  if (!c) return fail(SHARK_EINVAL, "null cfg");
  if (c->version != SHARK_CFG_VERSION) return fail(SHARK_EINVAL, "cfg version %d, expected %d", c->version, SHARK_CFG_VERSION);
  if (c->nx < 4 || c->ny < 1) return fail(SHARK_EINVAL, "nx=%d ny=%d: need nx>=4, ny>=1", c->nx, c->ny);
  if (c->ny > 1 && c->ny < 4) return fail(SHARK_EINVAL, "2D runs need ny>=4 (got %d)", c->ny);
  if (c->nghost != 2) return fail(SHARK_EUNSUP, "nghost=%d: only NGHOST=2 is built", c->nghost);
  if (c->ndust < 0 || c->ndust > SHARK_MAX_DUST) return fail(SHARK_EINVAL, "ndust=%d out of range", c->ndust);
  if (!shark::nd_supported(c->ndust))
    return fail(SHARK_EUNSUP, "ndust=%d: the step is built for 0 to %d dust species", c->ndust, SHARK_MAX_DUST);
  if (c->iso_cs > 1)
    return fail(SHARK_EINVAL, "iso_cs=%d: 1 is isothermal, below 1 adiabatic", c->iso_cs);
  if (c->iso_cs != 1 && ((c->geom != SHARK_GEOM_CARTESIAN && c->geom != SHARK_GEOM_SPHERE) || c->fargo))
    return fail(SHARK_EUNSUP, "iso_cs=%d: adiabatic gas is built on cartesian and spherical 1D grids without FARGO",
                c->iso_cs);
  if (c->bc_kind == SHARK_BC_SOD && (c->ndust > 0 || c->shock_dir < 0 || c->shock_dir > 1 ||
                                     (c->shock_dir == 1 && c->ny < 4)))
    return fail(SHARK_EUNSUP, "Sod boundaries need gas only and shock_dir 0 (x) or 1 (y, 2D)");
  if (c->geom != SHARK_GEOM_CARTESIAN && c->geom != SHARK_GEOM_POLAR && c->geom != SHARK_GEOM_SPHERE)
    return fail(SHARK_EUNSUP, "geom=%d: GEOM 0, 1 and 2 are built", c->geom);
  if (c->geom == SHARK_GEOM_SPHERE && c->ny != 1)
    return fail(SHARK_EUNSUP, "the spherical geometry (GEOM 1) is built in 1D (Collapse_1D)");
  if (c->gravity && c->geom != SHARK_GEOM_SPHERE)
    return fail(SHARK_EUNSUP, "self-gravity is built for the spherical 1D grid (GRAVITY=1 with NY=1)");
  if (c->eos_baro && c->iso_cs >= 1)
    return fail(SHARK_EUNSUP, "the barotropic EOS (non_standard_eos) is built with iso_cs < 1 (Collapse_1D)");
  if ((c->bc_kind == SHARK_BC_COLLAPSE || c->tstop_kind == SHARK_TSTOP_COLLAPSE) &&
      (c->ny != 1 || !c->eos_baro))
    return fail(SHARK_EUNSUP, "the Collapse_1D hooks need a 1D grid and the barotropic EOS");
  if (c->geom == SHARK_GEOM_POLAR && c->ny < 4) return fail(SHARK_EUNSUP, "polar geometry needs a 2D grid");
  if (c->solver < 0 || c->solver > 2) return fail(SHARK_EINVAL, "solver=%d", c->solver);
  if (c->solverdust < 0 || c->solverdust > 3) return fail(SHARK_EINVAL, "solverdust=%d", c->solverdust);
  if (c->slope_type < 1 || c->slope_type > 3) return fail(SHARK_EINVAL, "unknown slope limiter %d", c->slope_type);
  const bool cdisk = c->force_kind == SHARK_FORCE_CDISK || c->inloop_kind == SHARK_INLOOP_CDISK ||
                     c->tstop_kind == SHARK_TSTOP_EPSTEIN_CDISK;
  if (cdisk && (c->geom != SHARK_GEOM_CARTESIAN || c->ny < 4 || c->iso_cs != 1))
    return fail(SHARK_EUNSUP, "the cartesian disk hooks need a 2D cartesian grid and isothermal gas");
  if (c->viscosity && !(c->iso_cs == 1 && c->ny >= 4 &&
                        ((c->geom == SHARK_GEOM_CARTESIAN && c->inloop_kind == SHARK_INLOOP_CDISK) ||
                         (c->geom == SHARK_GEOM_POLAR && c->inloop_kind == SHARK_INLOOP_DISK))))
    return fail(SHARK_EUNSUP, "viscosity is built for the isothermal cartesian and polar disks (eta_visc = alpha_visc H cs)");
  if (c->fargo && (c->ny < 4 || c->geom != SHARK_GEOM_POLAR))
    return fail(SHARK_EUNSUP, "FARGO is built for 2D polar grids (rings of constant radius, the disk_polar setup)");
  if (c->force_kick && !c->drag && c->ndust > 0 && (c->force_kind != SHARK_FORCE_SI || c->tstop_kind != SHARK_TSTOP_SI))
    return fail(SHARK_EUNSUP, "explicit drag in the force (drag=.false. with force_kick) is built for the shearing "
                              "box, whose stopping times are constant");
  if (c->bc_kind < 0 || c->bc_kind > 5 || c->tstop_kind < 0 || c->tstop_kind > 5 ||
      c->force_kind < 0 || c->force_kind > 3 || c->inloop_kind < 0 || c->inloop_kind > 3)
    return fail(SHARK_EINVAL, "unknown setup hook kind");
  if (c->bc_kind == SHARK_BC_DISK && c->ny < 4) return fail(SHARK_EUNSUP, "disk boundaries need a 2D grid");
  if (c->force_kind == SHARK_FORCE_DISK && c->mplanet > 0.0 && (c->geom != SHARK_GEOM_POLAR || c->ny < 4))
    return fail(SHARK_EUNSUP, "the polar disk's planet is built for the 2D polar grid");
  if (c->force_kind == SHARK_FORCE_DISK && c->mplanet > 0.0 && !(c->rplanet > 0.0 && c->a_smooth_pl >= 0.0))
    return fail(SHARK_EINVAL, "the polar disk's planet needs rplanet > 0 and a_smooth_pl >= 0");
  if (c->self_gravity && (c->force_kind != SHARK_FORCE_DISK || c->geom != SHARK_GEOM_POLAR || c->ny < 4 || !c->force_kick))
    return fail(SHARK_EUNSUP, "self_gravity is the polar disk's (disk_polar/setup.f90:300-322): 2D polar grid with its forces");
  if (c->accretion && (c->bc_kind != SHARK_BC_DISK || c->iso_cs != 1))
    return fail(SHARK_EUNSUP, "accretion is built for the isothermal polar disk's boundaries");
  if (c->cfl <= 0) return fail(SHARK_EINVAL, "CFL must be positive");
  return SHARK_OK;
}

extern "C" int shark_gpu_create(const shark_cfg* cfg, shark_ctx** out) {
  if (!out) return fail(SHARK_EINVAL, "null ctx pointer");
  *out = nullptr;
  int rc = validate(cfg);
  if (rc) return rc;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1)
    return fail(SHARK_EHIP, "no HIP device visible");
  shark_ctx* c = new shark_ctx();
  c->cfg = *cfg;
  StepArgs& a = c->a;
  memset(&a, 0, sizeof(a));
  a.nx = cfg->nx;
  a.ny = cfg->ny > 1 ? cfg->ny : 1;
  a.nd = cfg->ndust;
  a.ndim = cfg->ny > 1 ? 2 : 1;
  a.nxm = cfg->nx + 4;
  a.nym = a.ndim == 2 ? cfg->ny + 4 : 1;
  a.nc = a.nxm * a.nym;
  a.nvar = nvar_of(cfg);
  a.geom = cfg->geom; a.solver = cfg->solver; a.solverdust = cfg->solverdust; a.slope = cfg->slope_type;
  a.force_kick = cfg->force_kick; a.force_kind = cfg->force_kick ? cfg->force_kind : 0;
  a.tstop_kind = cfg->tstop_kind; a.inloop_kind = cfg->inloop_kind; a.bc_kind = cfg->bc_kind;
  a.drag = cfg->drag; a.back_reaction = cfg->dust_back_reaction; a.fargo = cfg->fargo;
  a.disk_bc_mode = cfg->disk_bc_mode;
  a.gamma = cfg->gamma; a.cfl = cfg->cfl; a.smallr = cfg->smallr; a.dust_ratio_min = cfg->dust_ratio_min;
  a.q_shear = cfg->q_shear; a.omega_shear = cfg->omega_shear; a.box_l = cfg->box_l; a.rad0 = cfg->rad0;
  a.eta_stream = cfg->eta_stream; a.rho_init = cfg->rho_init; a.hoverr = cfg->hoverr;
  a.rho_l = cfg->rho_l; a.rho_r = cfg->rho_r; a.p_l = cfg->p_l; a.p_r = cfg->p_r;
  a.shock_dir = cfg->shock_dir;
  a.iso = cfg->iso_cs == 1;
  a.smallp = cfg->smallp;
  a.eps_k1 = sqrt((3.141592653589793238462643383279 * cfg->gamma) / 8.0) * cfg->rhograin;
  a.s2pi = sqrt(2.0 * 3.141592653589793238462643383279);
  a.rhograin = cfg->rhograin; a.mstar = cfg->mstar; a.r_relax = cfg->r_relax; a.n_rel = cfg->n_rel;
  a.smooth_r = cfg->smooth_r;
  a.box_l_y = cfg->box_l_y; a.mplanet = cfg->mplanet; a.rplanet = cfg->rplanet; a.sigma_r0 = cfg->sigma_r0;
  a.r0_disk = cfg->r0_disk; a.frac_r_disk = cfg->frac_r_disk; a.buffer_in = cfg->buffer_in;
  a.buffer_out = cfg->buffer_out; a.alpha_visc = cfg->alpha_visc; a.dust2gas = cfg->dust2gas;
  a.viscosity = cfg->viscosity;
  a.accretion = cfg->accretion; a.m_acc = cfg->m_acc; a.phi_mom = cfg->phi_mom; a.l_out = cfg->l_out;
  a.gravity = cfg->gravity; a.eos_baro = cfg->eos_baro; a.l_soft = cfg->l_soft; a.m_cent = cfg->m_cent;
  a.t0_cloud = cfg->t0_cloud; a.mu_gas = cfg->mu_gas; a.unit_v = cfg->unit_v; a.unit_nh = cfg->unit_nh;
  for (int k = 0; k < SHARK_MAX_DUST; ++k) a.K[k] = cfg->K[k];
  // launch shapes (shark_gpu_set_tune / shark_gpu_autotune change them)
  a.strip_h = 32;
  a.strip_g = 32;
  a.ngroups = (a.nx + 59) / 60;
  a.nstrips = a.ndim == 2 ? (a.ny + a.strip_h - 1) / a.strip_h : 1;
  a.nstrips_g = a.ndim == 2 ? (a.ny + a.strip_g - 1) / a.strip_g : 1;
  a.post_tw = shark::post_tw_default(a.nd);
  a.xcd_order = 1;

  const size_t nbytes = (size_t)a.nc * a.nvar * sizeof(double);
  hipError_t e = hipSuccess;
  if ((e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking)) != hipSuccess ||
      (e = hipMalloc(&c->dU, nbytes)) != hipSuccess || (e = hipMalloc(&c->dB, nbytes)) != hipSuccess ||
      (e = hipMalloc(&c->dcs, (size_t)a.nc * 8)) != hipSuccess ||
      (e = hipMalloc(&c->dgeom, ((size_t)7 * a.nxm + a.nym) * 8)) != hipSuccess ||
      (e = hipMalloc(&c->dsc, sizeof(DevScalars))) != hipSuccess ||
      (e = hipHostMalloc(&c->hsc, sizeof(DevScalars))) != hipSuccess) {
    rc = fail(e == hipErrorOutOfMemory ? SHARK_ENOMEM : SHARK_EHIP, "allocation failed: %s", hipGetErrorString(e));
    shark_gpu_destroy(c);
    return rc;
  }
  if (cfg->iso_cs != 1 && a.nd > 0 && cfg->solverdust == 3) {
    // godunov.f90:676-680, 706-712: the dust HLLC solver reads the gas face sound speeds
    if ((e = hipMalloc(&c->dfcs, (size_t)4 * a.nc * 8)) != hipSuccess) {
      rc = fail(e == hipErrorOutOfMemory ? SHARK_ENOMEM : SHARK_EHIP, "allocation failed: %s", hipGetErrorString(e));
      shark_gpu_destroy(c);
      return rc;
    }
    (void)hipMemsetAsync(c->dfcs, 0, (size_t)4 * a.nc * 8, c->stream);
  }
  a.fcs = c->dfcs;
  if (cfg->gravity) {
    if ((e = hipMalloc(&c->dmc, (size_t)a.nc * 8)) != hipSuccess ||
        (e = hipMalloc(&c->ddvol, (size_t)a.nc * 8)) != hipSuccess) {
      rc = fail(e == hipErrorOutOfMemory ? SHARK_ENOMEM : SHARK_EHIP, "allocation failed: %s", hipGetErrorString(e));
      shark_gpu_destroy(c);
      return rc;
    }
    (void)hipMemsetAsync(c->dmc, 0, (size_t)a.nc * 8, c->stream);
    (void)hipMemsetAsync(c->ddvol, 0, (size_t)a.nc * 8, c->stream);
  }
  a.mc = c->dmc;
  a.dvol = c->ddvol;
  if (cfg->force_kind == SHARK_FORCE_DISK && (cfg->mplanet > 0.0 || cfg->self_gravity)) {
    if ((e = hipMalloc(&c->dposxy, (size_t)6 * a.nc * 8)) != hipSuccess ||
        (cfg->self_gravity &&
         (e = hipMalloc(&c->dphisg, ((size_t)a.nc + (size_t)a.ny * a.ny) * 8)) != hipSuccess)) {
      rc = fail(e == hipErrorOutOfMemory ? SHARK_ENOMEM : SHARK_EHIP, "allocation failed: %s", hipGetErrorString(e));
      shark_gpu_destroy(c);
      return rc;
    }
    (void)hipMemsetAsync(c->dposxy, 0, (size_t)6 * a.nc * 8, c->stream);
  }
  a.posxy = c->dposxy;
  a.pgrad = c->dposxy ? c->dposxy + (size_t)2 * a.nc : nullptr;
  if (cfg->force_kick && !cfg->drag && cfg->ndust > 0) {
    if ((e = hipMalloc(&c->dfexp, (size_t)3 * (1 + cfg->ndust) * a.nc * 8)) != hipSuccess) {
      rc = fail(e == hipErrorOutOfMemory ? SHARK_ENOMEM : SHARK_EHIP, "allocation failed: %s", hipGetErrorString(e));
      shark_gpu_destroy(c);
      return rc;
    }
    (void)hipMemsetAsync(c->dfexp, 0, (size_t)3 * (1 + cfg->ndust) * a.nc * 8, c->stream);
  }
  a.fexp = c->dfexp;
  a.self_gravity = cfg->self_gravity ? 1 : 0;
  a.phi_sg = c->dphisg;
  a.sgcos = c->dphisg ? c->dphisg + a.nc : nullptr;
  if (c->dphisg) (void)hipMemsetAsync(c->dphisg, 0, (size_t)a.nc * 8, c->stream);
  a.a_smooth_pl = cfg->a_smooth_pl;
  (void)hipMemsetAsync(c->dU, 0, nbytes, c->stream);
  (void)hipMemsetAsync(c->dB, 0, nbytes, c->stream);
  memset(c->hsc, 0, sizeof(DevScalars));
  c->hsc->dtbits = 0;
  (void)hipMemcpyAsync(c->dsc, c->hsc, sizeof(DevScalars), hipMemcpyHostToDevice, c->stream);
  a.U = c->dU; a.B = c->dB; a.cs = c->dcs; a.sc = c->dsc;
  const double* gp = c->dgeom;
  a.g.dx1 = gp; a.g.dx2 = gp + a.nxm; a.g.surf1 = gp + 2 * a.nxm; a.g.surf2 = gp + 3 * a.nxm;
  a.g.vol = gp + 4 * a.nxm; a.g.rc = gp + 5 * a.nxm; a.g.pos1 = gp + 6 * a.nxm; a.g.pos2 = gp + 7 * a.nxm;
  if (hipStreamSynchronize(c->stream) != hipSuccess) {
    rc = fail(SHARK_EHIP, "stream sync failed");
    shark_gpu_destroy(c);
    return rc;
  }
  *out = c;
  return SHARK_OK;
}

extern "C" int shark_gpu_set_geometry(shark_ctx* c, const double* dx, const double* surf,
                                      const double* vol, const double* radii_c, const double* position) {
  if (!c) return fail(SHARK_EINVAL, "null ctx");
  if (!dx || !surf || !vol || !radii_c) return fail(SHARK_EINVAL, "null geometry array");
  const StepArgs& a = c->a;
  const int nxm = a.nxm, nym = a.nym, nc = a.nc;
  std::vector<double> col((size_t)7 * nxm + nym, 0.0);
  auto at = [&](const double* arr, int plane, int ix, int iy) {
    return arr[(size_t)plane * nc + (ix - 1) + (size_t)(iy - 1) * nxm];
  };
  for (int ix = 1; ix <= nxm; ++ix) {
    col[0 * nxm + ix - 1] = at(dx, 0, ix, 1);
    col[1 * nxm + ix - 1] = a.ndim == 2 ? at(dx, 1, ix, 1) : 0.0;
    col[2 * nxm + ix - 1] = at(surf, 0, ix, 1);
    col[3 * nxm + ix - 1] = a.ndim == 2 ? at(surf, 1, ix, 1) : 0.0;
    col[4 * nxm + ix - 1] = vol[ix - 1];
    col[5 * nxm + ix - 1] = radii_c[ix - 1];
    col[6 * nxm + ix - 1] = position ? position[ix - 1] : 0.0;
    // the kernels read one value per column: the grid must be separable
    for (int iy = 2; iy <= nym; ++iy) {
      const size_t ci = (ix - 1) + (size_t)(iy - 1) * nxm;
      bool ok = dx[ci] == col[ix - 1] && surf[ci] == col[2 * nxm + ix - 1] && vol[ci] == col[4 * nxm + ix - 1] &&
                radii_c[ci] == col[5 * nxm + ix - 1];
      if (a.ndim == 2) ok = ok && dx[nc + ci] == col[nxm + ix - 1] && surf[nc + ci] == col[3 * nxm + ix - 1];
      // position is read only by the shearing-box and cartesian-disk hooks
      const bool cdisk = c->cfg.force_kind == SHARK_FORCE_CDISK || c->cfg.inloop_kind == SHARK_INLOOP_CDISK ||
                         c->cfg.tstop_kind == SHARK_TSTOP_EPSTEIN_CDISK ||
                         (c->cfg.viscosity && c->cfg.geom == SHARK_GEOM_CARTESIAN);
      const bool need_pos = c->cfg.force_kind == SHARK_FORCE_SI || c->cfg.inloop_kind == SHARK_INLOOP_SI || cdisk;
      if (position && need_pos) ok = ok && position[ci] == col[6 * nxm + ix - 1];
      if (position && cdisk && a.ndim == 2) ok = ok && position[nc + ci] == position[nc + (size_t)(iy - 1) * nxm];
      if (!ok) return fail(SHARK_EUNSUP, "geometry varies along y at column %d: only separable grids are built", ix);
    }
  }
  if (position && a.ndim == 2)
    for (int iy = 1; iy <= nym; ++iy) col[7 * (size_t)nxm + iy - 1] = position[nc + (size_t)(iy - 1) * nxm];
  const bool need_pos = c->cfg.force_kind == SHARK_FORCE_SI || c->cfg.inloop_kind == SHARK_INLOOP_SI ||
                        c->cfg.force_kind == SHARK_FORCE_CDISK || c->cfg.inloop_kind == SHARK_INLOOP_CDISK ||
                        c->cfg.tstop_kind == SHARK_TSTOP_EPSTEIN_CDISK ||
                        (c->cfg.viscosity && c->cfg.geom == SHARK_GEOM_CARTESIAN);
  const bool planet = c->dposxy && c->cfg.mplanet > 0.0;
  if (!position && (need_pos || planet)) return fail(SHARK_EINVAL, "position is required by this setup's hooks");
  if (planet)
    HIPCHK(hipMemcpyAsync(c->dposxy, position, (size_t)2 * nc * 8, hipMemcpyHostToDevice, c->stream));
  if (c->dphisg) {
    // cos(phi_p - phi_loc) of the self-gravity sum (disk_polar/setup.f90:311-315) per (source
    // row, target row): phi = float(iy) * dx(:,2), with the host's cos, as the reference calls it
    const int ny = a.ny;
    const double d2 = col[nxm + 2];
    for (int ix = 3; ix <= a.nx + 2; ++ix)
      if (col[nxm + ix - 1] != d2) return fail(SHARK_EUNSUP, "self-gravity needs dx(:,2) equal in every column");
    std::vector<double> tab((size_t)ny * ny);
    for (int sy = 0; sy < ny; ++sy)
      for (int ty = 0; ty < ny; ++ty)
        tab[(size_t)sy * ny + ty] = cos((double)(sy + 3) * d2 - (double)(ty + 3) * d2);
    HIPCHK(hipMemcpyAsync(c->dphisg + nc, tab.data(), tab.size() * 8, hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));   // tab is released at the end of this block
  }
  HIPCHK(hipMemcpyAsync(c->dgeom, col.data(), col.size() * 8, hipMemcpyHostToDevice, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  c->hrc.assign(radii_c, radii_c + nc);
  if (position) c->hpos.assign(position, position + (size_t)nc * a.ndim);
  else c->hpos.clear();
  c->geom_set = true;
  return SHARK_OK;
}

extern "C" int shark_gpu_set_dvol(shark_ctx* c, const double* dvol) {
  if (!c) return fail(SHARK_EINVAL, "null ctx");
  if (!dvol) return fail(SHARK_EINVAL, "null dvol");
  if (!c->ddvol) return fail(SHARK_ESTATE, "dvol is read only with self-gravity (cfg.gravity = 1)");
  HIPCHK(hipMemcpyAsync(c->ddvol, dvol, (size_t)c->a.nc * 8, hipMemcpyHostToDevice, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  c->dvol_set = true;
  return SHARK_OK;
}

static void drop_graph(shark_ctx* c) {
  if (c->gexec) {
    (void)hipGraphExecDestroy(c->gexec);
    c->gexec = nullptr;
  }
}

extern "C" int shark_gpu_upload(shark_ctx* c, const double* u, const double* cs, const double* sdust,
                                const double* condinit, const double* fargo_velocity) {
  if (!c) return fail(SHARK_EINVAL, "null ctx");
  if (!c->geom_set) return fail(SHARK_ESTATE, "shark_gpu_set_geometry must precede upload");
  if (c->cfg.gravity && !c->dvol_set)
    return fail(SHARK_ESTATE, "self-gravity needs the cells' inner volumes: shark_gpu_set_dvol before upload");
  if (!u || !cs) return fail(SHARK_EINVAL, "u_prim and cs are required");
  StepArgs& a = c->a;
  const shark_cfg& cf = c->cfg;
  const size_t nc = a.nc;
  drop_graph(c);
  HIPCHK(hipMemcpyAsync(c->dU, u, nc * a.nvar * 8, hipMemcpyHostToDevice, c->stream));
  HIPCHK(hipMemcpyAsync(c->dcs, cs, nc * 8, hipMemcpyHostToDevice, c->stream));
  // grain sizes: one value per species when uniform over the grid (every reference setup); kept
  // for the outputs also where the stopping-time law does not read them
  a.sdust_uniform = 1;
  a.sdust = nullptr;
  for (int k = 0; k < SHARK_MAX_DUST; ++k) a.sdust_k[k] = 0.0;
  if (a.nd > 0 && !sdust && cf.tstop_kind != SHARK_TSTOP_DUSTYWAVE)
    return fail(SHARK_EINVAL, "sdust is required by this stopping-time law");
  if (a.nd > 0 && sdust) {
    for (int k = 0; k < a.nd; ++k) {
      const double s0 = sdust[(size_t)k * nc + (a.ndim == 2 ? 2 + 2 * (size_t)a.nxm : 2)];
      a.sdust_k[k] = s0;
      for (int iy = (a.ndim == 2 ? 3 : 1); iy <= (a.ndim == 2 ? a.ny + 2 : 1); ++iy)
        for (int ix = 3; ix <= a.nx + 2; ++ix)
          if (sdust[(size_t)k * nc + (ix - 1) + (size_t)(iy - 1) * a.nxm] != s0) a.sdust_uniform = 0;
    }
    if (!a.sdust_uniform) {
      if (!c->dsdust) HIPCHK(hipMalloc(&c->dsdust, nc * a.nd * 8));
      HIPCHK(hipMemcpyAsync(c->dsdust, sdust, nc * a.nd * 8, hipMemcpyHostToDevice, c->stream));
      a.sdust = c->dsdust;
    }
  }
  if (cf.inloop_kind == SHARK_INLOOP_DISK) {
    if (!condinit) return fail(SHARK_EINVAL, "the disk in-loop hook needs uprim_condinit");
    if (!c->dcond) HIPCHK(hipMalloc(&c->dcond, nc * a.nvar * 8));
    HIPCHK(hipMemcpyAsync(c->dcond, condinit, nc * a.nvar * 8, hipMemcpyHostToDevice, c->stream));
    a.condinit = c->dcond;
  }
  if (cf.fargo) {
    if (!fargo_velocity) return fail(SHARK_EINVAL, "fargo needs fargo_velocity");
    std::vector<double> vk(a.nxm);
    for (int ix = 1; ix <= a.nxm; ++ix) {
      vk[ix - 1] = fargo_velocity[ix - 1 + 2 * (size_t)a.nxm];
      for (int iy = 3; iy <= a.ny + 2; ++iy)
        if (fargo_velocity[(ix - 1) + (size_t)(iy - 1) * a.nxm] != vk[ix - 1] && ix >= 3 && ix <= a.nx + 2)
          return fail(SHARK_EUNSUP, "fargo_velocity varies along y at column %d", ix);
    }
    if (!c->dfargo) HIPCHK(hipMalloc(&c->dfargo, a.nxm * 8));
    HIPCHK(hipMemcpyAsync(c->dfargo, vk.data(), a.nxm * 8, hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    a.fargo_vk = c->dfargo;
  }
  // scalars: time 0, dt from the uploaded state (the first courant of the run)
  memset(c->hsc, 0, sizeof(DevScalars));
  c->hsc->dtbits = shark::dbits_host(2e44);
  c->hsc->active = 1;
  c->hsc->tend = INFINITY;
  HIPCHK(hipMemcpyAsync(c->dsc, c->hsc, sizeof(DevScalars), hipMemcpyHostToDevice, c->stream));
  shark::launch_courant_init(a, c->stream);
  HIPCHK(hipGetLastError());
  HIPCHK(hipStreamSynchronize(c->stream));
  c->uploaded = true;
  return SHARK_OK;
}

static void enqueue_step(shark_ctx* c) {
  shark::launch_bc(c->a, c->stream);
  shark::launch_hydro(c->a, c->stream);
  shark::launch_post(c->a, c->stream);
  shark::launch_finalize(c->a, c->stream);
}

static int build_graph(shark_ctx* c) {
  hipGraph_t g = nullptr;
  HIPCHK(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
  enqueue_step(c);
  HIPCHK(hipStreamEndCapture(c->stream, &g));
  hipError_t e = hipGraphInstantiate(&c->gexec, g, nullptr, nullptr, 0);
  (void)hipGraphDestroy(g);
  if (e != hipSuccess) {
    c->gexec = nullptr;
    return fail(SHARK_EHIP, "hipGraphInstantiate: %s", hipGetErrorString(e));
  }
  return SHARK_OK;
}

// Steps with dust growth between the drag and the in-loop hook (solve.f90:110-152): the post pass
// runs in two parts around the growth call, whose sub-step loop is driven from the host, so these
// steps are launched directly (no graph) and the run checks the device's active flag every step.
static int run_steps_growth(shark_ctx* c, int64_t n) {
  for (int64_t i = 0; i < n; ++i) {
    HIPCHK(hipMemcpyAsync(c->hsc, c->dsc, sizeof(DevScalars), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    if (!c->hsc->active) break;
    hipEvent_t e[KT_N + 2];
    if (c->timing)
      for (auto& x : e) HIPCHK(hipEventCreate(&x));
    if (c->timing) HIPCHK(hipEventRecord(e[0], c->stream));
    shark::launch_bc(c->a, c->stream);
    if (c->timing) HIPCHK(hipEventRecord(e[1], c->stream));
    shark::launch_hydro(c->a, c->stream);
    if (c->timing) HIPCHK(hipEventRecord(e[2], c->stream));
    shark::launch_post_part(c->a, 1, c->stream);
    if (c->timing) HIPCHK(hipEventRecord(e[3], c->stream));
    HIPCHK(hipGetLastError());
    int rc = shark::growth_step(c->growth, c->growth_mode);
    if (rc) return rc;
    if (c->timing) HIPCHK(hipEventRecord(e[4], c->stream));
    shark::launch_post_part(c->a, 2, c->stream);
    if (c->timing) HIPCHK(hipEventRecord(e[5], c->stream));
    shark::launch_finalize(c->a, c->stream);
    if (c->timing) HIPCHK(hipEventRecord(e[6], c->stream));
    HIPCHK(hipGetLastError());
    if (c->timing) {
      HIPCHK(hipStreamSynchronize(c->stream));
      float t[6];
      for (int k = 0; k < 6; ++k) HIPCHK(hipEventElapsedTime(&t[k], e[k], e[k + 1]));
      c->kt_ms[KT_BC] += t[0];
      c->kt_ms[KT_HYDRO] += t[1];
      c->kt_ms[KT_POST] += t[2] + t[4];
      c->kt_ms[KT_GROWTH] += t[3];
      c->kt_ms[KT_DT] += t[5];
      for (int k = 0; k < KT_N; ++k) c->kt_n[k] += 1;
      for (auto& x : e) (void)hipEventDestroy(x);
    }
  }
  return SHARK_OK;
}

static int run_steps(shark_ctx* c, int64_t n, double tend) {
  shark::launch_arm(c->a, tend, c->stream);
  HIPCHK(hipGetLastError());
  if (c->growth) return run_steps_growth(c, n);
  if (c->timing) {
    // bc, hydro, post, dt: four intervals between five events per step (no growth here)
    constexpr int NK = KT_DT + 1;
    std::vector<hipEvent_t> ev((size_t)n * (NK + 1));
    for (auto& e : ev) HIPCHK(hipEventCreate(&e));
    for (int64_t i = 0; i < n; ++i) {
      hipEvent_t* e = &ev[(size_t)i * (NK + 1)];
      HIPCHK(hipEventRecord(e[0], c->stream));
      shark::launch_bc(c->a, c->stream);
      HIPCHK(hipEventRecord(e[1], c->stream));
      shark::launch_hydro(c->a, c->stream);
      HIPCHK(hipEventRecord(e[2], c->stream));
      shark::launch_post(c->a, c->stream);
      HIPCHK(hipEventRecord(e[3], c->stream));
      shark::launch_finalize(c->a, c->stream);
      HIPCHK(hipEventRecord(e[4], c->stream));
    }
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(c->stream));
    for (int64_t i = 0; i < n; ++i) {
      hipEvent_t* e = &ev[(size_t)i * (NK + 1)];
      for (int k = 0; k < NK; ++k) {
        float ms = 0.f;
        HIPCHK(hipEventElapsedTime(&ms, e[k], e[k + 1]));
        c->kt_ms[k] += ms;
        c->kt_n[k] += 1;
      }
    }
    for (auto& e : ev) (void)hipEventDestroy(e);
    return SHARK_OK;
  }
  if (!c->gexec) {
    int rc = build_graph(c);
    if (rc) return rc;
  }
  for (int64_t i = 0; i < n; ++i) HIPCHK(hipGraphLaunch(c->gexec, c->stream));
  return SHARK_OK;
}

extern "C" int shark_gpu_run(shark_ctx* c, int64_t nsteps, double tend, int64_t* steps_done, double* time_out) {
  if (!c) return fail(SHARK_EINVAL, "null ctx");
  if (!c->uploaded) return fail(SHARK_ESTATE, "upload before stepping");
  if (nsteps < 0) return fail(SHARK_EINVAL, "nsteps < 0");
  HIPCHK(hipMemcpyAsync(c->hsc, c->dsc, sizeof(DevScalars), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  const int64_t before = c->hsc->nstep;
  int rc = run_steps(c, nsteps, tend);
  if (rc) return rc;
  HIPCHK(hipMemcpyAsync(c->hsc, c->dsc, sizeof(DevScalars), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  if (steps_done) *steps_done = c->hsc->nstep - before;
  if (time_out) *time_out = c->hsc->time;
  return SHARK_OK;
}

extern "C" int shark_gpu_step(shark_ctx* c, double* dt_out) {
  if (!c) return fail(SHARK_EINVAL, "null ctx");
  if (!c->uploaded) return fail(SHARK_ESTATE, "upload before stepping");
  int rc = run_steps(c, 1, INFINITY);
  if (rc) return rc;
  if (dt_out) {
    HIPCHK(hipMemcpyAsync(c->hsc, c->dsc, sizeof(DevScalars), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    *dt_out = c->hsc->dt_last;
  }
  return SHARK_OK;
}

extern "C" int shark_gpu_download(shark_ctx* c, double* u, double* q) {
  if (!c) return fail(SHARK_EINVAL, "null ctx");
  const size_t n = (size_t)c->a.nc * c->a.nvar;
  if (u) HIPCHK(hipMemcpyAsync(u, c->dU, n * 8, hipMemcpyDeviceToHost, c->stream));
  if (q) {
    shark::launch_ctoprim(c->a, c->dB, c->stream);
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(q, c->dB, n * 8, hipMemcpyDeviceToHost, c->stream));
  }
  HIPCHK(hipStreamSynchronize(c->stream));
  return SHARK_OK;
}

extern "C" int shark_gpu_get_time(shark_ctx* c, double* time, double* dt_last) {
  if (!c) return fail(SHARK_EINVAL, "null ctx");
  HIPCHK(hipMemcpyAsync(c->hsc, c->dsc, sizeof(DevScalars), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  if (time) *time = c->hsc->time;
  if (dt_last) *dt_last = c->hsc->dt_last;
  return SHARK_OK;
}

extern "C" int shark_gpu_set_time(shark_ctx* c, double time) {
  if (!c) return fail(SHARK_EINVAL, "null ctx");
  HIPCHK(hipMemcpyAsync(c->hsc, c->dsc, sizeof(DevScalars), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  c->hsc->time = time;
  HIPCHK(hipMemcpyAsync(c->dsc, c->hsc, sizeof(DevScalars), hipMemcpyHostToDevice, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  return SHARK_OK;
}

extern "C" int shark_gpu_set_timing(shark_ctx* c, int enable) {
  if (!c) return fail(SHARK_EINVAL, "null ctx");
  c->timing = enable != 0;
  for (int k = 0; k < KT_N; ++k) { c->kt_ms[k] = 0; c->kt_n[k] = 0; }
  return SHARK_OK;
}

extern "C" int shark_gpu_kernel_time(shark_ctx* c, const char* name, double* avg_ms, int64_t* count) {
  if (!c || !name) return fail(SHARK_EINVAL, "null argument");
  for (int k = 0; k < KT_N; ++k)
    if (strcmp(name, kKernelNames[k]) == 0) {
      if (avg_ms) *avg_ms = c->kt_n[k] ? c->kt_ms[k] / c->kt_n[k] : 0.0;
      if (count) *count = c->kt_n[k];
      return SHARK_OK;
    }
  return fail(SHARK_EINVAL, "unknown kernel name '%s'", name);
}

extern "C" int shark_gpu_attach_growth(shark_ctx* c, shark_growth* g, int mode) {
  if (!c) return fail(SHARK_EINVAL, "null ctx");
  if (!g) {
    c->growth = nullptr;
    return SHARK_OK;
  }
  if (!c->geom_set) return fail(SHARK_ESTATE, "set the geometry before attaching growth");
  if (mode < 0 || mode > 4) return fail(SHARK_EINVAL, "growth mode %d out of range", mode);
  if (shark::growth_nbins(g) != c->cfg.ndust)
    return fail(SHARK_EINVAL, "growth bins (%d) must equal ndust (%d)", shark::growth_nbins(g), c->cfg.ndust);
  std::vector<double> rc((size_t)c->a.nxm);
  HIPCHK(hipStreamSynchronize(c->stream));
  HIPCHK(hipMemcpy(rc.data(), c->dgeom + 5 * (size_t)c->a.nxm, rc.size() * 8, hipMemcpyDeviceToHost));
  int r = shark::growth_bind(g, c->dU, c->dcs, rc.data(), c->a.nxm, c->a.nym, c->a.nx, c->a.ny, c->a.ndim,
                             &c->dsc->dt, c->stream);
  if (r) return r;
  if (!c->dRold) {
    hipError_t e = hipMalloc(&c->dRold, sizeof(double) * (size_t)c->cfg.ndust * c->a.nc);
    if (e != hipSuccess) return fail(SHARK_ENOMEM, "growth scratch: %s", hipGetErrorString(e));
  }
  c->a.Rold = c->dRold;
  c->growth = g;
  c->growth_mode = mode;
  return SHARK_OK;
}

// ------------------------------------------------------------------ launch shapes and the tuner
static void apply_shape(shark_ctx* c, int strip_h, int strip_g, int tw, int xcd) {
  StepArgs& a = c->a;
  a.strip_h = strip_h;
  a.nstrips = a.ndim == 2 ? (a.ny + strip_h - 1) / strip_h : 1;
  a.strip_g = strip_g;
  a.nstrips_g = a.ndim == 2 ? (a.ny + strip_g - 1) / strip_g : 1;
  a.post_tw = tw;
  a.xcd_order = xcd;
  drop_graph(c);   // the captured step holds the old shape
}

extern "C" int shark_gpu_get_tune(shark_ctx* c, shark_tune* t) {
  if (!c || !t) return fail(SHARK_EINVAL, "null argument");
  memset(t, 0, sizeof(*t));
  t->strip_h = c->a.strip_h;
  t->strip_gas = c->a.strip_g;
  t->post_tw = c->a.post_tw;
  t->xcd_order = c->a.xcd_order;
  return SHARK_OK;
}

extern "C" int shark_gpu_set_tune(shark_ctx* c, const shark_tune* t) {
  if (!c || !t) return fail(SHARK_EINVAL, "null argument");
  if (t->strip_h < 1 || t->strip_h > 4096) return fail(SHARK_EINVAL, "strip_h %d outside 1..4096", t->strip_h);
  if (t->strip_gas < 1 || t->strip_gas > 4096)
    return fail(SHARK_EINVAL, "strip_gas %d outside 1..4096", t->strip_gas);
  if (!shark::post_tw_ok(c->a.nd, t->post_tw))
    return fail(SHARK_EINVAL, "post_tw %d not built for %d species", t->post_tw, c->a.nd);
  if (t->xcd_order != 0 && t->xcd_order != 1) return fail(SHARK_EINVAL, "xcd_order must be 0 or 1");
  HIPCHK(hipStreamSynchronize(c->stream));
  apply_shape(c, t->strip_h, t->strip_gas, t->post_tw, t->xcd_order);
  return SHARK_OK;
}

namespace {
struct TuneSave {
  double* U = nullptr;
  double* B = nullptr;
  DevScalars* sc = nullptr;
  ~TuneSave() {
    if (U) (void)hipFree(U);
    if (B) (void)hipFree(B);
    if (sc) (void)hipFree(sc);
  }
};

float median(std::vector<float>& v) {
  std::sort(v.begin(), v.end());
  return v[v.size() / 2];
}
}  // namespace

static int restore_state(shark_ctx* c, const TuneSave& sv, size_t nbytes) {
  HIPCHK(hipMemcpyAsync(c->dU, sv.U, nbytes, hipMemcpyDeviceToDevice, c->stream));
  HIPCHK(hipMemcpyAsync(c->dB, sv.B, nbytes, hipMemcpyDeviceToDevice, c->stream));
  HIPCHK(hipMemcpyAsync(c->dsc, sv.sc, sizeof(DevScalars), hipMemcpyDeviceToDevice, c->stream));
  return SHARK_OK;
}

// Median gas-hydro, dust-hydro and post times (ms) of `reps` steps of the current shape, each run
// from the saved state after one untimed step.
static int time_shape(shark_ctx* c, const TuneSave& sv, size_t nbytes, int reps, double* tg, double* td,
                      double* post) {
  int rc = restore_state(c, sv, nbytes);
  if (rc) return rc;
  shark::launch_arm(c->a, INFINITY, c->stream);
  std::vector<hipEvent_t> ev((size_t)4 * (reps + 1));
  for (auto& e : ev) HIPCHK(hipEventCreate(&e));
  for (int i = 0; i <= reps; ++i) {
    hipEvent_t* e = &ev[(size_t)4 * i];
    shark::launch_bc(c->a, c->stream);
    HIPCHK(hipEventRecord(e[0], c->stream));
    shark::launch_hydro_fluid(c->a, true, c->stream);
    HIPCHK(hipEventRecord(e[1], c->stream));
    shark::launch_hydro_fluid(c->a, false, c->stream);
    HIPCHK(hipEventRecord(e[2], c->stream));
    shark::launch_post(c->a, c->stream);
    HIPCHK(hipEventRecord(e[3], c->stream));
    shark::launch_finalize(c->a, c->stream);
  }
  HIPCHK(hipGetLastError());
  HIPCHK(hipStreamSynchronize(c->stream));
  std::vector<float> t0, t1, t2;
  for (int i = 1; i <= reps; ++i) {   // step 0 is the untimed warm-up
    hipEvent_t* e = &ev[(size_t)4 * i];
    float a = 0.f, b = 0.f, d = 0.f;
    HIPCHK(hipEventElapsedTime(&a, e[0], e[1]));
    HIPCHK(hipEventElapsedTime(&b, e[1], e[2]));
    HIPCHK(hipEventElapsedTime(&d, e[2], e[3]));
    t0.push_back(a);
    t1.push_back(b);
    t2.push_back(d);
  }
  for (auto& e : ev) (void)hipEventDestroy(e);
  *tg = median(t0);
  *td = median(t1);
  *post = median(t2);
  return SHARK_OK;
}

extern "C" int shark_gpu_autotune(shark_ctx* c, int reps, shark_tune* chosen) {
  if (!c) return fail(SHARK_EINVAL, "null ctx");
  if (!c->uploaded) return fail(SHARK_ESTATE, "upload before tuning");
  if (reps < 1 || reps > 1000) return fail(SHARK_EINVAL, "reps %d outside 1..1000", reps);
  StepArgs& a = c->a;
  const size_t nbytes = (size_t)a.nc * a.nvar * sizeof(double);
  TuneSave sv;
  hipError_t e;
  if ((e = hipMalloc(&sv.U, nbytes)) != hipSuccess || (e = hipMalloc(&sv.B, nbytes)) != hipSuccess ||
      (e = hipMalloc(&sv.sc, sizeof(DevScalars))) != hipSuccess)
    return fail(SHARK_ENOMEM, "autotune state copy: %s", hipGetErrorString(e));
  HIPCHK(hipMemcpyAsync(sv.U, c->dU, nbytes, hipMemcpyDeviceToDevice, c->stream));
  HIPCHK(hipMemcpyAsync(sv.B, c->dB, nbytes, hipMemcpyDeviceToDevice, c->stream));
  HIPCHK(hipMemcpyAsync(sv.sc, c->dsc, sizeof(DevScalars), hipMemcpyDeviceToDevice, c->stream));

  int best_h = a.strip_h, best_g = a.strip_g, best_tw = a.post_tw, best_x = a.xcd_order;
  double tg = 0, td = 0, tp = 0, bh = 0, bp = 0;
  int rc = 0;
  // 1. strip heights (2D only: a 1D run has one strip): dust by dust-hydro time, gas by gas-hydro
  //    time (the gas launch has one work item per strip and column group, so it wants short strips)
  if (a.ndim == 2) {
    static const int kStrip[] = {4, 8, 16, 24, 32, 48, 64, 96, 128, 256};
    double btd = INFINITY, btg = INFINITY;
    for (int h : kStrip) {
      if (h > a.ny && h != kStrip[0]) break;
      apply_shape(c, h, h, best_tw, best_x);
      if ((rc = time_shape(c, sv, nbytes, reps, &tg, &td, &tp))) break;
      if (a.nd > 0 && td < btd) { btd = td; best_h = h; }
      if (tg < btg) { btg = tg; best_g = h; }
    }
  }
  // 2. post segment width by post time
  if (!rc) {
    int tws[4];
    const int ntw = shark::post_tw_candidates(a.nd, tws);
    double bt = INFINITY;
    for (int k = 0; k < ntw && !rc; ++k) {
      apply_shape(c, best_h, best_g, tws[k], best_x);
      if ((rc = time_shape(c, sv, nbytes, reps, &tg, &td, &tp))) break;
      if (tp < bt) { bt = tp; best_tw = tws[k]; }
    }
  }
  // 3. XCD order by hydro + post time
  if (!rc) {
    double bt = INFINITY;
    for (int x = 1; x >= 0 && !rc; --x) {
      apply_shape(c, best_h, best_g, best_tw, x);
      if ((rc = time_shape(c, sv, nbytes, reps, &tg, &td, &tp))) break;
      if (tg + td + tp < bt) { bt = tg + td + tp; best_x = x; bh = tg + td; bp = tp; }
    }
  }
  // the run continues from the state it had before tuning, with the chosen shape
  const int rr = restore_state(c, sv, nbytes);
  apply_shape(c, best_h, best_g, best_tw, best_x);
  HIPCHK(hipStreamSynchronize(c->stream));
  if (rc) return rc;
  if (rr) return rr;
  if (chosen) {
    memset(chosen, 0, sizeof(*chosen));
    chosen->strip_h = best_h;
    chosen->strip_gas = best_g;
    chosen->post_tw = best_tw;
    chosen->xcd_order = best_x;
    chosen->hydro_ms = bh;
    chosen->post_ms = bp;
  }
  return SHARK_OK;
}

// ------------------------------------------------------------------ outputs and restart dumps
extern "C" int shark_gpu_write_output(shark_ctx* c, const shark_output_info* info, int iout) {
  if (!c) return fail(SHARK_EINVAL, "null ctx");
  if (!c->uploaded) return fail(SHARK_ESTATE, "upload before writing an output");
  std::string err;
  if (!shark::output_info_ok(info, &err)) return fail(SHARK_EINVAL, "%s", err.c_str());
  if (c->hpos.empty()) return fail(SHARK_ESTATE, "outputs need the position array at shark_gpu_set_geometry");
  const StepArgs& a = c->a;
  const size_t nc = a.nc;
  std::vector<double> q(nc * a.nvar);
  int rc = shark_gpu_download(c, nullptr, q.data());
  if (rc) return rc;
  std::vector<double> sd;
  if (a.nd > 0) {
    sd.resize(nc * a.nd);
    if (a.sdust_uniform) {
      for (int k = 0; k < a.nd; ++k) std::fill(sd.begin() + (size_t)k * nc, sd.begin() + (size_t)(k + 1) * nc, a.sdust_k[k]);
    } else if (a.sdust) {
      HIPCHK(hipMemcpy(sd.data(), a.sdust, sd.size() * 8, hipMemcpyDeviceToHost));
    } else {
      return fail(SHARK_ESTATE, "grain sizes were not uploaded");
    }
  }
  shark::OutputGrid g;
  g.nx = a.nx;
  g.ny = a.ny;
  g.ndim = a.ndim;
  g.nd = a.nd;
  g.nxm = a.nxm;
  g.geom = a.geom;
  g.nc = a.nc;
  if (shark::write_output_dir(info, iout, g, q.data(), a.nd > 0 ? sd.data() : nullptr, c->hpos.data(),
                              c->hrc.data(), &err))
    return fail(SHARK_EIO, "%s", err.c_str());
  return SHARK_OK;
}

namespace {
constexpr char kDumpMagic[8] = {'S', 'H', 'A', 'R', 'K', 'S', 'T', '2'};
struct DumpHead {
  char magic[8];
  int32_t nx, ny, ndust, nvar, geom, scalars_bytes;   // scalars_bytes: sizeof(DevScalars) written
  int64_t ncells;
};
}  // namespace

extern "C" int shark_gpu_save_state(shark_ctx* c, const char* path) {
  if (!c || !path) return fail(SHARK_EINVAL, "null argument");
  if (!c->uploaded) return fail(SHARK_ESTATE, "upload before saving a dump");
  const StepArgs& a = c->a;
  const size_t n = (size_t)a.nc * a.nvar;
  std::vector<double> u(n);
  HIPCHK(hipMemcpyAsync(u.data(), c->dU, n * 8, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipMemcpyAsync(c->hsc, c->dsc, sizeof(DevScalars), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  DumpHead h;
  memset(&h, 0, sizeof(h));
  memcpy(h.magic, kDumpMagic, 8);
  h.nx = a.nx; h.ny = a.ny; h.ndust = a.nd; h.nvar = a.nvar; h.geom = a.geom;
  h.scalars_bytes = (int32_t)sizeof(DevScalars);
  h.ncells = a.nc;
  FILE* f = fopen(path, "wb");
  if (!f) return fail(SHARK_EIO, "cannot open %s: %s", path, strerror(errno));
  bool ok = fwrite(&h, sizeof(h), 1, f) == 1 && fwrite(c->hsc, sizeof(DevScalars), 1, f) == 1 &&
            fwrite(u.data(), 8, n, f) == n;
  ok = (fclose(f) == 0) && ok;
  if (!ok) return fail(SHARK_EIO, "short write to %s", path);
  return SHARK_OK;
}

extern "C" int shark_gpu_load_state(shark_ctx* c, const char* path) {
  if (!c || !path) return fail(SHARK_EINVAL, "null argument");
  if (!c->uploaded) return fail(SHARK_ESTATE, "upload the setup's fields before loading a dump");
  const StepArgs& a = c->a;
  const size_t n = (size_t)a.nc * a.nvar;
  FILE* f = fopen(path, "rb");
  if (!f) return fail(SHARK_EIO, "cannot open %s: %s", path, strerror(errno));
  DumpHead h;
  DevScalars sc;
  std::vector<double> u(n);
  bool ok = fread(&h, sizeof(h), 1, f) == 1 && memcmp(h.magic, kDumpMagic, 8) == 0 &&
            h.scalars_bytes == (int32_t)sizeof(DevScalars);
  if (ok && (h.nx != a.nx || h.ny != a.ny || h.ndust != a.nd || h.nvar != a.nvar || h.geom != a.geom ||
             h.ncells != a.nc)) {
    fclose(f);
    return fail(SHARK_EINVAL, "%s holds a %dx%d grid with %d species (geometry %d), not this configuration",
                path, h.nx, h.ny, h.ndust, h.geom);
  }
  ok = ok && fread(&sc, sizeof(sc), 1, f) == 1 && fread(u.data(), 8, n, f) == n;
  fclose(f);
  if (!ok) return fail(SHARK_EIO, "%s is not a complete dump of this library", path);
  HIPCHK(hipStreamSynchronize(c->stream));
  HIPCHK(hipMemcpyAsync(c->dU, u.data(), n * 8, hipMemcpyHostToDevice, c->stream));
  *c->hsc = sc;
  HIPCHK(hipMemcpyAsync(c->dsc, c->hsc, sizeof(DevScalars), hipMemcpyHostToDevice, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  return SHARK_OK;
}

extern "C" int shark_gpu_destroy(shark_ctx* c) {
  if (!c) return SHARK_OK;
  if (c->stream) (void)hipStreamSynchronize(c->stream);
  drop_graph(c);
  // an attached growth state belongs to the caller (shark_growth_destroy after this call)
  (void)hipFree(c->dU); (void)hipFree(c->dB); (void)hipFree(c->dcs); (void)hipFree(c->dgeom); (void)hipFree(c->dsc);
  if (c->dRold) (void)hipFree(c->dRold);
  if (c->dfcs) (void)hipFree(c->dfcs);
  if (c->dmc) (void)hipFree(c->dmc);
  if (c->dposxy) (void)hipFree(c->dposxy);
  if (c->dphisg) (void)hipFree(c->dphisg);
  if (c->dfexp) (void)hipFree(c->dfexp);
  if (c->ddvol) (void)hipFree(c->ddvol);
  if (c->dsdust) (void)hipFree(c->dsdust);
  if (c->dcond) (void)hipFree(c->dcond);
  if (c->dfargo) (void)hipFree(c->dfargo);
  if (c->hsc) (void)hipHostFree(c->hsc);
  if (c->stream) (void)hipStreamDestroy(c->stream);
  delete c;
  return SHARK_OK;
}
