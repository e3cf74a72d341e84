/*
 * shark_gpu.h — C-ABI of libshark_gpu.so, the gfx950 HIP implementation of SHARK's
 * Godunov step for isothermal gas + N pressureless dust fluids.
 *
 * The reference has no FFI: its hot path is the parameterless Fortran subroutine
 * `solve(verbose,outputing)` (src/solver/solve.f90:10-181) acting on module-global arrays
 * allocated in allocate_init (src/commons/allocate_init.f90:3-282).  This header is the
 * boundary a Fortran driver binds with ISO_C_BINDING (shark_amd/fortran/shark_gpu_mod.f90):
 * every entry point takes plain pointers and sizes; arrays are the reference's own layout,
 * `real(dp) u_prim(1:ncells,1:nvar)` column-major (one contiguous plane of ncells doubles
 * per variable), cells x-fastest, icell = 1+(ix-1)+(iy-1)*nx_max (src/grid/grid.f90:12),
 * variable order rho, rho*vx, rho*vy, rho*vz, E, rhod(1..N), rhod*vdx(1..N),
 * rhod*vdy(1..N), rhod*vdz(1..N) (src/commons/allocate_init.f90:225-262).
 *
 * Error convention: 0 on success, a positive SHARK_E* code otherwise; the message is
 * available from shark_gpu_last_error() (the reference prints and stops,
 * src/commons/read_params.f90:18-33).  Calls are not thread-safe; one context owns one
 * HIP stream and its device buffers.
 */
#ifndef SHARK_GPU_H
#define SHARK_GPU_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SHARK_CFG_VERSION 2
#define SHARK_MAX_DUST 64

enum {
  SHARK_OK = 0,
  SHARK_EINVAL = 1,   /* bad configuration or argument */
  SHARK_EUNSUP = 2,   /* configuration outside the implemented scope */
  SHARK_EHIP = 3,     /* HIP runtime error */
  SHARK_ESTATE = 4,   /* call out of order (e.g. step before upload) */
  SHARK_ENOMEM = 5,
  SHARK_EIO = 6       /* file could not be written or read */
};

/* geometry (reference GEOM macro, bin/Makefile:14) */
enum { SHARK_GEOM_CARTESIAN = 0, SHARK_GEOM_SPHERE = 1, SHARK_GEOM_POLAR = 2 };
/* gas Riemann solver (SOLVER macro, src/solver/godunov.f90:783-791) */
enum { SHARK_SOLVER_LLF = 0, SHARK_SOLVER_HLL = 1, SHARK_SOLVER_HLLC = 2 };
/* dust Riemann solver (SOLVERDUST macro, src/solver/godunov.f90:797-813) */
enum { SHARK_DUST_HUANG_BAI = 0, SHARK_DUST_LLF = 1, SHARK_DUST_HLL = 2, SHARK_DUST_HLLC = 3 };
/* per-step setup hooks, resolved at link time in the reference (bin/Makefile:67) */
enum { SHARK_BC_PERIODIC = 0,   /* src/grid/boundary_types.f90:149-186 (DUSTYWAVE) */
       SHARK_BC_SHEAR_SI = 1,   /* src/setups/SI/boundaries.f90:10-85 */
       SHARK_BC_DISK = 2,       /* src/setups/disk_polar/boundaries.f90:10-62 */
       SHARK_BC_SOD = 3,        /* src/setups/SOD/boundaries.f90:10-80 */
       SHARK_BC_ZERO_GRADIENT = 4,   /* src/grid/boundary_types.f90:48-100 (cartesian disk) */
       SHARK_BC_COLLAPSE = 5 };      /* src/grid/boundary_types.f90:4-44 (Collapse_1D) */
enum { SHARK_TSTOP_DUSTYWAVE = 0,  /* src/setups/DUSTYWAVE/setup.f90:217 */
       SHARK_TSTOP_SI = 1,         /* src/setups/SI/setup.f90:392 */
       SHARK_TSTOP_EPSTEIN_DISK = 2, /* src/setups/disk_polar/setup.f90:436 */
       SHARK_TSTOP_EPSTEIN = 3,      /* src/setups/KH/setup.f90:238, SOD/setup.f90:182: uniform medium,
                                        sqrt(pi gamma/8) rhograin s / (rho cs), rhograin in code units */
       SHARK_TSTOP_EPSTEIN_CDISK = 4, /* src/setups/disk/setup.f90:322-345 (cartesian disk) */
       SHARK_TSTOP_COLLAPSE = 5 };    /* src/setups/Collapse_1D/setup.f90:296-324: Epstein with the
                                         barotropic cs */
enum { SHARK_FORCE_NONE = 0,
       SHARK_FORCE_SI = 1,          /* src/setups/SI/setup.f90:335-368 */
       SHARK_FORCE_DISK = 2,        /* src/setups/disk_polar/setup.f90:360-403 */
       SHARK_FORCE_CDISK = 3 };     /* src/setups/disk/setup.f90:265-318: softened star (+ planet) */
enum { SHARK_INLOOP_NONE = 0,
       SHARK_INLOOP_SI = 1,         /* src/setups/SI/setup.f90:299-332 */
       SHARK_INLOOP_DISK = 2,       /* src/setups/disk_polar/setup.f90:206-284 */
       SHARK_INLOOP_CDISK = 3 };    /* src/setups/disk/setup.f90:203-262: buffer zones reset */

/* Configuration: the reference's cpp macro set (bin/Makefile:5-36) plus the run-time
 * namelist scalars the step reads (&GRID_PARAMS src/commons/read_params.f90:10,
 * &DUST_PARAMS src/dust/dust_commons.f90, &SETUP_PARAMS of the setup). */
typedef struct shark_cfg {
  int32_t version;          /* SHARK_CFG_VERSION */
  int32_t nx, ny, ndust;    /* active cells per direction (ny=1 means 1D), dust species */
  int32_t nghost;           /* must be 2 (bin/Makefile:8) */
  int32_t geom, solver, solverdust, slope_type;
  int32_t iso_cs;           /* 1: isothermal (setup's cs); below 1: adiabatic (P from E, gamma) */
  int32_t force_kick, fargo, drag, dust_back_reaction, growth, viscosity;
  int32_t bc_kind, tstop_kind, force_kind, inloop_kind;
  int32_t disk_bc_mode;     /* 0: reference indexing (boundaries.f90:45-46), 1: per-row copy */
  int32_t shock_dir;        /* SOD: direction_shock (0: along x, 1: along y) */
  double gamma, cfl, smallr, smallc, smallp, dust_ratio_min;
  /* DUSTYWAVE */
  double K[SHARK_MAX_DUST];
  /* SI shearing box */
  double q_shear, omega_shear, box_l, rad0, eta_stream, rho_init, hoverr;
  double rhograin;          /* grain material density in code units (rhograin/unit_d) */
  /* disk_polar */
  double mstar, r_relax, n_rel, smooth_r;
  /* SOD (src/setups/SOD/setup_commons.f90): left and right states of the shock tube */
  double rho_l, rho_r, p_l, p_r;
  /* cartesian disk (src/setups/disk/setup_commons.f90, after setup's unit conversion; box_l,
   * mstar, smooth_r, hoverr above); dust2gas of &DUST_PARAMS; alpha_visc sets
   * eta_visc = alpha_visc H cs (setup.f90:60) when viscosity is on */
  double box_l_y, mplanet, rplanet, sigma_r0, r0_disk, frac_r_disk, buffer_in, buffer_out;
  double alpha_visc, dust2gas;
  double reserved[2];
  /* disk_polar accretion through the outer boundary (boundaries.f90:63-94): accretion = 1 sets
   * the two outer ghost columns and the last active column to the inflow state of the accretion
   * rate m_acc (code units: M_acc/(365.25*24*3600)*unit_t) at the local cs, with phi_mom
   * (only 2, every row) and L_out; box_l above is the disk's radial extent */
  int32_t accretion;
  /* 1D self-gravity (GRAVITY=1, src/gravity/gravity_utils.f90:1-51): the mass Mc enclosed within
   * each cell centre acts as -Mc/(r^2 + l_soft^2) in the predictor, the sources and the Courant
   * step (godunov.f90:243-246, 442-445; source_terms.f90:79-85; courant.f90:62-63) */
  int32_t gravity;
  double m_acc, phi_mom, l_out;
  /* non_standard_eos = 1 (Collapse_1D, src/setups/Collapse_1D/eos.f90): with iso_cs < 1 the
   * fluxes stay adiabatic but P = rho cs_eos(barotrop(rho))^2 in every cell and face */
  int32_t eos_baro;
  /* disk_polar self-gravity (&GRAVITY_PARAMS self_gravity, disk_polar/setup.f90:300-322): the
   * potential of the disk's own active cells, a direct sum over all of them for every cell */
  int32_t self_gravity;
  double l_soft;            /* gravity softening in code units (l_soft/unit_l) */
  double m_cent;            /* central point mass (M_cent) */
  double t0_cloud;          /* barotrop's T0_cloud (K) */
  double mu_gas;            /* cs_eos's mean molecular weight */
  double unit_v, unit_nh;   /* the setup's units (src/setups/Collapse_1D/units.f90) */
  /* disk_polar planet (setup.f90:323-378): with mplanet > 0 (mplanet, rplanet above, code units)
   * a softened point mass on a circular orbit adds the gradient of its potential to the forces */
  double a_smooth_pl;
} shark_cfg;

typedef struct shark_ctx shark_ctx;

/* Allocates device state for the configuration. */
int shark_gpu_create(const shark_cfg* cfg, shark_ctx** ctx);

/* Grid arrays as built by gridinit (src/grid/grid.f90:104-240): dx(ncells,ndim),
 * surf(ncells,ndim), vol(ncells), radii_c(ncells), position(ncells,ndim). */
int shark_gpu_set_geometry(shark_ctx* ctx, const double* dx, const double* surf,
                           const double* vol, const double* radii_c, const double* position);

/* Self-gravity on the spherical 1D grid: dvol(ncells) of gridinit_sphere1D
 * (src/grid/grid.f90:343-346), the part of each cell between its centre and its outer face,
 * which mtot (src/gravity/gravity_utils.f90:11-51) subtracts to get the mass within the cell
 * centre.  Required before shark_gpu_upload when cfg.gravity = 1; replaces the reference's
 * module array dvol. */
int shark_gpu_set_dvol(shark_ctx* ctx, const double* dvol);

/* State after setup: u_prim(ncells,nvar), cs(ncells), sdust(ncells,ndust) (may be NULL when
 * the stopping-time law does not read it), uprim_condinit(ncells,nvar) (disk setup only,
 * else NULL), fargo_velocity(ncells) (when fargo, else NULL). Copies; pointers are not kept. */
int shark_gpu_upload(shark_ctx* ctx, const double* u_prim, const double* cs,
                     const double* sdust, const double* uprim_condinit,
                     const double* fargo_velocity);

/* One call == one `solve` + `time = time + dt` of the reference time loop
 * (src/shark.f90:102-107).  dt_out (may be NULL) receives the step's dt; passing NULL lets
 * consecutive steps run without a host synchronisation. */
int shark_gpu_step(shark_ctx* ctx, double* dt_out);

/* Runs up to nsteps steps, stopping (on the device, without host round trips) once
 * time >= tend as flag_continue does (e.g. src/setups/DUSTYWAVE/setup.f90:136).
 * Writes the number of steps taken and the final time. */
int shark_gpu_run(shark_ctx* ctx, int64_t nsteps, double tend, int64_t* steps_done,
                  double* time_out);

/* Copies u_prim (ncells,nvar) and, if q != NULL, the primitive variables computed by
 * ctoprim (src/solver/solve.f90:195-253) from it. Either pointer may be NULL. */
int shark_gpu_download(shark_ctx* ctx, double* u_prim, double* q);

/* Current simulated time and the dt of the last step. */
int shark_gpu_get_time(shark_ctx* ctx, double* time, double* dt_last);
int shark_gpu_set_time(shark_ctx* ctx, double time);

/* Per-kernel average device time (ms) of the last shark_gpu_run, measured with HIP events
 * on the context's stream. names: "bc", "hydro", "post", "dt". Returns SHARK_EINVAL for an
 * unknown name. */
int shark_gpu_kernel_time(shark_ctx* ctx, const char* name, double* avg_ms, int64_t* count);
int shark_gpu_set_timing(shark_ctx* ctx, int enable);

/* Algorithmic HBM bytes per step of the plain step (one read + one write of u_prim plus cs,
 * plus a second read + write pass when FARGO is on). */
int64_t shark_gpu_algorithmic_bytes(const shark_cfg* cfg);

/* Launch shapes of the step kernels.  No counterpart in the reference (its loops have no launch
 * shape).  Every shape gives bit-identical results; only the speed differs.
 *   strip_h   rows per dust hydro work item (one wavefront marches strip_h rows of a 60-column
 *             group of one species; 2 extra rows are re-read at each end), 1..4096
 *   strip_gas the same for the gas launch, 1..4096
 *   post_tw   cells per post row segment: 16 or 32 up to 16 species, 8 or 16 for 64 species
 *   xcd_order 1: blocks are renumbered so each of the 8 XCDs takes a contiguous band of work
 *             (neighbouring rows share one L2); 0: hardware order */
typedef struct shark_tune {
  int32_t strip_h, post_tw, xcd_order, strip_gas;
  int32_t reserved_i[4];
  double hydro_ms, post_ms;   /* autotune: kernel times of the chosen shape (ms per step) */
  double reserved[6];
} shark_tune;

/* Current shape, and a new one (validated; takes effect from the next step). */
int shark_gpu_get_tune(shark_ctx* ctx, shark_tune* t);
int shark_gpu_set_tune(shark_ctx* ctx, const shark_tune* t);

/* The autotuner: times the step kernels of this context on its own uploaded state over a set of
 * candidate shapes (gas and dust strip heights 4..256, both segment widths, XCD order on and
 * off; coordinate search, `reps` timed steps per candidate, median kernel time) and keeps the
 * fastest.  The
 * state, time and dt are restored afterwards, so the run continues exactly as without tuning.
 * Needs device memory for one copy of u_prim and of the post-hydro state. */
int shark_gpu_autotune(shark_ctx* ctx, int reps, shark_tune* chosen);

int shark_gpu_destroy(shark_ctx* ctx);
const char* shark_gpu_last_error(void);

/* HBM bandwidth a plain streaming kernel attains on the current device (no counterpart in the
 * reference: the measurement the step's achieved bandwidth is rated against beside the vendor
 * peak).  Two buffers of nbytes (>= 1 MiB) are allocated and freed; each of write, copy and read
 * runs `reps` times after one untimed pass, 16-byte accesses, grid-stride over 8 blocks per CU.
 * gbs[0] write, gbs[1] copy (read + write bytes), gbs[2] read, in GB/s. */
int shark_hbm_probe(int64_t nbytes, int reps, double* gbs);
int shark_gpu_version(void);

#ifdef __cplusplus
}
#endif
#endif
