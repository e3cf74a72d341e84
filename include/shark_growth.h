/*
 * shark_growth.h — C-ABI of the dust-growth (Smoluchowski coagulation / fragmentation)
 * step in libshark_gpu.so.
 *
 * Replaces `dust_growth(verbose)` (src/dust/dust_growth.f90:3-256), called by `solve`
 * after the drag (src/solver/solve.f90:112).  The reference's growth runs per cell with an
 * explicit sub-cycled pair loop over the size bins; its disk branch is empty
 * (dust_growth.f90:80-82).  This library provides
 *   mode 0  FP64 direct pair loop (the plain coagulation; reference semantics per cell),
 *   mode 1  the pair-rate -> bin-gain contraction on MFMA in FP32,
 *   mode 2  the same contraction with BF16 operands and FP32 accumulation,
 * both MFMA modes with compensated FP32 accumulation and a per-cell FP64 mass rescale.
 * The state is the dust densities rhod(ncells, nbins), column-major like u_prim's dust
 * planes (src/commons/allocate_init.f90:259).
 */
#ifndef SHARK_GROWTH_H
#define SHARK_GROWTH_H
#include <stdint.h>
#ifdef __cplusplus
extern "C" {
#endif

#define SHARK_GROWTH_VERSION 1
#define SHARK_GROWTH_MAX_BINS 64

enum { SHARK_GROWTH_ENV_COLLAPSE = 0, SHARK_GROWTH_ENV_DISK = 1 };
enum { SHARK_GROWTH_FP64 = 0, SHARK_GROWTH_MFMA_FP32 = 1, SHARK_GROWTH_MFMA_BF16 = 2,
       SHARK_GROWTH_MFMA_BF16_PLAIN = 3, SHARK_GROWTH_SURROGATE = 4 };
/* return codes (0 = success; shark_gpu_last_error() holds the message) */
enum { SHARK_GROWTH_EINVAL = 1, SHARK_GROWTH_EHIP = 3, SHARK_GROWTH_ENUM = 6 };

/* &DUST_PARAMS growth switches and grain properties (src/dust/dust_commons.f90:85-143) */
typedef struct shark_growth_params {
  int32_t version;        /* SHARK_GROWTH_VERSION */
  int32_t nbins;          /* NDUST */
  int32_t env;            /* SHARK_GROWTH_ENV_* */
  int32_t kernel_type;    /* 0 physical, 1 constant, 2 additive (dust_growth.f90:150-162) */
  int32_t fragmentation;  /* fragmentation flag */
  int32_t frag_mode;      /* 0 energy threshold (dust_growth.f90:142-144), 1 velocity threshold (:145) */
  int32_t turb, brownian, drift, ambipolar;   /* *_in_growth flags */
  int32_t decomposed;     /* 1: bin targets independent of the fragmented fraction */
  int32_t max_substeps;   /* guard on the per-cell sub-step loop (dust_growth.f90:117) */
  double smin, smax, rhograin, ice_mantle, slope_mono, size_mono, abr, gamma_grains, estar_grains;
  double turb_coef;       /* turbulent velocity = turb_coef * cs (alpha_turb in the reference) */
  double vfrag;           /* cm/s */
  double cfl_growth, eps_threshold, eps_threshold_frag, dust_ratio_min, sticking, clustered;
  double delta_vambi;
  double unit_l, unit_d, unit_t, unit_m, unit_v;
  double mu_gas, gamma;
  double hoverr, mstar;   /* disk environment */
  double reserved[8];
} shark_growth_params;

typedef struct shark_growth shark_growth;
struct shark_mlp;

/* Growth state for ncells cells.  Rejects (SHARK_GROWTH_EINVAL) the collapse environment, and
 * fragmentation with decomposed=0, whose target bins depend on the cell. */
int shark_growth_create(const shark_growth_params* p, int64_t ncells, shark_growth** g);
/* rhod (ncells, nbins), gas density (ncells), sound speed (ncells), radius (ncells). */
int shark_growth_upload(shark_growth* g, const double* rhod, const double* rho_gas,
                        const double* cs, const double* radius);
/* One growth call over dt (dust_growth with time_growth from 0 to dt) in the given mode.
 * SHARK_GROWTH_ENUM when a density turns non-finite or a sub-step makes no progress. */
int shark_growth_run(shark_growth* g, int mode, double dt);
/* rhod after the last run, and per-cell sub-step counts (either may be NULL). */
int shark_growth_download(shark_growth* g, double* rhod, int64_t* nsub);
/* Grain charging of the current state (analytical charge, src/chemistry/ionisation.f90:809-922,
 * with the Newton-Raphson of src/commons/newton_raphson.f90): per cell the number-weighted mean
 * grain charge and the ion fraction zeta / sum_k k_i,k n_k, in the disk environment's cgs
 * midplane densities.  zeta: cosmic-ray ionisation rate (s^-1); mu_ions: ion mass in m_H;
 * electrons: include electron capture.  Outputs (ncells) may be NULL. */
/* Grain charging by the model of Marchand et al. 2021 (`charge`, src/chemistry/ionisation.f90:4-214):
 * per cell a Newton iteration on the grain surface potential psi (from the cell's psi_old when
 * non-zero, else from the global initial guess), then the grain charges zd, the ion and electron
 * densities, and with `resistivities` the Ohmic, Hall and ambipolar conductivities and
 * resistivities and the dust Hall factors.  Parameters are the reference's ionisation_commons
 * values and the setup's units. */
typedef struct shark_marchand_params {
  double zeta;             /* x: cosmic-ray ionisation rate (s^-1) */
  double mu_ions;          /* ion mass (m_H) */
  double stickeff_el;      /* electron sticking coefficient */
  double epsilon_ionis;    /* Newton tolerance on |f(psi)| */
  int32_t nitermax_ionis;  /* Newton iteration cap */
  int32_t resistivities;   /* res_Marchand */
  double B_0_lee;          /* B at n_H = 1e4 cm^-3 (G) */
  double rhograin;         /* grain material density (g cm^-3) */
  double gamma;            /* adiabatic index (stopping time in the resistivities) */
  double unit_l, unit_d, unit_m, unit_v, unit_nh;
} shark_marchand_params;

/* On the GPU over ncells cells and nbins species (host arrays, column-major (ncells, nbins) as the
 * reference's): rho, rhod, T (barotrop), cs (cs_eos(T), code units), sdust, mdust (code units);
 * psi_old (in/out).  Outputs: zd (ncells, nbins), ni, ne, and when resistivities is set
 * cond = eta_o | eta_h | eta_a | sigma_o | sigma_p | sigma_h (6 x ncells) and gamma_d
 * (ncells, nbins); cond and gamma_d may be NULL. */
int shark_charge_marchand(const shark_marchand_params* p, int64_t ncells, int nbins, const double* rho,
                          const double* rhod, const double* T, const double* cs, const double* sdust,
                          const double* mdust, double* psi_old, double* zd, double* ni, double* ne,
                          double* cond, double* gamma_d);

/* shark_charge_marchand without any allocation, for callers that cannot allocate device memory
 * (code under HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).  It has no
 * counterpart in the reference.  Every array is a device pointer the caller owns (same layouts as
 * shark_charge_marchand); the two kernels are enqueued on `stream` (a hipStream_t, NULL: the
 * default stream) and the call returns without waiting: synchronise before reading the outputs.
 * The per-(bin, cell) scratch is a static device array of SHARK_CHARGE_NOALLOC_MAX_PAIRS pairs
 * (24 MiB, held by each GPU the library runs on until the process ends); ncells * nbins beyond it
 * is rejected (SHARK_GROWTH_EINVAL: call on chunks of cells).  All calls on one GPU share that
 * array: calls that overlap in time on one GPU (on different streams, or from different host
 * threads) race.  cond and gamma_d are required when resistivities is set and unused otherwise. */
#define SHARK_CHARGE_NOALLOC_MAX_PAIRS (1 << 20)
int shark_charge_marchand_noalloc(const shark_marchand_params* p, int64_t ncells, int nbins,
                                  const double* rho, const double* rhod, const double* T, const double* cs,
                                  const double* sdust, const double* mdust, double* psi_old, double* zd,
                                  double* ni, double* ne, double* cond, double* gamma_d, void* stream);

int shark_growth_charge(shark_growth* g, double zeta, double mu_ions, int electrons, double* zmean,
                        double* xion);
/* Per-cell growth intervals (ncells, >= 0) used instead of run's dt; NULL returns to dt. */
int shark_growth_set_cell_dt(shark_growth* g, const double* dt_cells);
/* Launch shape of mode 0 (the per-cell sub-step loop of dust_growth.f90:117-241): chunks of at
 * most max_cells cells (the learned step's fallback cells are one chunk) run one wave per cell
 * (lane = bin); larger chunks run one lane per cell.  Both give the same bits.  Default 8192;
 * 0 puts every chunk on the one-lane-per-cell kernel. */
int shark_growth_set_wave_cells(shark_growth* g, int64_t max_cells);
/* Raw (unstandardised) network inputs of the current state, (ncells x (nbins+5)) row-major:
 * log10 midplane dust density per bin (g cm^-3), log10 T, log10 midplane gas density, log10 alpha,
 * log10 ion fraction (from the charge solve), log10 dt (or the per-cell interval). */
int shark_growth_features(shark_growth* g, double dt, float* feat);
/* Attach a trained network for SHARK_GROWTH_SURROGATE (NULL detaches).  norm holds
 * in_mean[nbins+5], in_std[nbins+5], out_mean[nbins+1], out_std[nbins+1]; outputs are the change of
 * the natural log of each bin's share of the dust mass and log10(-<Z>).  New shares are the
 * softmax of old log-shares + change (positive, summing to one).  A cell falls back to mode 0 and
 * the charge solve when the raw prediction's mass residual |sum_k share_k exp(change_k) - 1|
 * (before that normalisation; coagulation conserves mass, so a consistent prediction has none)
 * exceeds resid_max, when an output is not finite, or when an input lies beyond 8 standard
 * deviations of the training set.
 * alpha (<= 0: turb_coef^2), zeta and mu_ions feed the features and the fallback charge solve.
 * The network must outlive its use here. */
int shark_growth_set_surrogate(shark_growth* g, struct shark_mlp* m, const float* norm,
                               double alpha, double zeta, double mu_ions, double resid_max);
/* A second network on the same (standardised) inputs whose outputs, de-standardised with
 * out_norm2 = out_mean2[nbins+1], out_std2[nbins+1], are added to the first network's: it is
 * trained on what the first network misses (a residual stage).  Call after
 * shark_growth_set_surrogate (which detaches it); m2 == NULL detaches.  The network must outlive
 * its use here. */
int shark_growth_set_surrogate_residual(shark_growth* g, struct shark_mlp* m2, const float* out_norm2);
/* Of the last SHARK_GROWTH_SURROGATE run: cells that fell back, and the mean mass residual of the
 * accepted raw predictions before normalisation, |sum_k share_k exp(change_k) - 1|. */
int shark_growth_surrogate_stats(shark_growth* g, int64_t* n_fallback, double* raw_resid_mean);
/* Per-cell mean grain charge <Z> (ncells) of the last SHARK_GROWTH_SURROGATE run (the network's
 * charge output; the charge solve for cells that fell back) or of the last shark_growth_charge. */
int shark_growth_zmean(shark_growth* g, double* zmean);
/* Dust growth inside the step (src/solver/solve.f90:112): after the drag of every shark_gpu_step /
 * shark_gpu_run of ctx, a growth call of `mode` over the step's dt on the step's dust densities
 * (u_prim(:,irhod), active cells only; the gas density u_prim(:,irho) and cs of the step, the
 * cell radius of the grid).  g must have been created with the step's ncells (ghost cells
 * included) and nbins == ndust; it then belongs to the step (shark_growth_upload/run refuse it).
 * g == NULL detaches.  The caller keeps ownership: destroy the context (or detach) before g.
 * Each species keeps its velocity through growth (momentum scaled with its density; the reference
 * changes the densities only, dust_growth.f90:243-245, which collapses the next Courant step).
 * shark_gpu_kernel_time(ctx, "growth", ...) reports its time per step. */
struct shark_ctx;
int shark_gpu_attach_growth(struct shark_ctx* ctx, shark_growth* g, int mode);
/* Device time (ms) of the last run. */
int shark_growth_kernel_time(shark_growth* g, double* ms);
int shark_growth_destroy(shark_growth* g);

#ifdef __cplusplus
}
#endif
#endif
