// growth.hip — dust growth (Smoluchowski coagulation + fragmentation over size bins) on gfx950.
//
// Replaces dust_growth (src/dust/dust_growth.f90:3-256) for the disk environment.  Per cell the
// reference sub-cycles an explicit step of the pair loop over (i, j <= i) bins: a collision rate
// dndt from the relative velocity dv_ij (:87-114), a fragmented fraction f, losses of both
// partners, the coagulated mass split between the two bins around m1+m2 (:183-190), fragments
// redistributed over the bins below (:204-222), and a CFL-limited sub-step (:227-241).
//
// Pipeline per growth call, over chunks of cells (the per-pair scratch is bounded):
//   k_genv     per cell: temperature, eddy turnover time, Reynolds number, Epstein stopping times
//   k_gpairs   per (pair, cell): dv_ij and from it the collision prefactor A and the fragmented
//              fraction f (FP64), plus the mass-flux prefactor Ahat = A (m1+m2)/(m1 m2) in FP32
//   mode 0     k_gfp64: one lane per cell, the pair loop in the reference's order and rounding,
//              rates accumulated in LDS; bit-exact with the FP64 oracle
//   modes 1-3  k_gmfma: the rate vector of a cell is one dense contraction
//                drhodt = rho_d^2 * Gt [B x 2P] . X [2P]  (X = coagulated / fragmented mass flux
//              of every pair in units of the cell's dust density), run for 64-cell tiles as
//              32x32 MFMA blocks in FP32 (mode 1), BF16 hi+lo split with three products (mode 2)
//              or plain BF16 (mode 3), FP32 accumulation with a compensated (Kahan) sum over
//              K-chunks, then an FP64 update with a per-cell mass rescale.
// Every kernel launch runs a bounded number of sub-steps and keeps the per-cell time and
// sub-step count in device memory; the host relaunches until every cell reached dt.
#include <hip/hip_runtime.h>
#include <limits.h>
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <mutex>
#include <vector>

#include "../../include/shark_growth.h"
#include "growth.h"
#include "mlp.h"

namespace {

constexpr double kPi = 3.141592653589793238462643383279;
constexpr double kKB = 1.38e-16, kMH = 1.67e-24;
constexpr int MAXB = SHARK_GROWTH_MAX_BINS;
constexpr int KC = 128;       // K values (2 per pair) per MFMA chunk
constexpr int NCT = 64;       // cells per MFMA workgroup
// sub-steps per launch of the sub-step kernels: enough that a growth call normally finishes in
// one launch, so a tile that needs few sub-steps frees its slot for the next tile instead of
// waiting at a launch boundary for the slowest tile of the chunk
constexpr int SUB_PER_LAUNCH = 4096;

typedef float f32x16 __attribute__((ext_vector_type(16)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));

// ------------------------------------------------------------------ host tables (bins, pairs)
struct Tables {
  int nb;
  double zeta, eta, massmin, m_mono, ebr_mono;
  double aplus[MAXB], aminus[MAXB], mminus[MAXB], sdust[MAXB], mdust[MAXB];
  double redis[MAXB * MAXB];  // redis[j + nb*i]
};

struct PairC {
  int i, j, t1, t2;  // partner bins and the two coagulation target bins (0-based)
  int nfrag, rrow;   // fragments go to bins [0, nfrag); their weights start at redis[rrow]
  double eps, m1, m2, geom;  // geom = clustered^2 pi (s1+s2)^2
  double ym1, ym2;   // correctly rounded 1/m1, 1/m2 (exact quotients through exdiv)
};

// a / b correctly rounded from y = RN(1/b): q0 = a*y is within one ulp, the FMA residual is exact
// and q0 + r*y rounds to the correctly rounded quotient (Markstein's theorem) in the normal range;
// a zero dividend keeps the sign of a/b.
__host__ __device__ __forceinline__ double exdiv(double a, double b, double y) {
  const double q0 = a * y;
  const double r = fma(-b, q0, a);
  const double q = fma(r, y, q0);
  return a == 0.0 ? q0 : q;
}

static double powi(double a, int b) {  // x**n for integer n, binary powering
  int recip = b < 0;
  if (recip) b = -b;
  double r = 1.0;
  for (;;) {
    if (b & 1) r *= a;
    b /= 2;
    if (b == 0) break;
    a *= a;
  }
  return recip ? 1.0 / r : r;
}

static int ffloor(double x) {  // FLOOR to a default integer; -inf/NaN give INT_MIN
  double f = floor(x);
  if (!(f >= -2147483648.0 && f <= 2147483647.0)) return INT_MIN;
  return (int)f;
}

// Bin grid of distribution_dust (src/dust/dust_init.f90:22-39) and the growth constants of
// dust_growth.f90:41-59.  Integral real powers x**3. are products (as the reference's compiler
// lowers them), the mass grid uses the bare size (the ice mantle is added afterwards, :89-90).
static void build_tables(const shark_growth_params* p, Tables* t) {
  const int nb = p->nbins;
  memset(t, 0, sizeof(*t));
  t->nb = nb;
  t->zeta = pow(p->smax / p->smin, 1.0 / nb);
  t->eta = t->zeta * t->zeta * t->zeta;
  const double four3 = (double)(4.0f / 3.0f);
  const double smin3 = p->smin * p->smin * p->smin;
  double m_min = four3 * kPi * p->rhograin * smin3 / p->unit_m;
  if (p->kernel_type != 0) m_min = smin3;
  for (int k = 0; k < nb; ++k) {
    t->aminus[k] = p->smin * powi(t->zeta, k) / p->unit_l;
    t->aplus[k] = p->smin * powi(t->zeta, k + 1) / p->unit_l;
    t->mminus[k] = m_min * powi(t->eta, k);
    const double s = sqrt(t->aminus[k] * t->aplus[k]);
    t->mdust[k] = four3 * kPi * p->rhograin / p->unit_d * (s * s * s);
    t->sdust[k] = (p->kernel_type == 0) ? s + p->ice_mantle / p->unit_l : s;
  }
  t->massmin = t->mminus[0];
  const double a_mu = p->size_mono / 2.0;
  t->ebr_mono = p->abr * pow(p->gamma_grains, 5.0 / 3.0) * pow(a_mu, 4.0 / 3.0) /
                pow(p->estar_grains, 2.0 / 3.0) / (p->unit_v * p->unit_v * p->unit_m);
  t->m_mono = 4.0 / 3.0 * kPi * p->rhograin * (p->size_mono * p->size_mono * p->size_mono) / p->unit_m;
  const double e = 4.0 - p->slope_mono;
  for (int i = 0; i < nb; ++i)
    for (int j = 0; j < nb; ++j)
      t->redis[j + nb * i] = (j <= i) ? (pow(t->aplus[j], e) - pow(t->aminus[j], e)) /
                                            (pow(t->aplus[i], e) - pow(t->aminus[0], e))
                                      : 0.0;
  for (int i = 0; i < nb; ++i) {
    double s = 0.0;
    for (int j = 0; j < nb; ++j) s = s + t->redis[j + nb * i];
    for (int j = 0; j < nb; ++j) t->redis[j + nb * i] = t->redis[j + nb * i] / s;
  }
}

// Pair order (i outer, j <= i inner) as the reference's loop; the coagulation targets are binned
// by m1+m2 and the fragments by 0.1 (m1+m2) (dust_growth.f90:183-187, :220 with the fragmented
// fraction factored out), which makes them the same for every cell.
static void build_pairs(const shark_growth_params* p, const Tables* t, std::vector<PairC>* out) {
  const int nb = t->nb;
  const double log_eta = log(t->eta);
  out->clear();
  for (int i = 0; i < nb; ++i)
    for (int j = 0; j <= i; ++j) {
      PairC c;
      const double m1 = t->mdust[i], m2 = t->mdust[j], s1 = t->sdust[i], s2 = t->sdust[j];
      const double mt = m1 + m2;
      const int fl = ffloor(log(mt / t->massmin) / log_eta + 1);
      const int ic1 = std::max(std::min(fl, nb), 1);
      const int ic2 = std::max(std::min(fl == INT_MIN ? INT_MIN : fl + 1, nb), 1);
      double eps = 1.0;
      if (ic1 < ic2) eps = std::min((t->mdust[ic2 - 1] - mt) / (t->mdust[ic2 - 1] - t->mdust[ic1 - 1]), 1.0);
      const double ff = 1.0;
      const int g = ffloor(log(0.1 * ff * (m1 + m2) / t->massmin) / log_eta + 1);
      const int ifrag2 = std::max(std::min(g, ic1), 1);
      c.i = i;
      c.j = j;
      c.t1 = ic1 - 1;
      c.t2 = ic2 - 1;
      c.nfrag = ifrag2;
      c.rrow = nb * (ifrag2 - 1);
      c.eps = eps;
      c.m1 = m1;
      c.m2 = m2;
      c.ym1 = 1.0 / m1;
      c.ym2 = 1.0 / m2;
      c.geom = p->clustered * p->clustered * kPi * ((s1 + s2) * (s1 + s2));
      out->push_back(c);
    }
}

// ------------------------------------------------------------------ device arguments
struct GArgs {
  int nb, npairs, nch, c0;
  int64_t ld;  // leading dimension (ncells) of rhod
  int kernel_type, frag, frag_mode, brownian, turb, drift, ambipolar, max_sub, sub_per_launch;
  int kpad;  // padded K (2 per pair) for the MFMA modes
  int want64, want32;  // which per-pair prefactors k_gpairs writes
  double dt, cfl, eps_thr, eps_thr_frag, ratio_min, sticking, vfrag_code, m_mono, ebr_mono;
  double turb_coef, unit_v, unit_m, unit_d, mu_gas, gamma, hoverr, mstar;
  double kts;   // sqrt(pi gamma/8) * (rhograin/unit_d): Epstein prefactor
  double s2pi;  // sqrt(2 pi)
  const PairC* pairs;
  const int* pij;       // [kpad/2]: partner bins i | j << 8 of each pair (k_gmfma's X generation)
  const double* redis;  // [nb*nb]
  const double* mdust;
  const double* sdust;
  double* rhod;  // [nb][ld]
  const double* rho_gas;
  const double* cs;
  const double* radius;
  double* env;  // [4][nch]: T, t_L, t_eta, Re
  double* ts;   // [nb][nch]
  double* A;    // [npairs][nch]
  double* F;    // [npairs][nch]
  float* Ah;    // [npairs][nch]
  float* Fh;    // [npairs][nch]
  const double* dtc;  // [ld] per-cell growth interval, or null for dt everywhere
  const double* dt_dev;  // device scalar holding dt (bound to a step), or null for dt
  const int* cmask;      // [ld] 0 for cells that do not grow (ghost cells of a step), or null
  double* tg;   // [ld] time reached inside the growth call
  int* nit;     // [ld] sub-steps taken
  int* unfinished;
  int* bad;     // set when a rate or density is not finite
  const int* tile_order;  // launch order of the 64-cell tiles (slowest first), or null: blockIdx
  const int* cells;       // [nch] cell of each chunk index (a list of cells), or null: c0 + index
  const float* Gf32;     // MFMA A operand, FP32 layout
  const u32x4* Gbhi;     // MFMA A operand, BF16 hi
  const u32x4* Gblo;     // MFMA A operand, BF16 lo
};

__device__ __forceinline__ double fmn(double a, double b) { return a < b ? a : b; }
// growth interval of a cell: the per-cell interval when set, else dt (or the step's device dt);
// zero for masked cells
template <class Args>
__device__ __forceinline__ double cell_dt(const Args& a, int64_t c) {
  if (a.cmask && !a.cmask[c]) return 0.0;
  if (a.dtc) return a.dtc[c];
  return a.dt_dev ? *a.dt_dev : a.dt;
}
__device__ __forceinline__ double fmx(double a, double b) { return a > b ? a : b; }
// cell of chunk index cc: from the chunk's cell list when it has one, else c0 + cc
template <class Args>
__device__ __forceinline__ int64_t chunk_cell(const Args& a, int cc) {
  return a.cells ? (int64_t)a.cells[cc] : (int64_t)a.c0 + cc;
}

// Disk environment of one cell: T from the local isothermal sound speed, t_L = 1/Omega, the
// Reynolds number of the midplane density (the Reynolds formula of dust_growth.f90:77), and the
// Epstein stopping times of compute_tstop (src/setups/disk_polar/setup.f90:436).
__global__ void k_genv(GArgs a) {
  const int cc = blockIdx.x * blockDim.x + threadIdx.x;
  if (cc >= a.nch) return;
  const int64_t c = chunk_cell(a, cc);
  const double r = a.radius[c], sigma = a.rho_gas[c], cs = a.cs[c];
  const double H = a.hoverr * r;
  const double om = sqrt(a.mstar / (r * r * r));
  const double rho_mid = sigma / (a.s2pi * H);
  const double csp = cs * a.unit_v;
  const double T = csp * csp * (a.mu_gas * kMH) / (a.gamma * kKB);
  const double t_L = 1.0 / om;
  const double Re = 6.2e7 * sqrt(rho_mid * a.unit_d / (a.mu_gas * kMH) / 1e5) * sqrt(T / 10.0);
  const double t_eta = t_L / sqrt(Re);
  a.env[cc] = T;
  a.env[a.nch + cc] = t_L;
  a.env[2 * a.nch + cc] = t_eta;
  a.env[3 * a.nch + cc] = Re;
  for (int k = 0; k < a.nb; ++k)
    a.ts[(int64_t)k * a.nch + cc] = a.kts * a.sdust[k] / ((sigma / (a.s2pi * H)) * cs);
}

// Relative velocity of a pair (dust_growth.f90:87-114): turbulent (Ormel & Cuzzi regimes by
// the stopping time against t_eta and t_L), Brownian, and the drift/ambipolar terms, which are
// zero in the disk environment but still enter the quadrature sums as in the reference.
__device__ double pair_dv(const GArgs& a, double tsi, double tsj, double mi, double mj, double T,
                          double cs_T, double t_L, double t_eta, double reynolds) {
  const double x = tsj / tsi;
  const double fst = (double)3.2f - 1.0 - x + 2.0 / (1. + x) * ((double)(1.0f / 2.6f) + (x * x * x) / ((double)1.6f + x));
  const double St1 = tsi / t_L, St2 = tsj / t_L;
  const double rem = 1.0 / sqrt(reynolds);
  const double v1 = a.turb_coef * cs_T * sqrt((St1 - St2) / (St1 + St2)) *
                    sqrt(St1 * St1 / (St1 + rem) - St2 * St2 / (St2 + rem));
  const double v2 = a.turb_coef * cs_T * sqrt(fst * St1);
  const double v3 = a.turb_coef * cs_T * sqrt(1.0 / (1.0 + St1) + 1.0 / (1.0 + St2));
  double vt = v2;
  if (tsi < t_eta) vt = v1;
  if (tsi > t_L) vt = v3;
  const double um = a.unit_m;
  const double inner = (8.0 * kKB * T / kPi) * (mi * um + mj * um) / (mi * mj * (um * um)) /
                       (a.unit_v * a.unit_v);
  const double sq = sqrt(inner);
  const double vb = sqrt(sq * sq);
  const double vh = 0.0;
  const double ad = 0.0 / a.unit_v * (0.0 - 0.0);
  const double va = sqrt(ad * ad);
  double d = 0.0;
  if (a.turb) d = vt;
  if (a.brownian) d = sqrt(d * d + vb * vb);
  if (a.drift) d = sqrt(d * d + vh * vh);
  if (a.ambipolar) d = sqrt(d * d + va * va);
  return d;
}

// Per (pair, cell): collision prefactor A = clustered^2 pi (s1+s2)^2 dv, fragmented fraction
// (energy threshold dust_growth.f90:142-144 or velocity threshold :145), and the FP32 copies the
// contraction reads.
__global__ void k_gpairs(GArgs a) {
  const int cc = blockIdx.x * blockDim.x + threadIdx.x;
  const int p = blockIdx.y;
  if (cc >= a.nch) return;
  const PairC P = a.pairs[p];
  const int64_t o = (int64_t)p * a.nch + cc;
  const double m1 = P.m1, m2 = P.m2;
  double A, f = 0.0, ahat;
  if (a.kernel_type == 0) {
    const double T = a.env[cc], t_L = a.env[a.nch + cc], t_eta = a.env[2 * a.nch + cc];
    const double Re = a.env[3 * a.nch + cc];
    const double cs_T = a.cs[chunk_cell(a, cc)];
    const double dv = pair_dv(a, a.ts[(int64_t)P.i * a.nch + cc], a.ts[(int64_t)P.j * a.nch + cc],
                              m1, m2, T, cs_T, t_L, t_eta, Re);
    if (a.frag) {
      if (a.frag_mode == 0) {
        const double Ecol = 0.5 * (m1 * m2) / (m1 + m2) * (dv * dv);
        const double Ebr = (m1 + m2) / a.m_mono * a.ebr_mono;
        f = fmx(fmn((Ecol - 0.1 * Ebr) / (4.9 * Ebr), 1.0), 0.0);
      } else {
        f = fmx(fmn((fabs(dv) - 0.1 * a.vfrag_code) / a.vfrag_code, 1.0), 0.0);
      }
    }
    A = P.geom * dv;
    ahat = A * ((m1 + m2) / m1 / m2);
  } else {
    A = 0.0;
    ahat = (a.kernel_type == 1) ? (m1 + m2) / m1 / m2 : (m1 + m2) * (m1 + m2) / m1 / m2;
  }
  if (a.want64) {
    a.A[o] = A;
    a.F[o] = f;
  }
  if (a.want32) {
    a.Ah[o] = (float)ahat;
    a.Fh[o] = (float)f;
  }
}

// ------------------------------------------------------------------ mode 0: FP64 pair loop
// One lane per cell, rates in LDS as [bin][lane]; the pair loop is uniform across the wave so
// the pair constants come through scalar loads.  Statement order and rounding follow the
// reference (dust_growth.f90:117-241), so the result is bit-exact with the FP64 oracle.
__global__ __launch_bounds__(64) void k_gfp64(GArgs a) {
  extern __shared__ double dd[];
  const int lane = threadIdx.x;
  const int tile = a.tile_order ? a.tile_order[blockIdx.x] : (int)blockIdx.x;
  const int cc = tile * 64 + lane;
  const bool valid = cc < a.nch;
  const int64_t c = chunk_cell(a, valid ? cc : 0);
  const int nb = a.nb;
  double* __restrict__ rh = a.rhod + c;
  const double rg = valid ? a.rho_gas[c] : 1.0;
  double tgrow = valid ? a.tg[c] : 0.0;
  int niter = valid ? a.nit[c] : 0;
  const double dtl = cell_dt(a, c);
  bool act = valid && tgrow < dtl && !(a.max_sub > 0 && niter >= a.max_sub);
  const int kt = a.kernel_type;
  const double st = a.sticking;
  for (int s = 0; s < a.sub_per_launch; ++s) {
    if (!__any(act)) break;
    for (int k = 0; k < nb; ++k) dd[k * 64 + lane] = 0.0;
    if (act) {
      niter = niter + 1;
      // flat pair loop (same order as i outer, j <= i inner); the loads of pair p+1 are issued
      // before pair p is processed so their latency overlaps its arithmetic and LDS updates
      const int np = a.npairs;
      const double* __restrict__ Ap = a.A + cc;
      const double* __restrict__ Fp = a.F + cc;
      double An = (kt == 0) ? Ap[0] : 0.0;
      double Fn = a.frag ? Fp[0] : 0.0;
      double rin = rh[0], rjn = rh[0];
      for (int p = 0; p < np; ++p) {
        const PairC P = a.pairs[p];
        const double Ac = An, Fc = Fn, ri = rin, rj = rjn;
        if (p + 1 < np) {
          const PairC Q = a.pairs[p + 1];
          const int64_t on = (int64_t)(p + 1) * a.nch;
          An = (kt == 0) ? Ap[on] : 0.0;
          Fn = a.frag ? Fp[on] : 0.0;
          rin = rh[(int64_t)Q.i * a.ld];
          rjn = rh[(int64_t)Q.j * a.ld];
        }
        const int i = P.i, j = P.j;
        const double m1 = P.m1, m2 = P.m2;
        double f = Fc, pc = 1.0;
        double dndt;
        if (kt == 1) {
          pc = 1.0; f = 0.0;
          dndt = exdiv(exdiv(ri * rj, m1, P.ym1), m2, P.ym2);
        } else if (kt == 2) {
          pc = 1.0; f = 0.0;
          dndt = exdiv(exdiv((m1 + m2) * ri * rj, m1, P.ym1), m2, P.ym2);
        } else {
          dndt = exdiv(exdiv(Ac * ri * rj, m1, P.ym1), m2, P.ym2);
        }
        if (i == j) dndt = dndt / 2.0;
        if (kt == 0) {
          if (a.eps_thr > 0.0 && (ri < a.eps_thr * rg)) pc = 0.0;
          if (a.eps_thr > 0.0 && (rj < a.eps_thr * rg)) pc = 0.0;
          if (a.eps_thr_frag > 0.0 && (ri < a.eps_thr_frag * rg)) f = 0.0;
          if (a.eps_thr_frag > 0.0 && (rj < a.eps_thr_frag * rg)) f = 0.0;
        }
        const double eps = P.eps;
        dd[i * 64 + lane] = dd[i * 64 + lane] - pc * (1.0 - f) * m1 * dndt * st - pc * f * m1 * dndt;
        dd[j * 64 + lane] = dd[j * 64 + lane] - pc * (1.0 - f) * m2 * dndt * st - pc * f * m2 * dndt;
        if (f < 1.0) {
          dd[P.t1 * 64 + lane] = dd[P.t1 * 64 + lane] + eps * pc * (1.0 - f) * (m1 + m2) * dndt * st;
          dd[P.t2 * 64 + lane] = dd[P.t2 * 64 + lane] + (1. - eps) * pc * (1.0 - f) * (m1 + m2) * dndt * st;
        }
        if ((a.frag && f > 0.0) && pc > 0.) {
          const double base = pc * f * (m1 + m2) * dndt;
          const double* __restrict__ rr = a.redis + P.rrow;
#pragma unroll 4
          for (int k = 0; k < P.nfrag; ++k) dd[k * 64 + lane] = dd[k * 64 + lane] + base * rr[k];
        }
      }
      double dtg = 1e32;
      for (int k = 0; k < nb; ++k) {
        const double d = dd[k * 64 + lane];
        if (fabs(d) != 0.0) dtg = fmn(dtg, rh[(int64_t)k * a.ld] / fabs(d) * a.cfl);
      }
      dtg = fmn(dtg, dtl - tgrow);
      for (int k = 0; k < nb; ++k) {
        const double v = fmx(rh[(int64_t)k * a.ld] + dd[k * 64 + lane] * dtg, rg * a.ratio_min);
        rh[(int64_t)k * a.ld] = v;
      }
      tgrow = tgrow + dtg;
      if (!(tgrow == tgrow) || !(dtg > 0.0)) {  // NaN time or a sub-step that makes no progress
        atomicOr(a.bad, 1);
        act = false;
      } else {
        act = tgrow < dtl && !(a.max_sub > 0 && niter >= a.max_sub);
      }
    }
  }
  if (valid) {
    a.tg[c] = tgrow;
    a.nit[c] = niter;
  }
  // the host needs only whether any cell is unfinished: one atomic per wave
  if (__any(valid && act) && lane == 0) atomicAdd(a.unfinished, 1);
}

// ------------------------------------------------------------------ mode 0, one wave per cell
// The same pair loop with lane k holding bin k's density and rate (one cell per wave, four cells
// per workgroup).  The loop is uniform across the wave, so the pair constants come through scalar
// loads, the partner densities are read across lanes, and the cell's pair prefactors are loaded
// 64 pairs at a time (one pair per lane) one block ahead.  Each lane applies, in the serial loop's
// order, the updates whose target is its bin: every bin's rate is the same chain of FP64
// operations as in k_gfp64, and the CFL minimum folds the bins in order, so the two kernels agree
// bit for bit.  k_gfp64's one lane per cell waits on its serial chain of LDS updates every pair;
// here a sub-step costs the pair loop's instructions once per cell, which suits few cells (small
// chunks, the cells where the learned step falls back).  LDS: the redistribution table.
__device__ __forceinline__ double rdlane(double v, int l) {
  const long long b = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_readlane((int)b, l);
  const int hi = __builtin_amdgcn_readlane((int)(b >> 32), l);
  return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}

__global__ __launch_bounds__(256) void k_gfp64w(GArgs a) {
  extern __shared__ double rds[];   // redis [nb * nb]
  const int nb = a.nb;
  for (int t = threadIdx.x; t < nb * nb; t += 256) rds[t] = a.redis[t];
  __syncthreads();
  const int lane = threadIdx.x & 63;
  const int cc = blockIdx.x * 4 + (threadIdx.x >> 6);
  if (cc >= a.nch) return;   // wave-uniform, after the only barrier
  const int64_t c = chunk_cell(a, cc);
  const bool mine = lane < nb;
  double* __restrict__ rh = a.rhod + c;
  double r = mine ? rh[(int64_t)lane * a.ld] : 0.0;
  const double rg = a.rho_gas[c];
  double tgrow = a.tg[c];
  int niter = a.nit[c];
  const double dtl = cell_dt(a, c);
  bool act = tgrow < dtl && !(a.max_sub > 0 && niter >= a.max_sub);
  const int kt = a.kernel_type;
  const double st = a.sticking;
  const int np = a.npairs;
  typedef const __attribute__((address_space(4))) PairC* PairPtr;
  const PairPtr pairs = (PairPtr)a.pairs;
  for (int s = 0; s < a.sub_per_launch && act; ++s) {
    niter = niter + 1;
    double dd = 0.0;
    // prefactors of pairs p0 + lane, the next block loaded while this one is processed
    double Ab = 0.0, Fb = 0.0;
    {
      const int q = lane;
      if (kt == 0 && q < np) Ab = a.A[(int64_t)q * a.nch + cc];
      if (a.frag && q < np) Fb = a.F[(int64_t)q * a.nch + cc];
    }
    for (int p0 = 0; p0 < np; p0 += 64) {
      double An = 0.0, Fn = 0.0;
      {
        const int q = p0 + 64 + lane;
        if (kt == 0 && q < np) An = a.A[(int64_t)q * a.nch + cc];
        if (a.frag && q < np) Fn = a.F[(int64_t)q * a.nch + cc];
      }
      const int pe = min(np, p0 + 64);
      for (int p = p0; p < pe; ++p) {
        const int i = pairs[p].i, j = pairs[p].j, t1 = pairs[p].t1, t2 = pairs[p].t2;
        const int nfrag = pairs[p].nfrag, rrow = pairs[p].rrow;
        const double m1 = pairs[p].m1, m2 = pairs[p].m2, ym1 = pairs[p].ym1, ym2 = pairs[p].ym2;
        const double eps = pairs[p].eps;
        const double rr = (lane < nfrag) ? rds[rrow + lane] : 0.0;
        const double ri = rdlane(r, i), rj = rdlane(r, j);
        const double Ac = (kt == 0) ? rdlane(Ab, p - p0) : 0.0;
        double f = a.frag ? rdlane(Fb, p - p0) : 0.0, pc = 1.0;
        double dndt;
        if (kt == 1) {
          pc = 1.0; f = 0.0;
          dndt = exdiv(exdiv(ri * rj, m1, ym1), m2, ym2);
        } else if (kt == 2) {
          pc = 1.0; f = 0.0;
          dndt = exdiv(exdiv((m1 + m2) * ri * rj, m1, ym1), m2, ym2);
        } else {
          dndt = exdiv(exdiv(Ac * ri * rj, m1, ym1), m2, ym2);
        }
        if (i == j) dndt = dndt / 2.0;
        if (kt == 0) {
          if (a.eps_thr > 0.0 && (ri < a.eps_thr * rg)) pc = 0.0;
          if (a.eps_thr > 0.0 && (rj < a.eps_thr * rg)) pc = 0.0;
          if (a.eps_thr_frag > 0.0 && (ri < a.eps_thr_frag * rg)) f = 0.0;
          if (a.eps_thr_frag > 0.0 && (rj < a.eps_thr_frag * rg)) f = 0.0;
        }
        if (lane == i) dd = dd - pc * (1.0 - f) * m1 * dndt * st - pc * f * m1 * dndt;
        if (lane == j) dd = dd - pc * (1.0 - f) * m2 * dndt * st - pc * f * m2 * dndt;
        if (f < 1.0) {
          if (lane == t1) dd = dd + eps * pc * (1.0 - f) * (m1 + m2) * dndt * st;
          if (lane == t2) dd = dd + (1. - eps) * pc * (1.0 - f) * (m1 + m2) * dndt * st;
        }
        if ((a.frag && f > 0.0) && pc > 0.) {
          const double base = pc * f * (m1 + m2) * dndt;
          if (lane < nfrag) dd = dd + base * rr;
        }
      }
      Ab = An;
      Fb = Fn;
    }
    // CFL-limited sub-step: the minimum over the bins with a non-zero rate, folded in bin order
    const bool use = mine && fabs(dd) != 0.0;
    const double v = use ? r / fabs(dd) * a.cfl : 0.0;
    const unsigned long long m = __ballot(use);
    double dtg = 1e32;
    for (int k = 0; k < nb; ++k)
      if ((m >> k) & 1ull) dtg = fmn(dtg, rdlane(v, k));
    dtg = fmn(dtg, dtl - tgrow);
    if (mine) r = fmx(r + dd * dtg, rg * a.ratio_min);
    tgrow = tgrow + dtg;
    if (!(tgrow == tgrow) || !(dtg > 0.0)) {  // NaN time or a sub-step that makes no progress
      if (lane == 0) atomicOr(a.bad, 1);
      act = false;
    } else {
      act = tgrow < dtl && !(a.max_sub > 0 && niter >= a.max_sub);
    }
  }
  if (mine) rh[(int64_t)lane * a.ld] = r;
  if (lane == 0) {
    a.tg[c] = tgrow;
    a.nit[c] = niter;
    if (act) atomicAdd(a.unfinished, 1);
  }
}

// Compacts the cells whose flag is set into a list (order unspecified; each cell's arithmetic
// does not depend on it): one atomic per wave.
__global__ __launch_bounds__(256) void k_flag_list(const int* __restrict__ flag, int64_t n, int* list, int* count) {
  const int64_t c = (int64_t)blockIdx.x * 256 + threadIdx.x;
  const bool on = c < n && flag[c] != 0;
  const unsigned long long m = __ballot(on);
  if (!m) return;
  const int lane = threadIdx.x & 63;
  int base = 0;
  if (lane == 0) base = atomicAdd(count, __popcll(m));
  base = __shfl(base, 0);
  if (on) list[base + __popcll(m & ((1ull << lane) - 1ull))] = (int)c;
}

// ------------------------------------------------------------------ modes 1-3: MFMA contraction
__device__ __forceinline__ unsigned short bf16_rne(float x) {
  unsigned int u = __float_as_uint(x);
  u += 0x7FFFu + ((u >> 16) & 1u);
  return (unsigned short)(u >> 16);
}
__device__ __forceinline__ float bf16_to_f(unsigned short h) { return __uint_as_float((unsigned int)h << 16); }

template <int MODE>
struct XTile;  // LDS layout of one K-chunk of X for 64 cells

template <>
struct XTile<1> {
  float x[KC][NCT];
};
template <>
struct XTile<2> {
  unsigned short hi[KC / 8][NCT][8];
  unsigned short lo[KC / 8][NCT][8];
};
template <>
struct XTile<3> {
  unsigned short hi[KC / 8][NCT][8];
};

template <int MODE>
struct GSmem {
  union {
    XTile<MODE> xt;
    double d[MAXB][NCT];  // rates after the contraction
  } u;
  float epsf[MAXB][NCT];  // rho_k / rho_d,total in FP32
  unsigned long long okc[NCT], okf[NCT];
  double rho2[NCT];       // rho_d,total^2
  int act[NCT];
  int any;
};

// Workgroup: 4 waves over a 64-bin x 64-cell rate tile; wave w owns bins 32*(w&1).. and cells
// 32*(w>>1)..  The X chunk (KC values x 64 cells) is generated into LDS by all 256 lanes, then
// each wave multiplies its 32 x KC slice of Gt (pre-laid-out in global memory, read through L2)
// with it.  Chunk results are added with a Kahan sum.
template <int MODE>
__global__ __launch_bounds__(256) void k_gmfma(GArgs a) {
  __shared__ GSmem<MODE> sm;
  const int tid = threadIdx.x;
  const int lane = tid & 63, w = __builtin_amdgcn_readfirstlane(tid >> 6);  // wave-uniform
  const int rb = w & 1, ch = w >> 1;
  const int cbase = (a.tile_order ? a.tile_order[blockIdx.x] : (int)blockIdx.x) * NCT;
  const int nb = a.nb;
  const int kt = a.kernel_type;
  // per-cell state lives with thread `tid` for tid < 64
  int64_t c = 0;
  double tgrow = 0.0, rg = 1.0, dtl = 0.0;
  int niter = 0;
  bool valid = false;
  if (tid < NCT) {
    valid = cbase + tid < a.nch;
    c = a.c0 + (valid ? cbase + tid : 0);
    if (valid) {
      tgrow = a.tg[c];
      niter = a.nit[c];
      rg = a.rho_gas[c];
      dtl = cell_dt(a, c);
    }
  }
  const int nchunks = a.kpad / KC;
  for (int s = 0; s < a.sub_per_launch; ++s) {
    if (tid == 0) sm.any = 0;
    __syncthreads();
    if (tid < NCT) {
      const bool act = valid && tgrow < dtl && !(a.max_sub > 0 && niter >= a.max_sub);
      sm.act[tid] = act;
      if (act) sm.any = 1;  // benign race: every writer stores 1
      double tot = 0.0;
      unsigned long long mc = 0, mf = 0;
      for (int k = 0; k < nb; ++k) {
        const double r = a.rhod[(int64_t)k * a.ld + c];
        tot += r;
        const bool okc = !(kt == 0 && a.eps_thr > 0.0 && r < a.eps_thr * rg);
        const bool okf = !(kt == 0 && a.eps_thr_frag > 0.0 && r < a.eps_thr_frag * rg);
        mc |= (unsigned long long)okc << k;
        mf |= (unsigned long long)okf << k;
      }
      const double inv = act ? 1.0 / tot : 0.0;
      for (int k = 0; k < nb; ++k) sm.epsf[k][tid] = (float)(a.rhod[(int64_t)k * a.ld + c] * inv);
      sm.okc[tid] = mc;
      sm.okf[tid] = a.frag ? mf : 0ull;
      sm.rho2[tid] = tot * tot;
    }
    __syncthreads();
    if (!sm.any) break;

    f32x16 sum, comp;
    for (int r = 0; r < 16; ++r) { sum[r] = 0.f; comp[r] = 0.f; }
    for (int q = 0; q < nchunks; ++q) {
      // ---- X generation: wave w makes pairs pp = 16 w + n of the chunk (X entries 32 w ..
      // 32 w + 31), cell = lane.  The partner bins come from one scalar load of the packed pair
      // table, the 32 prefactor loads of a lane are issued together, and the BF16 modes write
      // each group of 8 entries of a cell as one 16-byte LDS store.
      {
        const int cell = lane;
        const bool cact = sm.act[cell] != 0;
        const unsigned long long mc = sm.okc[cell], mf = sm.okf[cell];
        const int p0 = q * (KC / 2) + 16 * w;
        const __attribute__((address_space(4))) int* pt =
            (const __attribute__((address_space(4))) int*)(a.pij + p0);
        int pvs[16];
#pragma unroll
        for (int n = 0; n < 16; ++n) pvs[n] = pt[n];   // the table is padded to kpad / 2
        const bool full = p0 + 16 <= a.npairs;          // every chunk but the last
        float ahv[16], fhv[16];
#pragma unroll
        for (int n = 0; n < 16; ++n) { ahv[n] = 0.f; fhv[n] = 0.f; }
        if (cact) {
          const float* ap = a.Ah + (int64_t)p0 * a.nch + cbase;   // wave-uniform
          const float* fp = a.Fh + (int64_t)p0 * a.nch + cbase;
          if (full) {
#pragma unroll
            for (int n = 0; n < 16; ++n) {
              ahv[n] = ap[(int64_t)n * a.nch + cell];
              fhv[n] = fp[(int64_t)n * a.nch + cell];
            }
          } else {
#pragma unroll
            for (int n = 0; n < 16; ++n)
              if (p0 + n < a.npairs) {
                ahv[n] = ap[(int64_t)n * a.nch + cell];
                fhv[n] = fp[(int64_t)n * a.nch + cell];
              }
          }
        }
        float xv[32];
#pragma unroll
        for (int n = 0; n < 16; ++n) {
          const int pv = pvs[n];
          const int pi = pv & 0xff, pj = (pv >> 8) & 0xff;
          float xc = 0.f, xf = 0.f;
          if ((full || p0 + n < a.npairs) && cact) {
            const float ah = ahv[n];
            float f = fhv[n];
            const bool pc = ((mc >> pi) & (mc >> pj) & 1ull) != 0;
            if (!(((mf >> pi) & (mf >> pj) & 1ull) != 0)) f = 0.f;
            float e = sm.epsf[pi][cell] * sm.epsf[pj][cell] * ah;
            if (pi == pj) e *= 0.5f;
            if (pc) {
              xc = (1.f - f) * e;
              xf = f * e;
            }
          }
          xv[2 * n] = xc;
          xv[2 * n + 1] = xf;
        }
        if constexpr (MODE == 1) {
#pragma unroll
          for (int kk = 0; kk < 32; ++kk) sm.u.xt.x[32 * w + kk][cell] = xv[kk];
        } else {
#pragma unroll
          for (int g8 = 0; g8 < 4; ++g8) {
            unsigned short h[8], l[8];
#pragma unroll
            for (int e = 0; e < 8; ++e) {
              h[e] = bf16_rne(xv[8 * g8 + e]);
              if constexpr (MODE == 2) l[e] = bf16_rne(xv[8 * g8 + e] - bf16_to_f(h[e]));
            }
            u32x4 hv, lv;
#pragma unroll
            for (int e = 0; e < 4; ++e) {
              hv[e] = (unsigned int)h[2 * e] | ((unsigned int)h[2 * e + 1] << 16);
              if constexpr (MODE == 2) lv[e] = (unsigned int)l[2 * e] | ((unsigned int)l[2 * e + 1] << 16);
            }
            *reinterpret_cast<u32x4*>(&sm.u.xt.hi[4 * w + g8][cell][0]) = hv;
            if constexpr (MODE == 2) *reinterpret_cast<u32x4*>(&sm.u.xt.lo[4 * w + g8][cell][0]) = lv;
          }
        }
      }
      __syncthreads();
      // ---- contraction of this chunk
      f32x16 acc;
      for (int r = 0; r < 16; ++r) acc[r] = 0.f;
      const int cl = ch * 32 + (lane & 31);
      if constexpr (MODE == 1) {
        const float* g = a.Gf32 + ((int64_t)(rb * (a.kpad / 2)) + q * (KC / 2)) * 64 + lane;
        for (int st = 0; st < KC / 2; ++st) {
          const float av = g[st * 64];
          const float bv = sm.u.xt.x[2 * st + (lane >> 5)][cl];
          acc = __builtin_amdgcn_mfma_f32_32x32x2f32(av, bv, acc, 0, 0, 0);
        }
      } else {
        const int64_t gb = ((int64_t)(rb * (a.kpad / 16)) + q * (KC / 16)) * 64 + lane;
        for (int st = 0; st < KC / 16; ++st) {
          const u32x4 ah = a.Gbhi[gb + st * 64];
          const int kg = 2 * st + (lane >> 5);
          const u32x4 bh = *reinterpret_cast<const u32x4*>(&sm.u.xt.hi[kg][cl][0]);
          acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(__builtin_bit_cast(bf16x8, ah),
                                                        __builtin_bit_cast(bf16x8, bh), acc, 0, 0, 0);
          if constexpr (MODE == 2) {
            const u32x4 al = a.Gblo[gb + st * 64];
            const u32x4 bl = *reinterpret_cast<const u32x4*>(&sm.u.xt.lo[kg][cl][0]);
            acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(__builtin_bit_cast(bf16x8, ah),
                                                          __builtin_bit_cast(bf16x8, bl), acc, 0, 0, 0);
            acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(__builtin_bit_cast(bf16x8, al),
                                                          __builtin_bit_cast(bf16x8, bh), acc, 0, 0, 0);
          }
        }
      }
      for (int r = 0; r < 16; ++r) {  // Kahan: sum += acc
        const float y = acc[r] - comp[r];
        const float t = sum[r] + y;
        comp[r] = (t - sum[r]) - y;
        sum[r] = t;
      }
      __syncthreads();
    }
    // ---- rates to LDS (FP64): lane holds column lane%32, rows 8*(r/4) + 4*(lane/32) + r%4
    {
      const int cl = ch * 32 + (lane & 31);
      for (int r = 0; r < 16; ++r) {
        const int row = rb * 32 + 8 * (r >> 2) + 4 * (lane >> 5) + (r & 3);
        sm.u.d[row][cl] = ((double)sum[r] - (double)comp[r]) * sm.rho2[cl];
      }
    }
    __syncthreads();
    // ---- FP64 update of each cell, CFL-limited sub-step (dust_growth.f90:227-241), mass rescale
    if (tid < NCT && sm.act[tid]) {
      niter = niter + 1;
      double dtg = 1e32, m_old = 0.0, m_new = 0.0;
      for (int k = 0; k < nb; ++k) {
        const double d = sm.u.d[k][tid];
        const double r = a.rhod[(int64_t)k * a.ld + c];
        m_old += r;
        if (fabs(d) != 0.0) dtg = fmn(dtg, r / fabs(d) * a.cfl);
      }
      dtg = fmn(dtg, dtl - tgrow);
      for (int k = 0; k < nb; ++k) {
        const double v = fmx(a.rhod[(int64_t)k * a.ld + c] + sm.u.d[k][tid] * dtg, rg * a.ratio_min);
        sm.u.d[k][tid] = v;
        m_new += v;
      }
      const double sc = m_old / m_new;
      for (int k = 0; k < nb; ++k) a.rhod[(int64_t)k * a.ld + c] = sm.u.d[k][tid] * sc;
      tgrow = tgrow + dtg;
      if (!(tgrow == tgrow) || !(dtg > 0.0) || !(sc == sc)) {
        atomicOr(a.bad, 1);
        valid = false;
        sm.act[tid] = 0;
      }
    }
    __syncthreads();
  }
  if (tid < NCT && valid) {
    a.tg[c] = tgrow;
    a.nit[c] = niter;
  }
  // the host needs only whether any cell is unfinished: one atomic per tile (wave 0 holds the cells)
  if (w == 0) {
    const bool unf = valid && tgrow < dtl && !(a.max_sub > 0 && niter >= a.max_sub);
    if (__any(unf) && lane == 0) atomicAdd(a.unfinished, 1);
  }
}

// Remaining sub-steps of each 64-cell tile after its first sub-steps, at the mean sub-step
// length so far ((dt - t) n / t), for its slowest cell; -1 for a finished tile.  Orders the
// tiles slowest first for the launches that follow, so the long tiles do not start last.
__global__ __launch_bounds__(64) void k_tile_est(GArgs a, float* est) {
  const int cc = blockIdx.x * 64 + threadIdx.x;
  float v = -1.f;
  if (cc < a.nch) {
    const int64_t c = chunk_cell(a, cc);
    const double tgrow = a.tg[c], dtl = cell_dt(a, c);
    const int niter = a.nit[c];
    if (tgrow < dtl && !(a.max_sub > 0 && niter >= a.max_sub))
      v = tgrow > 0.0 ? (float)((dtl - tgrow) / tgrow * (double)niter) : 3.0e38f;
  }
  for (int o = 32; o > 0; o >>= 1) v = fmaxf(v, __shfl_xor(v, o));
  if (threadIdx.x == 0) est[blockIdx.x] = v;
}

// ------------------------------------------------------------------ grain charging
// Analytical charge (src/chemistry/ionisation.f90:809-922, Fujii et al. 2011): per grain,
// Newton-Raphson from Z = -1 (src/commons/newton_raphson.f90:7-97) on f(Z) of
// functions_NR.f90:15-43 and its derivative :61-91, in cgs with the disk environment's midplane
// densities; stops at |f| < 1e-14 (the reference's test), a step below 1e-15 |Z|, or 100
// iterations.  Per cell: number-weighted mean charge and ion fraction zeta/sum_k k_i,k n_k.
struct QArgs {
  int nb, nch, c0;
  int64_t ld;
  int electrons, maxiter;
  double zeta, mu_ions, mu_gas, hoverr, unit_d, unit_m, unit_l, s2pi;
  const double* rhod;
  const double* rho_gas;
  const double* radius;
  const double* env;     // T at env[cc]
  const double* mdust;
  const double* sdust;
  double* zmean;         // [ncells]
  double* xion;          // [ncells]
  const int* mask;       // [ncells] or null: only cells with mask != 0 are charged
  const int* cells;      // [nch] cell of each chunk index (a list of cells), or null: c0 + index
};

constexpr double kEel = 4.8e-10, kMel = 9.109e-28;

__device__ double charge_f(double y, double zeta, double rho, double m_d, double rho_d, double s_d,
                           double T, double mu_ions, double mu_gas, int electrons) {
  const double m_i = mu_ions * kMH, m_e = kMel;
  const double e2 = kEel * kEel;
  const double k_e_d = kPi * (s_d * s_d) * sqrt(8.0 * kKB * T / (kPi * m_e)) * exp(e2 * y / (s_d * kKB * T));
  const double k_i_d = kPi * (s_d * s_d) * sqrt(8.0 * kKB * T / (kPi * m_i)) * (1.0 - e2 * y / (s_d * kKB * T));
  const double X = m_d * rho / (rho_d * mu_gas * kMH);
  if (!electrons) return mu_gas * kMH * zeta / rho * (X * X) * ((1.0 / k_i_d)) - y;
  return mu_gas * kMH * zeta / rho * (X * X) * ((1.0 / k_i_d) - (1.0 / k_e_d)) - y;
}

__device__ double charge_fp(double y, double zeta, double rho, double m_d, double rho_d, double s_d,
                            double T, double mu_ions, double mu_gas, int electrons) {
  const double m_i = mu_ions * kMH, m_e = kMel;
  const double e2 = kEel * kEel;
  const double k_e_d = kPi * (s_d * s_d) * sqrt(8.0 * kKB * T / (kPi * m_e)) * exp(e2 * y / (s_d * kKB * T));
  const double k_i_d = kPi * (s_d * s_d) * sqrt(8.0 * kKB * T / (kPi * m_i)) * (1.0 - e2 * y / (s_d * kKB * T));
  const double kp_e = kPi * s_d * sqrt(8.0 * kKB * T / kPi / m_e) * e2 / kKB / T * exp(e2 * (y) / s_d / kKB / T);
  const double kp_i = -(kPi * s_d * sqrt(8.0 * kKB * T / kPi / m_i) * e2 / kKB / T);
  const double X = m_d * rho / rho_d / mu_gas / kMH;
  if (!electrons) return mu_gas * kMH * zeta / rho * (X * X) * (-kp_i / (k_i_d * k_i_d)) - 1.0;
  return mu_gas * kMH * zeta / rho * (X * X) * (kp_e / (k_e_d * k_e_d) - kp_i / (k_i_d * k_i_d)) - 1.0;
}

__global__ void k_gcharge(QArgs a) {
  const int cc = blockIdx.x * blockDim.x + threadIdx.x;
  if (cc >= a.nch) return;
  const int64_t c = chunk_cell(a, cc);
  if (a.mask && !a.mask[c]) return;
  const double T = a.env[cc];
  const double H = a.hoverr * a.radius[c];
  const double rho = a.rho_gas[c] / (a.s2pi * H) * a.unit_d;
  const double m_i = a.mu_ions * kMH, e2 = kEel * kEel;
  double nz = 0.0, nn = 0.0, kin = 0.0;
  for (int k = 0; k < a.nb; ++k) {
    const double rd = a.rhod[(int64_t)k * a.ld + c] / (a.s2pi * H) * a.unit_d;
    const double m_d = a.mdust[k] * a.unit_m, s_d = a.sdust[k] * a.unit_l;
    double y = -1.0;
    for (int it = 1; it <= a.maxiter; ++it) {
      const double fy = charge_f(y, a.zeta, rho, m_d, rd, s_d, T, a.mu_ions, a.mu_gas, a.electrons);
      const double fyp = charge_fp(y, a.zeta, rho, m_d, rd, s_d, T, a.mu_ions, a.mu_gas, a.electrons);
      if (fabs(fy) < 1.e-14) break;
      const double deltay = fy / fyp;
      y = y - deltay;
      if (fabs(deltay) <= 1e-15 * fabs(y)) break;
    }
    const double n = rd / m_d;
    nz += n * y;
    nn += n;
    const double k_i_d = kPi * (s_d * s_d) * sqrt(8.0 * kKB * T / (kPi * m_i)) * (1.0 - e2 * y / (s_d * kKB * T));
    kin += k_i_d * n;
  }
  a.zmean[c] = nz / nn;
  a.xion[c] = a.zeta / kin;
}

// Gas temperature of each chunk cell (the T of k_genv), for the learned step's chunks.
__global__ void k_senv(GArgs a, double* Tout) {
  const int cc = blockIdx.x * blockDim.x + threadIdx.x;
  if (cc >= a.nch) return;
  const double csp = a.cs[a.c0 + cc] * a.unit_v;
  Tout[cc] = csp * csp * (a.mu_gas * kMH) / (a.gamma * kKB);
}

// ------------------------------------------------------------------ Marchand grain charging
// `charge` of src/chemistry/ionisation.f90:4-214 (Marchand et al. 2021), one thread per cell, in the
// reference's operand order with flang's lowering of the powers (x**2., x**3. as products, x**4.0
// and non-integral exponents through pow, default-REAL literals rounded to single precision).
struct MArgs {
  int64_t nc;
  int nb, nitermax, res;
  double thetai, psi0, x, mu_ions, eps_ionis, B0, rhograin, gamma, unit_l, unit_d, unit_m, unit_v, unit_nh;
  const double *rho, *rhod, *T, *cs, *sdust, *mdust;
  double *tau, *alpha, *nk;        // per (bin, cell) scratch
  double *psi, *zd, *ni, *ne, *cond, *gd;
};

constexpr double kCLight = 3e10;

__global__ void k_mcharge_pre(MArgs a) {
  const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= a.nc * a.nb) return;
  const int64_t i = j % a.nc;
  const double T = a.T[i];
  const double lg = a.sdust[j] * a.unit_l;
  const double tau = lg * kKB * T / (kEel * kEel);
  a.tau[j] = tau;
  a.alpha[j] = sqrt(8.0 / (kPi * tau));
  a.nk[j] = a.rhod[j] * a.unit_d / (a.mdust[j] * a.unit_m);
}

__global__ void k_mcharge(MArgs a) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= a.nc) return;
  const int nb = a.nb;
  const int64_t nc = a.nc;
  const double epsone = 0.99999, th = a.thetai, x = a.x;
  const double T = a.T[i];
  const double sigmav_ie = (2e-7) / sqrt(T / 300.0);
  const double vi = sqrt(8.0 * kKB * T / (kPi * a.mu_ions * kMH));
  const double nH = a.rho[i] * a.unit_nh;
  auto lg2 = [&](int k) {
    const double l = a.sdust[k * nc + i] * a.unit_l;
    return l * l;
  };
  double conv = 1e4, psi = a.psi0;
  int niter = 0, lowT = 0;
  if (a.psi[i] != 0.0) psi = a.psi[i];
  while (conv > a.eps_ionis && niter < a.nitermax) {
    const double eps = (1.0 - psi) / th * exp(-psi);
    const double et = eps * th;
    double n_i = 0.0, s2 = 0.0, s3 = 0.0, s4 = 0.0, s5 = 0.0, s6 = 0.0, s7 = 0.0;
    for (int k = 0; k < nb; ++k) {
      const double tau = a.tau[k * nc + i], al = a.alpha[k * nc + i], nk = a.nk[k * nc + i];
      n_i = n_i + -(psi * tau + (1.0 - et * et) / (1.0 + et * al + et * et)) * nk / (1.0 - eps);
    }
    for (int k = 0; k < nb; ++k) {
      const double tau = a.tau[k * nc + i], al = a.alpha[k * nc + i], nk = a.nk[k * nc + i];
      s2 = s2 + nk * kPi * lg2(k) * ((1.0 - psi) + (2.0 / tau) * (et * et + et) / (1.0 + et * al + et * et));
    }
    const double fpsi = sigmav_ie * eps * (n_i * n_i) / (x * nH) + n_i * vi / (x * nH) * s2 - 1.0;
    const double B = -(eps * (2.0 - psi) / (1.0 - psi));
    for (int k = 0; k < nb; ++k) {
      const double al = a.alpha[k * nc + i], nk = a.nk[k * nc + i];
      const double d0 = 1.0 + et * al + et * et;
      s3 = s3 + nk * (-pow(et, 4.0) + (4.0 * (th * th) - (th * th) * th * al) * (eps * eps) +
                      (2.0 * th * al - 4.0 * (th * th)) * eps + 1.0 - th * al) / (d0 * d0);
    }
    for (int k = 0; k < nb; ++k) s4 = s4 + a.nk[k * nc + i] * a.tau[k * nc + i];
    const double dni = -(1.0 / ((1.0 - eps) * (1.0 - eps)) * B * s3) - (psi / (1.0 - eps) * B + 1.0) * s4 / (1.0 - eps);
    for (int k = 0; k < nb; ++k) {
      const double tau = a.tau[k * nc + i], al = a.alpha[k * nc + i], nk = a.nk[k * nc + i];
      s5 = s5 + nk * kPi * lg2(k) / (tau * (1.0 + et * al + et * et));
    }
    for (int k = 0; k < nb; ++k) {
      const double tau = a.tau[k * nc + i], al = a.alpha[k * nc + i], nk = a.nk[k * nc + i];
      const double d0 = 1.0 + et * al + et * et;
      s6 = s6 + nk * kPi * lg2(k) * (2.0 * eps * (th * th) + th * al) / (tau * (d0 * d0));
    }
    for (int k = 0; k < nb; ++k) s7 = s7 + a.nk[k * nc + i] * kPi * lg2(k);
    const double dfdpsi = sigmav_ie * n_i / (x * nH) * (n_i * B + 2.0 * eps * dni) +
                          2.0 * n_i * vi / (x * nH) * (et * et + et) *
                              ((dni / n_i + B * (2.0 * et + 1.0) / (eps * (1.0 + et))) * s5 + B * s6) +
                          n_i * vi / (x * nH) * (dni / n_i * (1.0 - psi) - 1.0) * s7;
    psi = psi - fpsi / dfdpsi;
    if (psi < a.psi0) {
      psi = a.psi0 * (double)0.9999999f;
      lowT = 1;
    } else if (psi > 0.0) {
      psi = -1e-5;
    }
    conv = fabs(fpsi);
    niter = niter + 1;
  }
  double eps = (1.0 - psi) / th * exp(-psi);
  if (lowT > 0) eps = epsone;
  double n_i = 0.0;
  for (int k = 0; k < nb; ++k) {
    const double tau = a.tau[k * nc + i], al = a.alpha[k * nc + i], nk = a.nk[k * nc + i];
    const double z = psi * tau + (1.0 - (eps * eps) * (th * th)) / (1.0 + eps * th * al + (eps * eps) * (th * th));
    a.zd[k * nc + i] = z;
    n_i = n_i - z * nk / (1.0 - eps);
    if (lowT > 0) n_i = sqrt(eps * x * nH / sigmav_ie);
  }
  const double ne = eps * n_i;
  a.ni[i] = n_i;
  a.ne[i] = ne;
  a.psi[i] = lowT > 0 ? a.psi0 : psi;
  if (!a.res) return;
  const double Bg = a.B0 * sqrt(a.rho[i] * a.unit_nh / 1e4);
  const double mu_i = 2.0 * kMH * a.mu_ions * kMH / (2.0 * kMH + a.mu_ions * kMH);
  const double mu_e = 2.0 * kMH * kMel / (kMel + 2.0 * kMH);
  const double vrms_i = sqrt(8.0 * kKB * T / (kPi * mu_i)) * 1e-5;
  const double vrms_el = sqrt(8.0 * kKB * T / (kPi * mu_e)) * 1e-5;
  const double sigmav_el = 3.16e-11 * pow(vrms_el, (double)1.3f);
  const double sigmav_ions = 2.4e-9 * pow(vrms_i, (double)0.6f);
  const double t_sel = 1.0 / (double)1.16f * ((kMel + 2.0 * kMH) / (2.0 * kMH)) / sigmav_el / nH;
  const double t_sions = 1.0 / (double)1.14f * ((a.mu_ions * kMH + 2.0 * kMH) / (2.0 * kMH)) / sigmav_ions / nH;
  const double s_el = ne * (kEel * kEel) * t_sel / kMel;
  const double s_ions = n_i * (kEel * kEel) * t_sions / (a.mu_ions * kMH);
  const double o_el = -(kEel * Bg / kCLight / kMel);
  const double o_ions = kEel * Bg / kCLight / (a.mu_ions * kMH);
  double sum_sd = 0.0, sum_p = 0.0, sum_h = 0.0;
  for (int k = 0; k < nb; ++k) {
    const double tsd = sqrt(kPi * a.gamma / 8.0) * (a.rhograin) * (a.sdust[k * nc + i] * a.unit_l) /
                       (a.rho[i] * a.unit_d * a.cs[i] * a.unit_v);
    const double z = a.zd[k * nc + i], nk = a.nk[k * nc + i];
    const double mk = a.mdust[k * nc + i] * a.unit_m;
    const double sdk = nk * ((z * kEel) * (z * kEel)) * tsd / mk;
    const double odk = z * kEel * Bg / kCLight / mk;
    sum_sd = sum_sd + sdk;
    sum_p = sum_p + sdk / (1.0 + (odk * tsd) * (odk * tsd));
    sum_h = sum_h + sdk * (odk * tsd) / (1.0 + (odk * tsd) * (odk * tsd));
    if (a.gd) a.gd[k * nc + i] = tsd * odk;
  }
  const double so = sum_sd + s_ions + s_el;
  const double sp = s_ions / (1.0 + (o_ions * t_sions) * (o_ions * t_sions)) +
                    s_el / (1.0 + (o_el * t_sel) * (o_el * t_sel)) + sum_p;
  const double sh = -(s_ions * (o_ions * t_sions) / (1.0 + (o_ions * t_sions) * (o_ions * t_sions))) -
                    s_el * (o_el * t_sel) / (1.0 + (o_el * t_sel) * (o_el * t_sel)) - sum_h;
  if (a.cond) {
    a.cond[0 * nc + i] = 1.0 / so;
    a.cond[1 * nc + i] = sh / (sp * 2. + sh * sh);   // the reference's sigma_p*2. (ionisation.f90:197)
    a.cond[2 * nc + i] = sp / (sp * sp + sh * sh) - 1.0 / so;
    a.cond[3 * nc + i] = so;
    a.cond[4 * nc + i] = sp;
    a.cond[5 * nc + i] = sh;
  }
}

// Ion fraction of a cell for the learned step's features.  The same charge equation as k_gcharge
// with its terms regrouped (k_e = C_e exp(A y), k_i = C_i (1 - A y), A = e^2 / (s k T):
// f = Q (1/k_i - 1/k_e) - y, f' = Q (A/k_e + A C_i/k_i^2) - 1, one exp per iteration), each bin's
// Newton started from the previous bin's root scaled by the size ratio (the surface potential
// A y changes slowly from bin to bin), to |dy| <= 1e-12 |y|; a bin whose warm start does not
// converge in 30 iterations is solved again from y = -1.  The features are rounded to FP32, far
// above the difference to the reference-order solve of k_gcharge.
__global__ void k_fcharge(QArgs a) {
  const int cc = blockIdx.x * blockDim.x + threadIdx.x;
  if (cc >= a.nch) return;
  const int64_t c = a.c0 + cc;
  const double T = a.env[cc];
  const double H = a.hoverr * a.radius[c];
  const double rho = a.rho_gas[c] / (a.s2pi * H) * a.unit_d;
  const double kT = kKB * T, e2 = kEel * kEel;
  const double ve = sqrt(8.0 * kT / (kPi * kMel)), vi = sqrt(8.0 * kT / (kPi * (a.mu_ions * kMH)));
  const double base = a.mu_gas * kMH * a.zeta / rho;
  const double ee = a.electrons ? 1.0 : 0.0;
  double kin = 0.0, y = -1.0, s_prev = 0.0;
  for (int k = 0; k < a.nb; ++k) {
    const double rd = a.rhod[(int64_t)k * a.ld + c] / (a.s2pi * H) * a.unit_d;
    const double m_d = a.mdust[k] * a.unit_m, s_d = a.sdust[k] * a.unit_l;
    const double A = e2 / (s_d * kT);
    const double Ce = kPi * s_d * s_d * ve, Ci = kPi * s_d * s_d * vi;
    const double X = m_d * rho / (rd * a.mu_gas * kMH);
    const double Q = base * X * X;
    const double y_warm = k > 0 ? y * (s_d / s_prev) : -1.0;
    bool ok = false;
    for (int pass = 0; pass < 2 && !ok; ++pass) {
      y = pass == 0 ? y_warm : -1.0;
      const int lim = pass == 0 ? 30 : a.maxiter;
      for (int it = 0; it < lim; ++it) {
        const double yke = exp(-A * y) / Ce, yki = 1.0 / (Ci * (1.0 - A * y));
        const double f = Q * (yki - ee * yke) - y;
        const double fp = Q * (ee * A * yke + Ci * A * (yki * yki)) - 1.0;
        const double d = f / fp;
        y -= d;
        if (!(fabs(y) < 1e300)) break;
        if (fabs(d) <= 1e-12 * fabs(y)) {
          ok = true;
          break;
        }
      }
    }
    s_prev = s_d;
    kin += Ci * (1.0 - A * y) * (rd / m_d);
  }
  a.xion[c] = a.zeta / kin;
}

// ------------------------------------------------------------------ learned growth step
// Features of a cell (n_in = nbins + 5): log10 of each bin's midplane dust density (g cm^-3),
// log10 T, log10 midplane gas density, log10 alpha, log10 ion fraction, log10 dt; standardised
// with the training set's mean and spread.  Outputs (n_out = nbins + 1): the change over dt of the
// natural log of each bin's share of the cell's dust mass, and log10(-<Z>) of the new state,
// both standardised.
struct SArgs {
  int nb, nch, c0, nin, nout;
  int64_t ld;
  double s2pi, hoverr, unit_d, log10_alpha, dt, resid_max, ratio_min;
  float zmax;  // input envelope in standard deviations
  const double* dtc;
  const double* dt_dev;
  const int* cmask;
  double* rhod;
  const double* rho_gas;
  const double* radius;
  const double* env;
  const double* xion;
  float* feat;           // features, [j * fld + cc] (feature-major)
  const float* y;        // network outputs, [k * fld + cc]
  int64_t fld;
  const float* norm;  // in_mean[nin], in_std[nin], out_mean[nout], out_std[nout]; null: raw
  const float* y2;    // residual network's outputs (same layout as y), or null
  const float* norm2; // out_mean2[nout], out_std2[nout] of the residual network
  double* zmean;
  int* fb;            // [ncells] 1 = plain fallback
  double* tg;
  int* nit;
  int* nfb;           // fallback count
  double* resid_sum;  // sum over accepted cells of |sum_k share_k,old exp(change_k) - 1|
};

__global__ void k_surr_feat(SArgs a) {
  const int cc = blockIdx.x * blockDim.x + threadIdx.x;
  if (cc >= a.nch) return;
  const int64_t c = a.c0 + cc;
  const double H = a.hoverr * a.radius[c];
  const double conv = a.unit_d / (a.s2pi * H);
  float* f = a.feat + cc;
  const float* mu = a.norm;
  const float* sd = a.norm ? a.norm + a.nin : nullptr;
  auto put = [&](int j, double v) {
    float x = (float)v;
    if (mu) x = (x - mu[j]) / sd[j];
    f[(int64_t)j * a.fld] = x;
  };
  for (int k = 0; k < a.nb; ++k) put(k, log10(fmax(a.rhod[(int64_t)k * a.ld + c], 1e-300) * conv));
  const double dtl = cell_dt(a, c);
  put(a.nb, log10(a.env[cc]));
  put(a.nb + 1, log10(a.rho_gas[c] * conv));
  put(a.nb + 2, a.log10_alpha);
  put(a.nb + 3, log10(a.xion[c]));
  put(a.nb + 4, log10(dtl));
}

// New state from the network output.  The network predicts the change of the log of each bin's
// share of the cell's dust mass; the new shares are the softmax of (old log-shares + change),
// i.e. w_k = rho_k exp(change_k) normalised by sum_j w_j, so they are positive and sum to one by
// construction and the cell's dust mass is kept in FP64.  Coagulation and fragmentation conserve
// the dust mass, so the raw prediction should too: its mass residual |sum_k share_k exp(change_k)
// - 1| (before the normalisation) measures how far the network is from a consistent answer.  A
// cell falls back to the plain coagulation when that residual exceeds resid_max, when an output is
// not finite, when an old density is not positive, or when an input lies more than zmax standard
// deviations from the training set's mean (outside the envelope the network was trained on).  The
// per-bin values live in registers: every loop over bins is unrolled over MAXB with a bound check.
// Output j of the learned step: the network's de-standardised output, plus the residual
// network's when one is attached (it is trained on what the first network misses).
__device__ __forceinline__ double out_value(const SArgs& a, const float* y, int j, int64_t ld, const float* om,
                                            const float* os) {
  double v = (double)(y[j * ld] * os[j] + om[j]);
  if (a.y2) {
    const float* m2 = a.norm2;
    const float* s2 = a.norm2 + a.nout;
    v += (double)(a.y2[j * ld + (y - a.y)] * s2[j] + m2[j]);
  }
  return v;
}

// one cell of the learned step's new state; returns the cell's raw mass residual (0 for a cell that
// fell back or does not grow), which the kernel sums per wave before one atomic add.  The per-bin
// weights w_k = rho_k exp(change_k) (rounded to FP32: the network's prediction carries ~1e-3, and
// the cell's mass is restored in FP64 from the final densities) are recomputed in each of the
// three passes (sums, floored sum, new densities) rather than held in 64 registers per lane: the
// same operations in the same order, at four times the occupancy.
__device__ __forceinline__ float surr_w(const SArgs& a, const float* y, int k, int64_t ld, const float* om,
                                        const float* os, int64_t c, double* r_out) {
  const double r = a.rhod[(int64_t)k * a.ld + c];
  *r_out = r;
  return (float)(r * exp(out_value(a, y, k, ld, om, os)));
}

__device__ __forceinline__ double surr_post_cell(const SArgs& a, int cc) {
  const int64_t c = a.c0 + cc;
  const float* y = a.y + cc;
  const float* f = a.feat + cc;
  const int64_t ld = a.fld;
  const float* om = a.norm + 2 * a.nin;
  const float* os = om + a.nout;
  const int nb = a.nb;
  if (!(cell_dt(a, c) > 0.0)) {   // no growth interval (a masked ghost cell): unchanged
    a.fb[c] = 0;
    a.tg[c] = 0.0;
    a.nit[c] = 0;
    return 0.0;
  }
  double tot = 0.0, W = 0.0;
  bool fallback = false;
  for (int k = 0; k < nb; ++k) {
    double r;
    const float w = surr_w(a, y, k, ld, om, os, c, &r);
    tot += r;
    if (!(r > 0.0)) fallback = true;
    W += (double)w;
  }
  for (int j = 0; j < a.nin && !fallback; ++j)
    if (!(fabsf(f[j * ld]) <= a.zmax)) fallback = true;
  if (!(W > 0.0) || !(W < INFINITY)) fallback = true;
  if (!fallback && !(fabs(W / tot - 1.0) <= a.resid_max)) fallback = true;   // raw mass residual
  const double dtl = cell_dt(a, c);
  a.nit[c] = 0;
  if (fallback) {
    a.fb[c] = 1;
    a.tg[c] = 0.0;
    atomicAdd(a.nfb, 1);
    return 0.0;
  }
  a.fb[c] = 0;
  a.tg[c] = dtl;
  const double resid = fabs(W / tot - 1.0);   // |sum_k share_k,old exp(change_k) - 1|
  // new densities share * rho_tot, floored as the plain step floors, rescaled to the cell's mass
  const double floor_v = a.rho_gas[c] * a.ratio_min;
  const double sc0 = tot / W;
  double snew = 0.0;
  for (int k = 0; k < nb; ++k) {
    double r;
    snew += fmax((double)surr_w(a, y, k, ld, om, os, c, &r) * sc0, floor_v);
  }
  const double sc = tot / snew;
  for (int k = 0; k < nb; ++k) {
    double r;
    const double v = fmax((double)surr_w(a, y, k, ld, om, os, c, &r) * sc0, floor_v) * sc;
    a.rhod[(int64_t)k * a.ld + c] = v;
  }
  a.zmean[c] = -pow(10.0, out_value(a, y, nb, ld, om, os));
  return resid;
}

// launched with 256 threads: the residuals are summed per block before one atomic add, since the
// adds of all blocks land on one address and serialise there
__global__ __launch_bounds__(256) void k_surr_post(SArgs a) {
  __shared__ double part[4];
  const int cc = blockIdx.x * blockDim.x + threadIdx.x;
  double r = cc < a.nch ? surr_post_cell(a, cc) : 0.0;
  for (int o = 32; o > 0; o >>= 1) r += __shfl_xor(r, o);
  if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = r;
  __syncthreads();
  if (threadIdx.x == 0) {
    const double s = (part[0] + part[1]) + (part[2] + part[3]);
    if (s != 0.0) atomicAdd(a.resid_sum, s);
  }
}

}  // namespace

// ------------------------------------------------------------------ host state and C-ABI
struct shark_growth {
  shark_growth_params p;
  Tables t;
  std::vector<PairC> pairs;
  int64_t ncells = 0;
  int nch_max = 0;
  int kpad = 0;
  hipStream_t stream = nullptr;
  PairC* dpairs = nullptr;
  int* dpij = nullptr;
  double *dredis = nullptr, *dmdust = nullptr, *dsdust = nullptr;
  double *drhod = nullptr, *drho = nullptr, *dcs = nullptr, *drad = nullptr;
  double *denv = nullptr, *dts = nullptr, *dA = nullptr, *dF = nullptr;
  float *dAh = nullptr, *dFh = nullptr;
  double* dtg = nullptr;
  int* dnit = nullptr;
  int* dflags = nullptr;  // [0] unfinished, [1] bad
  int* hflags = nullptr;  // pinned
  float* dest = nullptr;  // per-tile remaining sub-step estimate [tiles of a chunk]
  int* dorder = nullptr;  // tile launch order [tiles of a chunk]
  float* hest = nullptr;  // pinned mirrors
  int* horder = nullptr;
  int* dcfb = nullptr;    // learned step: fallback count per chunk
  int* hcfb = nullptr;    // pinned mirror
  int64_t sch = 0;        // learned step: cells per chunk (features, outputs, temperatures)
  double* dsT = nullptr;  // learned step: gas temperature per chunk cell
  float* dGf32 = nullptr;
  u32x4 *dGbhi = nullptr, *dGblo = nullptr;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  double last_ms = 0.0;
  bool uploaded = false;
  double *dzmean = nullptr, *dxion = nullptr;  // grain charging outputs (allocated on first use)
  double* ddtc = nullptr;                      // per-cell growth interval
  bool use_dtc = false;
  // learned step
  shark_mlp* mlp = nullptr;
  float* dnorm = nullptr;
  float *dfeat = nullptr, *dy = nullptr;
  shark_mlp* mlp2 = nullptr;   // residual network (optional)
  float *dnorm2 = nullptr, *dy2 = nullptr;
  int* dfb = nullptr;
  int* dlist = nullptr;   // learned step: the cells that fell back, then their count
  double* dresid = nullptr;
  // bound to a step context (shark_gpu_attach_growth): rhod, gas density and cs are the step's
  // device arrays, dt is the step's device scalar, ghost cells are masked out
  bool ext = false;
  hipStream_t run_stream = nullptr;
  const double* dt_dev = nullptr;
  int* dmask = nullptr;
  double alpha = 0.0, zeta = 5e-17, mu_ions = 25.0, resid_max = 1e-6;
  float zmax = 8.0f;
  // the plain FP64 step runs one wave per cell (k_gfp64w) for chunks (or cell lists) of at most
  // this many cells, else one lane per cell (k_gfp64).  On the 64-bin disk one wave per cell is 3x
  // faster for 2,048 cells; for 16,384 cells it is faster at ~50 sub-steps per cell and slower at
  // ~500 (the one-lane kernel's latency is then hidden by its many waves)
  int64_t wave_cells = 8192;
  int64_t last_nfb = 0;
  double last_resid = 0.0;
};

void shark_growth_free(shark_growth* g) {
  if (!g) return;
  if (g->ext) g->drhod = g->drho = g->dcs = nullptr;  // the step context owns these
  void* ptrs[] = {g->dpairs, g->dpij, g->dredis, g->dmdust, g->dsdust, g->drhod, g->drho, g->dcs, g->drad,
                  g->denv, g->dts, g->dA, g->dF, g->dAh, g->dFh, g->dtg, g->dnit, g->dflags,
                  g->dGf32, g->dGbhi, g->dGblo, g->dzmean, g->dxion, g->ddtc, g->dnorm,
                  g->dfeat, g->dy, g->dfb, g->dlist, g->dresid, g->dmask, g->dsT, g->dnorm2, g->dy2};
  for (void* q : ptrs)
    if (q) (void)hipFree(q);
  if (g->hflags) (void)hipHostFree(g->hflags);
  if (g->dest) (void)hipFree(g->dest);
  if (g->dorder) (void)hipFree(g->dorder);
  if (g->hest) (void)hipHostFree(g->hest);
  if (g->horder) (void)hipHostFree(g->horder);
  if (g->dcfb) (void)hipFree(g->dcfb);
  if (g->hcfb) (void)hipHostFree(g->hcfb);
  if (g->e0) (void)hipEventDestroy(g->e0);
  if (g->e1) (void)hipEventDestroy(g->e1);
  if (g->stream) (void)hipStreamDestroy(g->stream);
  delete g;
}

#define GCHK(x)                                                                          \
  do {                                                                                   \
    hipError_t e_ = (x);                                                                 \
    if (e_ != hipSuccess) {                                                              \
      char m_[256];                                                                      \
      snprintf(m_, sizeof(m_), "%s failed: %s (%s:%d)", #x, hipGetErrorString(e_), __FILE__, \
               __LINE__);                                                                \
      return shark::fail_msg(SHARK_GROWTH_EHIP, m_);                                      \
    }                                                                                    \
  } while (0)

static int gfail(int code, const char* msg) { return shark::fail_msg(code, msg); }

static unsigned short bf16_host(float x) {
  unsigned int u;
  memcpy(&u, &x, 4);
  u += 0x7FFFu + ((u >> 16) & 1u);
  return (unsigned short)(u >> 16);
}
static float bf16_host_f(unsigned short h) {
  unsigned int u = (unsigned int)h << 16;
  float f;
  memcpy(&f, &u, 4);
  return f;
}

// Gt[k][kk]: kk = 2p (coagulated flux of pair p) or 2p+1 (fragmented flux), per unit mass flux
// (m1+m2) dndt.  Losses -m1/(m1+m2), -m2/(m1+m2); coagulation gains eps and 1-eps at the two
// target bins (times the sticking efficiency, which also scales the coagulation losses);
// fragment gains along the redistribution row of the pair.
static void build_gt(const shark_growth* g, std::vector<double>* gt) {
  const int nb = g->t.nb;
  const int kpad = g->kpad;
  gt->assign((size_t)MAXB * kpad, 0.0);
  for (size_t p = 0; p < g->pairs.size(); ++p) {
    const PairC& P = g->pairs[p];
    const double mt = P.m1 + P.m2;
    double* col_c = gt->data() + 2 * p;
    double* col_f = gt->data() + 2 * p + 1;
    auto at = [&](double* col, int k) -> double& { return col[(size_t)k * kpad]; };
    const double st = g->p.sticking;
    at(col_c, P.i) += -st * P.m1 / mt;
    at(col_c, P.j) += -st * P.m2 / mt;
    at(col_c, P.t1) += st * P.eps;
    at(col_c, P.t2) += st * (1.0 - P.eps);
    if (g->p.fragmentation && g->p.kernel_type == 0) {
      at(col_f, P.i) += -P.m1 / mt;
      at(col_f, P.j) += -P.m2 / mt;
      for (int k = 0; k < P.nfrag; ++k) at(col_f, k) += g->t.redis[P.rrow + k];
    }
  }
  (void)nb;
}

extern "C" {

int shark_growth_create(const shark_growth_params* p, int64_t ncells, shark_growth** out) {
  if (!out) return gfail(SHARK_GROWTH_EINVAL, "shark_growth_create: null output handle");
  *out = nullptr;
  if (!p) return gfail(SHARK_GROWTH_EINVAL, "shark_growth_create: null parameters");
  if (p->version != SHARK_GROWTH_VERSION)
    return gfail(SHARK_GROWTH_EINVAL, "shark_growth_params.version does not match SHARK_GROWTH_VERSION");
  if (p->nbins < 2 || p->nbins > MAXB)
    return gfail(SHARK_GROWTH_EINVAL, "nbins must lie in [2, 64]");
  if (ncells < 1) return gfail(SHARK_GROWTH_EINVAL, "ncells must be positive");
  if (p->env != SHARK_GROWTH_ENV_DISK)
    return gfail(SHARK_GROWTH_EINVAL,
                 "only the disk environment runs on the GPU (the collapse branch needs the "
                 "barotropic EOS and ambipolar velocities of its setup)");
  if (p->kernel_type < 0 || p->kernel_type > 2)
    return gfail(SHARK_GROWTH_EINVAL, "kernel_type must be 0, 1 or 2");
  if (p->kernel_type == 0 && p->fragmentation && !p->decomposed)
    return gfail(SHARK_GROWTH_EINVAL,
                 "fragmentation needs decomposed=1 on the GPU: otherwise the bin a pair feeds "
                 "depends on the cell's fragmented fraction");
  if (p->frag_mode != 0 && p->frag_mode != 1)
    return gfail(SHARK_GROWTH_EINVAL, "frag_mode must be 0 (energy) or 1 (velocity)");
  if (p->max_substeps < 0) return gfail(SHARK_GROWTH_EINVAL, "max_substeps must be >= 0");
  if (!(p->cfl_growth > 0.0)) return gfail(SHARK_GROWTH_EINVAL, "cfl_growth must be positive");
  shark_growth* g = new shark_growth();
  g->p = *p;
  g->ncells = ncells;
  build_tables(p, &g->t);
  build_pairs(p, &g->t, &g->pairs);
  const int np = (int)g->pairs.size();
  g->kpad = ((2 * np + KC - 1) / KC) * KC;
  // per-pair scratch (A, F in FP64 and their FP32 copies: 24 B per pair and cell) is bounded
  // to 16 GiB by processing the cells in chunks (a 64-bin, 512 x 512 disk is one chunk; fewer,
  // larger chunks keep the sub-step kernels' grids deep enough to hide the tiles' imbalance)
  const int64_t cap = ((int64_t)16 << 30) / ((int64_t)np * 24);
  int64_t nch = ncells < cap ? ncells : cap;
  nch = nch >= NCT ? (nch / NCT) * NCT : nch;
  if (nch < 1) nch = 1;
  g->nch_max = (int)nch;
  int rc = 0;
#define GC(x)                                     \
  do {                                            \
    hipError_t e_ = (x);                          \
    if (e_ != hipSuccess) {                       \
      char m_[256];                               \
      snprintf(m_, sizeof(m_), "%s failed: %s", #x, hipGetErrorString(e_)); \
      rc = shark::fail_msg(SHARK_GROWTH_EHIP, m_); \
      shark_growth_free(g);                       \
      return rc;                                  \
    }                                             \
  } while (0)
  GC(hipStreamCreateWithFlags(&g->stream, hipStreamNonBlocking));
  g->run_stream = g->stream;
  GC(hipEventCreate(&g->e0));
  GC(hipEventCreate(&g->e1));
  GC(hipMalloc(&g->dpairs, sizeof(PairC) * np));
  GC(hipMalloc(&g->dpij, sizeof(int) * (g->kpad / 2)));
  GC(hipMalloc(&g->dredis, sizeof(double) * MAXB * MAXB));
  GC(hipMalloc(&g->dmdust, sizeof(double) * MAXB));
  GC(hipMalloc(&g->dsdust, sizeof(double) * MAXB));
  GC(hipMalloc(&g->drhod, sizeof(double) * p->nbins * ncells));
  GC(hipMalloc(&g->drho, sizeof(double) * ncells));
  GC(hipMalloc(&g->dcs, sizeof(double) * ncells));
  GC(hipMalloc(&g->drad, sizeof(double) * ncells));
  GC(hipMalloc(&g->dtg, sizeof(double) * ncells));
  GC(hipMalloc(&g->dnit, sizeof(int) * ncells));
  GC(hipMalloc(&g->denv, sizeof(double) * 4 * nch));
  GC(hipMalloc(&g->dts, sizeof(double) * p->nbins * nch));
  GC(hipMalloc(&g->dA, sizeof(double) * np * nch));
  GC(hipMalloc(&g->dF, sizeof(double) * np * nch));
  GC(hipMalloc(&g->dAh, sizeof(float) * np * nch));
  GC(hipMalloc(&g->dFh, sizeof(float) * np * nch));
  GC(hipMalloc(&g->dflags, sizeof(int) * 2));
  GC(hipHostMalloc(&g->hflags, sizeof(int) * 2, hipHostMallocDefault));
  {
    const int64_t ntile = (nch + 63) / 64;
    GC(hipMalloc(&g->dest, sizeof(float) * ntile));
    GC(hipMalloc(&g->dorder, sizeof(int) * ntile));
    GC(hipHostMalloc(&g->hest, sizeof(float) * ntile, hipHostMallocDefault));
    GC(hipHostMalloc(&g->horder, sizeof(int) * ntile, hipHostMallocDefault));
  }
  GC(hipMemcpy(g->dpairs, g->pairs.data(), sizeof(PairC) * np, hipMemcpyHostToDevice));
  {
    std::vector<int> pij(g->kpad / 2, 0);   // pairs beyond np are never read (p < npairs)
    for (int q = 0; q < np; ++q) pij[q] = g->pairs[q].i | (g->pairs[q].j << 8);
    GC(hipMemcpy(g->dpij, pij.data(), sizeof(int) * pij.size(), hipMemcpyHostToDevice));
  }
  GC(hipMemcpy(g->dredis, g->t.redis, sizeof(double) * MAXB * MAXB, hipMemcpyHostToDevice));
  GC(hipMemcpy(g->dmdust, g->t.mdust, sizeof(double) * MAXB, hipMemcpyHostToDevice));
  GC(hipMemcpy(g->dsdust, g->t.sdust, sizeof(double) * MAXB, hipMemcpyHostToDevice));
  // contraction operands in the MFMA lane layouts
  std::vector<double> gt;
  build_gt(g, &gt);
  const int kpad = g->kpad;
  {
    // FP32, v_mfma_f32_32x32x2f32: lane l of step s supplies Gt[rb*32 + l%32][2s + l/32]
    std::vector<float> a32((size_t)2 * (kpad / 2) * 64);
    for (int rb = 0; rb < 2; ++rb)
      for (int s = 0; s < kpad / 2; ++s)
        for (int l = 0; l < 64; ++l)
          a32[((size_t)rb * (kpad / 2) + s) * 64 + l] = (float)gt[(size_t)(rb * 32 + (l & 31)) * kpad + 2 * s + (l >> 5)];
    GC(hipMalloc(&g->dGf32, sizeof(float) * a32.size()));
    GC(hipMemcpy(g->dGf32, a32.data(), sizeof(float) * a32.size(), hipMemcpyHostToDevice));
    // BF16, v_mfma_f32_32x32x16_bf16: lane l of step s supplies Gt[rb*32 + l%32][16s + 8(l/32) + e]
    const size_t n16 = (size_t)2 * (kpad / 16) * 64;
    std::vector<unsigned short> hi(n16 * 8), lo(n16 * 8);
    for (int rb = 0; rb < 2; ++rb)
      for (int s = 0; s < kpad / 16; ++s)
        for (int l = 0; l < 64; ++l)
          for (int e = 0; e < 8; ++e) {
            const float v = (float)gt[(size_t)(rb * 32 + (l & 31)) * kpad + 16 * s + 8 * (l >> 5) + e];
            const unsigned short h = bf16_host(v);
            const size_t o = (((size_t)rb * (kpad / 16) + s) * 64 + l) * 8 + e;
            hi[o] = h;
            lo[o] = bf16_host(v - bf16_host_f(h));
          }
    GC(hipMalloc(&g->dGbhi, 16 * n16));
    GC(hipMalloc(&g->dGblo, 16 * n16));
    GC(hipMemcpy(g->dGbhi, hi.data(), 16 * n16, hipMemcpyHostToDevice));
    GC(hipMemcpy(g->dGblo, lo.data(), 16 * n16, hipMemcpyHostToDevice));
  }
#undef GC
  *out = g;
  return 0;
}

int shark_growth_upload(shark_growth* g, const double* rhod, const double* rho_gas, const double* cs,
                        const double* radius) {
  if (!g || !rhod || !rho_gas || !cs || !radius)
    return gfail(SHARK_GROWTH_EINVAL, "shark_growth_upload: null argument");
  const int64_t n = g->ncells;
  GCHK(hipMemcpy(g->drhod, rhod, sizeof(double) * g->p.nbins * n, hipMemcpyHostToDevice));
  GCHK(hipMemcpy(g->drho, rho_gas, sizeof(double) * n, hipMemcpyHostToDevice));
  GCHK(hipMemcpy(g->dcs, cs, sizeof(double) * n, hipMemcpyHostToDevice));
  GCHK(hipMemcpy(g->drad, radius, sizeof(double) * n, hipMemcpyHostToDevice));
  g->uploaded = true;
  return 0;
}

static GArgs make_args(shark_growth* g, int mode, double dt) {
  const shark_growth_params& p = g->p;
  GArgs a;
  memset(&a, 0, sizeof(a));
  a.nb = p.nbins;
  a.npairs = (int)g->pairs.size();
  a.ld = g->ncells;
  a.kernel_type = p.kernel_type;
  a.frag = p.fragmentation ? 1 : 0;
  a.frag_mode = p.frag_mode;
  a.brownian = p.brownian;
  a.turb = p.turb;
  a.drift = p.drift;
  a.ambipolar = p.ambipolar;
  a.max_sub = p.max_substeps;
  a.sub_per_launch = SUB_PER_LAUNCH;
  a.kpad = g->kpad;
  a.want64 = mode == SHARK_GROWTH_FP64 || mode == SHARK_GROWTH_SURROGATE;
  a.want32 = !a.want64;
  a.dt = dt;
  a.cfl = p.cfl_growth;
  a.eps_thr = p.eps_threshold;
  a.eps_thr_frag = p.eps_threshold_frag;
  a.ratio_min = p.dust_ratio_min;
  a.sticking = p.sticking;
  a.vfrag_code = p.vfrag / p.unit_v;
  a.m_mono = g->t.m_mono;
  a.ebr_mono = g->t.ebr_mono;
  a.turb_coef = p.turb_coef;
  a.unit_v = p.unit_v;
  a.unit_m = p.unit_m;
  a.unit_d = p.unit_d;
  a.mu_gas = p.mu_gas;
  a.gamma = p.gamma;
  a.hoverr = p.hoverr;
  a.mstar = p.mstar;
  a.kts = sqrt(kPi * p.gamma / 8.0) * (p.rhograin / p.unit_d);
  a.s2pi = sqrt(2.0 * kPi);
  a.pairs = g->dpairs;
  a.pij = g->dpij;
  a.redis = g->dredis;
  a.mdust = g->dmdust;
  a.sdust = g->dsdust;
  a.rhod = g->drhod;
  a.rho_gas = g->drho;
  a.cs = g->dcs;
  a.radius = g->drad;
  a.env = g->denv;
  a.ts = g->dts;
  a.A = g->dA;
  a.F = g->dF;
  a.Ah = g->dAh;
  a.Fh = g->dFh;
  a.dtc = g->use_dtc ? g->ddtc : nullptr;
  a.dt_dev = g->dt_dev;
  a.cmask = g->dmask;
  a.tg = g->dtg;
  a.nit = g->dnit;
  a.unfinished = g->dflags;
  a.bad = g->dflags + 1;
  a.Gf32 = g->dGf32;
  a.Gbhi = g->dGbhi;
  a.Gblo = g->dGblo;
  return a;
}

static QArgs make_qargs(shark_growth* g) {
  const shark_growth_params& p = g->p;
  QArgs q;
  memset(&q, 0, sizeof(q));
  q.nb = p.nbins;
  q.ld = g->ncells;
  q.electrons = 1;
  q.maxiter = 100;
  q.zeta = g->zeta;
  q.mu_ions = g->mu_ions;
  q.mu_gas = p.mu_gas;
  q.hoverr = p.hoverr;
  q.unit_d = p.unit_d;
  q.unit_m = p.unit_m;
  q.unit_l = p.unit_l;
  q.s2pi = sqrt(2.0 * kPi);
  q.rhod = g->drhod;
  q.rho_gas = g->drho;
  q.radius = g->drad;
  q.env = g->denv;
  q.mdust = g->dmdust;
  q.sdust = g->dsdust;
  q.zmean = g->dzmean;
  q.xion = g->dxion;
  return q;
}

static int ensure_charge_buffers(shark_growth* g) {
  if (!g->dzmean) {
    GCHK(hipMalloc(&g->dzmean, sizeof(double) * g->ncells));
    GCHK(hipMalloc(&g->dxion, sizeof(double) * g->ncells));
  }
  return 0;
}

// pair prefactors for the chunk, then the sub-step kernel: one sub-step for every tile, then the
// unfinished tiles, slowest first (their remaining sub-steps estimated from the first), relaunched
// until every active cell of the chunk reached its interval.  The order changes which tiles run
// side by side, not any cell's arithmetic.
static int plain_chunk(shark_growth* g, GArgs& a, int mode) {
  hipStream_t s = g->run_stream;
  const int nblk = (a.nch + 255) / 256;
  hipLaunchKernelGGL(k_gpairs, dim3(nblk, a.npairs), dim3(256), 0, s, a);
  GCHK(hipGetLastError());
  if (mode == SHARK_GROWTH_FP64 && a.nch <= g->wave_cells) {
    // one wave per cell: every wave runs its cell to the end of the interval, so one launch
    // normally finishes the chunk
    a.tile_order = nullptr;
    a.sub_per_launch = SUB_PER_LAUNCH;
    for (;;) {
      GCHK(hipMemsetAsync(g->dflags, 0, sizeof(int), s));
      hipLaunchKernelGGL(k_gfp64w, dim3((a.nch + 3) / 4), dim3(256), sizeof(double) * a.nb * a.nb, s, a);
      GCHK(hipGetLastError());
      GCHK(hipMemcpyAsync(g->hflags, g->dflags, sizeof(int) * 2, hipMemcpyDeviceToHost, s));
      GCHK(hipStreamSynchronize(s));
      if (g->hflags[1]) return gfail(SHARK_GROWTH_ENUM, "growth produced a non-finite density or a zero sub-step");
      if (g->hflags[0] == 0) return 0;
    }
  }
  static_assert(NCT == 64, "both sub-step kernels take 64-cell tiles");
  const int ntile = (a.nch + 63) / 64;
  int grid = ntile;
  a.tile_order = nullptr;
  for (int launch = 0;; ++launch) {
    a.sub_per_launch = launch == 0 ? 1 : SUB_PER_LAUNCH;
    GCHK(hipMemsetAsync(g->dflags, 0, sizeof(int), s));
    if (mode == SHARK_GROWTH_FP64) {
      hipLaunchKernelGGL(k_gfp64, dim3(grid), dim3(64), sizeof(double) * 64 * a.nb, s, a);
    } else if (mode == SHARK_GROWTH_MFMA_FP32) {
      hipLaunchKernelGGL(k_gmfma<1>, dim3(grid), dim3(256), 0, s, a);
    } else if (mode == SHARK_GROWTH_MFMA_BF16) {
      hipLaunchKernelGGL(k_gmfma<2>, dim3(grid), dim3(256), 0, s, a);
    } else {
      hipLaunchKernelGGL(k_gmfma<3>, dim3(grid), dim3(256), 0, s, a);
    }
    GCHK(hipGetLastError());
    GCHK(hipMemcpyAsync(g->hflags, g->dflags, sizeof(int) * 2, hipMemcpyDeviceToHost, s));
    GCHK(hipStreamSynchronize(s));
    if (g->hflags[1]) return gfail(SHARK_GROWTH_ENUM, "growth produced a non-finite density or a zero sub-step");
    if (g->hflags[0] == 0) break;
    if (launch == 0) {
      hipLaunchKernelGGL(k_tile_est, dim3(ntile), dim3(64), 0, s, a, g->dest);
      GCHK(hipGetLastError());
      GCHK(hipMemcpyAsync(g->hest, g->dest, sizeof(float) * ntile, hipMemcpyDeviceToHost, s));
      GCHK(hipStreamSynchronize(s));
      int n = 0;
      for (int t = 0; t < ntile; ++t)
        if (g->hest[t] >= 0.f) g->horder[n++] = t;
      const float* est = g->hest;
      std::stable_sort(g->horder, g->horder + n, [est](int x, int y) { return est[x] > est[y]; });
      GCHK(hipMemcpyAsync(g->dorder, g->horder, sizeof(int) * n, hipMemcpyHostToDevice, s));
      a.tile_order = g->dorder;
      grid = n;
    }
  }
  a.tile_order = nullptr;
  a.sub_per_launch = SUB_PER_LAUNCH;
  return 0;
}

static SArgs make_sargs(shark_growth* g, double dt) {
  const shark_growth_params& p = g->p;
  SArgs sa;
  memset(&sa, 0, sizeof(sa));
  sa.nb = p.nbins;
  sa.ld = g->ncells;
  sa.nin = p.nbins + 5;
  sa.nout = p.nbins + 1;
  sa.s2pi = sqrt(2.0 * kPi);
  sa.hoverr = p.hoverr;
  sa.unit_d = p.unit_d;
  sa.log10_alpha = log10(g->alpha > 0.0 ? g->alpha : p.turb_coef * p.turb_coef);
  sa.dt = dt;
  sa.resid_max = g->resid_max;
  sa.zmax = g->zmax;
  sa.ratio_min = p.dust_ratio_min;
  sa.dtc = g->use_dtc ? g->ddtc : nullptr;
  sa.dt_dev = g->dt_dev;
  sa.cmask = g->dmask;
  sa.rhod = g->drhod;
  sa.rho_gas = g->drho;
  sa.radius = g->drad;
  sa.env = g->denv;
  sa.xion = g->dxion;
  sa.zmean = g->dzmean;
  sa.fb = g->dfb;
  sa.tg = g->dtg;
  sa.nit = g->dnit;
  sa.nfb = g->dflags;
  sa.resid_sum = g->dresid;
  return sa;
}

// learned step over all cells in chunks of up to 4M cells: temperature, ion fraction (an input),
// features, network (one fused kernel), new state.  One host synchronisation reads every chunk's
// fallback count; the cells that fell back then run the plain FP64 coagulation (in the plain
// step's own chunks) and the charge solve.
static int surrogate_run(shark_growth* g, double dt) {
  hipStream_t s = g->run_stream;
  GArgs a = make_args(g, SHARK_GROWTH_SURROGATE, dt);
  QArgs q = make_qargs(g);
  SArgs sa = make_sargs(g, dt);
  sa.norm = g->dnorm;
  sa.feat = g->dfeat;
  sa.y = g->dy;
  sa.y2 = g->mlp2 ? g->dy2 : nullptr;
  sa.norm2 = g->dnorm2;
  sa.env = g->dsT;
  sa.fld = g->sch;
  q.env = g->dsT;
  const int64_t B = g->sch;
  const int nchunk = (int)((g->ncells + B - 1) / B);
  GCHK(hipMemsetAsync(g->dresid, 0, sizeof(double), s));
  GCHK(hipMemsetAsync(g->dcfb, 0, sizeof(int) * nchunk, s));
  for (int i = 0; i < nchunk; ++i) {
    const int64_t c0 = (int64_t)i * B, rem = g->ncells - c0;
    const int nch = (int)(rem < B ? rem : B);
    a.c0 = q.c0 = sa.c0 = (int)c0;
    a.nch = q.nch = sa.nch = nch;
    sa.nfb = g->dcfb + i;
    const int nblk = (nch + 255) / 256;
    hipLaunchKernelGGL(k_senv, dim3(nblk), dim3(256), 0, s, a, g->dsT);
    GCHK(hipGetLastError());
    hipLaunchKernelGGL(k_fcharge, dim3(nblk), dim3(256), 0, s, q);
    GCHK(hipGetLastError());
    hipLaunchKernelGGL(k_surr_feat, dim3(nblk), dim3(256), 0, s, sa);
    GCHK(hipGetLastError());
    // the first network with hi + lo activations: the BF16 rounding of its activations would add
    // about as much error as the network's own.  The residual network's outputs are a few 1e-4 of
    // the first's, so its own BF16 rounding noise stays far below the target and it runs the
    // single-product kernel.
    int rc = shark::mlp_infer_device(g->mlp, g->dfeat, nch, g->dy, s, B, B, true);
    if (rc) return rc;
    if (g->mlp2) {
      rc = shark::mlp_infer_device(g->mlp2, g->dfeat, nch, g->dy2, s, B, B, false);
      if (rc) return rc;
    }
    hipLaunchKernelGGL(k_surr_post, dim3(nblk), dim3(256), 0, s, sa);
    GCHK(hipGetLastError());
  }
  double rs = 0.0;
  GCHK(hipMemcpyAsync(g->hcfb, g->dcfb, sizeof(int) * nchunk, hipMemcpyDeviceToHost, s));
  GCHK(hipMemcpyAsync(&rs, g->dresid, sizeof(double), hipMemcpyDeviceToHost, s));
  GCHK(hipStreamSynchronize(s));
  int64_t nfb_total = 0;
  for (int i = 0; i < nchunk; ++i) nfb_total += g->hcfb[i];
  if (nfb_total > 0) {
    // the cells that fell back, as a list: the plain FP64 coagulation (one wave per cell unless
    // the list is long), then the exact charge solve (plain environment), over pieces of the list
    int* cnt = g->dlist + g->ncells;
    GCHK(hipMemsetAsync(cnt, 0, sizeof(int), s));
    hipLaunchKernelGGL(k_flag_list, dim3((unsigned)((g->ncells + 255) / 256)), dim3(256), 0, s,
                       (const int*)g->dfb, g->ncells, g->dlist, cnt);
    GCHK(hipGetLastError());
    QArgs qf = make_qargs(g);
    for (int64_t o = 0; o < nfb_total; o += g->nch_max) {
      const int nch = (int)std::min<int64_t>(g->nch_max, nfb_total - o);
      a.c0 = qf.c0 = 0;
      a.cells = qf.cells = g->dlist + o;
      a.nch = qf.nch = nch;
      a.env = g->denv;
      qf.env = g->denv;
      const int nblk = (nch + 255) / 256;
      hipLaunchKernelGGL(k_genv, dim3(nblk), dim3(256), 0, s, a);
      GCHK(hipGetLastError());
      int rc = plain_chunk(g, a, SHARK_GROWTH_FP64);
      if (rc) return rc;
      hipLaunchKernelGGL(k_gcharge, dim3(nblk), dim3(256), 0, s, qf);
      GCHK(hipGetLastError());
    }
    a.cells = nullptr;
  }
  g->last_nfb = nfb_total;
  const int64_t nacc = g->ncells - nfb_total;
  g->last_resid = nacc > 0 ? rs / (double)nacc : 0.0;
  return 0;
}

static int run_impl(shark_growth* g, int mode, double dt) {
  if (!g) return gfail(SHARK_GROWTH_EINVAL, "shark_growth_run: null handle");
  if (!g->uploaded) return gfail(SHARK_GROWTH_EINVAL, "shark_growth_run: upload the state first");
  if (mode < SHARK_GROWTH_FP64 || mode > SHARK_GROWTH_SURROGATE)
    return gfail(SHARK_GROWTH_EINVAL, "shark_growth_run: unknown mode");
  if (!(dt >= 0.0)) return gfail(SHARK_GROWTH_EINVAL, "shark_growth_run: dt must be >= 0");
  if (mode == SHARK_GROWTH_SURROGATE && !g->mlp)
    return gfail(SHARK_GROWTH_EINVAL, "shark_growth_run: no surrogate attached (shark_growth_set_surrogate)");
  hipStream_t s = g->run_stream;
  GCHK(hipMemsetAsync(g->dtg, 0, sizeof(double) * g->ncells, s));
  GCHK(hipMemsetAsync(g->dnit, 0, sizeof(int) * g->ncells, s));
  GCHK(hipMemsetAsync(g->dflags, 0, sizeof(int) * 2, s));
  GCHK(hipEventRecord(g->e0, s));
  if (mode == SHARK_GROWTH_SURROGATE) {
    int rc = ensure_charge_buffers(g);
    if (rc) return rc;
    rc = surrogate_run(g, dt);
    if (rc) return rc;
  } else {
    GArgs a = make_args(g, mode, dt);
    for (int64_t c0 = 0; c0 < g->ncells; c0 += g->nch_max) {
      const int64_t rem = g->ncells - c0;
      a.c0 = (int)c0;
      a.nch = (int)(rem < g->nch_max ? rem : g->nch_max);
      hipLaunchKernelGGL(k_genv, dim3((a.nch + 255) / 256), dim3(256), 0, s, a);
      GCHK(hipGetLastError());
      int rc = plain_chunk(g, a, mode);
      if (rc) return rc;
    }
  }
  GCHK(hipEventRecord(g->e1, s));
  GCHK(hipEventSynchronize(g->e1));
  float ms = 0.f;
  GCHK(hipEventElapsedTime(&ms, g->e0, g->e1));
  g->last_ms = ms;
  return 0;
}

int shark_growth_run(shark_growth* g, int mode, double dt) {
  if (g && g->ext)
    return gfail(SHARK_GROWTH_EINVAL, "shark_growth_run: this state is attached to a step context and grows inside its steps");
  return run_impl(g, mode, dt);
}

}  // extern "C"

namespace shark {
// Binds a growth state to a step's device arrays: the dust-density planes of U (rhod, leading
// dimension ncells), the gas density plane, cs, the per-column radius, the step's device dt, and a
// mask of the active cells (ghost cells do not grow).
int growth_bind(shark_growth* g, double* dU, double* dcs, const double* rc_cols, int nxm, int nym, int nx,
                int ny, int ndim, const double* dt_dev, hipStream_t s) {
  const int64_t nc = (int64_t)nxm * nym;
  if (g->ncells != nc)
    return gfail(SHARK_GROWTH_EINVAL, "shark_gpu_attach_growth: the growth state must be created for the step's ncells (ghosts included)");
  std::vector<double> rad(nc);
  std::vector<int> mask(nc);
  for (int64_t c = 0; c < nc; ++c) {
    const int ix = (int)(c % nxm) + 1, iy = (ndim == 1) ? 1 : (int)(c / nxm) + 1;
    rad[c] = rc_cols[ix - 1];
    mask[c] = (ix >= 3 && ix <= nx + 2 && (ndim == 1 || (iy >= 3 && iy <= ny + 2))) ? 1 : 0;
  }
  if (!g->ext) {
    (void)hipFree(g->drhod);
    (void)hipFree(g->drho);
    (void)hipFree(g->dcs);
  }
  g->drhod = dU + 5 * nc;
  g->drho = dU;
  g->dcs = dcs;
  g->ext = true;
  if (!g->dmask) GCHK(hipMalloc(&g->dmask, sizeof(int) * nc));
  GCHK(hipMemcpy(g->drad, rad.data(), sizeof(double) * nc, hipMemcpyHostToDevice));
  GCHK(hipMemcpy(g->dmask, mask.data(), sizeof(int) * nc, hipMemcpyHostToDevice));
  g->dt_dev = dt_dev;
  g->run_stream = s;
  g->use_dtc = false;
  g->uploaded = true;
  return 0;
}

// One growth call of the bound state over the step's dt, on the step's stream.
int growth_step(shark_growth* g, int mode) { return run_impl(g, mode, 0.0); }

int growth_nbins(const shark_growth* g) { return g->p.nbins; }
}  // namespace shark

extern "C" {

int shark_growth_set_wave_cells(shark_growth* g, int64_t max_cells) {
  if (!g) return gfail(SHARK_GROWTH_EINVAL, "shark_growth_set_wave_cells: null handle");
  if (max_cells < 0) return gfail(SHARK_GROWTH_EINVAL, "shark_growth_set_wave_cells: max_cells must be >= 0");
  g->wave_cells = max_cells;
  return 0;
}

int shark_growth_set_cell_dt(shark_growth* g, const double* dt_cells) {
  if (!g) return gfail(SHARK_GROWTH_EINVAL, "shark_growth_set_cell_dt: null handle");
  if (!dt_cells) {
    g->use_dtc = false;
    return 0;
  }
  for (int64_t c = 0; c < g->ncells; ++c)
    if (!(dt_cells[c] >= 0.0)) return gfail(SHARK_GROWTH_EINVAL, "shark_growth_set_cell_dt: intervals must be >= 0");
  if (!g->ddtc) GCHK(hipMalloc(&g->ddtc, sizeof(double) * g->ncells));
  GCHK(hipMemcpy(g->ddtc, dt_cells, sizeof(double) * g->ncells, hipMemcpyHostToDevice));
  g->use_dtc = true;
  return 0;
}

int shark_growth_set_surrogate(shark_growth* g, shark_mlp* m, const float* norm, double alpha,
                               double zeta, double mu_ions, double resid_max) {
  if (!g) return gfail(SHARK_GROWTH_EINVAL, "shark_growth_set_surrogate: null handle");
  if (!m) {
    g->mlp = nullptr;
    g->mlp2 = nullptr;
    return 0;
  }
  const int nb = g->p.nbins;
  if (shark::mlp_n_in(m) != nb + 5 || shark::mlp_n_out(m) != nb + 1)
    return gfail(SHARK_GROWTH_EINVAL, "shark_growth_set_surrogate: the network needs nbins+5 inputs and nbins+1 outputs");
  g->mlp2 = nullptr;   // a residual network belongs to the network it was trained on
  if (!shark::mlp_fused_ok(m))
    return gfail(SHARK_GROWTH_EINVAL, "shark_growth_set_surrogate: the network needs <= 8 layers, widths <= 256 and hidden widths that are multiples of 32");
  if (!norm) return gfail(SHARK_GROWTH_EINVAL, "shark_growth_set_surrogate: null normalisation");
  if (!(resid_max > 0.0) || !(zeta > 0.0) || !(mu_ions > 0.0))
    return gfail(SHARK_GROWTH_EINVAL, "shark_growth_set_surrogate: resid_max, zeta and mu_ions must be positive");
  const int nn = 2 * (nb + 5) + 2 * (nb + 1);
  for (int i = 0; i < nn; ++i)
    if (!(norm[i] == norm[i])) return gfail(SHARK_GROWTH_EINVAL, "shark_growth_set_surrogate: non-finite normalisation");
  const int64_t B = std::min<int64_t>(g->ncells, (int64_t)1 << 22);   // learned step's chunk
  const int64_t nchunk = (g->ncells + B - 1) / B;
  if (!g->dnorm) GCHK(hipMalloc(&g->dnorm, sizeof(float) * 2 * (MAXB + 5) + sizeof(float) * 2 * (MAXB + 1)));
  if (!g->dcfb) {
    GCHK(hipMalloc(&g->dcfb, sizeof(int) * nchunk));
    GCHK(hipHostMalloc(&g->hcfb, sizeof(int) * nchunk, hipHostMallocDefault));
  }
  if (!g->dsT) GCHK(hipMalloc(&g->dsT, sizeof(double) * B));
  g->sch = B;
  if (g->dfeat) (void)hipFree(g->dfeat);
  if (g->dy) (void)hipFree(g->dy);
  g->dfeat = nullptr;
  g->dy = nullptr;
  GCHK(hipMalloc(&g->dfeat, sizeof(float) * B * (nb + 5)));
  GCHK(hipMalloc(&g->dy, sizeof(float) * B * (nb + 1)));
  if (!g->dfb) GCHK(hipMalloc(&g->dfb, sizeof(int) * g->ncells));
  if (!g->dlist) GCHK(hipMalloc(&g->dlist, sizeof(int) * (g->ncells + 1)));
  if (!g->dresid) GCHK(hipMalloc(&g->dresid, sizeof(double)));
  GCHK(hipMemcpy(g->dnorm, norm, sizeof(float) * nn, hipMemcpyHostToDevice));
  int rc = ensure_charge_buffers(g);
  if (rc) return rc;
  g->mlp = m;
  g->alpha = alpha;
  g->zeta = zeta;
  g->mu_ions = mu_ions;
  g->resid_max = resid_max;
  return 0;
}

int shark_growth_features(shark_growth* g, double dt, float* feat) {
  if (!g || !feat) return gfail(SHARK_GROWTH_EINVAL, "shark_growth_features: null argument");
  if (!g->uploaded) return gfail(SHARK_GROWTH_EINVAL, "shark_growth_features: upload the state first");
  int rc = ensure_charge_buffers(g);
  if (rc) return rc;
  const int nin = g->p.nbins + 5;
  hipStream_t s = g->stream;
  GArgs a = make_args(g, SHARK_GROWTH_FP64, dt);
  QArgs q = make_qargs(g);
  SArgs sa = make_sargs(g, dt);
  const int64_t B = g->nch_max;
  float* dfe = nullptr;
  GCHK(hipMalloc(&dfe, sizeof(float) * B * nin));
  sa.feat = dfe;
  sa.fld = B;
  sa.norm = nullptr;
  std::vector<float> tmp((size_t)B * nin);
  for (int64_t c0 = 0; c0 < g->ncells; c0 += B) {
    const int64_t rem = g->ncells - c0;
    const int nch = (int)(rem < B ? rem : B);
    a.c0 = q.c0 = sa.c0 = (int)c0;
    a.nch = q.nch = sa.nch = nch;
    const int nblk = (nch + 255) / 256;
    hipLaunchKernelGGL(k_genv, dim3(nblk), dim3(256), 0, s, a);
    hipLaunchKernelGGL(k_fcharge, dim3(nblk), dim3(256), 0, s, q);
    hipLaunchKernelGGL(k_surr_feat, dim3(nblk), dim3(256), 0, s, sa);
    hipError_t e = hipGetLastError();
    // feature-major on the device, row-major for the caller
    if (e == hipSuccess) e = hipMemcpy2DAsync(tmp.data(), sizeof(float) * nch, dfe, sizeof(float) * B,
                                              sizeof(float) * nch, nin, hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e != hipSuccess) {
      (void)hipFree(dfe);
      GCHK(e);
    }
    for (int j = 0; j < nin; ++j)
      for (int i = 0; i < nch; ++i) feat[(c0 + i) * nin + j] = tmp[(size_t)j * nch + i];
  }
  GCHK(hipFree(dfe));
  return 0;
}

int shark_growth_set_surrogate_residual(shark_growth* g, shark_mlp* m2, const float* out_norm2) {
  if (!g) return gfail(SHARK_GROWTH_EINVAL, "shark_growth_set_surrogate_residual: null handle");
  if (!m2) {
    g->mlp2 = nullptr;
    return 0;
  }
  if (!g->mlp) return gfail(SHARK_GROWTH_EINVAL, "shark_growth_set_surrogate_residual: attach the first network first");
  const int nb = g->p.nbins;
  if (shark::mlp_n_in(m2) != nb + 5 || shark::mlp_n_out(m2) != nb + 1 || !shark::mlp_fused_ok(m2))
    return gfail(SHARK_GROWTH_EINVAL, "shark_growth_set_surrogate_residual: the network needs nbins+5 inputs, nbins+1 outputs and the fused kernel's shape");
  if (!out_norm2) return gfail(SHARK_GROWTH_EINVAL, "shark_growth_set_surrogate_residual: null normalisation");
  for (int i = 0; i < 2 * (nb + 1); ++i)
    if (!(out_norm2[i] == out_norm2[i])) return gfail(SHARK_GROWTH_EINVAL, "shark_growth_set_surrogate_residual: non-finite normalisation");
  if (!g->dnorm2) GCHK(hipMalloc(&g->dnorm2, sizeof(float) * 2 * (MAXB + 1)));
  if (g->dy2) (void)hipFree(g->dy2);
  g->dy2 = nullptr;
  GCHK(hipMalloc(&g->dy2, sizeof(float) * g->sch * (nb + 1)));
  GCHK(hipMemcpy(g->dnorm2, out_norm2, sizeof(float) * 2 * (nb + 1), hipMemcpyHostToDevice));
  g->mlp2 = m2;
  return 0;
}

int shark_growth_zmean(shark_growth* g, double* zmean) {
  if (!g || !zmean) return gfail(SHARK_GROWTH_EINVAL, "shark_growth_zmean: null argument");
  if (!g->dzmean) return gfail(SHARK_GROWTH_EINVAL, "shark_growth_zmean: no learned step or charge solve has run");
  GCHK(hipStreamSynchronize(g->run_stream));
  GCHK(hipStreamSynchronize(g->stream));
  GCHK(hipMemcpy(zmean, g->dzmean, sizeof(double) * g->ncells, hipMemcpyDeviceToHost));
  return 0;
}

int shark_growth_surrogate_stats(shark_growth* g, int64_t* n_fallback, double* raw_resid_mean) {
  if (!g) return gfail(SHARK_GROWTH_EINVAL, "shark_growth_surrogate_stats: null handle");
  if (n_fallback) *n_fallback = g->last_nfb;
  if (raw_resid_mean) *raw_resid_mean = g->last_resid;
  return 0;
}

int shark_growth_download(shark_growth* g, double* rhod, int64_t* nsub) {
  if (!g) return gfail(SHARK_GROWTH_EINVAL, "shark_growth_download: null handle");
  GCHK(hipStreamSynchronize(g->stream));
  if (rhod) GCHK(hipMemcpy(rhod, g->drhod, sizeof(double) * g->p.nbins * g->ncells, hipMemcpyDeviceToHost));
  if (nsub) {
    std::vector<int> n(g->ncells);
    GCHK(hipMemcpy(n.data(), g->dnit, sizeof(int) * g->ncells, hipMemcpyDeviceToHost));
    for (int64_t c = 0; c < g->ncells; ++c) nsub[c] = n[c];
  }
  return 0;
}

// Scalar arguments of the Marchand kernels, with the global initial guess of psi
// (ionisation.f90:34-52) computed on the host; the array pointers are left null.
static void marchand_args(const shark_marchand_params* p, int64_t nc, int nb, MArgs* out) {
  const double epsone = 0.99999;
  const double thetai = p->stickeff_el * sqrt(p->mu_ions * kMH / kMel);
  double conv = 1e4, psi0 = -2.5;
  for (int it = 0; conv > 1e-6 && it < 1000; ++it) {
    double psi = psi0;
    const double fpsi = (1.0 - psi) / (epsone * thetai) * exp(-psi) - 1.0;
    const double dfdpsi = -(exp(-psi) / (epsone * thetai)) - (1.0 - psi) / (epsone * thetai) * exp(-psi);
    psi = psi - fpsi / dfdpsi;
    conv = fabs(fpsi);
    psi0 = psi;
  }
  MArgs& a = *out;
  memset(&a, 0, sizeof(a));
  a.nc = nc; a.nb = nb; a.nitermax = p->nitermax_ionis; a.res = p->resistivities;
  a.thetai = thetai; a.psi0 = psi0; a.x = p->zeta; a.mu_ions = p->mu_ions; a.eps_ionis = p->epsilon_ionis;
  a.B0 = p->B_0_lee; a.rhograin = p->rhograin; a.gamma = p->gamma;
  a.unit_l = p->unit_l; a.unit_d = p->unit_d; a.unit_m = p->unit_m; a.unit_v = p->unit_v; a.unit_nh = p->unit_nh;
}

int shark_charge_marchand(const shark_marchand_params* p, int64_t nc, int nb, const double* rho,
                          const double* rhod, const double* T, const double* cs, const double* sdust,
                          const double* mdust, double* psi_old, double* zd, double* ni, double* ne,
                          double* cond, double* gamma_d) {
  if (!p || !rho || !rhod || !T || !cs || !sdust || !mdust || !psi_old || !zd || !ni || !ne)
    return gfail(SHARK_GROWTH_EINVAL, "shark_charge_marchand: null argument");
  if (nc < 1 || nb < 1 || nb > MAXB) return gfail(SHARK_GROWTH_EINVAL, "shark_charge_marchand: need ncells >= 1 and 1 <= nbins <= 64");
  if (!(p->zeta > 0.0) || !(p->mu_ions > 0.0) || !(p->stickeff_el > 0.0) || p->nitermax_ionis < 1)
    return gfail(SHARK_GROWTH_EINVAL, "shark_charge_marchand: zeta, mu_ions, stickeff_el and nitermax_ionis must be positive");
  MArgs a;
  marchand_args(p, nc, nb, &a);
  const size_t n1 = (size_t)nc * 8, nn = (size_t)nc * nb * 8;
  // one device block: rho T cs psi ni ne | cond(6) | rhod sdust mdust zd gd tau alpha nk
  double* d = nullptr;
  const size_t total = 6 * n1 + 6 * n1 + 8 * nn;
  GCHK(hipMalloc(&d, total));
  char* b = (char*)d;
  auto take = [&](size_t bytes) { double* q = (double*)b; b += bytes; return q; };
  double *drho = take(n1), *dT = take(n1), *dcs = take(n1), *dpsi = take(n1), *dni = take(n1), *dne = take(n1);
  double* dcond = take(6 * n1);
  double *drhod = take(nn), *dsd = take(nn), *dmd = take(nn), *dzd = take(nn), *dgd = take(nn);
  a.tau = take(nn); a.alpha = take(nn); a.nk = take(nn);
  hipError_t e = hipSuccess;
  auto up = [&](double* dst, const double* src, size_t bytes) {
    if (e == hipSuccess) e = hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice);
  };
  up(drho, rho, n1); up(dT, T, n1); up(dcs, cs, n1); up(dpsi, psi_old, n1);
  up(drhod, rhod, nn); up(dsd, sdust, nn); up(dmd, mdust, nn);
  a.rho = drho; a.T = dT; a.cs = dcs; a.rhod = drhod; a.sdust = dsd; a.mdust = dmd;
  a.psi = dpsi; a.zd = dzd; a.ni = dni; a.ne = dne; a.cond = dcond; a.gd = dgd;
  if (e == hipSuccess) {
    const int64_t nj = nc * nb;
    hipLaunchKernelGGL(k_mcharge_pre, dim3((unsigned)((nj + 255) / 256)), dim3(256), 0, 0, a);
    hipLaunchKernelGGL(k_mcharge, dim3((unsigned)((nc + 63) / 64)), dim3(64), 0, 0, a);
    e = hipGetLastError();
  }
  if (e == hipSuccess) e = hipDeviceSynchronize();
  auto down = [&](double* dst, const double* src, size_t bytes) {
    if (dst && e == hipSuccess) e = hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost);
  };
  down(psi_old, dpsi, n1); down(zd, dzd, nn); down(ni, dni, n1); down(ne, dne, n1);
  if (p->resistivities) {
    down(cond, dcond, 6 * n1);
    down(gamma_d, dgd, nn);
  }
  (void)hipFree(d);
  GCHK(e);
  return 0;
}

// ---- Marchand charging that allocates nothing (shark_charge_marchand_noalloc)
// The per-(bin, cell) scratch of k_mcharge_pre / k_mcharge (tau, alpha, n_k: 24 B per pair) lives
// in this file-scope device array, 24 MiB, which every GPU the library runs on holds until the
// process ends.  Host code never passes the array's name to a kernel: in host code it is a
// stand-in, not a device address.  Each device's address is looked up once, with that device
// current, and kept in g_scratch_dev.
static __device__ __align__(256) double scratch[3 * SHARK_CHARGE_NOALLOC_MAX_PAIRS];
static constexpr int kMaxDevices = 64;
static double* g_scratch_dev[kMaxDevices] = {};
static std::mutex g_scratch_mu;

// The scratch array's address on the current device (looked up on the device's first call).
static int scratch_on_current_device(double** out) {
  int dev = -1;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return gfail(SHARK_GROWTH_EHIP, "shark_charge_marchand_noalloc: hipGetDevice failed");
  if (dev < 0 || dev >= kMaxDevices)
    return gfail(SHARK_GROWTH_EINVAL, "shark_charge_marchand_noalloc: device index beyond 63");
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_dev[dev]) {
    void* p = nullptr;
    e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) {
      char msg[256];
      snprintf(msg, sizeof(msg), "shark_charge_marchand_noalloc: hipGetSymbolAddress(scratch) failed: %s",
               hipGetErrorString(e));
      return gfail(SHARK_GROWTH_EHIP, msg);
    }
    g_scratch_dev[dev] = (double*)p;
  }
  *out = g_scratch_dev[dev];
  return 0;
}

int shark_charge_marchand_noalloc(const shark_marchand_params* p, int64_t nc, int nb, const double* rho,
                                  const double* rhod, const double* T, const double* cs, const double* sdust,
                                  const double* mdust, double* psi_old, double* zd, double* ni, double* ne,
                                  double* cond, double* gamma_d, void* stream) {
  if (!p || !rho || !rhod || !T || !cs || !sdust || !mdust || !psi_old || !zd || !ni || !ne)
    return gfail(SHARK_GROWTH_EINVAL, "shark_charge_marchand_noalloc: null argument");
  if (nc < 1 || nb < 1 || nb > MAXB)
    return gfail(SHARK_GROWTH_EINVAL, "shark_charge_marchand_noalloc: need ncells >= 1 and 1 <= nbins <= 64");
  if (nc * (int64_t)nb > SHARK_CHARGE_NOALLOC_MAX_PAIRS) {
    char msg[256];
    snprintf(msg, sizeof(msg),
             "shark_charge_marchand_noalloc: ncells * nbins = %lld exceeds the static scratch of %lld pairs "
             "(call it on chunks of cells, or use shark_charge_marchand)",
             (long long)(nc * (int64_t)nb), (long long)SHARK_CHARGE_NOALLOC_MAX_PAIRS);
    return gfail(SHARK_GROWTH_EINVAL, msg);
  }
  if (!(p->zeta > 0.0) || !(p->mu_ions > 0.0) || !(p->stickeff_el > 0.0) || p->nitermax_ionis < 1)
    return gfail(SHARK_GROWTH_EINVAL, "shark_charge_marchand_noalloc: zeta, mu_ions, stickeff_el and nitermax_ionis must be positive");
  if (p->resistivities && (!cond || !gamma_d))
    return gfail(SHARK_GROWTH_EINVAL, "shark_charge_marchand_noalloc: resistivities need cond and gamma_d");
  double* sp = nullptr;
  int rc = scratch_on_current_device(&sp);
  if (rc) return rc;
  MArgs a;
  marchand_args(p, nc, nb, &a);
  const int64_t nn = nc * nb;
  // k_mcharge_pre writes every pair's tau, alpha and n_k before k_mcharge reads them, so nothing
  // in the scratch has to start at zero
  a.tau = sp;
  a.alpha = sp + nn;
  a.nk = sp + 2 * nn;
  a.rho = rho; a.T = T; a.cs = cs; a.rhod = rhod; a.sdust = sdust; a.mdust = mdust;
  a.psi = psi_old; a.zd = zd; a.ni = ni; a.ne = ne;
  a.cond = p->resistivities ? cond : nullptr;
  a.gd = p->resistivities ? gamma_d : nullptr;
  hipStream_t s = (hipStream_t)stream;
  hipLaunchKernelGGL(k_mcharge_pre, dim3((unsigned)((nn + 255) / 256)), dim3(256), 0, s, a);
  hipLaunchKernelGGL(k_mcharge, dim3((unsigned)((nc + 63) / 64)), dim3(64), 0, s, a);
  GCHK(hipGetLastError());
  return 0;
}

int shark_growth_charge(shark_growth* g, double zeta, double mu_ions, int electrons, double* zmean,
                        double* xion) {
  if (!g) return gfail(SHARK_GROWTH_EINVAL, "shark_growth_charge: null handle");
  if (!g->uploaded) return gfail(SHARK_GROWTH_EINVAL, "shark_growth_charge: upload the state first");
  if (!(zeta > 0.0) || !(mu_ions > 0.0))
    return gfail(SHARK_GROWTH_EINVAL, "shark_growth_charge: zeta and mu_ions must be positive");
  if (!g->dzmean) {
    GCHK(hipMalloc(&g->dzmean, sizeof(double) * g->ncells));
    GCHK(hipMalloc(&g->dxion, sizeof(double) * g->ncells));
  }
  const shark_growth_params& p = g->p;
  GArgs a;
  memset(&a, 0, sizeof(a));
  a.nb = p.nbins;
  a.ld = g->ncells;
  a.unit_v = p.unit_v;
  a.unit_d = p.unit_d;
  a.mu_gas = p.mu_gas;
  a.gamma = p.gamma;
  a.hoverr = p.hoverr;
  a.mstar = p.mstar;
  a.kts = sqrt(kPi * p.gamma / 8.0) * (p.rhograin / p.unit_d);
  a.s2pi = sqrt(2.0 * kPi);
  a.sdust = g->dsdust;
  a.rho_gas = g->drho;
  a.cs = g->dcs;
  a.radius = g->drad;
  a.env = g->denv;
  a.ts = g->dts;
  QArgs q;
  memset(&q, 0, sizeof(q));
  q.nb = p.nbins;
  q.ld = g->ncells;
  q.electrons = electrons ? 1 : 0;
  q.maxiter = 100;
  q.zeta = zeta;
  q.mu_ions = mu_ions;
  q.mu_gas = p.mu_gas;
  q.hoverr = p.hoverr;
  q.unit_d = p.unit_d;
  q.unit_m = p.unit_m;
  q.unit_l = p.unit_l;
  q.s2pi = a.s2pi;
  q.rhod = g->drhod;
  q.rho_gas = g->drho;
  q.radius = g->drad;
  q.env = g->denv;
  q.mdust = g->dmdust;
  q.sdust = g->dsdust;
  q.zmean = g->dzmean;
  q.xion = g->dxion;
  hipStream_t s = g->stream;
  for (int64_t c0 = 0; c0 < g->ncells; c0 += g->nch_max) {
    const int64_t rem = g->ncells - c0;
    a.c0 = q.c0 = (int)c0;
    a.nch = q.nch = (int)(rem < g->nch_max ? rem : g->nch_max);
    const int nblk = (a.nch + 255) / 256;
    hipLaunchKernelGGL(k_genv, dim3(nblk), dim3(256), 0, s, a);
    GCHK(hipGetLastError());
    hipLaunchKernelGGL(k_gcharge, dim3(nblk), dim3(256), 0, s, q);
    GCHK(hipGetLastError());
  }
  GCHK(hipStreamSynchronize(s));
  if (zmean) GCHK(hipMemcpy(zmean, g->dzmean, sizeof(double) * g->ncells, hipMemcpyDeviceToHost));
  if (xion) GCHK(hipMemcpy(xion, g->dxion, sizeof(double) * g->ncells, hipMemcpyDeviceToHost));
  return 0;
}

int shark_growth_kernel_time(shark_growth* g, double* ms) {
  if (!g || !ms) return gfail(SHARK_GROWTH_EINVAL, "shark_growth_kernel_time: null argument");
  *ms = g->last_ms;
  return 0;
}

int shark_growth_destroy(shark_growth* g) {
  shark_growth_free(g);
  return 0;
}

}  // extern "C"
