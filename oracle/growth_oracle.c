/*
 * growth_oracle.c — FP64 CPU restatement of dust_growth (TEST INFRASTRUCTURE ONLY).
 *
 * Follows src/dust/dust_growth.f90:3-256 statement by statement, in the reference's
 * expression order, including its mixed-precision literals (3.2, 1./2.6, 1.6, 0.8 and
 * 4./3. are default-REAL constants promoted to double).  Two environments feed the per-cell
 * pair loop:
 *   collapse (dust_growth.f90:72-79): T = barotrop(rho), free-fall t_L, Reynolds number,
 *            ambipolar velocity from eta_a — the reference's own branch; pinned bit-exact
 *            against the flang-built reference in tests/test_growth_oracle.py;
 *   disk     the branch the reference leaves empty (:80-82): T from the local isothermal
 *            sound speed, t_L = 1/Omega, Reynolds number from the midplane density,
 *            turbulent velocity turb_coef*cs.
 * With `decomposed` set, the coagulated mass of a pair is binned by m1+m2 and its fragments
 * by 0.1*(m1+m2) (instead of (1-f)(m1+m2) and 0.1*f*(m1+m2), :183-187,:220), so every
 * pair's gain vector is independent of the cell — the form the MFMA contraction needs.
 * Identical to the reference whenever f_frag is 0 or 1.
 */
#include <math.h>
#include <limits.h>
#include <string.h>
#include "growth_oracle.h"

static const double PI = 3.141592653589793238462643383279;
static const double KB = 1.38e-16, MH = 1.67e-24, GRAV = 6.67e-8, CLIGHT = 3e10;

static double fmx(double a, double b) { return a > b ? a : b; }
static double fmn(double a, double b) { return a < b ? a : b; }

/* Fortran FLOOR to default integer; -inf/NaN (log of 0) convert to INT_MIN as on x86-64 */
static int ffloor(double x) {
  double f = floor(x);
  if (!(f >= -2147483648.0 && f <= 2147483647.0)) return INT_MIN;
  return (int)f;
}
static int imax(int a, int b) { return a > b ? a : b; }
static int imin(int a, int b) { return a < b ? a : b; }

/* x**n for integer n, by binary powering as Fortran compilers lower it */
static double powi(double a, int b) {
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

void growth_oracle_tables(const shark_growth_params* p, growth_tables* t) {
  const int nb = p->nbins;
  memset(t, 0, sizeof(*t));
  t->nb = nb;
  /* distribution_dust, dust_init.f90:22-39 */
  t->zeta = pow(p->smax / p->smin, 1.0 / nb);
  /* x**3. with an integral real exponent is lowered to multiplications by the reference's
   * compiler (checked against its mdust), so every such power is written as x*x*x */
  t->eta = t->zeta * t->zeta * t->zeta;
  const double four3 = (double)(4.0f / 3.0f);
  const double smin3 = p->smin * p->smin * p->smin;
  double m_min = four3 * PI * p->rhograin * smin3 / p->unit_m;
  if (p->kernel_type != 0) m_min = smin3;
  for (int k = 0; k < nb; ++k) {
    t->aminus[k] = p->smin * powi(t->zeta, k) / p->unit_l;
    t->aplus[k] = p->smin * powi(t->zeta, k + 1) / p->unit_l;
    t->mminus[k] = m_min * powi(t->eta, k);
    const double s = sqrt(t->aminus[k] * t->aplus[k]);
    /* mrn_distri (dust_init.f90:194) runs before the ice mantle is added (:89-90) */
    t->mdust[k] = four3 * PI * p->rhograin / p->unit_d * (s * s * s);
    t->sdust[k] = (p->kernel_type == 0) ? s + p->ice_mantle / p->unit_l : s;
  }
  t->massmin = t->mminus[0];
  /* dust_growth.f90:46-59 */
  const double a_mu = p->size_mono / 2.0;
  t->ebr_mono = p->abr * pow(p->gamma_grains, 5.0 / 3.0) * pow(a_mu, 4.0 / 3.0) /
                pow(p->estar_grains, 2.0 / 3.0) / (p->unit_v * p->unit_v * p->unit_m);
  t->m_mono = 4.0 / 3.0 * PI * p->rhograin * (p->size_mono * p->size_mono * p->size_mono) / p->unit_m;
  for (int i = 0; i < nb; ++i)
    for (int j = 0; j < nb; ++j)
      t->redis[j + nb * i] = (j <= i) ? (pow(t->aplus[j], 4.0 - p->slope_mono) - pow(t->aminus[j], 4.0 - p->slope_mono)) /
                                          (pow(t->aplus[i], 4.0 - p->slope_mono) - pow(t->aminus[0], 4.0 - p->slope_mono))
                                      : 0.0;
  for (int i = 0; i < nb; ++i) {
    double s = 0.0;
    for (int j = 0; j < nb; ++j) s = s + t->redis[j + nb * i];
    for (int j = 0; j < nb; ++j) t->redis[j + nb * i] = t->redis[j + nb * i] / s;
  }
}

/* relative velocities dvij (dust_growth.f90:87-114); dv[i*nb+j] for j <= i */
static void relative_velocities(const shark_growth_params* p, const double* sdust,
                                const double* mdust, const double* tstop, const double* vd,
                                double T, double cs_T, double t_L, double t_eta, double reynolds,
                                double vambi, int nb, double* dv) {
  (void)sdust;
  for (int i = 0; i < nb; ++i)
    for (int j = 0; j <= i; ++j) {
      double x = tstop[j] / tstop[i];
      double fst = (double)3.2f - 1.0 - x + 2.0 / (1. + x) * ((double)(1.0f / 2.6f) + (x * x * x) / ((double)1.6f + x));
      double St1 = tstop[i] / t_L, St2 = tstop[j] / t_L;
      /* Reynolds**(-0.5) is lowered to 1/sqrt by the reference's compiler (checked bitwise) */
      double rem = 1.0 / sqrt(reynolds);
      double v1 = p->turb_coef * cs_T * sqrt((St1 - St2) / (St1 + St2)) *
                  sqrt(St1 * St1 / (St1 + rem) - St2 * St2 / (St2 + rem));
      double v2 = p->turb_coef * cs_T * sqrt(fst * St1);
      double v3 = p->turb_coef * cs_T * sqrt(1.0 / (1.0 + St1) + 1.0 / (1.0 + St2));
      double vt = v2;
      if (tstop[i] < t_eta) vt = v1;
      if (tstop[i] > t_L) vt = v3;
      double um = p->unit_m;
      double inner = (8.0 * KB * T / PI) * (mdust[i] * um + mdust[j] * um) / (mdust[i] * mdust[j] * (um * um)) /
                     (p->unit_v * p->unit_v);
      double sq = sqrt(inner);
      double vb = sqrt(sq * sq);
      double vh = 0.0;
      if (vd) {
        double dx = vd[i] - vd[j], dy = vd[nb + i] - vd[nb + j], dz = vd[2 * nb + i] - vd[2 * nb + j];
        vh = sqrt(dx * dx + dy * dy + dz * dz);
      }
      double ad = vambi / p->unit_v * (0.0 - 0.0);  /* gamma_d = 0 without charging */
      double va = sqrt(ad * ad);
      double d = 0.0;
      if (p->turb) d = vt;
      if (p->brownian) d = sqrt(d * d + vb * vb);
      if (p->drift) d = sqrt(d * d + vh * vh);
      if (p->ambipolar) d = sqrt(d * d + va * va);
      dv[i * nb + j] = d;
    }
}

/* per-cell core shared by both environments */
static int64_t cell_core(const shark_growth_params* p, const growth_tables* t, const double* sdust,
                         const double* mdust, const double* dv, double rho_gas, double dt,
                         double* rhod, double* drhodt_first) {
  const int nb = t->nb;
  const int frag_test = p->fragmentation ? 1 : 0;
  const double log_eta = log(t->eta);
  double drhodt[SHARK_GROWTH_MAX_BINS];
  double time_growth = 0.0;
  int64_t niter = 0;
  while (time_growth < dt) {
    for (int k = 0; k < nb; ++k) drhodt[k] = 0.0;
    niter = niter + 1;
    for (int i = 0; i < nb; ++i) {
      const double m1 = mdust[i], s1 = sdust[i];
      for (int j = 0; j <= i; ++j) {
        const double m2 = mdust[j], s2 = sdust[j];
        const double dvij = dv[i * nb + j];
        double f_frag = 0.0, p_coag = 1.0;
        if (frag_test == 1) {
          if (p->frag_mode == 0) {
            double Ecol = 0.5 * (m1 * m2) / (m1 + m2) * (dvij * dvij);
            double Ebr = (m1 + m2) / t->m_mono * t->ebr_mono;
            f_frag = fmx(fmn((Ecol - 0.1 * Ebr) / (4.9 * Ebr), 1.0), 0.0);
          } else {
            double vf = p->vfrag / p->unit_v;
            f_frag = fmx(fmn((fabs(dvij) - 0.1 * vf) / vf, 1.0), 0.0);
          }
        }
        double dndt = p->clustered * p->clustered * PI * ((s1 + s2) * (s1 + s2)) * dvij * rhod[i] * rhod[j] / m1 / m2;
        if (p->kernel_type == 1) {
          p_coag = 1.0; f_frag = 0.0;
          dndt = rhod[i] * rhod[j] / m1 / m2;
        } else if (p->kernel_type == 2) {
          p_coag = 1.0; f_frag = 0.0;
          dndt = (m1 + m2) * rhod[i] * rhod[j] / m1 / m2;
        }
        if (i == j) dndt = dndt / 2.0;
        if (p->kernel_type == 0) {
          if (p->eps_threshold > 0.0 && (rhod[i] < p->eps_threshold * rho_gas)) p_coag = 0.0;
          if (p->eps_threshold > 0.0 && (rhod[j] < p->eps_threshold * rho_gas)) p_coag = 0.0;
          if (p->eps_threshold_frag > 0.0 && (rhod[i] < p->eps_threshold_frag * rho_gas)) f_frag = 0.0;
          if (p->eps_threshold_frag > 0.0 && (rhod[j] < p->eps_threshold_frag * rho_gas)) f_frag = 0.0;
        }
        const double mt = p->decomposed ? (m1 + m2) : (1.0 - f_frag) * (m1 + m2);
        const int fl = ffloor(log(mt / t->massmin) / log_eta + 1);
        const int ic1 = imax(imin(fl, nb), 1);
        const int ic2 = imax(imin(fl == INT_MIN ? INT_MIN : fl + 1, nb), 1);
        double epsilon_mass = 1.0;
        if (ic1 < ic2) epsilon_mass = fmn((mdust[ic2 - 1] - mt) / (mdust[ic2 - 1] - mdust[ic1 - 1]), 1.0);
        if (p_coag >= 0.0) {
          drhodt[i] = drhodt[i] - p_coag * (1.0 - f_frag) * m1 * dndt * p->sticking - p_coag * f_frag * m1 * dndt;
          drhodt[j] = drhodt[j] - p_coag * (1.0 - f_frag) * m2 * dndt * p->sticking - p_coag * f_frag * m2 * dndt;
          if (f_frag < 1.0) {
            drhodt[ic1 - 1] = drhodt[ic1 - 1] + epsilon_mass * p_coag * (1.0 - f_frag) * (m1 + m2) * dndt * p->sticking;
            drhodt[ic2 - 1] = drhodt[ic2 - 1] + (1. - epsilon_mass) * p_coag * (1.0 - f_frag) * (m1 + m2) * dndt * p->sticking;
          }
        }
        if ((frag_test == 1 && f_frag > 0.0) && p_coag > 0.) {
          const double ff = p->decomposed ? 1.0 : f_frag;
          const int g = ffloor(log(0.1 * ff * (m1 + m2) / t->massmin) / log_eta + 1);
          const int ifrag2 = imax(imin(g, ic1), 1);
          for (int k = 0; k < ifrag2; ++k)
            drhodt[k] = drhodt[k] + p_coag * f_frag * (m1 + m2) * dndt * t->redis[k + nb * (ifrag2 - 1)];
        }
      }
    }
    if (niter == 1 && drhodt_first)
      for (int k = 0; k < nb; ++k) drhodt_first[k] = drhodt[k];
    double dt_growth = 1e32;
    for (int k = 0; k < nb; ++k)
      if (fabs(drhodt[k]) != 0.0) dt_growth = fmn(dt_growth, rhod[k] / fabs(drhodt[k]) * p->cfl_growth);
    dt_growth = fmn(dt_growth, dt - time_growth);
    for (int k = 0; k < nb; ++k) rhod[k] = fmx(rhod[k] + drhodt[k] * dt_growth, rho_gas * p->dust_ratio_min);
    time_growth = time_growth + dt_growth;
    if (p->max_substeps > 0 && niter >= p->max_substeps) break;
  }
  return niter;
}

/* collapse environment (dust_growth.f90:72-79; barotrop and cs_eos of setups/Collapse_1D/eos.f90) */
int64_t growth_oracle_collapse(const shark_growth_params* p, const growth_tables* t, double T0,
                               double unit_nH, int64_t ncells, const double* q_rho,
                               const double* u_rho, const double* sdust, const double* mdust,
                               const double* tstop, const double* vd, const double* eta_a,
                               const unsigned char* active, double dt, double* rhod,
                               int64_t* nsub) {
  const int nb = p->nbins;
  int64_t total = 0;
  double sd[SHARK_GROWTH_MAX_BINS], md[SHARK_GROWTH_MAX_BINS], ts[SHARK_GROWTH_MAX_BINS];
  double rd[SHARK_GROWTH_MAX_BINS], v3[3 * SHARK_GROWTH_MAX_BINS];
  static double dv[SHARK_GROWTH_MAX_BINS * SHARK_GROWTH_MAX_BINS];
  for (int64_t c = 0; c < ncells; ++c) {
    if (active && !active[c]) { if (nsub) nsub[c] = 0; continue; }
    for (int k = 0; k < nb; ++k) {
      sd[k] = sdust[k * ncells + c]; md[k] = mdust[k * ncells + c]; ts[k] = tstop[k * ncells + c];
      rd[k] = rhod[k * ncells + c];
      if (vd) { v3[k] = vd[k * ncells + c]; v3[nb + k] = vd[(nb + k) * ncells + c]; v3[2 * nb + k] = vd[(2 * nb + k) * ncells + c]; }
    }
    const double rho = q_rho[c];
    const double nH = rho;
    const double T = T0 * sqrt(1.0 + pow(nH * unit_nH / 1e11, (double)0.8f)) * pow(1.0 + (nH * unit_nH / 1e16), -0.3) *
                     pow(1.0 + (nH * unit_nH / 1e21), 0.56667);
    const double cs_T = sqrt(p->gamma * KB * T / (p->mu_gas * MH)) / p->unit_v;
    const double lambJ = sqrt(3. * PI / (32. * GRAV * rho * p->unit_d)) * cs_T * p->unit_v;
    const double vambi = p->delta_vambi * (CLIGHT * CLIGHT) * (eta_a ? eta_a[c] : 0.0) / lambJ / (4.0 * PI);
    const double Re = 6.2e7 * sqrt(rho * p->unit_d / (p->mu_gas * MH) / 1e5) * sqrt(T / 10.0);
    const double t_L = sqrt(3. * PI / (32. * GRAV * rho * p->unit_d)) / p->unit_t;
    const double t_eta = t_L / sqrt(Re);
    relative_velocities(p, sd, md, ts, vd ? v3 : NULL, T, cs_T, t_L, t_eta, Re, vambi, nb, dv);
    int64_t n = cell_core(p, t, sd, md, dv, u_rho[c], dt, rd, NULL);
    for (int k = 0; k < nb; ++k) rhod[k * ncells + c] = rd[k];
    if (nsub) nsub[c] = n;
    total += n;
  }
  return total;
}

/* Disk environment inputs of one cell: T from cs, t_L = 1/Omega, Reynolds number of the
 * midplane density, Epstein stopping times of compute_tstop (disk_polar/setup.f90:436). */
void growth_oracle_disk_env(const shark_growth_params* p, const growth_tables* t, double sigma,
                            double cs, double r, double* T, double* t_L, double* t_eta,
                            double* Re, double* tstop) {
  const double H = p->hoverr * r;
  const double om = sqrt(p->mstar / (r * r * r));
  const double rho_mid = sigma / (sqrt(2.0 * PI) * H);
  const double csp = cs * p->unit_v;
  *T = csp * csp * (p->mu_gas * MH) / (p->gamma * KB);
  *t_L = 1.0 / om;
  *Re = 6.2e7 * sqrt(rho_mid * p->unit_d / (p->mu_gas * MH) / 1e5) * sqrt(*T / 10.0);
  *t_eta = *t_L / sqrt(*Re);
  for (int k = 0; k < t->nb; ++k)
    tstop[k] = sqrt(PI * p->gamma / 8.0) * (p->rhograin / p->unit_d) * t->sdust[k] / ((sigma / (sqrt(2.0 * PI) * H)) * cs);
}

int64_t growth_oracle_disk(const shark_growth_params* p, const growth_tables* t, int64_t ncells,
                           const double* sigma, const double* cs, const double* r, double dt,
                           double* rhod, int64_t* nsub, double* drhodt_first) {
  const int nb = p->nbins;
  int64_t total = 0;
  #pragma omp parallel for schedule(dynamic, 16) reduction(+ : total)
  for (int64_t c = 0; c < ncells; ++c) {
    double rd[SHARK_GROWTH_MAX_BINS], ts[SHARK_GROWTH_MAX_BINS], d1[SHARK_GROWTH_MAX_BINS];
    double dv[SHARK_GROWTH_MAX_BINS * SHARK_GROWTH_MAX_BINS];
    double T, tL, teta, Re;
    growth_oracle_disk_env(p, t, sigma[c], cs[c], r[c], &T, &tL, &teta, &Re, ts);
    for (int k = 0; k < nb; ++k) rd[k] = rhod[k * ncells + c];
    relative_velocities(p, t->sdust, t->mdust, ts, NULL, T, cs[c], tL, teta, Re, 0.0, nb, dv);
    int64_t n = cell_core(p, t, t->sdust, t->mdust, dv, sigma[c], dt, rd, drhodt_first ? d1 : NULL);
    for (int k = 0; k < nb; ++k) {
      rhod[k * ncells + c] = rd[k];
      if (drhodt_first) drhodt_first[k * ncells + c] = d1[k];
    }
    if (nsub) nsub[c] = n;
    total += n;
  }
  return total;
}

/* ------------------------------------------------------------------------------- charging
 * Grain charge of the analytical prescription (src/chemistry/ionisation.f90:809-922, Fujii et
 * al. 2011): Newton-Raphson (src/commons/newton_raphson.f90:7-97, tolerance 1e-14, start -1) on
 * f(Z) of functions_NR.f90:15-43 with derivative :61-91.  Constants of phys_const.f90. */
static const double E_EL = 4.8e-10, M_EL = 9.109e-28;

static double charge_f(double y, double zeta, double rho, double m_d, double rho_d, double s_d,
                       double T, double mu_ions, double mu_gas, int electrons) {
  const double m_i = mu_ions * MH, m_e = M_EL;
  const double e2 = E_EL * E_EL;
  const double k_e_d = PI * (s_d * s_d) * sqrt(8.0 * KB * T / (PI * m_e)) * exp(e2 * y / (s_d * KB * T));
  const double k_i_d = PI * (s_d * s_d) * sqrt(8.0 * KB * T / (PI * m_i)) * (1.0 - e2 * y / (s_d * KB * T));
  const double X = m_d * rho / (rho_d * mu_gas * MH);
  if (!electrons) return mu_gas * MH * zeta / rho * (X * X) * ((1.0 / k_i_d)) - y;
  return mu_gas * MH * zeta / rho * (X * X) * ((1.0 / k_i_d) - (1.0 / k_e_d)) - y;
}

static double charge_fp(double y, double zeta, double rho, double m_d, double rho_d, double s_d,
                        double T, double mu_ions, double mu_gas, int electrons) {
  const double m_i = mu_ions * MH, m_e = M_EL;
  const double e2 = E_EL * E_EL;
  const double k_e_d = PI * (s_d * s_d) * sqrt(8.0 * KB * T / (PI * m_e)) * exp(e2 * y / (s_d * KB * T));
  const double k_i_d = PI * (s_d * s_d) * sqrt(8.0 * KB * T / (PI * m_i)) * (1.0 - e2 * y / (s_d * KB * T));
  const double kp_e = PI * s_d * sqrt(8.0 * KB * T / PI / m_e) * e2 / KB / T * exp(e2 * (y) / s_d / KB / T);
  const double kp_i = -(PI * s_d * sqrt(8.0 * KB * T / PI / m_i) * e2 / KB / T);
  const double X = m_d * rho / rho_d / mu_gas / MH;
  if (!electrons) return mu_gas * MH * zeta / rho * (X * X) * (-kp_i / (k_i_d * k_i_d)) - 1.0;
  return mu_gas * MH * zeta / rho * (X * X) * (kp_e / (k_e_d * k_e_d) - kp_i / (k_i_d * k_i_d)) - 1.0;
}

/* solve_newton_raphson with y0 = -1; returns Z, *iters the reference's iteration count.
 * steptol > 0 adds a stop once a step is below steptol*|y| (the disk environment's criterion:
 * the reference's absolute 1e-14 on f cannot be met by charges of hundreds of electrons, where
 * it would run all maxiter iterations); steptol = 0 is the reference's loop. */
double growth_oracle_charge_root(double zeta, double rho, double m_d, double rho_d, double s_d,
                                 double T, double mu_ions, double mu_gas, int electrons,
                                 int maxiter, double steptol, int* iters) {
  const double tol = 1.e-14;
  double y = -1.0;
  int k;
  for (k = 1; k <= maxiter; ++k) {
    const double fy = charge_f(y, zeta, rho, m_d, rho_d, s_d, T, mu_ions, mu_gas, electrons);
    const double fyp = charge_fp(y, zeta, rho, m_d, rho_d, s_d, T, mu_ions, mu_gas, electrons);
    if (fabs(fy) < tol) break;
    const double deltay = fy / fyp;
    y = y - deltay;
    if (steptol > 0.0 && fabs(deltay) <= steptol * fabs(y)) { ++k; break; }
  }
  if (iters) *iters = k - 1;
  return y;
}

/* analytical_charge over the collapse cells: T = barotrop(q_rho/(mu_gas mH)) with unit_nH = 1
 * (setups/Collapse_1D/eos.f90:8), raw code-unit densities, masses and sizes as the reference
 * passes them; ni/ne of the last bin survive (the reference overwrites them per bin). */
void growth_oracle_charge_collapse(int64_t nc, int nb, const double* q_rho, const double* rhod,
                                   const double* mdust, const double* sdust, double T0,
                                   double zeta, double mu_ions, double mu_gas, int electrons,
                                   int maxiter, double* zd, double* ni, double* ne) {
  const double m_i = mu_ions * MH, m_e = M_EL, e2 = E_EL * E_EL;
  for (int k = 0; k < nb; ++k)
    for (int64_t c = 0; c < nc; ++c) {
      const double nH = q_rho[c] / (mu_gas * MH);
      const double T = T0 * sqrt(1.0 + pow(nH * 1.0 / 1e11, (double)0.8f)) * pow(1.0 + (nH * 1.0 / 1e16), -0.3) *
                       pow(1.0 + (nH * 1.0 / 1e21), 0.56667);
      const double rho = q_rho[c], m_d = mdust[k * nc + c], rd = rhod[k * nc + c], s_d = sdust[k * nc + c];
      const double y = growth_oracle_charge_root(zeta, rho, m_d, rd, s_d, T, mu_ions, mu_gas, electrons, maxiter, 0.0, NULL);
      zd[k * nc + c] = y;
      const double k_e_d = PI * (s_d * s_d) * sqrt(8.0 * KB * T / (PI * m_e)) * exp(e2 * y / (s_d * KB * T));
      const double k_i_d = PI * (s_d * s_d) * sqrt(8.0 * KB * T / (PI * m_i)) * (1.0 - e2 * y / (s_d * KB * T));
      ni[c] = zeta * rho * m_d / (k_i_d * mu_gas * MH * rd);
      ne[c] = zeta * rho * m_d / (k_e_d * mu_gas * MH * rd);
      if (!electrons) ne[c] = 0.0;
    }
}

/* Disk environment in cgs: midplane gas density sigma/(sqrt(2 pi) H), dust midplane densities
 * rhod/(sqrt(2 pi) H), T from cs.  Per cell: number-weighted mean charge sum n_k Z_k / sum n_k
 * and ion fraction n_i/n_H with n_i = zeta n_H / sum_k k_i,k n_k (ion balance on all grains). */
void growth_oracle_charge_disk(const shark_growth_params* p, const growth_tables* t, int64_t nc,
                               const double* sigma, const double* cs, const double* r,
                               const double* rhod, double zeta, double mu_ions, int electrons,
                               int maxiter, double* zd, double* zmean, double* xion) {
  const int nb = t->nb;
  const double m_i = mu_ions * MH, e2 = E_EL * E_EL;
  #pragma omp parallel for schedule(dynamic, 16)
  for (int64_t c = 0; c < nc; ++c) {
    double T, tL, teta, Re, ts[SHARK_GROWTH_MAX_BINS];
    growth_oracle_disk_env(p, t, sigma[c], cs[c], r[c], &T, &tL, &teta, &Re, ts);
    const double H = p->hoverr * r[c];
    const double rho = sigma[c] / (sqrt(2.0 * PI) * H) * p->unit_d;
    double nz = 0.0, nn = 0.0, kin = 0.0;
    for (int k = 0; k < nb; ++k) {
      const double rd = rhod[k * nc + c] / (sqrt(2.0 * PI) * H) * p->unit_d;
      const double m_d = t->mdust[k] * p->unit_m, s_d = t->sdust[k] * p->unit_l;
      const double y = growth_oracle_charge_root(zeta, rho, m_d, rd, s_d, T, mu_ions, p->mu_gas, electrons, maxiter, 1e-15, NULL);
      if (zd) zd[k * nc + c] = y;
      const double n = rd / m_d;
      nz += n * y;
      nn += n;
      const double k_i_d = PI * (s_d * s_d) * sqrt(8.0 * KB * T / (PI * m_i)) * (1.0 - e2 * y / (s_d * KB * T));
      kin += k_i_d * n;
    }
    zmean[c] = nz / nn;
    xion[c] = zeta / kin;
  }
}
