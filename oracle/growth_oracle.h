/* growth_oracle.h — FP64 CPU restatement of dust_growth (test infrastructure; see
 * growth_oracle.c).  Shared by the oracle library and, for the parameter struct only, by
 * the product header include/shark_growth.h which declares the same layout. */
#ifndef SHARK_GROWTH_ORACLE_H
#define SHARK_GROWTH_ORACLE_H
#include <stdint.h>
#include "../include/shark_growth.h"

/* Per-bin tables derived from the parameters (dust_init.f90:22-34, dust_growth.f90:41-59). */
typedef struct {
  int nb;
  double zeta, eta, massmin, m_mono, ebr_mono;
  double aplus[SHARK_GROWTH_MAX_BINS], aminus[SHARK_GROWTH_MAX_BINS], mminus[SHARK_GROWTH_MAX_BINS];
  double sdust[SHARK_GROWTH_MAX_BINS], mdust[SHARK_GROWTH_MAX_BINS];
  double redis[SHARK_GROWTH_MAX_BINS * SHARK_GROWTH_MAX_BINS];  /* redis[j + nb*i] */
} growth_tables;

void growth_oracle_tables(const shark_growth_params* p, growth_tables* t);
/* Collapse environment over ncells cells (arrays (ncells, nbins) column-major; vd is
 * (ncells, 3*nbins) vdx|vdy|vdz or NULL).  Returns the total number of sub-steps. */
int64_t growth_oracle_collapse(const shark_growth_params* p, const growth_tables* t, double T0,
                               double unit_nH, int64_t ncells, const double* q_rho,
                               const double* u_rho, const double* sdust, const double* mdust,
                               const double* tstop, const double* vd, const double* eta_a,
                               const unsigned char* active, double dt, double* rhod,
                               int64_t* nsub);
void growth_oracle_disk_env(const shark_growth_params* p, const growth_tables* t, double sigma,
                            double cs, double r, double* T, double* t_L, double* t_eta,
                            double* Re, double* tstop);
/* Disk environment (sigma, cs, r per cell; bins from the tables). drhodt_first (ncells, nbins)
 * receives the rates of the first sub-step when not NULL. */
int64_t growth_oracle_disk(const shark_growth_params* p, const growth_tables* t, int64_t ncells,
                           const double* sigma, const double* cs, const double* r, double dt,
                           double* rhod, int64_t* nsub, double* drhodt_first);
/* charging (src/chemistry/ionisation.f90:809-922, functions_NR.f90, newton_raphson.f90) */
double growth_oracle_charge_root(double zeta, double rho, double m_d, double rho_d, double s_d,
                                 double T, double mu_ions, double mu_gas, int electrons,
                                 int maxiter, double steptol, int* iters);
void growth_oracle_charge_collapse(int64_t nc, int nb, const double* q_rho, const double* rhod,
                                   const double* mdust, const double* sdust, double T0,
                                   double zeta, double mu_ions, double mu_gas, int electrons,
                                   int maxiter, double* zd, double* ni, double* ne);
/* disk environment: stops at maxiter or once a Newton step is below 1e-15 |Z| */
void growth_oracle_charge_disk(const shark_growth_params* p, const growth_tables* t, int64_t nc,
                               const double* sigma, const double* cs, const double* r,
                               const double* rhod, double zeta, double mu_ions, int electrons,
                               int maxiter, double* zd, double* zmean, double* xion);
#endif
