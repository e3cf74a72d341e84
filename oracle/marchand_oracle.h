/* marchand_oracle.h — CPU restatement of the Marchand et al. 2021 grain charging
 * (src/chemistry/ionisation.f90:4-214).  Test infrastructure only. */
#ifndef SHARK_MARCHAND_ORACLE_H
#define SHARK_MARCHAND_ORACLE_H
#include <stdint.h>
#include "../include/shark_growth.h"

/* global initial guess of the surface potential (ionisation.f90:34-52) */
double marchand_oracle_psi0(const shark_marchand_params* p);
/* arrays (nbins, ncells) plane-major, as the reference's (ncells, nbins) column-major;
 * psi_old in/out; cond (6, ncells) = eta_o, eta_h, eta_a, sigma_o, sigma_p, sigma_h (or NULL);
 * gamma_d (nbins, ncells) or NULL */
void marchand_oracle_charge(const shark_marchand_params* p, int64_t nc, int nb, const double* u_rho,
                            const double* rhod, const double* T, const double* cs, const double* sdust,
                            const double* mdust, double* psi_old, double* zd, double* ni, double* ne,
                            double* cond, double* gamma_d);
#endif
