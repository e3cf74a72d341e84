/* marchand_oracle.c — CPU restatement of the reference's Marchand et al. 2021 grain charging with
 * resistivities (`charge`, src/chemistry/ionisation.f90:4-214).  Test infrastructure: only the
 * tests use it, as the checker of the GPU kernel.  Every expression keeps the reference's operand
 * order and flang's lowering of powers (x**2., x**3. and real**2 as products, x**4.0 and
 * x**<non-integral> through pow, default-REAL literals 1.3, 0.6, 1.16, 1.14, 1.28, 0.9999999
 * rounded to single precision), so the results equal the flang build bit for bit
 * (tests/golden/ref_marchand.npz). */
#include <math.h>
#include <stdint.h>

#include "marchand_oracle.h"

static const double PI = 3.141592653589793238462643383279;
static const double KB = 1.38e-16, MH = 1.67e-24, E_EL = 4.8e-10, M_EL = 9.109e-28, CLIGHT = 3e10;

/* the global initial guess psi0 (ionisation.f90:34-52) */
double marchand_oracle_psi0(const shark_marchand_params* p) {
  const double epsone = 0.99999;
  const double thetai = p->stickeff_el * sqrt(p->mu_ions * MH / M_EL);
  double conv = 1e4, psi0 = -2.5;
  int niter = 0;
  while (conv > 1e-6 && niter < 1000) {
    double psi = psi0;
    const double fpsi = (1.0 - psi) / (epsone * thetai) * exp(-psi) - 1.0;
    const double dfdpsi = -(exp(-psi) / (epsone * thetai)) - (1.0 - psi) / (epsone * thetai) * exp(-psi);
    psi = psi - fpsi / dfdpsi;
    conv = fabs(fpsi);
    niter = niter + 1;
    psi0 = psi;
  }
  return psi0;
}

void marchand_oracle_charge(const shark_marchand_params* p, int64_t nc, int nb, const double* u_rho,
                            const double* rhod, const double* T_in, const double* cs_in, const double* sdust,
                            const double* mdust, double* psi_old, double* zd, double* ni, double* ne,
                            double* cond, double* gamma_d) {
  const double epsone = 0.99999;
  const double as_he_ions = (double)1.14f, as_he_el = (double)1.16f;
  const double thetai = p->stickeff_el * sqrt(p->mu_ions * MH / M_EL);
  const double psi0 = marchand_oracle_psi0(p);
  const double x = p->zeta;
  double tau[SHARK_GROWTH_MAX_BINS], alpha[SHARK_GROWTH_MAX_BINS], nk[SHARK_GROWTH_MAX_BINS],
      lg[SHARK_GROWTH_MAX_BINS];
  for (int64_t i = 0; i < nc; ++i) {
    const double T = T_in[i];
    for (int k = 0; k < nb; ++k) lg[k] = sdust[k * nc + i] * p->unit_l;
    const double sigmav_ie = (2e-7) / sqrt(T / 300.0);
    for (int k = 0; k < nb; ++k) tau[k] = lg[k] * KB * T / (E_EL * E_EL);
    const double vi = sqrt(8.0 * KB * T / (PI * p->mu_ions * MH));
    for (int k = 0; k < nb; ++k) alpha[k] = sqrt(8.0 / (PI * tau[k]));
    for (int k = 0; k < nb; ++k) nk[k] = rhod[k * nc + i] * p->unit_d / (mdust[k * nc + i] * p->unit_m);
    const double nH = u_rho[i] * p->unit_nh;
    double conv = 1e4;
    int niter = 0;
    double psi = psi0;
    if (psi_old[i] != 0.0) psi = psi_old[i];
    int lowT = 0;
    while (conv > p->epsilon_ionis && niter < p->nitermax_ionis) {
      const double eps = (1.0 - psi) / thetai * exp(-psi);
      const double et = eps * thetai;
      double n_i = 0.0;
      for (int k = 0; k < nb; ++k)
        n_i = n_i + -(psi * tau[k] + (1.0 - et * et) / (1.0 + et * alpha[k] + et * et)) * nk[k] / (1.0 - eps);
      double s2 = 0.0;
      for (int k = 0; k < nb; ++k)
        s2 = s2 + nk[k] * PI * (lg[k] * lg[k]) *
                      ((1.0 - psi) + (2.0 / tau[k]) * (et * et + et) / (1.0 + et * alpha[k] + et * et));
      const double fpsi = sigmav_ie * eps * (n_i * n_i) / (x * nH) + n_i * vi / (x * nH) * s2 - 1.0;
      const double B = -(eps * (2.0 - psi) / (1.0 - psi));
      double s3 = 0.0, s4 = 0.0;
      for (int k = 0; k < nb; ++k) {
        const double d0 = 1.0 + et * alpha[k] + et * et;
        s3 = s3 + nk[k] *
                      (-pow(et, 4.0) + (4.0 * (thetai * thetai) - (thetai * thetai) * thetai * alpha[k]) * (eps * eps) +
                       (2.0 * thetai * alpha[k] - 4.0 * (thetai * thetai)) * eps + 1.0 - thetai * alpha[k]) /
                      (d0 * d0);
      }
      for (int k = 0; k < nb; ++k) s4 = s4 + nk[k] * tau[k];
      const double dni = -(1.0 / ((1.0 - eps) * (1.0 - eps)) * B * s3) - (psi / (1.0 - eps) * B + 1.0) * s4 / (1.0 - eps);
      double s5 = 0.0, s6 = 0.0, s7 = 0.0;
      for (int k = 0; k < nb; ++k) {
        const double d0 = 1.0 + et * alpha[k] + et * et;
        s5 = s5 + nk[k] * PI * (lg[k] * lg[k]) / (tau[k] * d0);
      }
      for (int k = 0; k < nb; ++k) {
        const double d0 = 1.0 + et * alpha[k] + et * et;
        s6 = s6 + nk[k] * PI * (lg[k] * lg[k]) * (2.0 * eps * (thetai * thetai) + thetai * alpha[k]) / (tau[k] * (d0 * d0));
      }
      for (int k = 0; k < nb; ++k) s7 = s7 + nk[k] * PI * (lg[k] * lg[k]);
      const double dfdpsi =
          sigmav_ie * n_i / (x * nH) * (n_i * B + 2.0 * eps * dni) +
          2.0 * n_i * vi / (x * nH) * (et * et + et) *
              ((dni / n_i + B * (2.0 * et + 1.0) / (eps * (1.0 + et))) * s5 + B * s6) +
          n_i * vi / (x * nH) * (dni / n_i * (1.0 - psi) - 1.0) * s7;
      psi = psi - fpsi / dfdpsi;
      if (psi < psi0) {
        psi = psi0 * (double)0.9999999f;
        lowT = 1;
      } else if (psi > 0.0) {
        psi = -1e-5;
      }
      conv = fabs(fpsi);
      niter = niter + 1;
    }
    double eps = (1.0 - psi) / thetai * exp(-psi);
    if (lowT > 0) eps = epsone;
    double n_i = 0.0;
    for (int k = 0; k < nb; ++k) {
      const double z =
          psi * tau[k] + (1.0 - (eps * eps) * (thetai * thetai)) / (1.0 + eps * thetai * alpha[k] + (eps * eps) * (thetai * thetai));
      zd[k * nc + i] = z;
      n_i = n_i - z * nk[k] / (1.0 - eps);
      if (lowT > 0) n_i = sqrt(eps * x * nH / sigmav_ie);
    }
    ni[i] = n_i;
    ne[i] = eps * n_i;
    psi_old[i] = lowT > 0 ? psi0 : psi;
    if (!p->resistivities) continue;
    const double Bg = p->B_0_lee * sqrt(u_rho[i] * p->unit_nh / 1e4);
    const double mu_i = 2.0 * MH * p->mu_ions * MH / (2.0 * MH + p->mu_ions * MH);
    const double mu_e = 2.0 * MH * M_EL / (M_EL + 2.0 * MH);
    const double vrms_i = sqrt(8.0 * KB * T / (PI * mu_i)) * 1e-5;
    const double vrms_el = sqrt(8.0 * KB * T / (PI * mu_e)) * 1e-5;
    const double sigmav_el = 3.16e-11 * pow(vrms_el, (double)1.3f);
    const double sigmav_ions = 2.4e-9 * pow(vrms_i, (double)0.6f);
    const double t_sel = 1.0 / as_he_el * ((M_EL + 2.0 * MH) / (2.0 * MH)) / sigmav_el / nH;
    const double t_sions = 1.0 / as_he_ions * ((p->mu_ions * MH + 2.0 * MH) / (2.0 * MH)) / sigmav_ions / nH;
    const double s_el = ne[i] * (E_EL * E_EL) * t_sel / M_EL;
    const double s_ions = ni[i] * (E_EL * E_EL) * t_sions / (p->mu_ions * MH);
    const double o_el = -(E_EL * Bg / CLIGHT / M_EL);
    const double o_ions = E_EL * Bg / CLIGHT / (p->mu_ions * MH);
    const double cs_phys = cs_in[i];
    double sum_sd = 0.0, sum_p = 0.0, sum_h = 0.0;
    for (int k = 0; k < nb; ++k) {
      const double tsd = sqrt(PI * p->gamma / 8.0) * (p->rhograin) * (sdust[k * nc + i] * p->unit_l) /
                         (u_rho[i] * p->unit_d * cs_phys * p->unit_v);
      const double z = zd[k * nc + i];
      const double sdk = nk[k] * ((z * E_EL) * (z * E_EL)) * tsd / (mdust[k * nc + i] * p->unit_m);
      const double odk = z * E_EL * Bg / CLIGHT / (mdust[k * nc + i] * p->unit_m);
      sum_sd = sum_sd + sdk;
      sum_p = sum_p + sdk / (1.0 + (odk * tsd) * (odk * tsd));
      sum_h = sum_h + sdk * (odk * tsd) / (1.0 + (odk * tsd) * (odk * tsd));
      if (gamma_d) gamma_d[k * nc + i] = tsd * odk;
    }
    const double so = sum_sd + s_ions + s_el;
    const double sp = s_ions / (1.0 + (o_ions * t_sions) * (o_ions * t_sions)) +
                      s_el / (1.0 + (o_el * t_sel) * (o_el * t_sel)) + sum_p;
    const double sh = -(s_ions * (o_ions * t_sions) / (1.0 + (o_ions * t_sions) * (o_ions * t_sions))) -
                      s_el * (o_el * t_sel) / (1.0 + (o_el * t_sel) * (o_el * t_sel)) - sum_h;
    if (cond) {
      cond[0 * nc + i] = 1.0 / so;
      cond[1 * nc + i] = sh / (sp * 2. + sh * sh);
      cond[2 * nc + i] = sp / (sp * sp + sh * sh) - 1.0 / so;
      cond[3 * nc + i] = so;
      cond[4 * nc + i] = sp;
      cond[5 * nc + i] = sh;
    }
  }
}
