// physics.h — device-side building blocks of the SHARK step (gfx950, FP64).
//
// Every function keeps the reference's expression order (Fortran left-to-right evaluation,
// x**2 as x*x) and the library is compiled with -ffp-contract=off, so results are bitwise
// those of the reference's x86-64 build.  Citations are to the reference sources.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace shark {

constexpr double kHalf = 0.5;
constexpr double kPi = 3.141592653589793238462643383279;  // src/commons/phys_const.f90:14

__device__ __forceinline__ double fmaxf_(double a, double b) { return a > b ? a : b; }
__device__ __forceinline__ double fminf_(double a, double b) { return a < b ? a : b; }

// a / b, correctly rounded, for a divisor whose correctly rounded reciprocal y = 1/b is at hand
// (one IEEE division shared by several quotients).  q0 = a*y is within one ulp of a/b, the
// residual a - b*q0 is exact in an FMA, and q0 + (a - b*q0)*y rounds to the correctly rounded
// quotient (Markstein, IBM J. Res. Dev. 34 (1990) 111-119; Muller et al., Handbook of
// Floating-Point Arithmetic, division with FMA) for quotients in the normal range.  The result
// carries the sign of q0 = a*y, which is the quotient's sign: a zero dividend then gives the zero of
// a/b's sign (the FMA chain alone can turn -0 into +0), at the cost of one bit insert.
__device__ __forceinline__ double exdiv(double a, double b, double y) {
  const double q0 = a * y;
  const double r = fma(-b, q0, a);
  const double q = fma(r, y, q0);
  return copysign(q, q0);
}

// slope_limit, src/solver/slope_limiters.f90:5-37
__device__ __forceinline__ double slope_limit(int type, double sl, double sr) {
  double s;
  if (type == 1) {
    s = sl;
    if (fabs(sr) < fabs(sl)) s = sr;
    if (sr * sl < 0.0) s = 0.0;
  } else if (type == 2) {
    s = ((2.0 * sl) * sr) / (sl + sr);
    if (sr * sl <= 0.0) s = 0.0;
  } else {
    double dcen = kHalf * (sl + sr);
    double dsgn = copysign(1.0, dcen);
    double slop = fminf_(1.5 * fabs(sl), 1.5 * fabs(sr));
    double dlim = slop;
    if ((sl * sr) <= 0.0) dlim = 0.0;
    s = dsgn * fminf_(dlim, fabs(dcen));
  }
  return s;
}

// Face state of one fluid in the direction of the flux: density, normal velocity,
// transverse velocity, z velocity and (gas) pressure.
struct Face {
  double r, un, ut, w, p;
};
struct Flux5 {
  double r, un, ut, w, e;
};

// Gas fluxes: solver_llf :4-155, solver_hll :157-306, solver_hllc :308-500 of
// src/solver/solvers.f90 (MHD=0).
__device__ __forceinline__ Flux5 gas_flux(int solver, double gamma, const Face& L, const Face& R,
                                          double csl, double csr) {
  const double rho_r = R.r, rho_l = L.r, u_r = R.un, u_l = L.un, v_r = R.ut, v_l = L.ut;
  const double w_r = R.w, w_l = L.w, P_r = R.p, P_l = L.p;
  Flux5 f;
  if (solver == 2) {
    double E_r = P_r / (gamma - 1.0) + kHalf * rho_r * (u_r * u_r) + kHalf * rho_r * (v_r * v_r) +
                 kHalf * rho_r * (w_r * w_r);
    double E_l = P_l / (gamma - 1.0) + kHalf * rho_l * (u_l * u_l) + kHalf * rho_l * (v_l * v_l) +
                 kHalf * rho_l * (w_l * w_l);
    double S_l = fminf_(fminf_(u_l, u_r) - fmaxf_(csl, csr), 0.0);
    double S_r = fmaxf_(fmaxf_(u_l, u_r) + fmaxf_(csl, csr), 0.0);
    double hl = rho_l * (u_l - S_l), hr = rho_r * (S_r - u_r);
    double ustar = (hr * u_r + hl * u_l + (P_l - P_r)) / (hr + hl);
    double Pstar = (hr * P_l + hl * P_r + hl * hr * (u_l - u_r)) / (hr + hl);
    double r_o, u_o, P_o, e_o;
    if (S_l > 0.0) {
      r_o = rho_l; u_o = u_l; P_o = P_l; e_o = E_l;
    } else if (ustar > 0.0) {
      r_o = rho_l * (S_l - u_l) / (S_l - ustar); u_o = ustar; P_o = Pstar;
      e_o = ((S_l - u_l) * E_l - P_l * u_l + Pstar * ustar) / (S_l - ustar);
    } else if (S_r > 0.0) {
      r_o = rho_r * (S_r - u_r) / (S_r - ustar); u_o = ustar; P_o = Pstar;
      e_o = ((S_r - u_r) * E_r - P_r * u_r + Pstar * ustar) / (S_r - ustar);
    } else {
      r_o = rho_r; u_o = u_r; P_o = P_r; e_o = E_r;
    }
    f.r = r_o * u_o;
    f.un = r_o * u_o * u_o + P_o;
    f.e = (e_o + P_o) * u_o;
    if (ustar > 0.0) { f.ut = r_o * u_o * v_l; f.w = r_o * u_o * w_l; }
    else { f.ut = r_o * u_o * v_r; f.w = r_o * u_o * w_r; }
    return f;
  }
  const double mu_r = rho_r * u_r, mu_l = rho_l * u_l;
  const double mv_r = rho_r * v_r, mv_l = rho_l * v_l, mw_r = rho_r * w_r, mw_l = rho_l * w_l;
  double E_r = P_r / (gamma - 1.0) + kHalf * rho_r * (u_r * u_r);
  double E_l = P_l / (gamma - 1.0) + kHalf * rho_l * (u_l * u_l);
  E_r = E_r + kHalf * rho_r * (v_r * v_r);
  E_l = E_l + kHalf * rho_l * (v_l * v_l);
  E_r = E_r + kHalf * rho_r * (w_r * w_r);
  E_l = E_l + kHalf * rho_l * (w_l * w_l);
  const double fr_r = rho_r * u_r, fr_l = rho_l * u_l;
  const double fu_r = rho_r * (u_r * u_r) + P_r, fu_l = rho_l * (u_l * u_l) + P_l;
  const double fv_r = rho_r * u_r * v_r, fv_l = rho_l * u_l * v_l;
  const double fw_r = rho_r * u_r * w_r, fw_l = rho_l * u_l * w_l;
  const double fP_r = (E_r + P_r) * u_r, fP_l = (E_l + P_l) * u_l;
  if (solver == 0) {
    double lam = fmaxf_(fabs(u_l) + csl, fabs(u_r) + csr);
    f.r = kHalf * (fr_l + fr_r) - kHalf * lam * (rho_r - rho_l);
    f.un = kHalf * (fu_l + fu_r) - kHalf * lam * (mu_r - mu_l);
    f.ut = kHalf * (fv_l + fv_r) - kHalf * lam * (mv_r - mv_l);
    f.w = kHalf * (fw_l + fw_r) - kHalf * lam * (mw_r - mw_l);
    f.e = kHalf * (fP_l + fP_r) - kHalf * lam * (E_r - E_l);
    return f;
  }
  const double S_l = fminf_(fminf_(u_l, u_r) - fmaxf_(csl, csr), 0.0);
  const double S_r = fmaxf_(fmaxf_(u_l, u_r) + fmaxf_(csl, csr), 0.0);
  const double den = S_r - S_l, srsl = S_r * S_l;
  f.r = (S_r * fr_l - S_l * fr_r + srsl * (rho_r - rho_l)) / den;
  f.un = (S_r * fu_l - S_l * fu_r + srsl * (mu_r - mu_l)) / den;
  f.ut = (S_r * fv_l - S_l * fv_r + srsl * (mv_r - mv_l)) / den;
  f.w = (S_r * fw_l - S_l * fw_r + srsl * (mw_r - mw_l)) / den;
  f.e = (S_r * fP_l - S_l * fP_r + srsl * (E_r - E_l)) / den;
  return f;
}

// Dust fluxes for one species: solver_dust_Huang_Bai :506-698, solver_dust_llf :840-1001,
// solver_dust_hll :1010-1194, solver_hllc_dust :704-836 of src/solver/solvers.f90.
__device__ __forceinline__ Flux5 dust_flux(int solver, const Face& L, const Face& R, double csl,
                                           double csr) {
  const double rho_r = R.r, rho_l = L.r, u_r = R.un, u_l = L.un, v_r = R.ut, v_l = L.ut;
  const double w_r = R.w, w_l = L.w;
  const double fr_r = rho_r * u_r, fr_l = rho_l * u_l;
  const double fu_r = rho_r * (u_r * u_r), fu_l = rho_l * (u_l * u_l);
  const double fv_r = rho_r * u_r * v_r, fv_l = rho_l * u_l * v_l;
  const double fw_r = rho_r * u_r * w_r, fw_l = rho_l * u_l * w_l;
  Flux5 f;
  f.r = f.un = f.ut = f.w = 0.0;
  f.e = 0.0;
  if (solver == 0) {
    if (u_r > 0.0 && u_l > 0.0) {
      f.r = fr_l; f.un = fu_l; f.ut = fv_l; f.w = fw_l;
    } else if (u_l < 0.0 && u_r < 0.0) {
      f.r = fr_r; f.un = fu_r; f.ut = fv_r; f.w = fw_r;
    } else if (u_l < 0.0 && u_r > 0.0) {
      f.r = 0.0; f.un = 0.0; f.ut = 0.0; f.w = 0.0;
    } else if (u_l > 0.0 && u_r < 0.0) {
      f.r = fr_l + fr_r; f.un = fu_l + fu_r; f.ut = fv_l + fv_r; f.w = fw_l + fw_r;
    }
    return f;
  }
  const double mu_r = rho_r * u_r, mu_l = rho_l * u_l, mv_r = rho_r * v_r, mv_l = rho_l * v_l;
  const double mw_r = rho_r * w_r, mw_l = rho_l * w_l;
  if (solver == 1 || (solver == 2 && u_l == u_r)) {
    double lam = fmaxf_(fabs(u_l), fabs(u_r));
    f.r = kHalf * (fr_l + fr_r) - kHalf * lam * (rho_r - rho_l);
    f.un = kHalf * (fu_l + fu_r) - kHalf * lam * (mu_r - mu_l);
    f.ut = kHalf * (fv_l + fv_r) - kHalf * lam * (mv_r - mv_l);
    f.w = kHalf * (fw_l + fw_r) - kHalf * lam * (mw_r - mw_l);
    return f;
  }
  if (solver == 2) {
    double S_r = fmaxf_(fmaxf_(u_l, u_r), 0.0), S_l = fminf_(fminf_(u_l, u_r), 0.0);
    f.r = (S_r * fr_l - S_l * fr_r + S_r * S_l * (rho_r - rho_l)) / (S_r - S_l);
    f.un = (S_r * fu_l - S_l * fu_r + S_r * S_l * (mu_r - mu_l)) / (S_r - S_l);
    f.ut = (S_r * fv_l - S_l * fv_r + S_r * S_l * (mv_r - mv_l)) / (S_r - S_l);
    f.w = (S_r * fw_l - S_l * fw_r + S_r * S_l * (mw_r - mw_l)) / (S_r - S_l);
    return f;
  }
  // HLLC for pressureless dust (P=0)
  const double P_r = 0.0, P_l = 0.0;
  double S_l = fminf_(fminf_(u_l, u_r) - fmaxf_(csl, csr), 0.0);
  double S_r = fmaxf_(fmaxf_(u_l, u_r) + fmaxf_(csl, csr), 0.0);
  double hl = rho_l * (u_l - S_l), hr = rho_r * (S_r - u_r);
  double ustar = (hr * u_r + hl * u_l + (P_l - P_r)) / (hr + hl);
  double Pstar = (hr * P_l + hl * P_r + hl * hr * (u_l - u_r)) / (hr + hl);
  double r_o, u_o, P_o;
  if (S_l > 0.0) { r_o = rho_l; u_o = u_l; P_o = P_l; }
  else if (ustar > 0.0) { r_o = rho_l * (S_l - u_l) / (S_l - ustar); u_o = ustar; P_o = Pstar; }
  else if (S_r > 0.0) { r_o = rho_r * (S_r - u_r) / (S_r - ustar); u_o = ustar; P_o = Pstar; }
  else { r_o = rho_r; u_o = u_r; P_o = P_r; }
  f.r = r_o * u_o;
  f.un = r_o * u_o * u_o + P_o;
  if (ustar > 0.0) { f.ut = r_o * u_o * v_l; f.w = r_o * u_o * w_l; }
  else { f.ut = r_o * u_o * v_r; f.w = r_o * u_o * w_r; }
  return f;
}

// Ordered-integer encoding of a non-negative double, so a min over doubles is an integer min
// (exact and order independent; courant.f90:99 is a plain min over cells).
__device__ __forceinline__ unsigned long long dbits(double x) {
  return (unsigned long long)__double_as_longlong(x);
}
inline unsigned long long dbits_host(double x) {
  unsigned long long u;
  __builtin_memcpy(&u, &x, sizeof u);
  return u;
}

}  // namespace shark
