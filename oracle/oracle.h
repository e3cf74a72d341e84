/* oracle.h — CPU restatement of SHARK's step (test infrastructure; see oracle.c). */
#ifndef SHARK_ORACLE_H
#define SHARK_ORACLE_H
#include <stdint.h>
#include "../include/shark_gpu.h"
typedef struct oracle oracle;
oracle* oracle_create(const shark_cfg* cfg);
void oracle_destroy(oracle* o);
/* dust growth after the drag in every step (disk environment); NULL switches it off */
struct shark_growth_params;
int oracle_set_growth(oracle* o, const struct shark_growth_params* gp);
int oracle_ncells(const oracle* o);
int oracle_nvar(const oracle* o);
void oracle_set_geometry(oracle* o, const double* dx, const double* surf, const double* vol,
                         const double* radii_c, const double* position);
void oracle_set_dvol(oracle* o, const double* dvol);
double oracle_barotrop(const shark_cfg* p, double nH);
double oracle_cs_eos(const shark_cfg* p, double T);
void oracle_upload(oracle* o, const double* u, const double* cs, const double* sdust,
                   const double* condinit, const double* fargo_v);
double oracle_step(oracle* o);
int64_t oracle_run(oracle* o, int64_t nsteps, double tend);
void oracle_download(const oracle* o, double* u, double* q);
void oracle_primitives(oracle* o, double* q);
void oracle_get_aux(const oracle* o, double* tstop, double* force, double* force_dust, double* eps_dust);
double oracle_time(const oracle* o);
double oracle_dt(const oracle* o);
double oracle_slope_limit(int type, double sl, double sr);
void oracle_fargo_remap(int nx, int ny, int nvar, int slope, double dt, const double* dx2,
                        const double* radii_c, const double* fargo_v, const double* uin, double* uout);
void oracle_gas_flux(int solver, double gamma, const double* ql, const double* qr, double* flx,
                     double csl, double csr, int vn, int vt);
void oracle_dust_flux(int solver, double gamma, const double ql[4], const double qr[4],
                      double f[4], double csl, double csr);
#endif
