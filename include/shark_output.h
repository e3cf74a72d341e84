/*
 * shark_output.h — output snapshots and restart dumps of libshark_gpu.so.
 *
 * shark_output_write / shark_gpu_write_output replace the reference's `output(iout)`
 * (src/commons/write_output.f90:2-300, called by check_output, e.g.
 * src/setups/disk_polar/setup.f90:188, and at the end of the run, src/shark.f90:114).  They write
 * the same directory `output_NNNNN/` (NNNNN = iout, zero-padded as `title`,
 * write_output.f90:302-330): `info.dat` with the reference's formatted lines, then one
 * stream-unformatted FP64 file per field over the active cells in cell order, in physical units:
 * x (position(:,1), or radii_c in polar geometry), rho, v, [y], vy, vz, P, and with dust rhod_tot,
 * rhod, sd, vd, vdy, vdz (species-major).  The bytes equal the reference's for the same state.
 * MHD, turbulence, dust passive scalars and charging outputs are not written (those builds are
 * rejected).
 *
 * shark_gpu_save_state / shark_gpu_load_state have no counterpart in the reference (its
 * read_output, src/commons/read_output.f90:1-11, is empty: "restart is not implemented").  They
 * dump and restore the device state of a context (u_prim with ghosts, time, dt of the next step,
 * step count), so a run continued from a dump takes exactly the steps the uninterrupted run takes.
 */
#ifndef SHARK_OUTPUT_H
#define SHARK_OUTPUT_H
#include <stdint.h>
#include "shark_gpu.h"
#ifdef __cplusplus
extern "C" {
#endif

#define SHARK_OUTPUT_VERSION 1

typedef struct shark_output_info {
  int32_t version;      /* SHARK_OUTPUT_VERSION */
  int32_t gridspace;    /* GRIDSPACE macro, written to info.dat */
  int32_t mhd, turb, ndustpscal, charging;   /* must be 0 */
  double time;          /* simulated time, code units */
  /* the setup's units module (e.g. src/setups/disk_polar/units.f90) */
  double unit_t, unit_d, unit_l, unit_v, unit_p, unit_dcol;
  const char* base_dir; /* output_NNNNN/ is created here (NULL: the working directory) */
} shark_output_info;

/* From host arrays: q(ncells,nvar) primitives as ctoprim (solve.f90:195-253) leaves them,
 * sdust(ncells,ndust) grain sizes (code units; may be NULL without dust), position(ncells,ndim),
 * radii_c(ncells), with the geometry of cfg (nx, ny, ndust, geom).  Needs no GPU. */
int shark_output_write(const shark_output_info* info, int iout, const shark_cfg* cfg, const double* q,
                       const double* sdust, const double* position, const double* radii_c);

/* From a context's device state (ctoprim on the device, then the same writer).  position must
 * have been given to shark_gpu_set_geometry. */
int shark_gpu_write_output(shark_ctx* ctx, const shark_output_info* info, int iout);

/* Restart dump: writes / reads `path` (binary, this library's own format).  load_state needs a
 * context created, given its geometry and uploaded (the setup's static fields) with the same
 * configuration; it then replaces u_prim, time, dt and the step count. */
int shark_gpu_save_state(shark_ctx* ctx, const char* path);
int shark_gpu_load_state(shark_ctx* ctx, const char* path);

#ifdef __cplusplus
}
#endif
#endif
