// output.h — the host writer of output directories (output.hip), shared with the step context.
#pragma once
#include <stdint.h>

#include <string>

#include "../../include/shark_output.h"

namespace shark {
struct OutputGrid {
  int nx, ny, ndim, nd, nxm, geom;
  int64_t nc;
};
// 1 when info can be written (version, no MHD/turbulence/scalars/charging), else 0 with *err set
int output_info_ok(const shark_output_info* info, std::string* err);
// Fortran Ew.d as flang formats it
std::string fortran_e(double v, int w, int d);
int write_output_dir(const shark_output_info* info, int iout, const OutputGrid& g, const double* q,
                     const double* sdust, const double* position, const double* radii_c, std::string* err);
}  // namespace shark
