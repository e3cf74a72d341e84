// growth.h — dust-growth state owned by a step context (defined in growth.hip).
#pragma once
struct shark_growth;
void shark_growth_free(shark_growth* g);
#include <hip/hip_runtime.h>
namespace shark {
int fail_msg(int code, const char* msg);
int growth_bind(shark_growth* g, double* dU, double* dcs, const double* rc_cols, int nxm, int nym, int nx,
                int ny, int ndim, const double* dt_dev, hipStream_t s);
int growth_step(shark_growth* g, int mode);
int growth_nbins(const shark_growth* g);
}
