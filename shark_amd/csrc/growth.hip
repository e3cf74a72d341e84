// growth.hip — dust-growth entry points (kernels not built yet: every call reports so).
#include <hip/hip_runtime.h>
#include "../../include/shark_growth.h"
#include "growth.h"

struct shark_growth { int dummy; };
void shark_growth_free(shark_growth* g) { delete g; }

extern "C" {
static int not_built() { return shark::fail_msg(1, "dust growth kernels are not built in this library"); }
int shark_growth_create(const shark_growth_params*, int64_t, shark_growth** g) { if (g) *g = nullptr; return not_built(); }
int shark_growth_upload(shark_growth*, const double*, const double*, const double*, const double*) { return not_built(); }
int shark_growth_run(shark_growth*, int, double) { return not_built(); }
int shark_growth_download(shark_growth*, double*, int64_t*) { return not_built(); }
int shark_growth_kernel_time(shark_growth*, double*) { return not_built(); }
int shark_growth_destroy(shark_growth* g) { shark_growth_free(g); return 0; }
}
