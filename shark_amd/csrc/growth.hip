#include <hip/hip_runtime.h>
#include "growth.h"
struct shark_growth { int dummy; };
void shark_growth_free(shark_growth* g) { delete g; }
