// mlp.h — device-side entry points of the MLP shared with the growth surrogate (mlp.hip).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
struct shark_mlp;
namespace shark {
// y (n x n_out) = net(x (n x n_in)) for device pointers, n <= max_batch, on stream s
int mlp_forward_device(shark_mlp* m, const float* x, int64_t n, float* y, hipStream_t s);
// inference of any number of rows (device pointers) on stream s: one fused kernel for widths <= 256;
// ldx / ldy > 0 select feature-major x / y (element (row, feature) at [feature * ld + row])
// hilo: activations carried as BF16 hi + lo pairs (two products per weight fragment)
int mlp_infer_device(shark_mlp* m, const float* x, int64_t n, float* y, hipStream_t s, int64_t ldx = 0,
                     int64_t ldy = 0, bool hilo = false);
// the fused kernel applies: <= 8 layers, widths <= 256, hidden widths multiples of 32
bool mlp_fused_ok(const shark_mlp* m);
int64_t mlp_max_batch(const shark_mlp* m);
int mlp_n_in(const shark_mlp* m);
int mlp_n_out(const shark_mlp* m);
}  // namespace shark
