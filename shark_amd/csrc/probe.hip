// probe.hip — on-device HBM bandwidth a plain streaming kernel attains on this GPU (copy, read,
// write of FP64 buffers with 16-byte vector accesses), timed with HIP events.  The step kernels'
// achieved bandwidth is rated against this figure beside the vendor peak.
#include <hip/hip_runtime.h>

#include <cstdio>

#include "../../include/shark_gpu.h"

namespace {

__global__ __launch_bounds__(256) void k_copy(const double2* __restrict__ a, double2* __restrict__ b, int64_t n) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) b[i] = a[i];
}

__global__ __launch_bounds__(256) void k_read(const double2* __restrict__ a, int64_t n, double* out) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  double s = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
    const double2 v = a[i];
    s += v.x + v.y;
  }
  // one value per block keeps the loads live; the sum itself is not used
  __shared__ double red[256];
  red[threadIdx.x] = s;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[blockIdx.x] = red[0];
}

__global__ __launch_bounds__(256) void k_write(double2* __restrict__ b, int64_t n, double x) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) b[i] = make_double2(x, x);
}

}  // namespace

extern "C" int shark_hbm_probe(int64_t nbytes, int reps, double* gbs) {
  if (!gbs || nbytes < (1 << 20) || reps < 1) return SHARK_EINVAL;
  const int64_t n = nbytes / 16;   // double2 elements per buffer
  double2 *a = nullptr, *b = nullptr;
  double* part = nullptr;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  int dev = 0, ncu = 0;
  int rc = SHARK_OK;
  if (hipGetDevice(&dev) != hipSuccess ||
      hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess)
    return SHARK_EHIP;
  const int nblk = ncu * 8;   // eight blocks of four waves per CU, grid-stride
  if (hipMalloc(&a, n * 16) != hipSuccess || hipMalloc(&b, n * 16) != hipSuccess ||
      hipMalloc(&part, sizeof(double) * nblk) != hipSuccess || hipEventCreate(&e0) != hipSuccess ||
      hipEventCreate(&e1) != hipSuccess) {
    rc = SHARK_ENOMEM;
  }
  for (int which = 0; which < 3 && rc == SHARK_OK; ++which) {
    auto launch = [&] {
      if (which == 0) hipLaunchKernelGGL(k_write, dim3(nblk), dim3(256), 0, 0, a, n, 1.0);
      else if (which == 1) hipLaunchKernelGGL(k_copy, dim3(nblk), dim3(256), 0, 0, a, b, n);
      else hipLaunchKernelGGL(k_read, dim3(nblk), dim3(256), 0, 0, b, n, part);
    };
    launch();   // untimed first pass
    float ms = 0.f;
    if (hipEventRecord(e0, 0) != hipSuccess) { rc = SHARK_EHIP; break; }
    for (int r = 0; r < reps; ++r) launch();
    if (hipEventRecord(e1, 0) != hipSuccess || hipEventSynchronize(e1) != hipSuccess ||
        hipEventElapsedTime(&ms, e0, e1) != hipSuccess || hipGetLastError() != hipSuccess) {
      rc = SHARK_EHIP;
      break;
    }
    const double moved = (which == 1 ? 2.0 : 1.0) * (double)n * 16.0 * reps;   // write, copy, read
    gbs[which] = moved / (ms * 1e-3) / 1e9;
  }
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  if (a) (void)hipFree(a);
  if (b) (void)hipFree(b);
  if (part) (void)hipFree(part);
  return rc;
}
