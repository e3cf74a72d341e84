// mlp.hip — multilayer perceptron on gfx950: forward, backward and Adam in hand-written kernels.
//
// Every product is one kernel, k_gemm: C[M x N] = A[M x K] . B[K x N] for FP32 operands in
// global memory with arbitrary strides (so W, W^T, dY^T and H all load without copies), staged
// into LDS as BF16 in the lane layout of v_mfma_f32_32x32x16_bf16 and accumulated in FP32.  A
// 256-lane workgroup owns a 64 x 64 tile (four waves, one 32 x 32 MFMA block each) and walks K in
// steps of 32.  The epilogue fuses bias, SiLU (storing the pre-activation for the backward pass)
// or the SiLU derivative.  Weight gradients reduce over the batch with split-K partials summed in
// a fixed order, so training is deterministic for a given seed.
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <random>
#include <vector>

#include "../../include/shark_mlp.h"
#include "growth.h"
#include "mlp.h"

namespace {

typedef float f32x16 __attribute__((ext_vector_type(16)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));

constexpr int BM = 64, BN = 64, BK = 32;
enum { EPI_STORE = 0, EPI_BIAS = 1, EPI_BIAS_SILU = 2, EPI_DSILU = 3 };

struct GemmArgs {
  int M, N, K;
  const float* A;
  int64_t sAm, sAk;
  const float* B;
  int64_t sBk, sBn;
  float* C;
  int64_t ldc;
  const float* bias;  // [N]
  float* Z;           // pre-activation: written by EPI_BIAS_SILU, read by EPI_DSILU (ldc)
  int kchunk;         // K range per blockIdx.z
  int64_t zstride;    // C offset per blockIdx.z (split-K partials)
};

__device__ __forceinline__ unsigned short to_bf16(float x) {
  unsigned int u = __float_as_uint(x);
  u += 0x7FFFu + ((u >> 16) & 1u);
  return (unsigned short)(u >> 16);
}
__device__ __forceinline__ float from_bf16(unsigned short u) { return __uint_as_float((unsigned int)u << 16); }

// SiLU from the hardware exp2 and reciprocal (v_exp_f32, v_rcp_f32: about 1 ulp each), shared
// by the training forward and the fused inference so the two stay bit-identical; the activations
// are rounded to BF16 right after, far coarser than these approximations.
__device__ __forceinline__ float silu(float z) {
  return z * __builtin_amdgcn_rcpf(1.0f + __builtin_amdgcn_exp2f(-z * 1.44269504088896341f));
}
__device__ __forceinline__ float dsilu(float z) {
  const float s = 1.0f / (1.0f + expf(-z));
  return s * (1.0f + z * (1.0f - s));
}

template <int EPI>
__global__ __launch_bounds__(256) void k_gemm(GemmArgs g) {
  __shared__ unsigned short As[BK / 8][BM][8];
  __shared__ unsigned short Bs[BK / 8][BN][8];
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const int m0 = blockIdx.y * BM, n0 = blockIdx.x * BN;
  const int wr = (w & 1) * 32, wc = (w >> 1) * 32;
  const int kbeg = blockIdx.z * g.kchunk;
  const int kend = min(g.K, kbeg + g.kchunk);
  const bool a_kfast = g.sAk == 1;
  const bool b_nfast = g.sBn == 1;
  f32x16 acc;
  for (int r = 0; r < 16; ++r) acc[r] = 0.f;
  for (int k0 = kbeg; k0 < kend; k0 += BK) {
    for (int i = 0; i < (BM * BK) / 256; ++i) {
      const int e = tid + 256 * i;
      int m, k;
      if (a_kfast) { m = e >> 5; k = e & 31; } else { m = e & 63; k = e >> 6; }
      const int gm = m0 + m, gk = k0 + k;
      float v = 0.f;
      if (gm < g.M && gk < kend) v = g.A[(int64_t)gm * g.sAm + (int64_t)gk * g.sAk];
      As[k >> 3][m][k & 7] = to_bf16(v);
    }
    for (int i = 0; i < (BN * BK) / 256; ++i) {
      const int e = tid + 256 * i;
      int n, k;
      if (b_nfast) { n = e & 63; k = e >> 6; } else { n = e >> 5; k = e & 31; }
      const int gn = n0 + n, gk = k0 + k;
      float v = 0.f;
      if (gn < g.N && gk < kend) v = g.B[(int64_t)gk * g.sBk + (int64_t)gn * g.sBn];
      Bs[k >> 3][n][k & 7] = to_bf16(v);
    }
    __syncthreads();
    for (int s = 0; s < BK / 16; ++s) {
      const int kg = 2 * s + (lane >> 5);
      const u32x4 a = *reinterpret_cast<const u32x4*>(&As[kg][wr + (lane & 31)][0]);
      const u32x4 b = *reinterpret_cast<const u32x4*>(&Bs[kg][wc + (lane & 31)][0]);
      acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(__builtin_bit_cast(bf16x8, a),
                                                    __builtin_bit_cast(bf16x8, b), acc, 0, 0, 0);
    }
    __syncthreads();
  }
  float* C = g.C + blockIdx.z * g.zstride;
  for (int r = 0; r < 16; ++r) {
    const int row = m0 + wr + 8 * (r >> 2) + 4 * (lane >> 5) + (r & 3);
    const int col = n0 + wc + (lane & 31);
    if (row >= g.M || col >= g.N) continue;
    const int64_t o = (int64_t)row * g.ldc + col;
    float v = acc[r];
    if constexpr (EPI == EPI_BIAS) {
      v += g.bias[col];
    } else if constexpr (EPI == EPI_BIAS_SILU) {
      v += g.bias[col];
      g.Z[o] = v;
      v = silu(v);
    } else if constexpr (EPI == EPI_DSILU) {
      v *= dsilu(g.Z[o]);
    }
    C[o] = v;
  }
}

// out[i] = sum_z P[z*n + i], z in order
__global__ void k_sum_parts(const float* P, int parts, int64_t n, float* out) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  float s = 0.f;
  for (int z = 0; z < parts; ++z) s += P[(int64_t)z * n + i];
  out[i] = s;
}

// out[n] = sum_m X[m*ld + n]; 64 columns per block, 4 row phases reduced in order
__global__ __launch_bounds__(256) void k_colsum(const float* X, int M, int N, int64_t ld, float* out) {
  __shared__ float part[4][64];
  const int c = threadIdx.x & 63, ph = threadIdx.x >> 6;
  const int n = blockIdx.x * 64 + c;
  float s = 0.f;
  if (n < N)
    for (int m = ph; m < M; m += 4) s += X[(int64_t)m * ld + n];
  part[ph][c] = s;
  __syncthreads();
  if (ph == 0 && n < N) out[n] = ((part[0][c] + part[1][c]) + part[2][c]) + part[3][c];
}

// dY = 2 (Y - T) / (n n_out); per-block partial sums of (Y - T)^2 / (n n_out)
// loss w (y - t)^2 * scale and its gradient; W == nullptr: w = 1
__global__ __launch_bounds__(256) void k_lossgrad(const float* Y, const float* T, const float* W,
                                                  int64_t total, float scale, float* dY, float* partial) {
  __shared__ float red[256];
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  float e2 = 0.f;
  if (i < total) {
    const float d = Y[i] - T[i];
    const float ws = W ? W[i] * scale : scale;
    dY[i] = 2.f * d * ws;
    e2 = d * d * ws;
  }
  red[threadIdx.x] = e2;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if (threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) partial[blockIdx.x] = red[0];
}

__global__ void k_adam(float* p, const float* g, float* m, float* v, int64_t n, float lr, float b1,
                       float b2, float eps, float c1, float c2) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float gi = g[i];
  const float mi = b1 * m[i] + (1.f - b1) * gi;
  const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
  m[i] = mi;
  v[i] = vi;
  p[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
}

__global__ void k_gather(const float* src, const int* idx, int64_t rows, int cols, float* dst) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= rows * cols) return;
  const int64_t r = i / cols;
  const int c = (int)(i - r * cols);
  dst[i] = src[(int64_t)idx[r] * cols + c];
}

// ------------------------------------------------------------------ fused inference
// The whole network in one kernel for widths <= 256: a workgroup of 256 lanes takes FM = 128
// rows (samples), keeps their activations in LDS as BF16 between layers, and reads the weights
// from a copy packed in the operand order of v_mfma_f32_32x32x16_bf16 (one 1 KB fragment per wave
// and k-step, fetched FPF k-steps ahead through L2).  Each product is computed transposed,
// out^T = W . act^T: the weights are the A operand and the activations the B operand, so a lane's
// accumulators hold four consecutive features of one sample, which go to LDS as one 8-byte store
// and to a feature-major output as coalesced rows.  Wave w computes the 32-feature blocks w, w+4
// of a layer for all 128 samples.  Operand values, k order and epilogue equal k_gemm's, so the
// outputs equal the layer-by-layer forward bit for bit.
constexpr int FM = 128;      // samples per workgroup
constexpr int FW = 256;      // widest layer of the fused path
constexpr int FLD = FW + 8;  // LDS row pitch in BF16 elements (16-byte rows, 4-bank shift per row)
constexpr int FMAXL = 8;     // layers
constexpr int FPF = 4;       // k-steps of weight fragments in flight per wave

typedef __bf16 bf16x2 __attribute__((ext_vector_type(2)));
typedef float f32x2 __attribute__((ext_vector_type(2)));

struct InferArgs {
  const float* x;
  float* y;
  int64_t n;
  int64_t xr, xc;            // element strides of x: per sample, per feature
  int64_t yr, yc;            // element strides of y
  int n_in, n_out, L;
  const unsigned short* wf;  // packed BF16 weights
  const float* p;            // FP32 parameters (biases)
  int64_t foff[FMAXL];       // fragment offset of each layer (BF16 elements)
  int64_t boff[FMAXL];       // bias offset of each layer (floats)
  int kpad[FMAXL];           // input width rounded up to 16
  int nbn[FMAXL];            // output width in 32-feature blocks
  int dout[FMAXL];
};

// fragment (nb, ks) of layer l, lane j, element e  <-  W[n = 32 nb + j % 32][k = 16 ks + 8 (j / 32) + e]
__global__ void k_pack_frag(const float* W, int din, int dout, int nks, int nbn, unsigned short* out) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t total = (int64_t)nbn * nks * 64 * 8;
  if (i >= total) return;
  const int e = (int)(i & 7), j = (int)((i >> 3) & 63);
  const int64_t f = i >> 9;
  const int ks = (int)(f % nks), nb = (int)(f / nks);
  const int n = 32 * nb + (j & 31), k = 16 * ks + 8 * (j >> 5) + e;
  out[i] = (n < dout && k < din) ? to_bf16(W[(int64_t)n * din + k]) : (unsigned short)0;
}

__device__ __forceinline__ f32x2 bf16x2_to_f32(unsigned int u) {
  return f32x2{__uint_as_float(u << 16), __uint_as_float(u & 0xffff0000u)};
}

__device__ __forceinline__ unsigned int pack_bf16x2(float a, float b) {
  const f32x2 v = {a, b};
  return __builtin_bit_cast(unsigned int, __builtin_convertvector(v, bf16x2));
}

// One layer for all FMT samples of the workgroup (FMT / 32 blocks of 32 samples per wave).  Hidden
// layers (every width a multiple of 32): bias + SiLU, BF16 back into act once every wave has read
// the layer's input.  Last layer: bias, FP32 out to y.  HL: activations carried as BF16 pairs
// hi + lo (lo = BF16(v - hi)), two products per weight fragment, so the activations keep about 16
// significant bits; the weights stay BF16.
template <bool LAST, int FMT, bool HL>
__device__ __forceinline__ void infer_layer(const InferArgs& g, int l, unsigned short (*act)[FLD],
                                            unsigned short (*alo)[FLD], int64_t r0, int lane, int w) {
  constexpr int MB = FMT / 32;
  const int nks = g.kpad[l] >> 4, nbn = g.nbn[l];
  const float* bias = g.p + g.boff[l];
  const int dout = g.dout[l];
  const u32x4* wf = reinterpret_cast<const u32x4*>(g.wf + g.foff[l]);
  // two passes, feature blocks w and w + 4; a hidden layer's outputs wait in registers as packed
  // BF16 until every wave has finished reading the layer's input
  uint2 pk[2][MB][4];
  uint2 pl[2][HL ? MB : 1][4];
#pragma unroll
  for (int pass = 0; pass < 2; ++pass) {
    const int nb = w + 4 * pass;
    if (nb >= nbn) break;
    f32x16 acc[MB];
#pragma unroll
    for (int mb = 0; mb < MB; ++mb)
#pragma unroll
      for (int r = 0; r < 16; ++r) acc[mb][r] = 0.f;
    const u32x4* wp = wf + (int64_t)nb * nks * 64 + lane;
    u32x4 b[FPF];
#pragma unroll
    for (int u = 0; u < FPF; ++u)
      if (u < nks) b[u] = wp[u * 64];
    for (int ks0 = 0; ks0 < nks; ks0 += FPF) {
#pragma unroll
      for (int u = 0; u < FPF; ++u) {
        const int ks = ks0 + u;
        if (ks < nks) {
          const u32x4 c = b[u];
          if (ks + FPF < nks) b[u] = wp[(ks + FPF) * 64];
#pragma unroll
          for (int mb = 0; mb < MB; ++mb) {
            const int row = 32 * mb + (lane & 31), col = 16 * ks + 8 * (lane >> 5);
            const u32x4 a = *reinterpret_cast<const u32x4*>(&act[row][col]);
            acc[mb] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(__builtin_bit_cast(bf16x8, c),
                                                              __builtin_bit_cast(bf16x8, a), acc[mb], 0, 0, 0);
            if constexpr (HL) {
              const u32x4 al = *reinterpret_cast<const u32x4*>(&alo[row][col]);
              acc[mb] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(__builtin_bit_cast(bf16x8, c),
                                                                __builtin_bit_cast(bf16x8, al), acc[mb], 0, 0, 0);
            }
          }
        }
      }
    }
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      // accumulators 4q..4q+3: features n0..n0+3 of sample 32 mb + lane % 32
      const int n0 = 32 * nb + 8 * q + 4 * (lane >> 5);
      if constexpr (!LAST) {
        const float4 bv = *reinterpret_cast<const float4*>(bias + n0);
#pragma unroll
        for (int mb = 0; mb < MB; ++mb) {
          float v0 = acc[mb][4 * q], v1 = acc[mb][4 * q + 1], v2 = acc[mb][4 * q + 2], v3 = acc[mb][4 * q + 3];
          v0 += bv.x;
          v1 += bv.y;
          v2 += bv.z;
          v3 += bv.w;
          v0 = silu(v0);
          v1 = silu(v1);
          v2 = silu(v2);
          v3 = silu(v3);
          pk[pass][mb][q] = uint2{pack_bf16x2(v0, v1), pack_bf16x2(v2, v3)};
          if constexpr (HL) {
            const f32x2 h01 = bf16x2_to_f32(pk[pass][mb][q].x), h23 = bf16x2_to_f32(pk[pass][mb][q].y);
            pl[pass][mb][q] = uint2{pack_bf16x2(v0 - h01.x, v1 - h01.y), pack_bf16x2(v2 - h23.x, v3 - h23.y)};
          }
        }
      } else {
#pragma unroll
        for (int i = 0; i < 4; ++i) {
          if (n0 + i >= dout) break;
          const float bv = bias[n0 + i];
          float* yc = g.y + (int64_t)(n0 + i) * g.yc;
#pragma unroll
          for (int mb = 0; mb < MB; ++mb) {
            const int64_t row = r0 + 32 * mb + (lane & 31);
            float v = acc[mb][4 * q + i];
            v += bv;
            if (row < g.n) yc[row * g.yr] = v;
          }
        }
      }
    }
  }
  if constexpr (!LAST) {
    __syncthreads();  // every wave has read this layer's input rows
#pragma unroll
    for (int pass = 0; pass < 2; ++pass) {
      const int nb = w + 4 * pass;
      if (nb >= nbn) break;
#pragma unroll
      for (int q = 0; q < 4; ++q)
#pragma unroll
        for (int mb = 0; mb < MB; ++mb) {
          const int row = 32 * mb + (lane & 31), col = 32 * nb + 8 * q + 4 * (lane >> 5);
          *reinterpret_cast<uint2*>(&act[row][col]) = pk[pass][mb][q];
          if constexpr (HL) *reinterpret_cast<uint2*>(&alo[row][col]) = pl[pass][mb][q];
        }
    }
    __syncthreads();
  }
}

template <int FMT, bool HL>
__global__ __launch_bounds__(256) void k_mlp_infer(InferArgs g) {
  __shared__ __attribute__((aligned(16))) unsigned short act[FMT][FLD];
  __shared__ __attribute__((aligned(16))) unsigned short alo[HL ? FMT : 1][FLD];
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const int64_t r0 = (int64_t)blockIdx.x * FMT;
  {
    const int kp = g.kpad[0];
    const bool feat_major = g.xc != 1;   // consecutive lanes take consecutive samples
    for (int e = tid; e < FMT * kp; e += 256) {
      int m, k;
      if (feat_major) { m = e % FMT; k = e / FMT; } else { m = e / kp; k = e - m * kp; }
      const int64_t row = r0 + m;
      const float v = (row < g.n && k < g.n_in) ? g.x[row * g.xr + k * g.xc] : 0.f;
      act[m][k] = to_bf16(v);
      if constexpr (HL) alo[m][k] = to_bf16(v - from_bf16(act[m][k]));
    }
  }
  __syncthreads();
  for (int l = 0; l < g.L - 1; ++l) infer_layer<false, FMT, HL>(g, l, act, alo, r0, lane, w);
  infer_layer<true, FMT, HL>(g, g.L - 1, act, alo, r0, lane, w);
}

}  // namespace

// ------------------------------------------------------------------ host object
struct shark_mlp {
  shark_mlp_cfg cfg;
  int L = 0;
  std::vector<int> din, dout;          // per layer
  std::vector<int64_t> woff, boff;     // parameter offsets
  int64_t np = 0;
  hipStream_t stream = nullptr;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  float *dp = nullptr, *dg = nullptr, *dm = nullptr, *dv = nullptr;  // params, grads, Adam moments
  float* dx = nullptr;                 // batch input [max_batch x n_in]
  float* dt = nullptr;                 // batch target [max_batch x n_out]
  std::vector<float*> Z, H;            // per hidden layer [max_batch x n_hidden]
  float *dy = nullptr, *ddy = nullptr; // output and its gradient [max_batch x n_out]
  float *dda = nullptr, *ddb = nullptr;  // backward ping-pong [max_batch x n_hidden]
  float* dpart = nullptr;              // split-K partials / loss partials
  int64_t part_cap = 0;
  float* hloss = nullptr;              // pinned
  float *dX = nullptr, *dT = nullptr;  // training set
  float* dW = nullptr;                 // training-set loss weights [ndata x n_out] (optional)
  float* dw = nullptr;                 // batch loss weights [max_batch x n_out]
  int* didx = nullptr;
  int64_t ndata = 0;
  int64_t adam_t = 0;
  double last_ms = 0.0;
  unsigned short* dwf = nullptr;       // BF16 weights packed for k_mlp_infer
  bool wf_valid = false;               // false after any change of the parameters
};

static int mfail(const char* msg) { return shark::fail_msg(1, msg); }

#define MCHK(x)                                                                              \
  do {                                                                                       \
    hipError_t e_ = (x);                                                                     \
    if (e_ != hipSuccess) {                                                                  \
      char m_[256];                                                                          \
      snprintf(m_, sizeof(m_), "%s failed: %s (%s:%d)", #x, hipGetErrorString(e_), __FILE__, \
               __LINE__);                                                                    \
      return shark::fail_msg(3, m_);                                                         \
    }                                                                                        \
  } while (0)

static void free_mlp(shark_mlp* m) {
  if (!m) return;
  float* bufs[] = {m->dp, m->dg, m->dm, m->dv, m->dx, m->dt, m->dy, m->ddy, m->dda, m->ddb, m->dpart,
                   m->dX, m->dT, m->dW, m->dw};
  for (float* b : bufs)
    if (b) (void)hipFree(b);
  for (float* b : m->Z)
    if (b) (void)hipFree(b);
  for (float* b : m->H)
    if (b) (void)hipFree(b);
  if (m->didx) (void)hipFree(m->didx);
  if (m->dwf) (void)hipFree(m->dwf);
  if (m->hloss) (void)hipHostFree(m->hloss);
  if (m->e0) (void)hipEventDestroy(m->e0);
  if (m->e1) (void)hipEventDestroy(m->e1);
  if (m->stream) (void)hipStreamDestroy(m->stream);
  delete m;
}

static int gemm(hipStream_t s, int epi, int M, int N, int K, const float* A, int64_t sAm, int64_t sAk,
                const float* B, int64_t sBk, int64_t sBn, float* C, int64_t ldc, const float* bias,
                float* Z, int splits = 1) {
  GemmArgs g;
  g.M = M;
  g.N = N;
  g.K = K;
  g.A = A;
  g.sAm = sAm;
  g.sAk = sAk;
  g.B = B;
  g.sBk = sBk;
  g.sBn = sBn;
  g.C = C;
  g.ldc = ldc;
  g.bias = bias;
  g.Z = Z;
  int kchunk = (K + splits - 1) / splits;
  kchunk = ((kchunk + BK - 1) / BK) * BK;
  g.kchunk = kchunk;
  g.zstride = (int64_t)M * ldc;
  const int nz = (K + kchunk - 1) / kchunk;
  dim3 grid((N + BN - 1) / BN, (M + BM - 1) / BM, nz > 0 ? nz : 1);
  switch (epi) {
    case EPI_STORE: hipLaunchKernelGGL(k_gemm<EPI_STORE>, grid, dim3(256), 0, s, g); break;
    case EPI_BIAS: hipLaunchKernelGGL(k_gemm<EPI_BIAS>, grid, dim3(256), 0, s, g); break;
    case EPI_BIAS_SILU: hipLaunchKernelGGL(k_gemm<EPI_BIAS_SILU>, grid, dim3(256), 0, s, g); break;
    default: hipLaunchKernelGGL(k_gemm<EPI_DSILU>, grid, dim3(256), 0, s, g); break;
  }
  MCHK(hipGetLastError());
  return 0;
}

namespace shark {

// Forward pass of n rows (device pointers), activations kept for the backward pass.
int mlp_forward_device(shark_mlp* m, const float* x, int64_t n, float* y, hipStream_t s) {
  const int L = m->L;
  const float* in = x;
  int64_t ld_in = m->din[0];
  for (int l = 0; l < L; ++l) {
    const float* W = m->dp + m->woff[l];
    const float* b = m->dp + m->boff[l];
    const int di = m->din[l], dout = m->dout[l];
    if (l < L - 1) {
      int rc = gemm(s, EPI_BIAS_SILU, (int)n, dout, di, in, ld_in, 1, W, 1, di, m->H[l], dout, b, m->Z[l]);
      if (rc) return rc;
      in = m->H[l];
      ld_in = dout;
    } else {
      int rc = gemm(s, EPI_BIAS, (int)n, dout, di, in, ld_in, 1, W, 1, di, y, dout, b, nullptr);
      if (rc) return rc;
    }
  }
  return 0;
}

// the fused kernel: at most FMAXL layers, every width <= 256, hidden widths multiples of 32 (so
// a hidden layer fills whole feature blocks and the next layer's padded input is exact)
static bool fused_ok(const shark_mlp* m) {
  if (m->L > FMAXL) return false;
  for (int l = 0; l < m->L; ++l) {
    if (m->din[l] > FW || m->dout[l] > FW) return false;
    if (l < m->L - 1 && m->dout[l] % 32 != 0) return false;
  }
  return true;
}

static int pack_weights(shark_mlp* m, hipStream_t s) {
  int64_t total = 0;
  for (int l = 0; l < m->L; ++l) total += (int64_t)((m->dout[l] + 31) / 32) * ((m->din[l] + 15) / 16) * 512;
  if (!m->dwf) MCHK(hipMalloc(&m->dwf, sizeof(unsigned short) * total));
  int64_t off = 0;
  for (int l = 0; l < m->L; ++l) {
    const int nks = (m->din[l] + 15) / 16, nbn = (m->dout[l] + 31) / 32;
    const int64_t cnt = (int64_t)nbn * nks * 512;
    hipLaunchKernelGGL(k_pack_frag, dim3((unsigned)((cnt + 255) / 256)), dim3(256), 0, s, m->dp + m->woff[l],
                       m->din[l], m->dout[l], nks, nbn, m->dwf + off);
    MCHK(hipGetLastError());
    off += cnt;
  }
  m->wf_valid = true;
  return 0;
}

// Inference of n rows (device pointers, any n): the fused kernel when every width is <= 256,
// else the layer-by-layer forward in chunks of max_batch rows (row-major data only).
// ldx / ldy > 0: x / y are feature-major, element (row, feature) at [feature * ld + row].
int mlp_infer_device(shark_mlp* m, const float* x, int64_t n, float* y, hipStream_t s, int64_t ldx,
                     int64_t ldy, bool hilo) {
  if (n <= 0) return 0;
  if (!fused_ok(m)) {
    if (ldx > 0 || ldy > 0) return mfail("mlp inference: feature-major data needs every width <= 256");
    for (int64_t r = 0; r < n; r += m->cfg.max_batch) {
      const int64_t k = std::min<int64_t>(m->cfg.max_batch, n - r);
      int rc = mlp_forward_device(m, x + r * m->cfg.n_in, k, y + r * m->cfg.n_out, s);
      if (rc) return rc;
    }
    return 0;
  }
  if (!m->wf_valid) {
    // the parameters are written on the MLP's own stream (training, set_params)
    MCHK(hipStreamSynchronize(m->stream));
    int rc = pack_weights(m, s);
    if (rc) return rc;
  }
  InferArgs g;
  memset(&g, 0, sizeof(g));
  g.x = x;
  g.y = y;
  g.n = n;
  g.xr = ldx > 0 ? 1 : m->cfg.n_in;
  g.xc = ldx > 0 ? ldx : 1;
  g.yr = ldy > 0 ? 1 : m->cfg.n_out;
  g.yc = ldy > 0 ? ldy : 1;
  g.n_in = m->cfg.n_in;
  g.n_out = m->cfg.n_out;
  g.L = m->L;
  g.wf = m->dwf;
  g.p = m->dp;
  int64_t off = 0;
  for (int l = 0; l < m->L; ++l) {
    g.kpad[l] = ((m->din[l] + 15) / 16) * 16;
    g.nbn[l] = (m->dout[l] + 31) / 32;
    g.dout[l] = m->dout[l];
    g.foff[l] = off;
    g.boff[l] = m->boff[l];
    off += (int64_t)g.nbn[l] * (g.kpad[l] / 16) * 512;
  }
  // hi + lo activations take twice the LDS per sample: 64 samples per workgroup keep two
  // workgroups per CU
  const int fm = hilo ? FM / 2 : FM;
  const int64_t nblk = (n + fm - 1) / fm;
  if (nblk > 0x7fffffff) return mfail("mlp inference: too many rows");
  if (hilo) hipLaunchKernelGGL((k_mlp_infer<FM / 2, true>), dim3((unsigned)nblk), dim3(256), 0, s, g);
  else hipLaunchKernelGGL((k_mlp_infer<FM, false>), dim3((unsigned)nblk), dim3(256), 0, s, g);
  MCHK(hipGetLastError());
  return 0;
}

bool mlp_fused_ok(const shark_mlp* m) { return fused_ok(m); }
int64_t mlp_max_batch(const shark_mlp* m) { return m->cfg.max_batch; }
int mlp_n_in(const shark_mlp* m) { return m->cfg.n_in; }
int mlp_n_out(const shark_mlp* m) { return m->cfg.n_out; }

}  // namespace shark

// gradients of the loss on rows already in m->dx / m->dt (n rows); returns loss via hloss[0]
static int grads_device(shark_mlp* m, int64_t n, bool weighted) {
  hipStream_t s = m->stream;
  const int L = m->L;
  int rc = shark::mlp_forward_device(m, m->dx, n, m->dy, s);
  if (rc) return rc;
  const int no = m->cfg.n_out;
  const int64_t total = n * no;
  const int nblk = (int)((total + 255) / 256);
  hipLaunchKernelGGL(k_lossgrad, dim3(nblk), dim3(256), 0, s, m->dy, m->dt, weighted ? m->dw : nullptr,
                     total, 1.0f / (float)total, m->ddy, m->dpart);
  MCHK(hipGetLastError());
  // reduce the loss partials in a fixed order on the device into hloss (pinned, mapped by copy)
  hipLaunchKernelGGL(k_sum_parts, dim3(1), dim3(1), 0, s, m->dpart, nblk, (int64_t)1, m->dpart + nblk);
  MCHK(hipGetLastError());
  MCHK(hipMemcpyAsync(m->hloss, m->dpart + nblk, sizeof(float), hipMemcpyDeviceToHost, s));
  // backward: layer l has input a_{l-1} (x for l = 0) and output gradient dZ_l
  const float* dZ = m->ddy;  // gradient at the output of the last layer (no activation)
  int64_t ld_dz = no;
  float* bufs[2] = {m->dda, m->ddb};
  const int splits = (int)std::min<int64_t>(64, std::max<int64_t>(1, n / 256));
  for (int l = L - 1; l >= 0; --l) {
    const int di = m->din[l], dout = m->dout[l];
    const float* a_in = (l == 0) ? m->dx : m->H[l - 1];
    const int64_t ld_in = (l == 0) ? m->cfg.n_in : m->cfg.n_hidden;
    // dW_l (dout x di) = dZ^T (dout x n) . a_in (n x di), split over the batch
    const int64_t mn = (int64_t)dout * di;
    if ((int64_t)splits * mn > m->part_cap) return mfail("mlp: split-K partial buffer too small");
    rc = gemm(s, EPI_STORE, dout, di, (int)n, dZ, 1, ld_dz, a_in, ld_in, 1, m->dpart, di, nullptr,
              nullptr, splits);
    if (rc) return rc;
    const int kchunk = (((int)((n + splits - 1) / splits) + BK - 1) / BK) * BK;
    const int nz = (int)((n + kchunk - 1) / kchunk);
    hipLaunchKernelGGL(k_sum_parts, dim3((unsigned)((mn + 255) / 256)), dim3(256), 0, s, m->dpart, nz,
                       mn, m->dg + m->woff[l]);
    MCHK(hipGetLastError());
    hipLaunchKernelGGL(k_colsum, dim3((dout + 63) / 64), dim3(256), 0, s, dZ, (int)n, dout, ld_dz,
                       m->dg + m->boff[l]);
    MCHK(hipGetLastError());
    if (l > 0) {
      // dZ_{l-1} = (dZ_l . W_l) * silu'(Z_{l-1})
      float* out = bufs[l & 1];
      const float* W = m->dp + m->woff[l];
      rc = gemm(s, EPI_DSILU, (int)n, di, dout, dZ, ld_dz, 1, W, di, 1, out, di, nullptr, m->Z[l - 1]);
      if (rc) return rc;
      dZ = out;
      ld_dz = di;
    }
  }
  return 0;
}

extern "C" {

int shark_mlp_create(const shark_mlp_cfg* cfg, shark_mlp** out) {
  if (!out) return mfail("shark_mlp_create: null output handle");
  *out = nullptr;
  if (!cfg) return mfail("shark_mlp_create: null configuration");
  if (cfg->version != SHARK_MLP_VERSION) return mfail("shark_mlp_cfg.version does not match SHARK_MLP_VERSION");
  if (cfg->n_in < 1 || cfg->n_out < 1 || cfg->n_hidden < 1 || cfg->n_layers < 2)
    return mfail("shark_mlp_create: n_in, n_out, n_hidden >= 1 and n_layers >= 2 required");
  if (cfg->n_in > 4096 || cfg->n_out > 4096 || cfg->n_hidden > 4096)
    return mfail("shark_mlp_create: layer widths are limited to 4096");
  if (cfg->max_batch < 1 || cfg->max_batch > (1 << 22))
    return mfail("shark_mlp_create: max_batch must lie in [1, 4194304]");
  shark_mlp* m = new shark_mlp();
  m->cfg = *cfg;
  m->L = cfg->n_layers;
  for (int l = 0; l < m->L; ++l) {
    m->din.push_back(l == 0 ? cfg->n_in : cfg->n_hidden);
    m->dout.push_back(l == m->L - 1 ? cfg->n_out : cfg->n_hidden);
    m->woff.push_back(m->np);
    m->np += (int64_t)m->din[l] * m->dout[l];
    m->boff.push_back(m->np);
    m->np += m->dout[l];
  }
  const int64_t B = cfg->max_batch;
  const int64_t wmax = std::max<int64_t>(cfg->n_hidden, std::max(cfg->n_in, cfg->n_out));
  m->part_cap = 64 * wmax * wmax + ((B * cfg->n_out + 255) / 256) + 16;
#define ALLOC(p, n)                                       \
  do {                                                    \
    hipError_t e_ = hipMalloc(&(p), sizeof(float) * (n)); \
    if (e_ != hipSuccess) {                               \
      free_mlp(m);                                        \
      return shark::fail_msg(3, "shark_mlp_create: device allocation failed"); \
    }                                                     \
  } while (0)
  if (hipStreamCreateWithFlags(&m->stream, hipStreamNonBlocking) != hipSuccess ||
      hipEventCreate(&m->e0) != hipSuccess || hipEventCreate(&m->e1) != hipSuccess) {
    free_mlp(m);
    return shark::fail_msg(3, "shark_mlp_create: stream/event creation failed");
  }
  ALLOC(m->dp, m->np);
  ALLOC(m->dg, m->np);
  ALLOC(m->dm, m->np);
  ALLOC(m->dv, m->np);
  ALLOC(m->dx, B * cfg->n_in);
  ALLOC(m->dt, B * cfg->n_out);
  ALLOC(m->dy, B * cfg->n_out);
  ALLOC(m->ddy, B * cfg->n_out);
  ALLOC(m->dw, B * cfg->n_out);
  ALLOC(m->dda, B * cfg->n_hidden);
  ALLOC(m->ddb, B * cfg->n_hidden);
  ALLOC(m->dpart, m->part_cap);
  for (int l = 0; l < m->L - 1; ++l) {
    float *z = nullptr, *h = nullptr;
    ALLOC(z, B * cfg->n_hidden);
    m->Z.push_back(z);
    ALLOC(h, B * cfg->n_hidden);
    m->H.push_back(h);
  }
#undef ALLOC
  if (hipHostMalloc(&m->hloss, sizeof(float) * 4, hipHostMallocDefault) != hipSuccess) {
    free_mlp(m);
    return shark::fail_msg(3, "shark_mlp_create: pinned allocation failed");
  }
  if (hipMemset(m->dp, 0, sizeof(float) * m->np) != hipSuccess ||
      hipMemset(m->dm, 0, sizeof(float) * m->np) != hipSuccess ||
      hipMemset(m->dv, 0, sizeof(float) * m->np) != hipSuccess) {
    free_mlp(m);
    return shark::fail_msg(3, "shark_mlp_create: memset failed");
  }
  *out = m;
  return 0;
}

int64_t shark_mlp_num_params(const shark_mlp* m) { return m ? m->np : -1; }

int shark_mlp_set_params(shark_mlp* m, const float* params) {
  if (!m || !params) return mfail("shark_mlp_set_params: null argument");
  m->wf_valid = false;
  MCHK(hipMemcpy(m->dp, params, sizeof(float) * m->np, hipMemcpyHostToDevice));
  MCHK(hipMemset(m->dm, 0, sizeof(float) * m->np));
  MCHK(hipMemset(m->dv, 0, sizeof(float) * m->np));
  m->adam_t = 0;
  return 0;
}

int shark_mlp_get_params(shark_mlp* m, float* params) {
  if (!m || !params) return mfail("shark_mlp_get_params: null argument");
  MCHK(hipStreamSynchronize(m->stream));
  MCHK(hipMemcpy(params, m->dp, sizeof(float) * m->np, hipMemcpyDeviceToHost));
  return 0;
}

static int forward_host(shark_mlp* m, const float* x, int64_t n, float* y, bool fused, bool hilo = false) {
  if (!m || !x || !y) return mfail("shark_mlp_forward: null argument");
  if (n < 1 || n > m->cfg.max_batch) return mfail("shark_mlp_forward: n must lie in [1, max_batch]");
  hipStream_t s = m->stream;
  MCHK(hipMemcpyAsync(m->dx, x, sizeof(float) * n * m->cfg.n_in, hipMemcpyHostToDevice, s));
  MCHK(hipEventRecord(m->e0, s));
  int rc = fused ? shark::mlp_infer_device(m, m->dx, n, m->dy, s, 0, 0, hilo)
                 : shark::mlp_forward_device(m, m->dx, n, m->dy, s);
  if (rc) return rc;
  MCHK(hipEventRecord(m->e1, s));
  MCHK(hipMemcpyAsync(y, m->dy, sizeof(float) * n * m->cfg.n_out, hipMemcpyDeviceToHost, s));
  MCHK(hipStreamSynchronize(s));
  float ms = 0.f;
  MCHK(hipEventElapsedTime(&ms, m->e0, m->e1));
  m->last_ms = ms;
  return 0;
}

int shark_mlp_forward(shark_mlp* m, const float* x, int64_t n, float* y) { return forward_host(m, x, n, y, true); }
int shark_mlp_forward_hilo(shark_mlp* m, const float* x, int64_t n, float* y) {
  if (m && !shark::mlp_fused_ok(m)) return mfail("shark_mlp_forward_hilo: needs the fused kernel's shape");
  return forward_host(m, x, n, y, true, true);
}
int shark_mlp_forward_layered(shark_mlp* m, const float* x, int64_t n, float* y) {
  return forward_host(m, x, n, y, false);
}

int shark_mlp_grads(shark_mlp* m, const float* x, const float* t, int64_t n, float* loss, float* grads) {
  return shark_mlp_grads_weighted(m, x, t, nullptr, n, loss, grads);
}

int shark_mlp_grads_weighted(shark_mlp* m, const float* x, const float* t, const float* w, int64_t n,
                             float* loss, float* grads) {
  if (!m || !x || !t) return mfail("shark_mlp_grads: null argument");
  if (n < 1 || n > m->cfg.max_batch) return mfail("shark_mlp_grads: n must lie in [1, max_batch]");
  hipStream_t s = m->stream;
  MCHK(hipMemcpyAsync(m->dx, x, sizeof(float) * n * m->cfg.n_in, hipMemcpyHostToDevice, s));
  MCHK(hipMemcpyAsync(m->dt, t, sizeof(float) * n * m->cfg.n_out, hipMemcpyHostToDevice, s));
  if (w) MCHK(hipMemcpyAsync(m->dw, w, sizeof(float) * n * m->cfg.n_out, hipMemcpyHostToDevice, s));
  int rc = grads_device(m, n, w != nullptr);
  if (rc) return rc;
  MCHK(hipStreamSynchronize(s));
  if (loss) *loss = m->hloss[0];
  if (grads) MCHK(hipMemcpy(grads, m->dg, sizeof(float) * m->np, hipMemcpyDeviceToHost));
  return 0;
}

int shark_mlp_set_data(shark_mlp* m, const float* x, const float* t, int64_t n) {
  return shark_mlp_set_data_weighted(m, x, t, nullptr, n);
}

int shark_mlp_set_data_weighted(shark_mlp* m, const float* x, const float* t, const float* w, int64_t n) {
  if (!m || !x || !t || n < 1) return mfail("shark_mlp_set_data: null argument or empty set");
  if (n > (int64_t)1 << 30) return mfail("shark_mlp_set_data: at most 2^30 rows");
  if (m->dX) (void)hipFree(m->dX);
  if (m->dT) (void)hipFree(m->dT);
  if (m->dW) (void)hipFree(m->dW);
  if (m->didx) (void)hipFree(m->didx);
  m->dX = m->dT = m->dW = nullptr;
  m->didx = nullptr;
  m->ndata = 0;
  MCHK(hipMalloc(&m->dX, sizeof(float) * n * m->cfg.n_in));
  MCHK(hipMalloc(&m->dT, sizeof(float) * n * m->cfg.n_out));
  if (w) MCHK(hipMalloc(&m->dW, sizeof(float) * n * m->cfg.n_out));
  MCHK(hipMalloc(&m->didx, sizeof(int) * n));
  MCHK(hipMemcpy(m->dX, x, sizeof(float) * n * m->cfg.n_in, hipMemcpyHostToDevice));
  MCHK(hipMemcpy(m->dT, t, sizeof(float) * n * m->cfg.n_out, hipMemcpyHostToDevice));
  if (w) MCHK(hipMemcpy(m->dW, w, sizeof(float) * n * m->cfg.n_out, hipMemcpyHostToDevice));
  m->ndata = n;
  return 0;
}

int shark_mlp_train(shark_mlp* m, int epochs, int batch, float lr, uint64_t seed, float* loss_out) {
  if (!m) return mfail("shark_mlp_train: null handle");
  if (!m->ndata) return mfail("shark_mlp_train: no training set (shark_mlp_set_data)");
  if (batch < 1 || batch > m->cfg.max_batch) return mfail("shark_mlp_train: batch must lie in [1, max_batch]");
  if (epochs < 0) return mfail("shark_mlp_train: epochs must be >= 0");
  hipStream_t s = m->stream;
  m->wf_valid = false;   // Adam rewrites the parameters
  std::mt19937_64 rng(seed);
  std::vector<int> perm(m->ndata);
  const int ni = m->cfg.n_in, no = m->cfg.n_out;
  const float b1 = (float)m->cfg.beta1, b2 = (float)m->cfg.beta2, eps = (float)m->cfg.adam_eps;
  MCHK(hipEventRecord(m->e0, s));
  for (int e = 0; e < epochs; ++e) {
    for (int64_t i = 0; i < m->ndata; ++i) perm[i] = (int)i;
    for (int64_t i = m->ndata - 1; i > 0; --i) {
      const int64_t j = (int64_t)(rng() % (uint64_t)(i + 1));
      std::swap(perm[i], perm[j]);
    }
    MCHK(hipMemcpyAsync(m->didx, perm.data(), sizeof(int) * m->ndata, hipMemcpyHostToDevice, s));
    MCHK(hipStreamSynchronize(s));  // perm is reused by the next epoch
    double lsum = 0.0;
    int nb = 0;
    for (int64_t b0 = 0; b0 + batch <= m->ndata; b0 += batch) {
      const int64_t rows = batch;
      hipLaunchKernelGGL(k_gather, dim3((unsigned)((rows * ni + 255) / 256)), dim3(256), 0, s, m->dX,
                         m->didx + b0, rows, ni, m->dx);
      hipLaunchKernelGGL(k_gather, dim3((unsigned)((rows * no + 255) / 256)), dim3(256), 0, s, m->dT,
                         m->didx + b0, rows, no, m->dt);
      if (m->dW)
        hipLaunchKernelGGL(k_gather, dim3((unsigned)((rows * no + 255) / 256)), dim3(256), 0, s, m->dW,
                           m->didx + b0, rows, no, m->dw);
      MCHK(hipGetLastError());
      int rc = grads_device(m, rows, m->dW != nullptr);
      if (rc) return rc;
      ++m->adam_t;
      const float c1 = 1.f - powf(b1, (float)m->adam_t), c2 = 1.f - powf(b2, (float)m->adam_t);
      hipLaunchKernelGGL(k_adam, dim3((unsigned)((m->np + 255) / 256)), dim3(256), 0, s, m->dp, m->dg,
                         m->dm, m->dv, m->np, lr, b1, b2, eps, c1, c2);
      MCHK(hipGetLastError());
      MCHK(hipStreamSynchronize(s));
      lsum += m->hloss[0];
      ++nb;
    }
    if (loss_out) loss_out[e] = nb ? (float)(lsum / nb) : 0.f;
  }
  MCHK(hipEventRecord(m->e1, s));
  MCHK(hipEventSynchronize(m->e1));
  float ms = 0.f;
  MCHK(hipEventElapsedTime(&ms, m->e0, m->e1));
  m->last_ms = ms;
  return 0;
}

int shark_mlp_kernel_time(shark_mlp* m, double* ms) {
  if (!m || !ms) return mfail("shark_mlp_kernel_time: null argument");
  *ms = m->last_ms;
  return 0;
}

int shark_mlp_destroy(shark_mlp* m) {
  free_mlp(m);
  return 0;
}

}  // extern "C"
