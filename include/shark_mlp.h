/*
 * shark_mlp.h — C-ABI of the learned growth-and-charging source step in libshark_gpu.so.
 *
 * No counterpart in the reference: SHARK has no learned components.  The network stands in for
 * the growth-and-charging source step (dust_growth, src/dust/dust_growth.f90:3-256, and
 * analytical_charge, src/chemistry/ionisation.f90:809-922) and is driven through
 * shark_growth_set_surrogate / SHARK_GROWTH_SURROGATE of shark_growth.h.
 *
 * A multilayer perceptron n_in -> n_hidden (x n_layers-1) -> n_out with SiLU activations.  Master
 * weights are FP32; every product runs on MFMA with BF16 operands (weights and activations
 * rounded to BF16 at the point they enter a product) and FP32 accumulation.  Forward, backward
 * and Adam run in this library's own HIP kernels.
 *
 * Parameter vector layout (FP32): for each layer l: W_l (n_out_l x n_in_l, row-major), then
 * b_l (n_out_l).  Data are row-major (rows = samples).
 * Loss: L = 1/(n n_out) sum_{rows, outputs} w (y - t)^2, with per-element weights w (1 unless
 * given, row-major like t).
 */
#ifndef SHARK_MLP_H
#define SHARK_MLP_H
#include <stdint.h>
#ifdef __cplusplus
extern "C" {
#endif

#define SHARK_MLP_VERSION 1

typedef struct shark_mlp_cfg {
  int32_t version;    /* SHARK_MLP_VERSION */
  int32_t n_in;       /* inputs per sample */
  int32_t n_hidden;   /* width of the hidden layers */
  int32_t n_layers;   /* weight layers (>= 2) */
  int32_t n_out;      /* outputs per sample */
  int32_t max_batch;  /* largest row count of one forward / gradient call */
  int32_t reserved_i[2];
  double beta1, beta2, adam_eps; /* Adam moments */
  double reserved[8];
} shark_mlp_cfg;

typedef struct shark_mlp shark_mlp;

int shark_mlp_create(const shark_mlp_cfg* cfg, shark_mlp** m);
int64_t shark_mlp_num_params(const shark_mlp* m);
/* FP32 master parameters in and out (host arrays of shark_mlp_num_params floats). */
int shark_mlp_set_params(shark_mlp* m, const float* params);
int shark_mlp_get_params(shark_mlp* m, float* params);
/* y (n x n_out) = net(x (n x n_in)); n <= max_batch; host arrays.  Runs the fused inference
 * kernel (all layers in one launch, activations in LDS) when every width is <= 256. */
int shark_mlp_forward(shark_mlp* m, const float* x, int64_t n, float* y);
/* The fused forward with the activations carried as BF16 hi + lo pairs (two products per weight
 * fragment): the weights stay BF16, the activations keep about 16 significant bits, so the network
 * is evaluated without the BF16 rounding noise of its activations.  Needs <= 8 layers, widths <= 256
 * and hidden widths that are multiples of 32.  The learned growth step uses it. */
int shark_mlp_forward_hilo(shark_mlp* m, const float* x, int64_t n, float* y);
/* The same forward as the training path computes it (one product kernel per layer).  Its outputs
 * equal shark_mlp_forward's bit for bit. */
int shark_mlp_forward_layered(shark_mlp* m, const float* x, int64_t n, float* y);
/* Loss and its gradient w.r.t. every parameter on one batch (n <= max_batch). */
int shark_mlp_grads(shark_mlp* m, const float* x, const float* t, int64_t n, float* loss,
                    float* grads);
int shark_mlp_grads_weighted(shark_mlp* m, const float* x, const float* t, const float* w, int64_t n,
                             float* loss, float* grads);
/* Training set kept on the device (copied in); replaces any earlier set.  The weighted form adds
 * per-element loss weights (n x n_out; NULL: all 1). */
int shark_mlp_set_data(shark_mlp* m, const float* x, const float* t, int64_t n);
int shark_mlp_set_data_weighted(shark_mlp* m, const float* x, const float* t, const float* w, int64_t n);
/* Adam over the device training set: `epochs` passes in batches of `batch` rows (<= max_batch)
 * in a seeded random order, learning rate lr.  loss_out[e] = mean batch loss of epoch e. */
int shark_mlp_train(shark_mlp* m, int epochs, int batch, float lr, uint64_t seed, float* loss_out);
/* Device time (ms) of the last forward / train call. */
int shark_mlp_kernel_time(shark_mlp* m, double* ms);
int shark_mlp_destroy(shark_mlp* m);

#ifdef __cplusplus
}
#endif
#endif
