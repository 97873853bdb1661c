/* gpucell/surrogate.h -- C ABI of the learned membrane substepper in
 * libgpucell_hip.so (no counterpart in the reference, whose membrane moves
 * with the fluid time step).
 *
 * Part 1: the per-vertex MLP, 36 -> 128 -> 128 -> 128 -> 24 with tanh
 * hidden units, BF16 weights and activations and FP32 accumulation on the
 * matrix cores (v_mfma_f32_16x16x32_bf16).  Parameters travel as one flat
 * FP32 vector in this order (row-major, out x in):
 *   W1[128][36] b1[128] W2[128][128] b2[128] W3[128][128] b3[128]
 *   W4[24][128] b4[24]
 * The network computes, per sample row x (36 floats),
 *   h1 = tanh(W1 x + b1), h2 = tanh(W2 h1 + b2), h3 = tanh(W3 h2 + b3),
 *   y  = W4 h3 + b4,
 * with x, h1, h2, h3 and the weights rounded to BF16 (round to nearest
 * even) before each product.  Device pointers; launches go to the stream of
 * gc_SetStream.
 */
#ifndef GPUCELL_SURROGATE_H
#define GPUCELL_SURROGATE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define GC_MLP_IN 36
#define GC_MLP_HIDDEN 128
#define GC_MLP_OUT 24

typedef struct gc_mlp gc_mlp;

/* Number of FP32 parameters (40856). */
int gc_MlpParamCount(void);
/* New network with Glorot-uniform weights from `seed` and zero biases; NULL
 * on error (*status set). */
gc_mlp* gc_MlpCreate(uint64_t seed, int* status);
void gc_MlpDestroy(gc_mlp* m);
/* Host copies of the flat FP32 parameter vector (n = gc_MlpParamCount()). */
int gc_MlpSetParams(gc_mlp* m, const float* params, int n);
int gc_MlpGetParams(gc_mlp* m, float* params, int n);
/* Inference: y[batch][24] from x[batch][36] (fused kernel, weights held in
 * LDS, activations kept on chip). */
int gc_MlpForward(gc_mlp* m, const float* x, float* y, int batch);
/* Gradient of L = sum_{b,o} dy[b][o] * y[b][o] with respect to the flat
 * parameters (dy rounded to BF16 like every product input); grad has
 * gc_MlpParamCount() floats. */
int gc_MlpGradient(gc_mlp* m, const float* x, const float* dy, int batch, float* grad);
/* Mean-squared-error training with Adam (beta 0.9 / 0.999, eps 1e-8) on
 * nsamples rows (x[n][36], y[n][24], device); `epochs` passes over a
 * permutation drawn from `seed` each epoch, minibatches of `batch` rows,
 * learning rate lr decayed by cosine to lr/20.  loss_per_epoch (host, may be
 * NULL) receives each epoch's mean training loss. */
int gc_MlpTrain(gc_mlp* m, const float* x, const float* y, int nsamples, int batch, int epochs,
                float lr, uint64_t seed, float* loss_per_epoch);

#ifdef __cplusplus
} /* extern "C" */
#endif

#endif /* GPUCELL_SURROGATE_H */
