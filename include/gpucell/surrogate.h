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

/* Part 2: membrane time stepping of an engine (one suspension, not a batch).
 *
 * mode 0: the reference's: membrane vertices move with the fluid step.
 * mode 1: plain substepping: per engine step of length dt, the membrane
 *         takes k velocity-Verlet substeps of dt/k under the fluid force of
 *         the step (SDPD pair force, Morse, body force; held fixed) plus
 *         the membrane force re-evaluated at every substep (FP64 positions,
 *         FP64 cell area/volume reductions).
 * mode 2: surrogate: the MLP predicts the k = 8 substep positions of every
 *         vertex from its 1-ring in a local frame; the membrane force is then
 *         evaluated once at the predicted end state, and the residual
 *           R = rms_v |d2x_v - dts^2 a_v| / l_v
 *         (d2x = x8 - 2 x7 + x6, a = acceleration at x8, l = the vertex's
 *         rest edge length: the velocity-Verlet relation of the last
 *         substeps, which an exact trajectory meets to O(dts^3), in edge
 *         lengths) decides: R > threshold -> the step's membrane part is
 *         recomputed by plain substepping (the fallback).
 * The fluid keeps the engine's step in every mode.  In modes 1 and 2 the
 * exported force of a membrane vertex is the fluid part (the membrane part
 * is in defforce).  norm: 120 floats (feature mean[36], feature std[36],
 * target mean[24], target std[24]) or NULL for none.  With the per-cell
 * membrane kernel (tuning variant 3) the fallback is decided on the device:
 * the residual kernel sets a flag that gates the plain substeps from the
 * saved start state, so no value is read back per step and gc_EngineStepGraph
 * replays modes 1 and 2 from its graph.  It steps eagerly (as
 * gc_EngineStep) while samples are being recorded, and in mode 2 with the
 * three-pass membrane variants, whose fallback is decided on the host. */
#ifndef GPUCELL_GPUCELL_H
typedef struct gc_engine gc_engine;
#endif
int gc_EngineSetMembraneMode(gc_engine* e, int mode, int k, gc_mlp* mlp, const float* norm,
                             float threshold);
/* Mode 1 with k = 8: every step appends one row per membrane vertex to the
 * device buffers x[capacity][36] (raw features) and y[capacity][24] (raw
 * targets: the 8 substep displacements minus the ballistic part s dts v0,
 * in the vertex frame, in units of its rest edge length) until capacity
 * rows are written.  x = NULL stops recording. */
int gc_EngineRecordSamples(gc_engine* e, float* x, float* y, int capacity);
typedef struct {
    long steps;          /* membrane steps in mode 1/2 */
    long fallbacks;      /* surrogate steps recomputed by plain substepping */
    long samples;        /* rows recorded */
    float last_residual; /* R of the last surrogate step */
    float plain_residual;/* R of the last plain 8-substep trajectory (the
                            same formula on the exact trajectory: the
                            scale a threshold is set against) */
} gc_membrane_stats;
int gc_EngineMembraneStats(gc_engine* e, gc_membrane_stats* out);

#ifdef __cplusplus
} /* extern "C" */
#endif

#endif /* GPUCELL_SURROGATE_H */
