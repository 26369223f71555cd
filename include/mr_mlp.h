/* Warm-start MLP: training entry points (C ABI, device pointers, no torch
 * types). The network maps per-contact features
 *   [normal velocity, penetration depth, |J_n|, |J_u|, |J_v|, lambda_n, lambda_u, lambda_v]
 * (all at the start of a microstep's cone solve; lambda is the previous-step
 * warm start) to the initial impulse of that solve, as a correction of the
 * previous-step impulse:
 *   8 -> 256 (GELU) -> 256 (GELU) -> 3, BF16 weights on the matrix cores,
 *   FP32 accumulation; initial impulse = lambda + out_scale * y_hat, projected
 *   onto the contact's cone. Trained on the impulses the plain solve converges
 *   to (targets: converged impulse minus lambda).
 *
 * This subsystem has no counterpart in the reference: it is the learned
 * warm start that stands beside the plain temporalCone solve, which stays
 * its fallback: a microstep whose MLP-warm-started sweeps miss the residual
 * tolerance is solved again from the plain warm start. The stepping side is enabled per world through
 * mr_world_set_warm_start_mlp (mr_c_api.h).
 *
 * All pointers are device pointers (HIP); `stream` is a hipStream_t or
 * null for the default stream. Functions return 0 on success and -1 with
 * mr_last_error() set on failure, like the rest of the C ABI. Weight
 * matrices are out-major ([out][in]) BF16 stored as uint16_t. */
#ifndef MR_MLP_H
#define MR_MLP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define MR_MLP_INPUTS 8u
#define MR_MLP_HIDDEN 256u
#define MR_MLP_OUTPUTS 3u

typedef struct MRMlpWeightsC {
    const uint16_t* w1;      /* [256][8]   bf16 */
    const float* b1;         /* [256] */
    const uint16_t* w2;      /* [256][256] bf16 */
    const float* b2;         /* [256] */
    const uint16_t* w3;      /* [3][256]   bf16 */
    const float* b3;         /* [3] */
    const float* in_shift;   /* [8] feature mean */
    const float* in_scale;   /* [8] 1 / feature std */
    const float* out_scale;  /* [3] target std */
    /* In-solver use only: a contact's correction is applied when its largest
     * component exceeds this (impulse units); smaller ones keep the plain warm
     * start, which is already as close as the network's precision. */
    float min_correction;
    /* In-solver use only: a microstep whose learned solve reaches the residual
     * tolerance but leaves some contact's normal residual above this is solved
     * again from the plain warm start, like one that misses the tolerance
     * (0 disables the check). */
    float max_normal_residual;
} MRMlpWeightsC;

/* Activations a training forward pass records for the backward pass. */
typedef struct MRMlpTraceC {
    uint16_t* xq;  /* [rows][8]   normalized input, bf16 */
    float* z1;     /* [rows][256] */
    uint16_t* h1;  /* [rows][256] bf16 */
    float* z2;     /* [rows][256] */
    uint16_t* h2;  /* [rows][256] bf16 */
} MRMlpTraceC;

typedef struct MRMlpGradsC {
    float* w1;  /* [256][8] */
    float* b1;  /* [256] */
    float* w2;  /* [256][256] */
    float* b2;  /* [256] */
    float* w3;  /* [3][256] */
    float* b3;  /* [3] */
} MRMlpGradsC;

/* y_hat[rows][3] of features x[rows][8]; trace may be null (inference). */
int mr_mlp_forward(const MRMlpWeightsC* weights, const float* x, uint32_t rows, float* y_hat,
                   const MRMlpTraceC* trace, void* stream);

/* Floats of device scratch mr_mlp_loss needs for `rows` rows. */
uint32_t mr_mlp_loss_scratch_floats(uint32_t rows);

/* Loss in normalized target units, d = y_hat - target / out_scale: with
 * huber_delta = 0 the mean squared error mean(d^2); with huber_delta > 0 the
 * Huber loss (d^2 / 2 within delta, linear beyond). loss is a device scalar,
 * dy = d loss / d y_hat. */
int mr_mlp_loss(const float* y_hat, const float* target, const float* out_scale, uint32_t rows, float huber_delta,
                float* dy, float* loss, float* scratch, void* stream);

/* Gradients of the loss with respect to every weight and bias, from dy and a
 * trace of the same rows; dz2 and dz1 are [rows][256] FP32 scratch. */
int mr_mlp_backward(const MRMlpWeightsC* weights, const MRMlpTraceC* trace, const float* dy, uint32_t rows,
                    float* dz2, float* dz1, const MRMlpGradsC* grads, void* stream);

/* One Adam step (bias-corrected, step >= 1) on `count` FP32 master values;
 * param_bf16 (may be null) receives the rounded copy the matrix cores read. */
int mr_mlp_adam(float* param, uint16_t* param_bf16, const float* grad, float* m1, float* m2, uint32_t count,
                float lr, float beta1, float beta2, float eps, uint32_t step, void* stream);

#ifdef __cplusplus
}
#endif

#endif /* MR_MLP_H */
