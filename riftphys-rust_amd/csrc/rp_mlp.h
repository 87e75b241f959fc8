// Learned PGS warm start: a 12 -> 64 -> 64 -> 3 MLP (ReLU) in BF16 with FP32 accumulation on
// gfx950 MFMA (v_mfma_f32_16x16x32_bf16).  It maps 12 per-contact features to a correction of the
// previous-frame warm start: the contact starts from (matched jn, matched jt projected on the
// current tangents) + prediction, 0 where nothing matched, and the solver still iterates to its
// residual.  A model that outputs zeros is therefore the previous-frame warm start.
//
// Gate: a contact takes the prediction only when every normalised input lies in
// [-RP_MLP_ZMAX, RP_MLP_ZMAX] and every normalised output in [-RP_MLP_YMAX, RP_MLP_YMAX]; any
// other contact (out of the training distribution, or a non-finite value) gets the previous-frame
// warm start exactly.
//
// Canonical parameter vector (f32, RP_MLP_NPARAM entries, the ABI's layout):
//   W1[64][12] b1[64] W2[64][64] b2[64] W3[3][64] b3[3] x_mean[12] x_std[12] y_scale[3]
// Inputs are normalised as (x - x_mean) / x_std; outputs are in units of y_scale.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#define RP_MLP_IN 12
#define RP_MLP_H 64
#define RP_MLP_OUT 3
#define RP_MLP_NPARAM (64 * 12 + 64 + 64 * 64 + 64 + 3 * 64 + 3 + 12 + 12 + 3)
// offsets into the canonical vector
#define RP_MLP_W1 0
#define RP_MLP_B1 (RP_MLP_W1 + 64 * 12)
#define RP_MLP_W2 (RP_MLP_B1 + 64)
#define RP_MLP_B2 (RP_MLP_W2 + 64 * 64)
#define RP_MLP_W3 (RP_MLP_B2 + 64)
#define RP_MLP_B3 (RP_MLP_W3 + 3 * 64)
#define RP_MLP_XMEAN (RP_MLP_B3 + 3)
#define RP_MLP_XSTD (RP_MLP_XMEAN + 12)
#define RP_MLP_YSCALE (RP_MLP_XSTD + 12)
#define RP_MLP_ZMAX 8.0f
#define RP_MLP_YMAX 8.0f

// Device copies used by the MFMA tiles (bf16, zero-padded): W1p[64][32], W2p[64][64] (out x in),
// W2t[64][64] (in x out), W3p[16][64], W3t[64][32] (in x out, out padded to 32).
struct MlpWeights {
    __bf16 *w1p, *w2p, *w2t, *w3p, *w3t;
    const float *p;  // canonical f32 parameters (biases, normalisation)
};
#define RP_MLP_W1P (64 * 32)
#define RP_MLP_W2P (64 * 64)
#define RP_MLP_W3P (16 * 64)
#define RP_MLP_W3T (64 * 32)
#define RP_MLP_NBF16 (RP_MLP_W1P + 2 * RP_MLP_W2P + RP_MLP_W3P + RP_MLP_W3T)

// rphys_mlp_forward_noalloc: the most rows one call takes
#define RP_MLP_NOALLOC_MAX_ROWS (1u << 30)

// Training: one workgroup of 4 waves, RP_MLP_TILES tiles of 32 contacts per wave per Adam step.
#define RP_MLP_TILES 4
#define RP_MLP_STEP_CONTACTS (4 * RP_MLP_TILES * 32)
#define RP_MLP_NW (RP_MLP_B3 + 3)  // trained parameters (weights and biases): the vector's prefix

struct TrainArgs {
    float *p;        // canonical parameters (updated in place)
    float *m, *v;    // Adam moments (RP_MLP_NW each)
    const float *feat, *targ;  // [n][12], [n][3]
    uint64_t n;
    const uint64_t *perm;  // per epoch (a, b): sample s is contact (a * s + b) mod n
    uint32_t epochs, steps_per_epoch;
    float lr, beta1, beta2, eps;
    float *loss;     // per epoch: mean loss over its steps
    float *grad;     // gradient-check mode: the first step's gradient (RP_MLP_NW), no update
    uint32_t grad_only;
};
