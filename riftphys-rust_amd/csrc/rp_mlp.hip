// MLP kernels of the learned warm start (layout and math in rp_mlp.h): bf16 weight preparation,
// tiled MFMA inference, and a persistent single-workgroup Adam trainer with MFMA forward and
// backward passes.  Deterministic: fixed tile order, fixed-order cross-wave reductions.
#include <stdint.h>

#include <mutex>
#include <string>

#include "../../include/riftphys_c.h"
#include "rp_mlp.h"

#define RP_LDS __attribute__((address_space(3)))
typedef __attribute__((__vector_size__(8 * sizeof(__bf16)))) __bf16 bf16x8;
typedef __attribute__((__vector_size__(4 * sizeof(float)))) float f32x4;

// D(16x16, f32) += A(16x32) B(32x16).  Lane l holds A row l%16, k = 8*(l/16)..+7; B column l%16,
// the same k; D rows 4*(l/16)+r (r = 0..3) of column l%16.
static __device__ __forceinline__ f32x4 mfma(bf16x8 a, bf16x8 b, f32x4 c) { return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0); }
static __device__ __forceinline__ f32x4 zero4() { f32x4 z = {0.0f, 0.0f, 0.0f, 0.0f}; return z; }
static __device__ __forceinline__ __bf16 bf(float x) { return (__bf16)x; }  // round to nearest even

template <class P> static __device__ __forceinline__ bf16x8 ld8(P p) { return *(const bf16x8 *)p; }
static __device__ __forceinline__ bf16x8 ld8(const RP_LDS __bf16 *p) { return *(const RP_LDS bf16x8 *)p; }
// 8 elements of a column (stride in elements) as one operand
static __device__ __forceinline__ bf16x8 col8(const RP_LDS __bf16 *p, uint32_t stride) {
    bf16x8 v;
    for (int j = 0; j < 8; j++) v[j] = p[j * stride];
    return v;
}

// ------------------------------------------------------------------ weights
// Fills the padded bf16 copies from the canonical f32 parameters (one workgroup).
template <class BP>
static __device__ void prep_weights(const float *p, BP w1p, BP w2p, BP w2t, BP w3p, BP w3t) {
    for (uint32_t i = threadIdx.x; i < 64 * 32; i += blockDim.x) {
        uint32_t o = i / 32, k = i % 32;
        w1p[i] = bf(k < 12 ? p[RP_MLP_W1 + o * 12 + k] : 0.0f);
    }
    for (uint32_t i = threadIdx.x; i < 64 * 64; i += blockDim.x) {
        uint32_t o = i / 64, k = i % 64;
        w2p[i] = bf(p[RP_MLP_W2 + o * 64 + k]);
        w2t[k * 64 + o] = bf(p[RP_MLP_W2 + o * 64 + k]);
    }
    for (uint32_t i = threadIdx.x; i < 16 * 64; i += blockDim.x) {
        uint32_t o = i / 64, k = i % 64;
        float v = o < 3 ? p[RP_MLP_W3 + o * 64 + k] : 0.0f;
        w3p[i] = bf(v);
    }
    for (uint32_t i = threadIdx.x; i < 64 * 32; i += blockDim.x) {
        uint32_t k = i / 32, o = i % 32;  // w3t[in k][out o]
        w3t[i] = bf(o < 3 ? p[RP_MLP_W3 + o * 64 + k] : 0.0f);
    }
}
__global__ void __launch_bounds__(256) k_mlp_prep(const float *p, __bf16 *w1p, __bf16 *w2p, __bf16 *w2t, __bf16 *w3p, __bf16 *w3t) {
    prep_weights(p, w1p, w2p, w2t, w3p, w3t);
}

// Normalised input operand of the tile: contact c (row l%16), features 8*(l/16)..+7.
static __device__ __forceinline__ bf16x8 input_operand(const float *feat, const float *p, uint64_t c, bool valid) {
    const uint32_t grp = (threadIdx.x & 63u) >> 4;
    bf16x8 x;
    for (int j = 0; j < 8; j++) {
        uint32_t k = 8u * grp + j;
        float v = 0.0f;
        if (valid && k < 12u) v = (feat[c * 12 + k] - p[RP_MLP_XMEAN + k]) / p[RP_MLP_XSTD + k];
        x[j] = bf(v);
    }
    return x;
}

// Forward pass of 16 contacts (one wave).  h1s / h2s: the wave's [16][64] bf16 activation tiles
// (row stride 64); returns the output block (columns 0..2 real, normalised units).
template <class WP>
static __device__ f32x4 tile_forward(bf16x8 x, WP w1p, WP w2p, WP w3p, const float *p, RP_LDS __bf16 *h1s, RP_LDS __bf16 *h2s) {
    const uint32_t l = threadIdx.x & 63u, col = l & 15u, grp = l >> 4;
    for (uint32_t nb = 0; nb < 4; nb++) {
        f32x4 d = mfma(x, ld8(w1p + (16 * nb + col) * 32 + 8 * grp), zero4());
        const float bias = p[RP_MLP_B1 + 16 * nb + col];
        for (int r = 0; r < 4; r++) h1s[(4 * grp + r) * 64 + 16 * nb + col] = bf(fmaxf(d[r] + bias, 0.0f));
    }
    __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
    f32x4 acc[4] = {zero4(), zero4(), zero4(), zero4()};
    for (uint32_t ks = 0; ks < 2; ks++) {
        bf16x8 a = ld8(h1s + col * 64 + 32 * ks + 8 * grp);
        for (uint32_t nb = 0; nb < 4; nb++) acc[nb] = mfma(a, ld8(w2p + (16 * nb + col) * 64 + 32 * ks + 8 * grp), acc[nb]);
    }
    for (uint32_t nb = 0; nb < 4; nb++) {
        const float bias = p[RP_MLP_B2 + 16 * nb + col];
        for (int r = 0; r < 4; r++) h2s[(4 * grp + r) * 64 + 16 * nb + col] = bf(fmaxf(acc[nb][r] + bias, 0.0f));
    }
    __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
    f32x4 y = zero4();
    for (uint32_t ks = 0; ks < 2; ks++) y = mfma(ld8(h2s + col * 64 + 32 * ks + 8 * grp), ld8(w3p + col * 64 + 32 * ks + 8 * grp), y);
    const float b3 = col < 3 ? p[RP_MLP_B3 + col] : 0.0f;
    for (int r = 0; r < 4; r++) y[r] += b3;
    return y;
}

// ------------------------------------------------------------------ inference
// out[c][0..2] = prediction for feat[c] (denormalised), out[c][3] = 1 when the contact passes the
// gate (inputs within RP_MLP_ZMAX, outputs within RP_MLP_YMAX, in normalised units) and 0
// otherwise, c < min(*n_ptr, cap).  Each block takes 64 contacts per iteration (one 16-contact
// tile per wave).  n_trusted (optional) gains the number of contacts that pass.
__global__ void __launch_bounds__(256) k_mlp_infer(const float *p, const __bf16 *w1p, const __bf16 *w2p, const __bf16 *w3p,
                                                   const float *feat, const uint32_t *n_ptr, uint32_t cap, float *out, uint32_t *n_trusted) {
    __shared__ __bf16 act[4][2][16 * 64];
    __shared__ uint32_t bad[4][16];
    __shared__ uint32_t block_trusted;
    const uint32_t n = n_ptr ? min(*n_ptr, cap) : cap;
    const uint32_t wave = threadIdx.x >> 6, l = threadIdx.x & 63u, col = l & 15u, grp = l >> 4;
    RP_LDS __bf16 *h1s = (RP_LDS __bf16 *)act[wave][0], *h2s = (RP_LDS __bf16 *)act[wave][1];
    RP_LDS uint32_t *bw = (RP_LDS uint32_t *)bad[wave];
    if (threadIdx.x == 0) block_trusted = 0;
    uint32_t trusted = 0;
    for (uint64_t base = (uint64_t)blockIdx.x * 64; base < n; base += (uint64_t)gridDim.x * 64) {
        const uint64_t t0 = base + 16 * wave;
        const uint64_t c = t0 + col;
        if (l < 16) bw[l] = 0u;
        __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
        // input gate: lanes of groups 0 and 1 hold features 0..7 and 8..11 of row `col`
        if (grp < 2 && c < n) {
            bool ok = true;
            for (int j = 0; j < 8; j++) {
                uint32_t k = 8u * grp + j;
                if (k < 12u) {
                    float z = (feat[c * 12 + k] - p[RP_MLP_XMEAN + k]) / p[RP_MLP_XSTD + k];
                    ok = ok && fabsf(z) <= RP_MLP_ZMAX;
                }
            }
            if (!ok) bw[col] = 1u;
        }
        bf16x8 x = input_operand(feat, p, c, c < n);
        f32x4 y = tile_forward(x, w1p, w2p, w3p, p, h1s, h2s);
        if (col < 3) {
            const float sc = p[RP_MLP_YSCALE + col];
            for (int r = 0; r < 4; r++) {
                uint64_t cr = t0 + 4 * grp + r;
                if (cr < n) out[cr * 4 + col] = y[r] * sc;
                if (!(fabsf(y[r]) <= RP_MLP_YMAX)) bw[4 * grp + r] = 1u;
            }
        }
        __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
        if (l < 16 && c < n) {
            const bool ok = bw[l] == 0u;
            out[c * 4 + 3] = ok ? 1.0f : 0.0f;
            trusted += ok ? 1u : 0u;
        }
        __syncthreads();
    }
    if (n_trusted) {
        if (trusted) atomicAdd(&block_trusted, trusted);
        __syncthreads();
        if (threadIdx.x == 0 && block_trusted) atomicAdd(n_trusted, block_trusted);
    }
}

// ------------------------------------------------------------------ inference without allocation
// The bf16 weight tiles of rphys_mlp_forward_noalloc live in this file-scope device array, one
// copy per GPU, held until the process ends.  Host code never passes `scratch` itself (on the host
// it only names the symbol): each device's address comes from hipGetSymbolAddress, taken once per
// device with that device current.
static __device__ __align__(256) __bf16 scratch[RP_MLP_NBF16];
#define RP_MAX_DEVICES 64
static std::mutex g_scratch_mu;
static __bf16 *g_scratch_ptr[RP_MAX_DEVICES];

extern "C" void rp_set_last_error(const char *msg);

// The scratch tiles of the current device (nullptr, with the error set, on failure).
static __bf16 *device_scratch() {
    int dev = -1;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= RP_MAX_DEVICES) {
        rp_set_last_error("rphys_mlp_forward_noalloc: no current device");
        return nullptr;
    }
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void *ptr = nullptr;
        hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !ptr) {
            rp_set_last_error((std::string("rphys_mlp_forward_noalloc: hipGetSymbolAddress: ") + hipGetErrorString(e)).c_str());
            return nullptr;
        }
        g_scratch_ptr[dev] = (__bf16 *)ptr;
    }
    return g_scratch_ptr[dev];
}

RPhysResult rphys_mlp_forward_noalloc(const float *params, const float *feat12, uint32_t n, float *out4, void *stream) {
    if (!params || (n && (!feat12 || !out4))) { rp_set_last_error("rphys_mlp_forward_noalloc: null arg"); return RPHYS_ERR_NULL; }
    if (n > RP_MLP_NOALLOC_MAX_ROWS) { rp_set_last_error("rphys_mlp_forward_noalloc: n exceeds 2^30 rows"); return RPHYS_ERR_BAD_ARG; }
    if (n == 0) return RPHYS_OK;
    __bf16 *s = device_scratch();
    if (!s) return RPHYS_ERR_INTERNAL;
    __bf16 *w1p = s, *w2p = w1p + RP_MLP_W1P, *w2t = w2p + RP_MLP_W2P, *w3p = w2t + RP_MLP_W2P, *w3t = w3p + RP_MLP_W3P;
    hipStream_t st = (hipStream_t)stream;
    hipLaunchKernelGGL(k_mlp_prep, dim3(1), dim3(256), 0, st, params, w1p, w2p, w2t, w3p, w3t);
    uint32_t blocks = (n + 63) / 64;
    if (blocks > 4096) blocks = 4096;
    hipLaunchKernelGGL(k_mlp_infer, dim3(blocks), dim3(256), 0, st, params, w1p, w2p, w3p, feat12, (const uint32_t *)nullptr, n, out4,
                       (uint32_t *)nullptr);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) {
        rp_set_last_error((std::string("rphys_mlp_forward_noalloc: launch: ") + hipGetErrorString(e)).c_str());
        return RPHYS_ERR_INTERNAL;
    }
    return RPHYS_OK;
}

extern "C" void rp_launch_mlp_prep(const float *p, __bf16 *w1p, __bf16 *w2p, __bf16 *w2t, __bf16 *w3p, __bf16 *w3t, hipStream_t st) {
    hipLaunchKernelGGL(k_mlp_prep, dim3(1), dim3(256), 0, st, p, w1p, w2p, w2t, w3p, w3t);
}
extern "C" void rp_launch_mlp_infer(const float *p, const __bf16 *w1p, const __bf16 *w2p, const __bf16 *w3p, const float *feat, const uint32_t *n_ptr,
                                    uint32_t cap, float *out, uint32_t *n_trusted, hipStream_t st) {
    uint32_t blocks = (cap + 63) / 64;
    if (blocks < 1) blocks = 1;
    if (blocks > 4096) blocks = 4096;
    hipLaunchKernelGGL(k_mlp_infer, dim3(blocks), dim3(256), 0, st, p, w1p, w2p, w3p, feat, n_ptr, cap, out, n_trusted);
}

// ------------------------------------------------------------------ training
// Loss per step: mean over the step's contacts of sum_o (y_o - t_o / y_scale_o)^2.  One
// workgroup of 4 waves; a step is RP_MLP_TILES tiles of 32 contacts per wave.  Sample s of epoch
// e is contact (a_e * s + b_e) mod n (a_e coprime to n): a fixed-seed shuffle.

struct TrainLds {
    __bf16 w1p[64 * 32], w2p[64 * 64], w2t[64 * 64], w3p[16 * 64], w3t[64 * 32];
    __bf16 xs[4][32 * 32], h1[4][32 * 64], h2[4][32 * 64], dy[4][32 * 32], g2[4][32 * 64], g1[4][32 * 64];
    float gsum[RP_MLP_NW];
    float bpart[4][4][64];  // per wave, per lane group: bias-gradient partials
    float lsum[4];
};

// Bias gradient of one 16-contact half: the lane's 4 rows summed, then the 4 lane groups in order.
static __device__ __forceinline__ void bias_partial(RP_LDS float *bp, f32x4 g, uint32_t unit) {
    const uint32_t grp = (threadIdx.x & 63u) >> 4;
    bp[grp * 64 + unit] = ((g[0] + g[1]) + g[2]) + g[3];
}

__global__ void __launch_bounds__(256) k_mlp_train(TrainArgs A) {
    __shared__ TrainLds S_;
    RP_LDS TrainLds &S = *(RP_LDS TrainLds *)&S_;
    const uint32_t wave = threadIdx.x >> 6, l = threadIdx.x & 63u, col = l & 15u, grp = l >> 4;
    const uint32_t B = RP_MLP_STEP_CONTACTS;
    RP_LDS __bf16 *xs = S.xs[wave], *h1 = S.h1[wave], *h2 = S.h2[wave], *dys = S.dy[wave], *g2 = S.g2[wave], *g1 = S.g1[wave];
    uint32_t t = 0;  // Adam step
    for (uint32_t e = 0; e < A.epochs; e++) {
        float eloss = 0.0f;
        const uint64_t pa = A.perm[2 * e], pb = A.perm[2 * e + 1];
        for (uint32_t st = 0; st < A.steps_per_epoch; st++) {
            prep_weights(A.p, (RP_LDS __bf16 *)S.w1p, (RP_LDS __bf16 *)S.w2p, (RP_LDS __bf16 *)S.w2t, (RP_LDS __bf16 *)S.w3p, (RP_LDS __bf16 *)S.w3t);
            if (threadIdx.x < 4) S.lsum[threadIdx.x] = 0.0f;
            __syncthreads();
            f32x4 dw1[4], dw2[16], dw3[4];
            for (int i = 0; i < 4; i++) { dw1[i] = zero4(); dw3[i] = zero4(); }
            for (int i = 0; i < 16; i++) dw2[i] = zero4();
            float db1 = 0.0f, db2 = 0.0f, db3 = 0.0f, lw = 0.0f;  // this lane's unit (l) / output (l < 3)
            for (uint32_t tile = 0; tile < RP_MLP_TILES; tile++) {
                const uint64_t s0 = (uint64_t)st * B + (uint64_t)(wave * RP_MLP_TILES + tile) * 32;
                // forward, two halves of 16 contacts; keep x, h1, h2 of all 32 rows
                f32x4 dyv[2];
                for (uint32_t half = 0; half < 2; half++) {
                    const uint64_t c = (pa * (s0 + 16 * half + col) + pb) % A.n;
                    bf16x8 x = input_operand(A.feat, A.p, c, true);
                    for (int j = 0; j < 8; j++) xs[(16 * half + col) * 32 + 8 * grp + j] = x[j];
                    f32x4 y = tile_forward(x, (const RP_LDS __bf16 *)S.w1p, (const RP_LDS __bf16 *)S.w2p, (const RP_LDS __bf16 *)S.w3p, A.p,
                                           h1 + 16 * half * 64, h2 + 16 * half * 64);
                    for (int r = 0; r < 4; r++) {
                        float d = 0.0f;
                        if (col < 3) {
                            const uint64_t cr = (pa * (s0 + 16 * half + 4 * grp + r) + pb) % A.n;
                            float tv = A.targ[cr * 3 + col] / A.p[RP_MLP_YSCALE + col];
                            float diff = y[r] - tv;
                            lw += diff * diff;
                            d = 2.0f * diff / (float)B;
                        }
                        dyv[half][r] = d;
                        dys[(16 * half + 4 * grp + r) * 32 + col] = bf(d);
                        dys[(16 * half + 4 * grp + r) * 32 + 16 + col] = bf(0.0f);
                    }
                }
                __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
                // db3 (both halves' rows in one partial)
                {
                    f32x4 dsum;
                    for (int r = 0; r < 4; r++) dsum[r] = dyv[0][r] + dyv[1][r];
                    bias_partial((RP_LDS float *)S.bpart[wave], dsum, col);
                }
                __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
                if (l < 3) db3 += ((S.bpart[wave][0][l] + S.bpart[wave][1][l]) + S.bpart[wave][2][l]) + S.bpart[wave][3][l];
                __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
                // g2 = (dy W3) * [h2 > 0]
                for (uint32_t half = 0; half < 2; half++) {
                    bf16x8 a = ld8(dys + (16 * half + col) * 32 + 8 * grp);
                    for (uint32_t nb = 0; nb < 4; nb++) {
                        f32x4 d = mfma(a, ld8((const RP_LDS __bf16 *)S.w3t + (16 * nb + col) * 32 + 8 * grp), zero4());
                        for (int r = 0; r < 4; r++) {
                            uint32_t row = 16 * half + 4 * grp + r;
                            float g = (float)h2[row * 64 + 16 * nb + col] > 0.0f ? d[r] : 0.0f;
                            d[r] = g;
                            g2[row * 64 + 16 * nb + col] = bf(g);
                        }
                        bias_partial((RP_LDS float *)S.bpart[wave], d, 16 * nb + col);
                    }
                    __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
                    db2 += ((S.bpart[wave][0][l] + S.bpart[wave][1][l]) + S.bpart[wave][2][l]) + S.bpart[wave][3][l];
                    __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
                }
                // g1 = (g2 W2) * [h1 > 0]
                for (uint32_t half = 0; half < 2; half++) {
                    f32x4 acc[4] = {zero4(), zero4(), zero4(), zero4()};
                    for (uint32_t ks = 0; ks < 2; ks++) {
                        bf16x8 a = ld8(g2 + (16 * half + col) * 64 + 32 * ks + 8 * grp);
                        for (uint32_t nb = 0; nb < 4; nb++) acc[nb] = mfma(a, ld8((const RP_LDS __bf16 *)S.w2t + (16 * nb + col) * 64 + 32 * ks + 8 * grp), acc[nb]);
                    }
                    for (uint32_t nb = 0; nb < 4; nb++) {
                        for (int r = 0; r < 4; r++) {
                            uint32_t row = 16 * half + 4 * grp + r;
                            float g = (float)h1[row * 64 + 16 * nb + col] > 0.0f ? acc[nb][r] : 0.0f;
                            acc[nb][r] = g;
                            g1[row * 64 + 16 * nb + col] = bf(g);
                        }
                        bias_partial((RP_LDS float *)S.bpart[wave], acc[nb], 16 * nb + col);
                    }
                    __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
                    db1 += ((S.bpart[wave][0][l] + S.bpart[wave][1][l]) + S.bpart[wave][2][l]) + S.bpart[wave][3][l];
                    __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
                }
                // weight gradients, contacts as K: dW[o][i] += sum_c g[c][o] a[c][i]
                for (uint32_t ib = 0; ib < 4; ib++) dw3[ib] = mfma(col8(dys + 8 * grp * 32 + col, 32), col8(h2 + 8 * grp * 64 + 16 * ib + col, 64), dw3[ib]);
                for (uint32_t ob = 0; ob < 4; ob++) {
                    bf16x8 a = col8(g2 + 8 * grp * 64 + 16 * ob + col, 64);
                    for (uint32_t ib = 0; ib < 4; ib++) dw2[ob * 4 + ib] = mfma(a, col8(h1 + 8 * grp * 64 + 16 * ib + col, 64), dw2[ob * 4 + ib]);
                    dw1[ob] = mfma(col8(g1 + 8 * grp * 64 + 16 * ob + col, 64), col8(xs + 8 * grp * 32 + col, 32), dw1[ob]);
                }
            }
            // loss of the step
            for (int off = 32; off > 0; off >>= 1) lw += __shfl_xor(lw, off);
            if (l == 0) S.lsum[wave] = lw;
            // fixed-order reduction of the four waves' gradients into gsum (canonical layout)
            for (uint32_t w = 0; w < 4; w++) {
                __syncthreads();
                if (wave != w) continue;
                const bool first = w == 0;
                for (uint32_t ob = 0; ob < 4; ob++)
                    for (int r = 0; r < 4; r++) {
                        uint32_t o = 16 * ob + 4 * grp + r;
                        if (col < 12) { uint32_t k = RP_MLP_W1 + o * 12 + col; S.gsum[k] = first ? dw1[ob][r] : S.gsum[k] + dw1[ob][r]; }
                        for (uint32_t ib = 0; ib < 4; ib++) {
                            uint32_t k = RP_MLP_W2 + o * 64 + 16 * ib + col;
                            S.gsum[k] = first ? dw2[ob * 4 + ib][r] : S.gsum[k] + dw2[ob * 4 + ib][r];
                        }
                    }
                for (uint32_t ib = 0; ib < 4; ib++)
                    for (int r = 0; r < 4; r++) {
                        uint32_t o = 4 * grp + r;
                        if (o < 3) { uint32_t k = RP_MLP_W3 + o * 64 + 16 * ib + col; S.gsum[k] = first ? dw3[ib][r] : S.gsum[k] + dw3[ib][r]; }
                    }
                { uint32_t k = RP_MLP_B1 + l; S.gsum[k] = first ? db1 : S.gsum[k] + db1; }
                { uint32_t k = RP_MLP_B2 + l; S.gsum[k] = first ? db2 : S.gsum[k] + db2; }
                if (l < 3) { uint32_t k = RP_MLP_B3 + l; S.gsum[k] = first ? db3 : S.gsum[k] + db3; }
            }
            __syncthreads();
            eloss += ((S.lsum[0] + S.lsum[1]) + S.lsum[2]) + S.lsum[3];
            if (A.grad_only) {
                for (uint32_t i = threadIdx.x; i < RP_MLP_NW; i += blockDim.x) A.grad[i] = S.gsum[i];
                if (threadIdx.x == 0) A.loss[0] = eloss / (float)B;
                return;
            }
            // Adam
            t++;
            const float c1 = 1.0f - __builtin_powf(A.beta1, (float)t), c2 = 1.0f - __builtin_powf(A.beta2, (float)t);
            for (uint32_t i = threadIdx.x; i < RP_MLP_NW; i += blockDim.x) {
                const float g = S.gsum[i];
                const float m = A.beta1 * A.m[i] + (1.0f - A.beta1) * g;
                const float v = A.beta2 * A.v[i] + (1.0f - A.beta2) * (g * g);
                A.m[i] = m;
                A.v[i] = v;
                A.p[i] -= A.lr * (m / c1) / (sqrtf(v / c2) + A.eps);
            }
            __threadfence_block();
            __syncthreads();
        }
        if (threadIdx.x == 0) A.loss[e] = eloss / ((float)A.steps_per_epoch * (float)B);
    }
}

extern "C" void rp_launch_mlp_train(const TrainArgs &a, hipStream_t st) { hipLaunchKernelGGL(k_mlp_train, dim3(1), dim3(256), 0, st, a); }
