// Per-world step hash on the device: the byte stream of World::step_hash
// (crates/riftphys-world/src/world/provenance.rs:6-24 + hash_carried world.rs:882-909) is laid
// out by prefix sums, each 1 KiB chunk is compressed by one lane, and the chaining values are
// merged pairwise, level by level, into BLAKE3's left-complete tree (BLAKE3 specification,
// section 2.1: a pairwise merge that carries an odd last node up a level builds the same tree).
#include "rp_dev.h"
#include "rp_prims.h"

struct HashDev {
    uint32_t *keyflag;    // cap_warm + 1: warm-key group starts, then exclusive scan
    uint32_t *w_chunk;    // nw + 1: chunks per world, then exclusive scan
    uint32_t *w_len;      // nw: stream bytes per world
    uint64_t *w_epoch, *w_rng;
    uint8_t *stream;      // chunk-aligned world streams
    uint32_t *cv;         // 8 u32 per chunk
    uint32_t *cv2;        // merge scratch
    uint8_t *digest;      // 32 per world
    uint32_t cap_chunks;
    uint32_t *w_jbytes;   // nw: bytes of the world's warm-joint entries
    uint32_t *j_off;      // nj + 1: bytes of each joint's entry, then their exclusive scan
};

static __constant__ uint32_t B3_IV[8] = {0x6A09E667u, 0xBB67AE85u, 0x3C6EF372u, 0xA54FF53Au,
                                        0x510E527Fu, 0x9B05688Cu, 0x1F83D9ABu, 0x5BE0CD19u};
static __constant__ uint8_t B3_PERM[16] = {2, 6, 3, 10, 7, 0, 4, 13, 1, 11, 12, 5, 9, 14, 15, 8};
enum { B3_START = 1, B3_END = 2, B3_PARENT = 4, B3_ROOT = 8 };

static __device__ __forceinline__ uint32_t rotr(uint32_t x, int n) { return (x >> n) | (x << (32 - n)); }
static __device__ __forceinline__ void b3g(uint32_t *s, int a, int b, int c, int d, uint32_t x, uint32_t y) {
    s[a] = s[a] + s[b] + x; s[d] = rotr(s[d] ^ s[a], 16);
    s[c] = s[c] + s[d];     s[b] = rotr(s[b] ^ s[c], 12);
    s[a] = s[a] + s[b] + y; s[d] = rotr(s[d] ^ s[a], 8);
    s[c] = s[c] + s[d];     s[b] = rotr(s[b] ^ s[c], 7);
}
static __device__ void b3_compress(const uint32_t *cv, const uint32_t *blk, uint64_t ctr, uint32_t len, uint32_t flags, uint32_t *out) {
    uint32_t s[16], m[16], t[16];
    for (int i = 0; i < 8; i++) s[i] = cv[i];
    for (int i = 0; i < 4; i++) s[8 + i] = B3_IV[i];
    s[12] = (uint32_t)ctr; s[13] = (uint32_t)(ctr >> 32); s[14] = len; s[15] = flags;
    for (int i = 0; i < 16; i++) m[i] = blk[i];
    for (int r = 0; r < 7; r++) {
        b3g(s, 0, 4, 8, 12, m[0], m[1]);  b3g(s, 1, 5, 9, 13, m[2], m[3]);
        b3g(s, 2, 6, 10, 14, m[4], m[5]); b3g(s, 3, 7, 11, 15, m[6], m[7]);
        b3g(s, 0, 5, 10, 15, m[8], m[9]); b3g(s, 1, 6, 11, 12, m[10], m[11]);
        b3g(s, 2, 7, 8, 13, m[12], m[13]); b3g(s, 3, 4, 9, 14, m[14], m[15]);
        if (r < 6) { for (int i = 0; i < 16; i++) t[i] = m[B3_PERM[i]]; for (int i = 0; i < 16; i++) m[i] = t[i]; }
    }
    for (int i = 0; i < 8; i++) { out[i] = s[i] ^ s[i + 8]; out[i + 8] = s[i + 8] ^ cv[i]; }
}

#define GS(i, n) for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < (n); i += gridDim.x * blockDim.x)

static __device__ __forceinline__ void put_u32(uint8_t *p, uint32_t v) { p[0] = (uint8_t)v; p[1] = (uint8_t)(v >> 8); p[2] = (uint8_t)(v >> 16); p[3] = (uint8_t)(v >> 24); }
static __device__ __forceinline__ void put_u64(uint8_t *p, uint64_t v) { put_u32(p, (uint32_t)v); put_u32(p + 4, (uint32_t)(v >> 32)); }
// hash.rs:28-34 canon_f32_bits
static __device__ __forceinline__ uint32_t canon(float f) {
    if (!isfinite(f)) return 0x7fc00000u;
    uint32_t b = __float_as_uint(f);
    return b == 0x80000000u ? 0u : b;
}
static __device__ __forceinline__ void put_f(uint8_t *p, float f) { put_u32(p, canon(f)); }

__global__ void k_hash_keyflags(Dev d, HashDev h, uint32_t cur) {
    const uint32_t tot = d.wc_off[cur][d.nw];
    GS(e, d.cap_warm) {
        uint32_t f = 0;
        if (e < tot) {
            uint32_t w = 0, lo = 0, hi = d.nw;
            while (lo < hi) { uint32_t mid = (lo + hi) >> 1; if (d.wc_off[cur][mid] <= e) lo = mid + 1; else hi = mid; }
            w = lo - 1;
            f = (e == d.wc_off[cur][w] || d.wk_key[cur][e] != d.wk_key[cur][e - 1]) ? 1u : 0u;
        }
        h.keyflag[e] = f;
    }
}

// stream layout per world: header 48, bodies 56 n, warm 8 + 20 K + 28 E, warm patches 8 + 32 P,
// warm joints 8 + JB, rotors 8 (none), last_lin 8 + 12 n, awake 8 + cap, sleep ticks 8 + 2 cap
// carried patch entries of world w (world.rs:896-897: key, jt, tw)
static __device__ __forceinline__ uint32_t world_patches_n(const Dev &d, uint32_t w, uint32_t cur) {
    return d.cm_on ? d.wpc_off[cur][w + 1] - d.wpc_off[cur][w] : 0u;
}
static __device__ __forceinline__ uint32_t world_len(const Dev &d, const HashDev &h, uint32_t w, uint32_t cur) {
    uint32_t n = d.w_body_off[w + 1] - d.w_body_off[w], cap = d.w_cap[w];
    uint32_t e0 = d.wc_off[cur][w], e1 = d.wc_off[cur][w + 1];
    uint32_t K = h.keyflag[e1] - h.keyflag[e0], E = e1 - e0;
    return 48u + 56u * n + 8u + 20u * K + 28u * E + 24u + 32u * world_patches_n(d, w, cur) + h.w_jbytes[w] + 8u + 12u * n + 8u + cap + 8u +
           2u * cap;
}
// A stored joint's entry (world.rs:896-900): index u32, row count u32, then (tag u8, impulse f32)
// per row, rows in axis order with tags axis << 3.
static __device__ __forceinline__ uint32_t joint_entry_bytes(uint32_t m) { return (m >> 31) ? 8u + 5u * (uint32_t)__popc(m & 7u) : 0u; }
// whether joint q has stored rows, and how many (a ball joint's mask bits, a D6 joint's count word).
// Both words are loaded and the values selected: a branch between the two loads has been compiled
// into one load through a base pointer chosen for the whole wavefront.
static __device__ __forceinline__ bool joint_stored(const Dev &d, uint32_t q, uint32_t &n) {
    const uint32_t m = d.wj_mask[q];
    const uint32_t wn = d.d6_on ? d.d6_wn[q] : 0u;
    const bool d6 = d.d6_on && d.jt_kind[q] == RP_JK_D6;
    n = d6 ? (wn & 0xFFu) : (uint32_t)__popc(m & 7u);
    return ((d6 ? wn : m) >> 31) != 0u;
}

__global__ void k_hash_jbytes(Dev d, HashDev h) {
    GS(q, d.nj) {
        uint32_t n;
        h.j_off[q] = joint_stored(d, q, n) ? 8u + 5u * n : 0u;
    }
}

__global__ void k_hash_sizes(Dev d, HashDev h, uint32_t cur) {
    GS(w, d.nw) {
        h.w_jbytes[w] = d.nj ? h.j_off[d.w_joint_off[w + 1]] - h.j_off[d.w_joint_off[w]] : 0u;
        uint32_t len = world_len(d, h, w, cur);
        h.w_len[w] = len;
        h.w_chunk[w] = (len + 1023u) / 1024u;
    }
}

__global__ void k_hash_write(Dev d, HashDev h, uint32_t cur) {
    const uint32_t tot_warm = d.wc_off[cur][d.nw];
    // header + bodies + last_lin + awake/ticks of live bodies
    GS(b, d.nb) {
        uint32_t w = d.bworld[b], i = b - d.w_body_off[w];
        uint32_t n = d.w_body_off[w + 1] - d.w_body_off[w], cap = d.w_cap[w];
        uint8_t *S = h.stream + (size_t)h.w_chunk[w] * 1024u;
        uint32_t e0 = d.wc_off[cur][w], e1 = d.wc_off[cur][w + 1];
        uint32_t K = h.keyflag[e1] - h.keyflag[e0], E = e1 - e0;
        if (i == 0) {
            put_u64(S, h.w_epoch[w]);
            put_u64(S + 8, h.w_rng[w]);
            // BLAKE3 of the empty schedule (viz.rs:13, schedule.rs:14-20)
            const uint8_t empty[32] = {0xaf, 0x13, 0x49, 0xb9, 0xf5, 0xf9, 0xa1, 0xa6, 0xa0, 0x40, 0x4d, 0xea, 0x36, 0xdc, 0xc9, 0x49,
                                       0x9b, 0xcb, 0x25, 0xc9, 0xad, 0xc1, 0x12, 0xb7, 0xcc, 0x9a, 0x93, 0xca, 0xe4, 0x1f, 0x32, 0x62};
            for (int k = 0; k < 32; k++) S[16 + k] = empty[k];
            uint32_t o = 48u + 56u * n;
            put_u64(S + o, K);
            uint32_t z = o + 8u + 20u * K + 28u * E;
            const uint32_t JB = h.w_jbytes[w], P = world_patches_n(d, w, cur), PB = 32u * P;
            put_u64(S + z, P);  // the patch entries themselves are written per entry below
            uint32_t nstored = 0;  // joints with an entry (8 bytes or more); the entries are written per joint below
            for (uint32_t q = d.w_joint_off[w]; q < d.w_joint_off[w + 1]; q++) nstored += h.j_off[q + 1] != h.j_off[q] ? 1u : 0u;
            put_u64(S + z + 8 + PB, nstored);
            put_u64(S + z + 16 + PB + JB, 0);
            put_u64(S + z + 24 + PB + JB, n);
            uint32_t ao = z + 32u + PB + JB + 12u * n;
            put_u64(S + ao, cap);
            put_u64(S + ao + 8u + cap, cap);
        }
        uint8_t *p = S + 48u + 56u * i;
        float4 ps = d.pos[b], q = d.rot[b], l = d.lin[b], a = d.ang[b];
        put_u32(p, i);
        put_f(p + 4, ps.x); put_f(p + 8, ps.y); put_f(p + 12, ps.z);
        put_f(p + 16, q.x); put_f(p + 20, q.y); put_f(p + 24, q.z); put_f(p + 28, q.w);
        put_f(p + 32, l.x); put_f(p + 36, l.y); put_f(p + 40, l.z);
        put_f(p + 44, a.x); put_f(p + 48, a.y); put_f(p + 52, a.z);
        uint32_t z = 48u + 56u * n + 8u + 20u * K + 28u * E + 32u * world_patches_n(d, w, cur) + h.w_jbytes[w];
        uint8_t *ll = S + z + 32u + 12u * i;
        float4 lv = d.last_lin[b];
        put_f(ll, lv.x); put_f(ll + 4, lv.y); put_f(ll + 8, lv.z);
        uint32_t ao = z + 32u + 12u * n;
        S[ao + 8u + i] = (uint8_t)d.awake[b];
        uint32_t so = ao + 8u + cap + 8u + 2u * i;
        uint32_t t = d.sticks[b];
        S[so] = (uint8_t)t; S[so + 1] = (uint8_t)(t >> 8);
    }
    // warm-joint entries, one lane per joint, at the joint's offset within its world's entries
    GS(q, d.nj) {
        const uint32_t m = d.wj_mask[q];
        uint32_t nrows;
        if (!joint_stored(d, q, nrows)) continue;
        const uint32_t w = d.bworld[d.jt_a[q]], q0 = d.w_joint_off[w];
        const uint32_t n = d.w_body_off[w + 1] - d.w_body_off[w];
        const uint32_t e0 = d.wc_off[cur][w], e1 = d.wc_off[cur][w + 1];
        const uint32_t K = h.keyflag[e1] - h.keyflag[e0], E = e1 - e0;
        uint8_t *jp = h.stream + (size_t)h.w_chunk[w] * 1024u + 48u + 56u * n + 8u + 20u * K + 28u * E + 16u + 32u * world_patches_n(d, w, cur) +
                      (h.j_off[q] - h.j_off[q0]);
        put_u32(jp, q - q0); put_u32(jp + 4, nrows);
        jp += 8;
        if (d.d6_on && d.jt_kind[q] == RP_JK_D6) {
            for (uint32_t r = 0; r < nrows; r++) {
                jp[0] = (uint8_t)d.d6_wtag[32 * q + r];
                put_u32(jp + 1, canon(d.d6_wlam[32 * q + r]));
                jp += 5;
            }
            continue;
        }
        for (uint32_t i = 0; i < 3; i++) {
            if (!((m >> i) & 1u)) continue;
            jp[0] = (uint8_t)(i << 3);
            put_u32(jp + 1, canon(d.wj_lam[3 * q + i]));
            jp += 5;
        }
    }
    // carried patch entries, key order: key (4 u32), jt (3 f32), tw
    if (d.cm_on) {
        GS(e, d.wpc_off[cur][d.nw]) {
            uint32_t lo = 0, hi = d.nw;
            while (lo < hi) { uint32_t mid = (lo + hi) >> 1; if (d.wpc_off[cur][mid] <= e) lo = mid + 1; else hi = mid; }
            const uint32_t w = lo - 1;
            const uint32_t n = d.w_body_off[w + 1] - d.w_body_off[w];
            const uint32_t e0 = d.wc_off[cur][w], e1 = d.wc_off[cur][w + 1];
            const uint32_t K = h.keyflag[e1] - h.keyflag[e0], E = e1 - e0;
            uint8_t *p = h.stream + (size_t)h.w_chunk[w] * 1024u + 48u + 56u * n + 8u + 20u * K + 28u * E + 8u + 32u * (e - d.wpc_off[cur][w]);
            const uint64_t k = d.wp_key[cur][e];
            const float4 jt = d.wp_jt[cur][e];
            put_u32(p, (uint32_t)(k >> 32)); put_u32(p + 4, (uint32_t)k); put_u32(p + 8, 0); put_u32(p + 12, 0);
            put_f(p + 16, jt.x); put_f(p + 20, jt.y); put_f(p + 24, jt.z); put_f(p + 28, jt.w);
        }
    }
    // warm entries (key headers written by the group's first entry)
    GS(e, tot_warm) {
        uint32_t lo = 0, hi = d.nw;
        while (lo < hi) { uint32_t mid = (lo + hi) >> 1; if (d.wc_off[cur][mid] <= e) lo = mid + 1; else hi = mid; }
        uint32_t w = lo - 1;
        uint32_t n = d.w_body_off[w + 1] - d.w_body_off[w];
        uint8_t *S = h.stream + (size_t)h.w_chunk[w] * 1024u;
        uint32_t e0 = d.wc_off[cur][w], e1 = d.wc_off[cur][w + 1];
        uint32_t j = e - e0;
        uint32_t g = h.keyflag[e] - h.keyflag[e0];  // groups before this entry's (exclusive scan)
        bool start = (e == e0) || d.wk_key[cur][e] != d.wk_key[cur][e - 1];
        uint32_t base = 48u + 56u * n + 8u;
        if (start) {
            uint8_t *p = S + base + 20u * g + 28u * j;
            uint64_t k = d.wk_key[cur][e];
            uint32_t cnt = 1;
            while (e + cnt < e1 && d.wk_key[cur][e + cnt] == k) cnt++;
            put_u32(p, (uint32_t)(k >> 32)); put_u32(p + 4, (uint32_t)k); put_u32(p + 8, 0); put_u32(p + 12, 0);
            put_u32(p + 16, cnt);
        }
        uint32_t gi = start ? g : g - 1;  // index of this entry's group
        uint8_t *p = S + base + 20u * (gi + 1u) + 28u * j;
        float4 la = d.wk_la[cur][e], jt = d.wk_jt[cur][e];
        put_f(p, la.x); put_f(p + 4, la.y); put_f(p + 8, la.z);
        put_f(p + 12, d.wk_jn[cur][e]);
        put_f(p + 16, jt.x); put_f(p + 20, jt.y); put_f(p + 24, jt.z);
    }
}

// Sleep slots past the live bodies (capacity > bodies): awake 1, ticks 0 (sleep.rs:15-24).
__global__ void k_hash_tail(Dev d, HashDev h, uint32_t cur, uint32_t max_extra) {
    GS(x, max_extra * d.nw) {
        uint32_t w = x / max_extra, k = x % max_extra;
        uint32_t n = d.w_body_off[w + 1] - d.w_body_off[w], cap = d.w_cap[w];
        if (n + k >= cap) continue;
        uint8_t *S = h.stream + (size_t)h.w_chunk[w] * 1024u;
        uint32_t e0 = d.wc_off[cur][w], e1 = d.wc_off[cur][w + 1];
        uint32_t K = h.keyflag[e1] - h.keyflag[e0], E = e1 - e0;
        uint32_t z = 48u + 56u * n + 8u + 20u * K + 28u * E + 32u * world_patches_n(d, w, cur) + h.w_jbytes[w];
        uint32_t ao = z + 32u + 12u * n;
        uint32_t i = n + k;
        S[ao + 8u + i] = 1;
        uint32_t so = ao + 8u + cap + 8u + 2u * i;
        S[so] = 0; S[so + 1] = 0;
    }
}

// One lane per 1 KiB chunk.  A world with one chunk finishes here with the ROOT flag.
__global__ void k_hash_chunks(Dev d, HashDev h) {
    const uint32_t nch = h.w_chunk[d.nw];
    GS(g, nch) {
        uint32_t lo = 0, hi = d.nw;
        while (lo < hi) { uint32_t mid = (lo + hi) >> 1; if (h.w_chunk[mid] <= g) lo = mid + 1; else hi = mid; }
        uint32_t w = lo - 1;
        uint32_t local = g - h.w_chunk[w];
        uint32_t nw_ch = h.w_chunk[w + 1] - h.w_chunk[w];
        uint32_t len = h.w_len[w];
        uint32_t clen = len - local * 1024u;
        if (clen > 1024u) clen = 1024u;
        const uint8_t *src = h.stream + (size_t)g * 1024u;
        uint32_t cv[8], out[16], blk[16];
        for (int i = 0; i < 8; i++) cv[i] = B3_IV[i];
        uint32_t nblocks = (clen + 63u) / 64u;
        if (nblocks == 0) nblocks = 1;
        for (uint32_t b = 0; b < nblocks; b++) {
            uint32_t blen = clen - b * 64u;
            if (blen > 64u) blen = 64u;
            for (int i = 0; i < 16; i++) {
                uint32_t v = 0;
                for (int k = 0; k < 4; k++) {
                    uint32_t off = b * 64u + (uint32_t)(4 * i + k);
                    if (off < b * 64u + blen) v |= (uint32_t)src[off] << (8 * k);
                }
                blk[i] = v;
            }
            uint32_t flags = (b == 0 ? B3_START : 0u) | (b + 1 == nblocks ? B3_END : 0u);
            if (b + 1 == nblocks && nw_ch == 1) flags |= B3_ROOT;
            b3_compress(cv, blk, local, blen, flags, out);
            for (int i = 0; i < 8; i++) cv[i] = out[i];
        }
        for (int i = 0; i < 8; i++) h.cv[8 * g + i] = cv[i];
        if (nw_ch == 1)
            for (int i = 0; i < 8; i++) {
                h.digest[32 * w + 4 * i] = (uint8_t)cv[i]; h.digest[32 * w + 4 * i + 1] = (uint8_t)(cv[i] >> 8);
                h.digest[32 * w + 4 * i + 2] = (uint8_t)(cv[i] >> 16); h.digest[32 * w + 4 * i + 3] = (uint8_t)(cv[i] >> 24);
            }
    }
}

// Pairwise merge, one workgroup per world; the last merge of two nodes is the root.
__global__ void __launch_bounds__(256) k_hash_tree(Dev d, HashDev h) {
    const uint32_t w = blockIdx.x;
    const uint32_t c0 = h.w_chunk[w];
    uint32_t m = h.w_chunk[w + 1] - c0;
    if (m <= 1) return;
    uint32_t *a = h.cv + 8ull * c0, *b = h.cv2 + 8ull * c0;
    while (m > 2) {
        uint32_t half = m / 2;
        for (uint32_t i = threadIdx.x; i < (m + 1) / 2; i += blockDim.x) {
            if (i < half) {
                uint32_t blk[16], out[16];
                for (int k = 0; k < 8; k++) { blk[k] = a[8 * (2 * i) + k]; blk[8 + k] = a[8 * (2 * i + 1) + k]; }
                b3_compress(B3_IV, blk, 0, 64, B3_PARENT, out);
                for (int k = 0; k < 8; k++) b[8 * i + k] = out[k];
            } else {
                for (int k = 0; k < 8; k++) b[8 * i + k] = a[8 * (m - 1) + k];
            }
        }
        __threadfence_block();
        __syncthreads();
        uint32_t *t = a; a = b; b = t;
        m = (m + 1) / 2;
    }
    if (threadIdx.x == 0) {
        uint32_t blk[16], out[16];
        for (int k = 0; k < 8; k++) { blk[k] = a[k]; blk[8 + k] = a[8 + k]; }
        b3_compress(B3_IV, blk, 0, 64, B3_PARENT | B3_ROOT, out);
        for (int i = 0; i < 8; i++) {
            h.digest[32 * w + 4 * i] = (uint8_t)out[i]; h.digest[32 * w + 4 * i + 1] = (uint8_t)(out[i] >> 8);
            h.digest[32 * w + 4 * i + 2] = (uint8_t)(out[i] >> 16); h.digest[32 * w + 4 * i + 3] = (uint8_t)(out[i] >> 24);
        }
    }
}

__global__ void k_hash_store_total(uint32_t *a, uint32_t n, const uint32_t *total) { if (threadIdx.x == 0 && blockIdx.x == 0) a[n] = *total; }

static inline uint32_t gfor(uint32_t n, uint32_t b = 256, uint32_t mx = 8192) {
    uint32_t g = (n + b - 1) / b; if (g < 1) g = 1; if (g > mx) g = mx; return g;
}
static void hscan(uint32_t *a, uint32_t n, uint32_t *tmp, uint32_t *total, hipStream_t st) {
    uint32_t nb = (n + RP_SCAN_BLOCK - 1) / RP_SCAN_BLOCK;
    if (nb < 1) nb = 1;
    hipLaunchKernelGGL(k_scan_reduce, dim3(nb), dim3(256), 0, st, a, n, tmp);
    hipLaunchKernelGGL(k_scan_sums, dim3(1), dim3(256), 0, st, tmp, nb, total);
    hipLaunchKernelGGL(k_scan_apply, dim3(nb), dim3(256), 0, st, a, n, tmp, a);
}

// Digests of every world of the scene into h.digest (device).  cap_chunks bounds the stream.
extern "C" void rp_launch_hash(const Dev &d, const HashDev &h, uint32_t cur, uint32_t max_extra_slots, hipStream_t st) {
    hipLaunchKernelGGL(k_hash_keyflags, dim3(gfor(d.cap_warm)), dim3(256), 0, st, d, h, cur);
    hscan(h.keyflag, d.cap_warm, d.scan_tmp, d.cnt + CT_SCRATCH, st);
    hipLaunchKernelGGL(k_hash_store_total, dim3(1), dim3(1), 0, st, h.keyflag, d.cap_warm, d.cnt + CT_SCRATCH);
    if (d.nj) {
        hipLaunchKernelGGL(k_hash_jbytes, dim3(gfor(d.nj)), dim3(256), 0, st, d, h);
        hscan(h.j_off, d.nj, d.scan_tmp, d.cnt + CT_SCRATCH, st);
        hipLaunchKernelGGL(k_hash_store_total, dim3(1), dim3(1), 0, st, h.j_off, d.nj, d.cnt + CT_SCRATCH);
    }
    hipLaunchKernelGGL(k_hash_sizes, dim3(gfor(d.nw)), dim3(256), 0, st, d, h, cur);
    hscan(h.w_chunk, d.nw, d.scan_tmp, d.cnt + CT_SCRATCH, st);
    hipLaunchKernelGGL(k_hash_store_total, dim3(1), dim3(1), 0, st, h.w_chunk, d.nw, d.cnt + CT_SCRATCH);
    hipLaunchKernelGGL(k_hash_write, dim3(gfor(d.nb > d.cap_warm ? d.nb : d.cap_warm)), dim3(256), 0, st, d, h, cur);
    if (max_extra_slots) hipLaunchKernelGGL(k_hash_tail, dim3(gfor(max_extra_slots * d.nw)), dim3(256), 0, st, d, h, cur, max_extra_slots);
    hipLaunchKernelGGL(k_hash_chunks, dim3(gfor(h.cap_chunks)), dim3(256), 0, st, d, h);
    hipLaunchKernelGGL(k_hash_tree, dim3(d.nw), dim3(256), 0, st, d, h);
}

// Raw BLAKE3 of a device buffer of len bytes (one world-free stream), for the hash tests.
extern "C" void rp_launch_blake3_raw(const Dev &d, const HashDev &h, hipStream_t st) {
    hipLaunchKernelGGL(k_hash_chunks, dim3(gfor(h.cap_chunks)), dim3(256), 0, st, d, h);
    hipLaunchKernelGGL(k_hash_tree, dim3(d.nw), dim3(256), 0, st, d, h);
}
