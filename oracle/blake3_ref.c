/* Oracle (test infrastructure only): sequential BLAKE3-256, restated from the public BLAKE3
 * specification (Aumasson, Neves, O'Connor, Wilcox-O'Hearn, 2020; section 2: compression
 * function, chunk chaining, left-complete binary tree with the ROOT flag on the final
 * compression).  The reference hashes its step state with the `blake3` crate 1.8.2
 * (crates/riftphys-core/src/hash.rs:4-22); this file is the checker's copy of that function.
 *
 * Pinned by: BLAKE3("") and BLAKE3("abc") published digests (tests/test_oracle_core.py), and by
 * cross-checking against the independent tree-parallel device implementation on multi-chunk
 * inputs.  Only tests/, smoke() and bench.py's cpu_baseline leg may load this code.
 */
#include <stdint.h>
#include <string.h>
#include "oracle.h"

static const uint32_t IV[8] = {0x6A09E667u, 0xBB67AE85u, 0x3C6EF372u, 0xA54FF53Au,
                               0x510E527Fu, 0x9B05688Cu, 0x1F83D9ABu, 0x5BE0CD19u};
static const uint8_t PERM[16] = {2, 6, 3, 10, 7, 0, 4, 13, 1, 11, 12, 5, 9, 14, 15, 8};
enum { CHUNK_START = 1, CHUNK_END = 2, PARENT = 4, ROOT = 8 };

static uint32_t rotr(uint32_t x, int n) { return (x >> n) | (x << (32 - n)); }

static void g(uint32_t *s, int a, int b, int c, int d, uint32_t mx, uint32_t my) {
    s[a] = s[a] + s[b] + mx; s[d] = rotr(s[d] ^ s[a], 16);
    s[c] = s[c] + s[d];      s[b] = rotr(s[b] ^ s[c], 12);
    s[a] = s[a] + s[b] + my; s[d] = rotr(s[d] ^ s[a], 8);
    s[c] = s[c] + s[d];      s[b] = rotr(s[b] ^ s[c], 7);
}

static void compress(const uint32_t cv[8], const uint32_t block[16], uint64_t counter,
                     uint32_t block_len, uint32_t flags, uint32_t out[16]) {
    uint32_t s[16];
    uint32_t m[16], t[16];
    memcpy(s, cv, 32);
    s[8] = IV[0]; s[9] = IV[1]; s[10] = IV[2]; s[11] = IV[3];
    s[12] = (uint32_t)counter; s[13] = (uint32_t)(counter >> 32); s[14] = block_len; s[15] = flags;
    memcpy(m, block, 64);
    for (int r = 0; r < 7; r++) {
        g(s, 0, 4, 8, 12, m[0], m[1]);  g(s, 1, 5, 9, 13, m[2], m[3]);
        g(s, 2, 6, 10, 14, m[4], m[5]); g(s, 3, 7, 11, 15, m[6], m[7]);
        g(s, 0, 5, 10, 15, m[8], m[9]); g(s, 1, 6, 11, 12, m[10], m[11]);
        g(s, 2, 7, 8, 13, m[12], m[13]); g(s, 3, 4, 9, 14, m[14], m[15]);
        if (r < 6) { for (int i = 0; i < 16; i++) t[i] = m[PERM[i]]; memcpy(m, t, 64); }
    }
    for (int i = 0; i < 8; i++) { out[i] = s[i] ^ s[i + 8]; out[i + 8] = s[i + 8] ^ cv[i]; }
}

static void load_block(const uint8_t *p, uint32_t len, uint32_t w[16]) {
    uint8_t buf[64];
    memset(buf, 0, 64);
    memcpy(buf, p, len);
    for (int i = 0; i < 16; i++)
        w[i] = (uint32_t)buf[4 * i] | ((uint32_t)buf[4 * i + 1] << 8) | ((uint32_t)buf[4 * i + 2] << 16) |
               ((uint32_t)buf[4 * i + 3] << 24);
}

void or_blake3_init(OrBlake3 *h) {
    memset(h, 0, sizeof(*h));
    memcpy(h->cv, IV, 32);
}

/* Chunk chaining value of the current (full) chunk; used when a full chunk is followed by more
 * input, so it is never the root. */
static void chunk_cv(OrBlake3 *h, uint32_t cv_out[8]) {
    uint32_t w[16], out[16];
    uint32_t flags = (h->blocks_compressed == 0 ? CHUNK_START : 0) | CHUNK_END;
    load_block(h->buf, h->buf_len, w);
    compress(h->cv, w, h->chunk_counter, h->buf_len, flags, out);
    memcpy(cv_out, out, 32);
}

static void parent_cv(const uint32_t l[8], const uint32_t r[8], uint32_t flags, uint32_t out16[16]) {
    uint32_t block[16];
    memcpy(block, l, 32);
    memcpy(block + 8, r, 32);
    compress(IV, block, 0, 64, PARENT | flags, out16);
}

static void push_cv(OrBlake3 *h, uint32_t cv[8], uint64_t total_chunks) {
    /* Merge complete subtrees: one merge per trailing zero bit of the chunk count. */
    while ((total_chunks & 1) == 0) {
        uint32_t out[16];
        h->stack_len--;
        parent_cv(h->stack[h->stack_len], cv, 0, out);
        memcpy(cv, out, 32);
        total_chunks >>= 1;
    }
    memcpy(h->stack[h->stack_len], cv, 32);
    h->stack_len++;
}

void or_blake3_update(OrBlake3 *h, const void *data, size_t len) {
    const uint8_t *p = (const uint8_t *)data;
    while (len > 0) {
        if (h->blocks_compressed == 16 - 1 && h->buf_len == 64) {
            /* unreachable: a full chunk is finalised below before more bytes arrive */
        }
        if (h->buf_len == 64) {
            if (h->blocks_compressed == 15) {
                /* chunk full and more input follows: finish it */
                uint32_t cv[8];
                chunk_cv(h, cv);
                h->chunk_counter++;
                push_cv(h, cv, h->chunk_counter);
                memcpy(h->cv, IV, 32);
                h->blocks_compressed = 0;
                h->buf_len = 0;
            } else {
                uint32_t w[16], out[16];
                load_block(h->buf, 64, w);
                compress(h->cv, w, h->chunk_counter, 64, h->blocks_compressed == 0 ? CHUNK_START : 0, out);
                memcpy(h->cv, out, 32);
                h->blocks_compressed++;
                h->buf_len = 0;
            }
        }
        uint32_t take = 64 - h->buf_len;
        if (take > len) take = (uint32_t)len;
        memcpy(h->buf + h->buf_len, p, take);
        h->buf_len += take;
        p += take;
        len -= take;
    }
}

void or_blake3_final(const OrBlake3 *h0, uint8_t out32[32]) {
    OrBlake3 h = *h0;
    uint32_t w[16], out[16];
    uint32_t flags = (h.blocks_compressed == 0 ? CHUNK_START : 0) | CHUNK_END;
    load_block(h.buf, h.buf_len, w);
    if (h.stack_len == 0) {
        compress(h.cv, w, h.chunk_counter, h.buf_len, flags | ROOT, out);
    } else {
        uint32_t cv[8];
        compress(h.cv, w, h.chunk_counter, h.buf_len, flags, out);
        memcpy(cv, out, 32);
        int i = (int)h.stack_len - 1;
        while (i > 0) {
            parent_cv(h.stack[i], cv, 0, out);
            memcpy(cv, out, 32);
            i--;
        }
        parent_cv(h.stack[0], cv, ROOT, out);
    }
    for (int i = 0; i < 8; i++) {
        out32[4 * i] = (uint8_t)out[i];
        out32[4 * i + 1] = (uint8_t)(out[i] >> 8);
        out32[4 * i + 2] = (uint8_t)(out[i] >> 16);
        out32[4 * i + 3] = (uint8_t)(out[i] >> 24);
    }
}

void or_blake3(const void *data, size_t len, uint8_t out32[32]) {
    OrBlake3 h;
    or_blake3_init(&h);
    or_blake3_update(&h, data, len);
    or_blake3_final(&h, out32);
}
