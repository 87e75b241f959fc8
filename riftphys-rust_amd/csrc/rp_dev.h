// Device-side layout of a batch of worlds ("scene") and the per-tick scratch.  Bodies,
// colliders and contacts of all worlds are packed world-major; per-world offsets locate each
// world's segment.  State is SoA with float4 rows (xyz + pad) so one body's vector is one 16-byte
// load in the gather-heavy stages.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#define RP_NONE 0xFFFFFFFFu
#define RP_TERRAIN 0xFFFFFFFFu
#define RP_MAX_WORLD_COLLIDERS 16384u  // LDS sort width of the broadphase
#define RP_ISLAND_LDS_MAX 2048u       // islands up to this many contacts keep their chains in LDS (64 KiB)
#define RP_ISLAND_MAX (1u << 20)      // position bits of the shock-order key

enum { BF_DYN = 1u, BF_ALIVE = 2u, BF_SLEEP_OK = 4u };
enum { FL_CONSTRAINED = 1u, FL_LOADED = 2u, FL_GROUND = 4u };
enum { CLS_SAP = 0u, CLS_OVER = 1u, CLS_INVALID = 2u };

struct HfDesc {
    uint32_t nx, nz, off, mat;
    float sx, sz, ox, oz, hmin, hmax;
    uint32_t valid, pad;
};

// Counters on device (u32 slots)
enum {
    CT_PAIRS = 0, CT_CANDS, CT_CONTACTS, CT_ISLC, CT_ISLANDS, CT_ABORT, CT_NEED_PAIRS, CT_NEED_CANDS,
    CT_NEED_CONTACTS, CT_NEED_WARM, CT_BIG_ISLAND, CT_WARM_TOTAL, CT_NEW_TOTAL, CT_KEEP_TOTAL, CT_SCRATCH, CT_DONE, CT_COUNT
};

struct Dev {
    uint32_t nb, nc, nw;
    uint32_t cap_pairs, cap_cands, cap_contacts, cap_warm;
    // bodies
    float4 *pos, *rot, *lin, *ang, *last_lin;
    float *inv_mass;
    uint32_t *bflags;
    float *iil;  // 9 per body, column-major
    uint32_t *awake, *sticks;
    uint32_t *btype;
    float4 *bdims;
    uint32_t *bworld;
    // colliders
    uint32_t *col_body, *col_mat, *col_world;
    // worlds
    uint32_t *w_body_off, *w_col_off, *w_cap;
    HfDesc *hf;
    float *heights;
    float4 *pair_tab;  // 16 x 16 (mu_s, mu_k, restitution, stribeck_v)
    // pre-tick backup (overflow replay)
    float4 *bk_pos, *bk_rot, *bk_lin, *bk_ang, *bk_last;
    uint32_t *bk_awake, *bk_sticks;
    // per-body tick scratch
    float4 *pre_lin, *pre_ang;
    float *iiw;  // 9 per body
    uint32_t *flags, *ballf, *root;
    float4 *opos, *orot, *ppos, *prot, *slin, *sang;
    // colliders tick scratch
    float4 *amin, *amax, *smin, *smax;
    float *sweep;
    uint32_t *ccls, *bp_ord, *bp_rank;
    uint32_t *w_nsap, *w_nover, *w_solve;
    float *w_maxext;
    uint32_t *out_off, *cand_off, *c_off;  // nc + 1 each (counts, then exclusive scans)
    // pairs
    uint32_t *pair_i, *pair_j, *pair_n;
    float4 *pc_p, *pc_n;  // 4 slots per pair
    // terrain candidates
    uint32_t *tc_n;
    float4 *tc_p, *tc_nf;  // 4 slots per candidate
    uint32_t *tc_cell;
    float4 *tcul_p, *tcul_n;
    uint32_t *tcul_cell, *tcul_cnt;
    // contacts
    uint32_t *ct_a, *ct_b, *ct_cell;
    float4 *ct_p, *ct_n;  // point + depth, normal + feat bits
    float *ct_imp;
    // constraints
    uint32_t *cs_ai, *cs_bi, *cs_wa, *cs_wb;
    float2 *cs_inv;
    float4 *cs_ra, *cs_rb, *cs_n, *cs_t1, *cs_t2, *cs_m, *cs_mu;
    float *jn, *jn_pre;
    float2 *jt;
    uint64_t *ct_key;
    float4 *ct_anchor;
    int32_t *warm_idx;
    // islands
    uint32_t *isl_k, *isl_v, *isl_k2, *isl_v2, *isl_start, *isl_flag;
    // chain scratch for islands above the LDS size (sliced by island position)
    uint64_t *g_ent;  // 4 * cap_contacts
    uint32_t *g_order, *g_pa, *g_pb, *g_done;
    uint32_t *rs_hist;
    uint32_t rs_blocks;
    // projection contacts
    float4 *pc_la, *pc_lb;
    // warm cache (double buffer)
    uint64_t *wk_key[2];
    float4 *wk_la[2];
    float *wk_jn[2];
    float4 *wk_jt[2];
    uint32_t *wc_off[2];  // nw + 1
    uint32_t *newflag;   // cap_contacts + 1: new cache entry flags, then their exclusive scan
    uint32_t *keepflag;  // cap_warm + 1: kept old entry flags, then their exclusive scan
    uint32_t *scan_tmp;
    uint32_t *cnt;  // CT_COUNT counters
};

struct TickParams {
    float dt;
    float gx, gy, gz;      // q6-quantised gravity acceleration
    float upx, upy, upz;   // -normalize(gravity) (or +Y)
    uint32_t iv, ip;       // velocity / position iterations
    uint32_t cur;          // warm buffer holding last tick's cache
};
