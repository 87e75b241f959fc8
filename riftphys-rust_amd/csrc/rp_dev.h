// Device-side layout of a batch of worlds ("scene") and the per-tick scratch.  Bodies,
// colliders and contacts of all worlds are packed world-major; per-world offsets locate each
// world's segment.  State is SoA with float4 rows (xyz + pad) so one body's vector is one 16-byte
// load in the gather-heavy stages.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "rp_mlp.h"

#define RP_NONE 0xFFFFFFFFu
#define RP_TERRAIN 0xFFFFFFFFu
#define RP_MAX_WORLD_COLLIDERS 16384u  // LDS sort width of the broadphase's one-workgroup sort (larger worlds: radix sorts)

enum { BF_DYN = 1u, BF_ALIVE = 2u, BF_SLEEP_OK = 4u };
enum { FL_CONSTRAINED = 1u, FL_LOADED = 2u, FL_GROUND = 4u, FL_JOINTED = 8u, FL_JW = 16u, FL_DRIVEN = 32u };
// FL_JOINTED: in a joint (step.rs:164-173); FL_JW: written by a joint row this tick (a dynamic
// body of an active joint, asleep or not); FL_DRIVEN: in a D6 joint with a rotational velocity
// drive (its angular velocity is the velocity solve's, step.rs:353-365)
#define RP_JK_D6 2u    // jt_kind of a D6 joint (0: ball)
#define RP_JK_DIST 1u  // jt_kind of a distance joint

// Island-list entries are contact indices, or RP_JBIT | joint index for a ball joint's rows, or
// RP_PBIT | patch index for a friction patch (scenes with a contact-model pair: contact indices
// stay below RP_PBIT there).
#define RP_JBIT 0x80000000u
#define RP_PBIT 0x40000000u
// Contact-model flags of a contact (cs_soft[c].w bits): a compliant pair's contact, with an
// implicit soft row, a point of a friction patch
enum { CM_SOFT = 1u, CM_IMPLICIT = 2u, CM_PATCH_PT = 4u };
enum { CLS_SAP = 0u, CLS_OVER = 1u, CLS_INVALID = 2u };

struct HfDesc {
    uint32_t nx, nz, off, mat;
    float sx, sz, ox, oz, hmin, hmax;
    uint32_t valid, pad;
};

// Counters on device (u32 slots)
enum {
    CT_PAIRS = 0, CT_CANDS, CT_CONTACTS, CT_ISLC, CT_ISLANDS, CT_ABORT, CT_NEED_PAIRS, CT_NEED_CANDS,
    CT_NEED_CONTACTS, CT_NEED_WARM, CT_CHAIN_N, CT_WARM_TOTAL, CT_NEW_TOTAL, CT_KEEP_TOTAL, CT_SCRATCH, CT_DONE,
    CT_ISLC_GLB,  // contacts of the islands solved by the 1024-thread LDS-flag tier (this tick)
    CT_ITERS_MAX, CT_ITERS_SUM,  // PGS sweeps: most run by an island; sum of sweeps x island contacts
    CT_ISL_CONV, CT_ISLC_CONV,   // residual mode: islands (and their contacts) that reached the tolerance
    CT_WS_LEARNED,               // contacts whose learned warm start passed the gate
    CT_WIDE_N, CT_WIDE_MAX,      // islands with wide levels; most wide levels of one island
    CT_BIG_N,                    // islands of the 1024-thread tiers
    CT_WSH_MAX,                  // most leading wide levels of one island's shock order
    CT_TIER_N, CT_TIER_N1, CT_TIER_N2, CT_TIER_N3,  // islands per sweep tier
    CT_TIER_N4,                  // islands of the contact-model tier (compliant rows or patches)
    CT_ROWS,                     // island-list entries before the sort: joints + contacts
    CT_LOG_N,                    // contacts appended to the contact log this tick
    // sums over the completed ticks since the counters were last cleared
    CT_SUM_PAIRS, CT_SUM_CANDS, CT_SUM_CONTACTS, CT_SUM_ISLC, CT_SUM_ISLANDS, CT_SUM_ISLC_GLB, CT_SUM_ITERS_MAX, CT_SUM_ITERS_SUM,
    CT_SUM_ISL_CONV, CT_SUM_ISLC_CONV, CT_SUM_WS_LEARNED,
    CT_SUM_WIDE, CT_SUM_WIDE_SH,  // leading wide levels of the largest island (list, shock order)
    CT_SUM_ITERS_SUM_HI,  // high word of the CT_SUM_ITERS_SUM sum (sweeps x contacts outgrow 32 bits)
    CT_ITERS_SUM_CONV,    // residual mode: sweeps x entries of the islands that reached the tolerance
    CT_SUM_ITERS_CONV, CT_SUM_ITERS_CONV_HI,
    CT_LANE_N,            // islands of the lane tier this tick
    CT_NCOL,              // the scene's collider count (the radix broadphase order's length)
    CT_WAVE_K, CT_GRP_K, CT_LANE_K,  // with the lane tier on: where tier 0's sorted islands solved by
                                     // lane groups of four, lane pairs and single lanes begin
    CT_W6_TOI, CT_W6_SUPPORT, CT_W6_PSEUDO, CT_W6_FALLBACK,  // W6 this tick: bounce rows, support rows,
                                                             // pseudo rows, components swept
    CT_W6_RELAX,                                             // contacts bounced by the relax pass
    CT_NPATCH,            // friction patches this tick
    CT_WP_NEW, CT_WP_KEEP,  // carried patch entries stored: new ones, kept ones
    CT_COUNT
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
    uint32_t *bk_wj_mask;  // joints' carried warm rows (nj, 3 nj)
    float *bk_wj_lam;
    // per-body tick scratch
    float4 *pre_lin, *pre_ang;
    float *iiw;  // 9 per body
    uint32_t *flags, *ballf, *root;
    float4 *opos, *orot, *ppos, *prot, *slin, *sang;
    // colliders tick scratch
    float4 *amin, *amax, *smin, *smax;
    float *sweep;
    uint32_t *ccls, *bp_ord, *bp_rank;
    uint32_t bp_radix;                         // a world exceeds RP_MAX_WORLD_COLLIDERS: order by radix sorts
    uint32_t *bp_k0, *bp_v0, *bp_k1, *bp_v1;   // their buffers (nc each; allocated only then)
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
    uint32_t *rank_conv;  // residual mode: 1 where the entry's island reached the tolerance (by island-list rank)
    // sweep orders over island-list positions: chains of the list order, the shock order (contact
    // at each shock rank) and its chains, readiness flags of the largest islands
    uint32_t *ch_pa, *ch_pb, *ch_sa, *ch_sb, *sh_ord, *sh_pa, *sh_pb, *sh_sa, *sh_sb, *g_done;
    // level orders (execution slot -> island rank) of the islands larger than a workgroup, the
    // slots of the list order and of the shock order (predecessor slots, contact), and the
    // rank -> slot scratch
    uint32_t *ex_base, *ex_shock;
    uint32_t *bs_pa, *bs_pb, *bs_c, *ss_pa, *ss_pb, *ss_c, *slot_of;
    // wide levels: an island's leading levels of >= RP_WIDE_MIN ranks run as GPU-wide launches
    uint32_t *lvl_off;    // per island segment: start slot of each leading wide level, then their end
    uint32_t *isl_w;      // per island: number of leading wide levels (0: none)
    uint32_t *wide_list;  // islands with wide levels this tick (CT_WIDE_N of them)
    uint32_t *big_list;   // islands of the 1024-thread tiers this tick (CT_BIG_N of them)
    uint32_t *bp_fill;    // per collider: pairs appended to its list (broadphase write pass)
    uint32_t *tier_list[5];  // island indices of each sweep tier this tick (CT_TIER_N + tier of them)
    uint32_t *lvl_off_sh, *isl_wsh;  // as lvl_off / isl_w, for the shock order of the current shock pass
    unsigned long long *ph;  // RP_PH_COUNT sweep-phase clocks of the large-island tiers
    float4 *srow;            // RP_SROW * cap_contacts: constant rows of the current sweep order, per slot
    float4 *lrow;            // RP_LROW * cap_lane: rows and accumulators of the lane tier, in the lane layout
    size_t cap_lane;         // entries of lrow: cap_contacts + RP_LANE_PAD
    uint32_t *lane_order;    // tier 0's islands this tick by decreasing size (lane tier on)
    uint32_t *lid;           // 2 per island-list entry: LDS velocity slot of the body each side writes
    // learned warm start: per-contact features [cap][12] and predictions [cap][4], the model
    float *feat;
    float4 *pred;
    MlpWeights mlp;
    // contact log (training data): rows of features [log_cap][12] and converged impulses [log_cap][3]
    float *log_feat, *log_targ;
    uint32_t *log_n;  // [0] rows written, [1] rows dropped for lack of room (kept across ticks)
    uint32_t log_cap;
    // radix-sort ping-pong buffers (2 * cap_contacts each) and the shock keys' high bits
    uint32_t *sk0, *sv0, *sk1, *sv1, *sh_hi;
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
    // ball joints, world-major (articulation/src/bodies.rs:326-331): bodies, frames in each body's
    // space (position, rotation), this tick's rows (axis i of frame A and 1 / K in jr_u[3 q + i],
    // lever arms, bit i of jr_mask when the row exists) and their accumulators, the warm rows
    // carried between ticks (wj_mask: bit 31 stored, bits 0-2 the rows' axes; wj_lam[3 q + i]),
    // each body's jointed partners (CSR over bodies) and each world's first joint.  The island
    // lists hold contacts and joints, so cap_contacts bounds their sum.
    uint32_t nj;
    uint32_t *jt_a, *jt_b;
    float4 *jt_fap, *jt_far, *jt_fbp, *jt_fbr;
    float4 *jr_u, *jr_ra, *jr_rb;
    uint32_t *jr_mask;
    float *jr_acc;
    uint32_t *wj_mask;
    float *wj_lam;
    uint32_t *jadj_off, *jadj;
    uint32_t *w_joint_off;
    // D6 joints (allocated when the scene has one, d6_on; layout in rp_d6.h): kind and stored hinge
    // axis per joint, 18 axis words per joint, 32 rows per joint and their count and accumulators,
    // the carried rows (count word, tags, impulses) and their backups, 6 position multipliers
    uint32_t d6_on;
    uint32_t *jt_kind, *jt_hinge, *d6_nrows, *d6_wn, *d6_wtag, *bk_d6_wn, *bk_d6_wtag;
    float4 *jt_ax, *d6_row;
    float *d6_racc, *d6_wlam, *bk_d6_wlam, *d6_plam;
    // W6 restitution (solver.rs:569-588, 995-1245), allocated when a pair of the scene's materials
    // has restitution > 0 (w6_on): per world, a contact with a restitution bias ([w]) and
    // pseudo-velocities ([nw + w]); per contact with a bias (written by the prep), (bias, start gap,
    // e, time-of-impact row) and the sides' normal velocities at the step's start and after the
    // forces; the bounce passes' scratch over each world's contact segment (marks, lists, rows,
    // components); per body, a partner or component word and bits, the union-find parent, the
    // pseudo-velocities and their f64 sums; the component systems' workspace slots.
    uint32_t w6_on, w6_slots;
    uint32_t *w_w6;
    float4 *w6_toi, *w6_tv;
    uint32_t *w6_mark, *w6_rl, *w6_rc, *w6_pc, *w6_grp, *w6_comp, *w6_cst;
    float *w6_rt, *w6_racc, *w6_pu;
    uint32_t *w6_b, *w6_par;
    float4 *w6_pl, *w6_pa;
    double *w6_acc;
    void *w6_ws;
    // kinematic players (world.rs:44-69 PlayerController), world-major: each player's body,
    // (radius, height, speed, vertical velocity), input direction and grounded flag, the backups
    // of the state a tick changes (replayed ticks start again from them), each world's first
    // player, and per body the player it carries (RP_NONE); allocated when the scene has players
    uint32_t np;
    uint32_t *pl_body, *pl_grounded, *bk_pl_grounded, *w_player_off, *pl_of_body;
    float4 *pl_state, *bk_pl_state, *pl_input, *pl_bmin, *pl_bmax;
    // contact models (materials.rs:75-90), allocated when a pair of the scene has a compliant or a
    // patch-friction model (cm_on): per material pair (stiffness, dissipation, patch 0/1, 0); per
    // contact the implicit row's (gamma, bias, mass, CM_* bits) and its patch (RP_NONE: none); per
    // patch (at most one per two contacts) its first point and point count, the centroid
    // constraint (bodies, inverse masses, lever arms, normal, tangents, (m_n, m_t1, m_t2, m_tw),
    // (mu_s, mu_k, Stribeck speed, radius)), its accumulators (t1, t2, torsion) and its first
    // point's carried-cache key; per island the contact-model tier flag; the carried patch cache
    // (double buffer as the warm cache: keys, (jt, tw), per-world offsets) and its store's flags.
    uint32_t cm_on, cap_patch;
    float4 *pair_cm, *cs_soft;
    uint32_t *ct_patch, *pt_head, *pt_first, *pt_n, *pt_ai, *pt_bi, *isl_cm;
    float2 *pt_inv;
    float4 *pt_ra, *pt_rb, *pt_nn, *pt_t1, *pt_t2, *pt_m, *pt_mu, *pt_acc;
    uint64_t *pt_key;
    uint32_t *pt_soft;
    uint64_t *wp_key[2];
    float4 *wp_jt[2];
    uint32_t *wpc_off[2];
    uint32_t *wp_new, *wp_keep;
};

// A spatial query (k_query) and its answer.
struct RpQuery {
    uint32_t kind;    // 0 ray, 1 capsule sweep
    uint32_t world;   // world index in the scene
    uint32_t ignore;  // scene body index skipped, or RP_NONE
    uint32_t pad;
    float a[4], b[4];  // ray: origin, unit direction; sweep: from, to
    float max_d, radius, half_height, pad2;
};
struct RpQueryHit {
    uint32_t hit, body, started, pad;  // body: scene index, or RP_TERRAIN
    float fraction, pad2[3];
    float point[4], normal[4];
};

// Stage boundaries of one tick (events recorded by the profiled step; stage k spans marks k..k+1).
enum {
    RP_ST_PRE = 0,          // S1 forces, AABBs, swept bounds
    RP_ST_BROAD,            // SAP broadphase (sort, enumerate, scans)
    RP_ST_NARROW_PAIRS,     // body-pair narrowphase
    RP_ST_NARROW_TERRAIN,   // collider-triangle narrowphase
    RP_ST_CONTACTS,         // terrain cull, contact count/scan/write
    RP_ST_PREP,             // constraint prep, islands, island sort, warm match
    RP_ST_ORDERS,           // list-order chains, velocity shock order and its chains
    RP_ST_VEL_SMALL,        // velocity solve, islands of up to 4,096 entries (tiers 0 and 3)
    RP_ST_VEL_LARGE,        // velocity solve, islands of up to 131,072 contacts (1024 threads, LDS flags)
    RP_ST_VEL_HUGE,         // velocity solve, larger islands (1024 threads, global flags)
    RP_ST_WARM_STORE,       // warm cache store, loaded flags
    RP_ST_BALLISTIC,        // ballistic classification
    RP_ST_INTEGRATE,        // S4 integration
    RP_ST_POS_PREP,         // projection anchors
    RP_ST_PROJ_SMALL,       // projection iterations, small islands
    RP_ST_PROJ_LARGE,       // projection iterations, islands of up to 131,072 contacts
    RP_ST_PROJ_HUGE,        // projection iterations, larger islands
    RP_ST_PROJ_ORDER,       // projection shock order and its chains
    RP_ST_PROJ_SHOCK,       // projection shock pass
    RP_ST_S6,               // S6 velocity from corrections, S7 sleep
    RP_ST_END,
    RP_NMARKS
};

// Sweep phases of the large-island tiers (wall-clock ticks summed over workgroups), plus the
// level counts of their orders.
enum {
    RP_PH_VEL_ORDER = 0, RP_PH_WARM, RP_PH_PGS, RP_PH_VEL_SHOCK_ORDER, RP_PH_VEL_SHOCK, RP_PH_BALL, RP_PH_PROJ,
    RP_PH_PROJ_SHOCK_ORDER, RP_PH_PROJ_SHOCK, RP_PH_ROUNDS_BASE, RP_PH_ROUNDS_SHOCK, RP_PH_COUNT
};

#define RP_SROW 12u
#define RP_LROW 13u  // lane layout: the constant row (RP_SROW components), then (jn, jt, entry)
// lane layout room beyond the island lists' length: 64 lanes x the largest lane-tier island
#define RP_LANE_PAD (64u * 256u)
// defaults of tier 0's split with the lane tier on: islands of up to RP_LANE_MAX entries one per
// lane, then up to RP_PAIR_MAX by lane pairs (128 at most), up to RP_QUAD_MAX by groups of four
// lanes (256 at most), the larger ones a wave each
#define RP_LANE_MAX 150u
#define RP_PAIR_MAX 150u
#define RP_QUAD_MAX 150u
#define RP_WAVE_MIN 2048u  // default: at least this many of tier 0's largest islands a wave each
#define RP_LANE_POS_ITERS 4u  // the lane tier runs the projection too up to this many position iterations

// The position solve's SHAKE directions need both bodies turned by less than 0.5 rad this step:
// 2 acos(min(|w|, 1)) < 0.5 for w of q_start * conj(q_now), which (acosf, rounded as glibc does)
// holds exactly when min(|w|, 1) >= this float (0.968912482)
#define RP_SHAKE_COS_BITS 0x3f780aa6u
#define RP_WIDE_MIN 1024u         // default wide level: ranks one workgroup cannot take in one round
#define RP_WIDE_MAX_LEVELS 64u    // leading wide levels recorded per island

// Optional side streams for the island sweeps: the four size tiers of one stage solve disjoint
// islands, so they run concurrently (tier 0 on side[0], tier 1 on the tick's stream, tiers 3 and 2
// on side[1]), forked from and joined back to the tick's stream through ev (no timing).
struct TierStreams { hipStream_t side[2]; hipEvent_t ev[3]; };

struct TickParams {
    float dt;
    float gx, gy, gz;      // q6-quantised gravity acceleration
    float upx, upy, upz;   // -normalize(gravity) (or +Y)
    uint32_t iv, ip;       // velocity iterations (the most, in residual mode) / position iterations
    uint32_t cur;          // warm buffer holding last tick's cache
    uint32_t resid_on;     // residual mode: an island stops once its sweep residual <= resid_tol
    float resid_tol;
    uint32_t ws_mode;      // warm start: 0 previous frame (the reference), 1 zero, 2 learned
    uint32_t log_on;       // append (features, converged impulses) of solved contacts to the log
    uint32_t tier_max[3];  // largest island of sweep tiers 0, 3 and 1 (within their flag storage)
    uint32_t lane_max;     // largest island of the lane tier (0: no lane tier; tier 0 a wave per island)
    uint32_t pair_max;     // largest island of the lane pairs (at least lane_max, at most 128)
    uint32_t quad_max;     // largest island of the groups of four lanes (at least pair_max)
    uint32_t wave_min;     // at least this many of tier 0's largest islands solved a wave each
    uint32_t lane_pos;     // the lane tier's islands also run the ballistic and projection passes one per thread
    uint32_t wide_min;     // ranks of a leading level that make it a GPU-wide level launch
    uint32_t damp_start;   // compliant rows: damping linearised at the step-start penetration (world.rs:992)
};
