// The per-tick step of crates/riftphys-world/src/world/step.rs:40-420 as a chain of gfx950
// kernels over a batch of worlds.  Stage map (reference line -> kernel):
//   S1 forces step.rs:53-68                      k_bodies_pre
//   S2 AABB / sweep step.rs:71-81                k_col_aabb
//   broadphase broadphase.rs:29-99               k_bp_sort, k_bp_enum (count, write)
//   narrowphase step.rs:95-119                   k_narrow_pairs
//   terrain step.rs:121-142, terrain.rs:300-359  k_narrow_terrain, k_terrain_cull
//   cull/orient step.rs:499-542                  k_terrain_cull, k_contact_count/write
//   solver prep solver.rs:494-590                k_prep
//   warm start solver.rs:717-762                 k_warm_match + first sweep of k_sweep_vel
//   PGS solver.rs:788-880 + shock :964-993       k_sweep_vel (one workgroup per island)
//   warm store solver.rs:917-944                 k_warm_newflags / k_warm_keepflags / k_warm_merge
//   loaded + impulses solver.rs:1247-1270        k_loaded
//   S4 ballistic + integrate step.rs:154-280     k_sweep_ball, k_integrate
//   S5 projection step.rs:284-310, 423-497       k_posc_prep, k_sweep_proj
//   S6/S7 step.rs:312-384, sleep.rs:59-80        k_s6
// Ordered (Gauss-Seidel) stages run per island: constraints that write a common body keep the
// reference's order through per-body predecessor chains; constraints of different islands share
// no written body, so the result equals the reference's serial sweep bit for bit.
#include "rp_dev.h"
#include "rp_math.h"
#include "rp_narrow.h"
#include "rp_prims.h"

static __device__ __forceinline__ V3 L3(float4 v) { return mk3(v.x, v.y, v.z); }
static __device__ __forceinline__ Q4 LQ(float4 v) { return mkq(v.x, v.y, v.z, v.w); }
static __device__ __forceinline__ float4 S3(V3 v, float w = 0.0f) { return make_float4(v.x, v.y, v.z, w); }
static __device__ __forceinline__ float4 SQ(Q4 q) { return make_float4(q.x, q.y, q.z, q.w); }
static __device__ __forceinline__ uint32_t fbits(float f) { return __float_as_uint(f); }
static __device__ __forceinline__ V3 unit_axis(uint32_t i) { return mk3(i == 0 ? 1.0f : 0.0f, i == 1 ? 1.0f : 0.0f, i == 2 ? 1.0f : 0.0f); }
static __device__ __forceinline__ M3 load_m9(const float *p) {
    return mkm(mk3(p[0], p[1], p[2]), mk3(p[3], p[4], p[5]), mk3(p[6], p[7], p[8]));
}
static __device__ __forceinline__ void store_m9(float *p, M3 m) {
    p[0] = m.c0.x; p[1] = m.c0.y; p[2] = m.c0.z; p[3] = m.c1.x; p[4] = m.c1.y; p[5] = m.c1.z;
    p[6] = m.c2.x; p[7] = m.c2.y; p[8] = m.c2.z;
}
#define GRID_STRIDE(i, n) for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < (n); i += gridDim.x * blockDim.x)
static __device__ __forceinline__ bool aborted(const Dev &d) { return d.cnt[CT_ABORT] != 0; }

// dynamics.rs:158-165 inv_inertia_world = R I^-1 R^T (zero when inv_mass == 0)
static __device__ __forceinline__ M3 inv_inertia_world(const Dev &d, uint32_t b, Q4 q) {
    if (d.inv_mass[b] == 0.0f) return mzero();
    M3 r = mquat(q);
    return mmul(mmul(r, load_m9(d.iil + 9 * b)), mtrans(r));
}
// dynamics.rs:130-147 set_pose / set_vel sanitisation
static __device__ __forceinline__ void set_pose(const Dev &d, uint32_t b, V3 p, Q4 q) {
    if (fin3(p)) d.pos[b] = S3(fin_v(p));
    if (finq(q)) d.rot[b] = SQ(fin_q(q));
}
static __device__ __forceinline__ void set_vel(const Dev &d, uint32_t b, V3 l, V3 a) {
    if (fin3(l)) d.lin[b] = S3(fin_v(l));
    if (fin3(a)) d.ang[b] = S3(fin_v(a));
}

// With mu_s == mu_k the Stribeck term is (+0) * exp(..) = +0 for every slip that is not NaN (exp
// of a finite or infinite negative argument is finite), so mu = mu_k + 0 exactly: the exp is
// skipped, bit-identically (mu_k + 0.0f keeps the sum's sign of zero).
// A contact's contact-model bits (CM_*): 0 in scenes without a contact-model pair.
static __device__ __forceinline__ uint32_t cm_flags(const Dev &d, uint32_t c) { return d.cm_on ? __float_as_uint(d.cs_soft[c].w) : 0u; }
static __device__ __forceinline__ float mu_dyn(float4 mu4, float slip) {
    if (mu4.x == mu4.y && slip == slip) return mu4.y + 0.0f;
    float x = slip / mu4.z;
    return mu4.y + (mu4.x - mu4.y) * rp_expf(-(x * x));
}

// ------------------------------------------------------------------ S1
__global__ void k_bodies_pre(Dev d, TickParams tp) {
    if (aborted(d)) return;
    GRID_STRIDE(b, d.nb) {
        float4 p = d.pos[b], q = d.rot[b], l = d.lin[b], a = d.ang[b];
        if (!d.np) {  // with players the backup is taken before they move (k_player_backup)
            d.bk_pos[b] = p; d.bk_rot[b] = q; d.bk_lin[b] = l; d.bk_ang[b] = a;
            d.bk_last[b] = d.last_lin[b]; d.bk_awake[b] = d.awake[b]; d.bk_sticks[b] = d.sticks[b];
        }
        d.pre_lin[b] = l; d.pre_ang[b] = a;
        d.flags[b] = 0; d.root[b] = b; d.ballf[b] = 0;
        d.slin[b] = make_float4(0, 0, 0, 0); d.sang[b] = make_float4(0, 0, 0, 0);
        bool dyn = d.bflags[b] & BF_DYN;
        if (dyn && d.awake[b]) {
            V3 g = fin3(L3(p)) ? mk3(tp.gx, tp.gy, tp.gz) : splat3(0.0f);
            set_vel(d, b, L3(l) + g * tp.dt, L3(a));
        }
        store_m9(d.iiw + 9 * b, inv_inertia_world(d, b, LQ(q)));
    }
    // the joints' carried warm rows: the velocity solve rewrites them in place
    GRID_STRIDE(q, d.nj) {
        d.bk_wj_mask[q] = d.wj_mask[q];
        for (uint32_t i = 0; i < 3; i++) d.bk_wj_lam[3 * q + i] = d.wj_lam[3 * q + i];
        if (d.d6_on) {
            d.bk_d6_wn[q] = d.d6_wn[q];
            for (uint32_t r = 0; r < 32; r++) { d.bk_d6_wtag[32 * q + r] = d.d6_wtag[32 * q + r]; d.bk_d6_wlam[32 * q + r] = d.d6_wlam[32 * q + r]; }
        }
    }
}

// ------------------------------------------------------------------ S2 AABBs
// shape.rs:18-76 aabb_of + aabb.rs:21-31 (qf32 then min/max)
static __device__ void aabb_of(uint32_t type, float4 dims, V3 p, Q4 q, V3 &mn, V3 &mx) {
    V3 he;
    if (type == 0) he = splat3(fabsf(dims.x));
    else if (type == 1) {
        V3 e = mk3(fabsf(dims.x), fabsf(dims.y), fabsf(dims.z));
        M3 r = mquat(q);
        he = mv(mkm(v3abs(r.c0), v3abs(r.c1), v3abs(r.c2)), e);
    } else {
        V3 axw = qrot(q, mk3(0, 1, 0)) * fabsf(dims.y);
        he = v3abs(axw) + splat3(fabsf(dims.x));
    }
    V3 a = p - he, b = p + he;
    V3 qa = mk3(qf32(a.x), qf32(a.y), qf32(a.z)), qb = mk3(qf32(b.x), qf32(b.y), qf32(b.z));
    mn = v3min(qa, qb); mx = v3max(qa, qb);
}

// terrain.rs:300-319 cell range of a collider's (unswept) AABB; false if culled
static __device__ bool hf_range(const HfDesc &h, V3 mn, V3 mx, int &gx0, int &gx1, int &gz0, int &gz1) {
    if (!h.valid) return false;
    int nx = (int)h.nx, nz = (int)h.nz;
    if (nx < 2 || nz < 2) return false;
    float lminx = mn.x - h.ox, lmaxx = mx.x - h.ox, lminz = mn.z - h.oz, lmaxz = mx.z - h.oz;
    // `as i32` saturates; NaN becomes 0
    auto sat = [](float f) -> int { if (f != f) return 0; if (f >= 2147483647.0f) return 2147483647; if (f <= -2147483648.0f) return (-2147483647 - 1); return (int)f; };
    long long a0 = (long long)sat(floorf(lminx / h.sx)) - 1, a1 = (long long)sat(ceilf(lmaxx / h.sx)) + 1;
    long long b0 = (long long)sat(floorf(lminz / h.sz)) - 1, b1 = (long long)sat(ceilf(lmaxz / h.sz)) + 1;
    auto cl = [](long long v, long long lo, long long hi) -> int { return (int)(v < lo ? lo : (v > hi ? hi : v)); };
    gx0 = cl(a0, 0, nx - 2); gx1 = cl(a1, 0, nx - 2); gz0 = cl(b0, 0, nz - 2); gz1 = cl(b1, 0, nz - 2);
    if (mn.y > h.hmax + 1.0f || mx.y < h.hmin - 1.0f) return false;
    return gx1 >= gx0 && gz1 >= gz0;
}

__global__ void k_col_aabb(Dev d, TickParams tp) {
    if (aborted(d)) return;
    GRID_STRIDE(c, d.nc) {
        uint32_t b = d.col_body[c];
        V3 p = L3(d.pos[b]);
        Q4 q = LQ(d.rot[b]);
        V3 mn, mx;
        aabb_of(d.btype[b], d.bdims[b], p, q, mn, mx);
        d.amin[c] = S3(mn); d.amax[c] = S3(mx);
        float reach = 0.5f * len3(mx - mn);                              // world.rs:97-100
        float s = 0.05f + (len3(L3(d.lin[b])) + len3(L3(d.ang[b])) * reach) * tp.dt;  // step.rs:78
        d.sweep[c] = s;
        V3 smn = mn - splat3(s), smx = mx + splat3(s);
        d.smin[c] = S3(smn); d.smax[c] = S3(smx);
        uint32_t cls;
        if (!isfinite(smn.x)) cls = CLS_INVALID;
        else {
            float dx = smx.x - smn.x, dy = smx.y - smn.y, dz = smx.z - smn.z;
            const float lim = 2.5f * 5.0f;  // broadphase.rs:44 with cell 2.5 (world.rs:419)
            cls = (dx > lim || dy > lim || dz > lim) ? CLS_OVER : CLS_SAP;
        }
        d.ccls[c] = cls;
        int gx0, gx1, gz0, gz1;
        uint32_t cand = 0;
        if (hf_range(d.hf[d.col_world[c]], mn, mx, gx0, gx1, gz0, gz1))
            cand = 2u * (uint32_t)(gx1 - gx0 + 1) * (uint32_t)(gz1 - gz0 + 1);
        d.cand_off[c] = cand;
    }
}

// ------------------------------------------------------------------ broadphase
static __device__ __forceinline__ uint32_t sortable(float f) {
    uint32_t b = __float_as_uint(f);
    return (b & 0x80000000u) ? ~b : (b | 0x80000000u);
}

// One workgroup per world: bitonic sort of (key, local index) in LDS; SAP colliders by min.x,
// then oversized, then invalid colliders, each group in index order.
__global__ void __launch_bounds__(1024) k_bp_sort(Dev d) {
    if (aborted(d)) return;
    __shared__ uint64_t e[RP_MAX_WORLD_COLLIDERS];
    __shared__ float red[1024];
    __shared__ uint32_t cnt_sap, cnt_over;
    const uint32_t w = blockIdx.x, t = threadIdx.x;
    const uint32_t c0 = d.w_col_off[w], n = d.w_col_off[w + 1] - c0;
    uint32_t m = 1;
    while (m < n) m <<= 1;
    if (t == 0) { cnt_sap = 0; cnt_over = 0; }
    __syncthreads();
    float mext = 0.0f;
    for (uint32_t i = t; i < m; i += blockDim.x) {
        uint64_t key = ~0ull;
        if (i < n) {
            uint32_t c = c0 + i, cls = d.ccls[c];
            uint32_t k;
            if (cls == CLS_SAP) {
                k = sortable(d.smin[c].x);
                if (k == 0xFFFFFFFFu || k == 0xFFFFFFFEu) k = 0xFFFFFFFDu;
                atomicAdd(&cnt_sap, 1u);
                mext = fmaxf(mext, d.smax[c].x - d.smin[c].x);
            } else if (cls == CLS_OVER) { k = 0xFFFFFFFEu; atomicAdd(&cnt_over, 1u); }
            else k = 0xFFFFFFFFu;
            key = ((uint64_t)k << 32) | i;
        }
        e[i] = key;
    }
    red[t] = mext;
    __syncthreads();
    for (uint32_t k = 2; k <= m; k <<= 1) {
        for (uint32_t j = k >> 1; j > 0; j >>= 1) {
            for (uint32_t i = t; i < m; i += blockDim.x) {
                uint32_t ixj = i ^ j;
                if (ixj > i) {
                    uint64_t a = e[i], b = e[ixj];
                    bool asc = (i & k) == 0;
                    if ((a > b) == asc) { e[i] = b; e[ixj] = a; }
                }
            }
            __syncthreads();
        }
    }
    for (uint32_t s = 512; s > 0; s >>= 1) {
        if (t < s) red[t] = fmaxf(red[t], red[t + s]);
        __syncthreads();
    }
    for (uint32_t i = t; i < n; i += blockDim.x) {
        uint32_t li = (uint32_t)e[i];
        d.bp_ord[c0 + i] = li;
        d.bp_rank[c0 + li] = i;
    }
    if (t == 0) {
        d.w_nsap[w] = cnt_sap;
        d.w_nover[w] = cnt_over;
        // inflated so the backward scan never stops early on rounding
        d.w_maxext[w] = red[0] * 1.0001f + 1.0e-4f;
    }
}

// The same order for scenes with a world too large for k_bp_sort's LDS (more than
// RP_MAX_WORLD_COLLIDERS colliders): every collider's class key (k_bp_sort's), then two stable
// radix sorts over the whole scene, by key and then by world, which leave index order on equal
// keys, as the (key, index) pairs of the bitonic sort do.
__global__ void k_bp_keys(Dev d, uint32_t *keys, uint32_t *vals) {
    if (aborted(d)) return;
    GRID_STRIDE(c, d.nc) {
        const uint32_t cls = d.ccls[c], w = d.col_world[c];
        uint32_t k;
        if (cls == CLS_SAP) {
            k = sortable(d.smin[c].x);
            if (k == 0xFFFFFFFFu || k == 0xFFFFFFFEu) k = 0xFFFFFFFDu;
            atomicAdd(&d.w_nsap[w], 1u);
        } else if (cls == CLS_OVER) {
            k = 0xFFFFFFFEu;
            atomicAdd(&d.w_nover[w], 1u);
        } else {
            k = 0xFFFFFFFFu;
        }
        keys[c] = k;
        vals[c] = c;
    }
}
__global__ void k_bp_world_keys(Dev d, const uint32_t *vals, uint32_t *keys) {
    if (aborted(d)) return;
    GRID_STRIDE(p, d.nc) keys[p] = d.col_world[vals[p]];
}
// sorted position p holds collider vals[p]; world w's colliders take positions [c0, c0 + n)
__global__ void k_bp_ranks(Dev d, const uint32_t *vals) {
    if (aborted(d)) return;
    GRID_STRIDE(p, d.nc) {
        const uint32_t c = vals[p], c0 = d.w_col_off[d.col_world[c]];
        d.bp_ord[p] = c - c0;
        d.bp_rank[c] = p - c0;
    }
}

// aabb.rs:35-40 inclusive overlap (symmetric)
static __device__ __forceinline__ bool overlaps(float4 amn, float4 amx, float4 bmn, float4 bmx) {
    return !(amx.x < bmn.x || amn.x > bmx.x || amx.y < bmn.y || amn.y > bmx.y || amx.z < bmn.z || amn.z > bmx.z);
}

// The pairs of broadphase.rs:29-99, as per-collider lists of partners j > i (world-global
// indices, ascending j).  A collider of the sweep class scans forward in min.x order only, up to
// its own max.x, so each x-overlapping pair of the sweep class is found once, by its first member
// in that order, and appended to the list of its lower index; oversized colliders test every
// later collider, and every collider tests the oversized ones after it.  Mode 0 counts each
// list's length (out_off, zeroed before), mode 1 appends through per-list cursors (bp_fill,
// zeroed before); k_bp_sort_lists then puts each list in ascending order, so the lists do not
// depend on which lane appended first.
//
// RP_BP_G lanes scan one collider's run together, RP_BP_G positions a step: a ballot finds the
// step's first position past max.x, and only the lanes before it test and emit, so the pairs are
// exactly the serial scan's.  Groups are aligned within a wavefront (blockDim is a multiple of 64),
// and a group's lanes leave the loop together.
#define RP_BP_G 32u
template <int MODE>
__global__ void k_bp_enum(Dev d) {
    if (aborted(d)) return;
    const uint32_t cap = d.cap_pairs;
    auto emit = [&](uint32_t owner, uint32_t other) {
        if (MODE == 0) {
            atomicAdd(&d.out_off[owner], 1u);
        } else {
            const uint32_t o = d.out_off[owner] + atomicAdd(&d.bp_fill[owner], 1u);
            if (o < cap) { d.pair_i[o] = owner; d.pair_j[o] = other; }
        }
    };
    const uint32_t lane = threadIdx.x % RP_BP_G, shift = (threadIdx.x % 64u) - lane;
    GRID_STRIDE(t, d.nc * RP_BP_G) {
        const uint32_t c = t / RP_BP_G;
        const uint32_t w = d.col_world[c], c0 = d.w_col_off[w], n = d.w_col_off[w + 1] - c0, i = c - c0;
        const uint32_t nsap = d.w_nsap[w], nover = d.w_nover[w];
        const float4 mn = d.smin[c], mx = d.smax[c];
        const uint32_t cls = d.ccls[c];
        if (cls == CLS_OVER) {
            for (uint32_t j = i + 1 + lane; j < n; j += RP_BP_G)
                if (overlaps(mn, mx, d.smin[c0 + j], d.smax[c0 + j])) emit(c, c0 + j);
        } else {
            if (cls == CLS_SAP) {
                const uint32_t r = d.bp_rank[c];
                for (uint32_t q0 = r + 1; q0 < nsap; q0 += RP_BP_G) {
                    const uint32_t q = q0 + lane;
                    uint32_t j = 0;
                    float4 jmn = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
                    bool brk = q >= nsap;
                    if (!brk) {
                        j = d.bp_ord[c0 + q];
                        jmn = d.smin[c0 + j];
                        brk = jmn.x > mx.x;
                    }
                    const uint64_t gm = ((uint64_t)__ballot(brk) >> shift) & (RP_BP_G == 64u ? ~0ull : ((1ull << (RP_BP_G % 64u)) - 1ull));
                    const uint32_t first = gm ? (uint32_t)__ffsll((unsigned long long)gm) - 1u : RP_BP_G;  // the serial scan's stop
                    if (lane < first && overlaps(mn, mx, jmn, d.smax[c0 + j])) emit(min(c, c0 + j), max(c, c0 + j));
                    if (gm) break;
                }
            }
            for (uint32_t q = nsap + lane; q < nsap + nover; q += RP_BP_G) {
                const uint32_t j = d.bp_ord[c0 + q];
                if (j > i && overlaps(mn, mx, d.smin[c0 + j], d.smax[c0 + j])) emit(c, c0 + j);
            }
        }
    }
}
// Insertion sort of each collider's partner list (ascending j).
__global__ void k_bp_sort_lists(Dev d) {
    if (aborted(d)) return;
    GRID_STRIDE(c, d.nc) {
        const uint32_t base = d.out_off[c], count = d.bp_fill[c];
        if (base + count > d.cap_pairs) continue;
        for (uint32_t a = base + 1; a < base + count; a++) {
            const uint32_t v = d.pair_j[a];
            uint32_t b = a;
            while (b > base && d.pair_j[b - 1] > v) { d.pair_j[b] = d.pair_j[b - 1]; b--; }
            d.pair_j[b] = v;
        }
    }
}

// ------------------------------------------------------------------ narrowphase
static __device__ __forceinline__ Shape shape_of(const Dev &d, uint32_t b) {
    float4 dm = d.bdims[b];
    Shape s; s.type = (int)d.btype[b]; s.d0 = dm.x; s.d1 = dm.y; s.d2 = dm.z;
    return s;
}

__global__ void k_narrow_pairs(Dev d) {
    if (aborted(d)) return;
    const uint32_t np = min(d.cnt[CT_PAIRS], d.cap_pairs);
    GRID_STRIDE(p, np) {
        uint32_t i = d.pair_i[p], j = d.pair_j[p];
        uint32_t ba = d.col_body[i], bb = d.col_body[j];
        uint32_t cnt = 0;
        bool jointed = false;  // bodies joined directly by a joint do not collide (step.rs:85-93,102)
        if (d.nj)
            for (uint32_t k = d.jadj_off[ba]; k < d.jadj_off[ba + 1]; k++) jointed |= d.jadj[k] == bb;
        if (ba != bb && !jointed) {
            float reach = (d.sweep[i] + d.sweep[j]) - 0.05f;  // step.rs:105
            Manifold m;
            if (collide_pair(shape_of(d, ba), L3(d.pos[ba]), LQ(d.rot[ba]), shape_of(d, bb), L3(d.pos[bb]), LQ(d.rot[bb]), reach, m)) {
                for (int k = 0; k < m.len; k++) {
                    float depth = (m.p[k].pen - reach) + 0.05f;  // step.rs:109
                    d.pc_p[4 * p + k] = S3(m.p[k].pos, depth);
                    d.pc_n[4 * p + k] = S3(m.p[k].n, __uint_as_float(m.p[k].feat));
                }
                cnt = (uint32_t)m.len;
            }
        }
        d.pair_n[p] = cnt;
    }
}

static __device__ __forceinline__ uint32_t upper_bound_u32(const uint32_t *a, uint32_t n, uint32_t v) {
    uint32_t lo = 0, hi = n;
    while (lo < hi) { uint32_t mid = (lo + hi) >> 1; if (a[mid] <= v) lo = mid + 1; else hi = mid; }
    return lo;
}

// terrain.rs:321-356 triangle t of cell (gx, gz): tri1 = (p00, p01, p10), tri2 = (p10, p01, p11)
static __device__ void hf_tri(const Dev &d, const HfDesc &h, int gx, int gz, int tri, V3 &a, V3 &b, V3 &c) {
    const float *H = d.heights + h.off;
    int nx = (int)h.nx;
    float h00 = H[gx + gz * nx], h10 = H[gx + 1 + gz * nx], h01 = H[gx + (gz + 1) * nx], h11 = H[gx + 1 + (gz + 1) * nx];
    float x0 = (float)gx * h.sx + h.ox, x1 = x0 + h.sx;
    float z0 = (float)gz * h.sz + h.oz, z1 = z0 + h.sz;
    V3 p00 = mk3(x0, h00, z0), p10 = mk3(x1, h10, z0), p01 = mk3(x0, h01, z1), p11 = mk3(x1, h11, z1);
    if (tri == 0) { a = p00; b = p01; c = p10; } else { a = p10; b = p01; c = p11; }
}

__global__ void k_narrow_terrain(Dev d) {
    if (aborted(d)) return;
    const uint32_t nt = min(d.cnt[CT_CANDS], d.cap_cands);
    GRID_STRIDE(t, nt) {
        uint32_t c = upper_bound_u32(d.cand_off, d.nc, t) - 1;
        uint32_t local = t - d.cand_off[c];
        uint32_t b = d.col_body[c];
        const HfDesc h = d.hf[d.col_world[c]];
        int gx0, gx1, gz0, gz1;
        hf_range(h, L3(d.amin[c]), L3(d.amax[c]), gx0, gx1, gz0, gz1);
        uint32_t width = (uint32_t)(gx1 - gx0 + 1);
        uint32_t cell = local >> 1, tri = local & 1u;
        int gz = gz0 + (int)(cell / width), gx = gx0 + (int)(cell % width);
        V3 ta, tb, tc;
        hf_tri(d, h, gx, gz, (int)tri, ta, tb, tc);
        float s = d.sweep[c];
        Manifold m;
        uint32_t cnt = 0;
        if (collide_terrain(shape_of(d, b), L3(d.pos[b]), LQ(d.rot[b]), ta, tb, tc, s, m)) {
            for (int k = 0; k < m.len; k++) {
                float depth = (m.p[k].pen - s) + 0.05f;  // step.rs:129
                d.tc_p[4 * t + k] = S3(m.p[k].pos, depth);
                d.tc_nf[4 * t + k] = S3(m.p[k].n, __uint_as_float(m.p[k].feat));
            }
            cnt = (uint32_t)m.len;
        }
        d.tc_n[t] = cnt;
        d.tc_cell[t] = ((uint32_t)gz * (h.nx - 1u) + (uint32_t)gx) * 2u + tri;
    }
}

// step.rs:499-531 for a collider's terrain bucket: <= 4 kept; a box keeps reduce_to_4 of its
// (depth, point) list with the deepest point's normal; other shapes keep the 4 deepest
// (sort_key6 descending, ties in narrowphase order).  Streams over the candidate slots.
__global__ void k_terrain_cull(Dev d) {
    if (aborted(d)) return;
    GRID_STRIDE(c, d.nc) {
        uint32_t t0 = d.cand_off[c], t1 = d.cand_off[c + 1];
        if (t1 > d.cap_cands) t1 = d.cap_cands;
        if (t0 > t1) t0 = t1;
        uint32_t n = 0;
        for (uint32_t t = t0; t < t1; t++) n += d.tc_n[t];
        uint32_t out = 0;
        auto put = [&](uint32_t t, uint32_t k) {
            d.tcul_p[4 * c + out] = d.tc_p[4 * t + k];
            d.tcul_n[4 * c + out] = d.tc_nf[4 * t + k];
            d.tcul_cell[4 * c + out] = d.tc_cell[t];
            out++;
        };
        if (n <= 4) {
            for (uint32_t t = t0; t < t1; t++)
                for (uint32_t k = 0; k < d.tc_n[t]; k++) put(t, k);
        } else if (d.btype[d.col_body[c]] == 1) {
            // reduce_to_4 (narrowphase.rs:312-338) streamed; items in (candidate, point) order
            uint32_t bt0 = 0, bk0 = 0; int64_t best = INT64_MIN;
            for (uint32_t t = t0; t < t1; t++)
                for (uint32_t k = 0; k < d.tc_n[t]; k++) {
                    int64_t key = sort_key6(d.tc_p[4 * t + k].w);
                    if (key > best) { best = key; bt0 = t; bk0 = k; }
                }
            V3 nrm = L3(d.tc_nf[4 * bt0 + bk0]);
            V3 p0 = L3(d.tc_p[4 * bt0 + bk0]);
            put(bt0, bk0);
            uint32_t bt1 = 0, bk1 = 0; float d1 = -INFINITY; best = INT64_MIN;
            for (uint32_t t = t0; t < t1; t++)
                for (uint32_t k = 0; k < d.tc_n[t]; k++) {
                    float v = len2(L3(d.tc_p[4 * t + k]) - p0);
                    int64_t key = sort_key6(v);
                    if (key > best) { best = key; bt1 = t; bk1 = k; d1 = v; }
                }
            if (d1 > 1.0e-12f) {
                put(bt1, bk1);
                V3 e = L3(d.tc_p[4 * bt1 + bk1]) - p0;
                uint32_t bt2 = 0, bk2 = 0; float a2 = -INFINITY; best = INT64_MIN;
                for (uint32_t t = t0; t < t1; t++)
                    for (uint32_t k = 0; k < d.tc_n[t]; k++) {
                        float v = fabsf(dot3(cross3(L3(d.tc_p[4 * t + k]) - p0, e), nrm));
                        int64_t key = sort_key6(v);
                        if (key > best) { best = key; bt2 = t; bk2 = k; a2 = v; }
                    }
                if (a2 > 1.0e-12f) {
                    put(bt2, bk2);
                    float side2 = dot3(cross3(L3(d.tc_p[4 * bt2 + bk2]) - p0, e), nrm);
                    float s2 = (side2 != side2) ? side2 : (signbit(side2) ? -1.0f : 1.0f);
                    uint32_t bt3 = 0, bk3 = 0; float a3 = -INFINITY; best = INT64_MIN;
                    for (uint32_t t = t0; t < t1; t++)
                        for (uint32_t k = 0; k < d.tc_n[t]; k++) {
                            float v = -dot3(cross3(L3(d.tc_p[4 * t + k]) - p0, e), nrm) * s2;
                            int64_t key = sort_key6(v);
                            if (key > best) { best = key; bt3 = t; bk3 = k; a3 = v; }
                        }
                    if (a3 > 1.0e-12f) put(bt3, bk3);
                }
            }
        } else {
            // top 4 by (sort_key6(depth) desc, order asc)
            int64_t kk[4]; uint32_t tt[4], ks[4]; int m = 0;
            for (uint32_t t = t0; t < t1; t++)
                for (uint32_t k = 0; k < d.tc_n[t]; k++) {
                    int64_t key = sort_key6(d.tc_p[4 * t + k].w);
                    int pos = m;
                    while (pos > 0 && kk[pos - 1] < key) pos--;
                    if (pos >= 4) continue;
                    int last = m < 4 ? m : 3;
                    for (int z = last; z > pos; z--) { kk[z] = kk[z - 1]; tt[z] = tt[z - 1]; ks[z] = ks[z - 1]; }
                    kk[pos] = key; tt[pos] = t; ks[pos] = k;
                    if (m < 4) m++;
                }
            for (int z = 0; z < m; z++) put(tt[z], ks[z]);
        }
        d.tcul_cnt[c] = out;
    }
}

__global__ void k_contact_count(Dev d) {
    if (aborted(d)) return;
    GRID_STRIDE(c, d.nc) {
        uint32_t p0 = d.out_off[c], p1 = d.out_off[c + 1];
        uint32_t s = 0;
        for (uint32_t p = p0; p < p1 && p < d.cap_pairs; p++) s += d.pair_n[p];
        d.c_off[c] = s + d.tcul_cnt[c];
    }
}

// step.rs:532-540 orientation: body-pair normals flipped so n.(pb - pa) >= 0
__global__ void k_contact_write(Dev d) {
    if (aborted(d)) return;
    GRID_STRIDE(c, d.nc) {
        uint32_t o = d.c_off[c];
        uint32_t ba = d.col_body[c];
        V3 pa = L3(d.pos[ba]);
        for (uint32_t p = d.out_off[c]; p < d.out_off[c + 1] && p < d.cap_pairs; p++) {
            uint32_t j = d.pair_j[p];
            V3 pb = L3(d.pos[d.col_body[j]]);
            for (uint32_t k = 0; k < d.pair_n[p]; k++, o++) {
                if (o >= d.cap_contacts) continue;
                float4 nf = d.pc_n[4 * p + k];
                V3 n = L3(nf);
                if (dot3(n, pb - pa) < 0.0f) n = -n;
                d.ct_a[o] = c; d.ct_b[o] = j; d.ct_cell[o] = RP_NONE;
                d.ct_p[o] = d.pc_p[4 * p + k];
                d.ct_n[o] = S3(n, nf.w);
            }
        }
        for (uint32_t k = 0; k < d.tcul_cnt[c]; k++, o++) {
            if (o >= d.cap_contacts) continue;
            d.ct_a[o] = c; d.ct_b[o] = RP_TERRAIN; d.ct_cell[o] = d.tcul_cell[4 * c + k];
            d.ct_p[o] = d.tcul_p[4 * c + k];
            d.ct_n[o] = d.tcul_n[4 * c + k];
        }
    }
}

// ------------------------------------------------------------------ solver prep
static __device__ __forceinline__ uint32_t uf_find(uint32_t *root, uint32_t x) {
    uint32_t p = __hip_atomic_load(&root[x], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    while (p != x) {
        x = p;
        p = __hip_atomic_load(&root[x], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    return x;
}
static __device__ void uf_union(uint32_t *root, uint32_t a, uint32_t b) {
    while (true) {
        uint32_t ra = uf_find(root, a), rb = uf_find(root, b);
        if (ra == rb) return;
        uint32_t hi = ra > rb ? ra : rb, lo = ra > rb ? rb : ra;
        uint32_t expected = hi;
        if (__hip_atomic_compare_exchange_strong(&root[hi], &expected, lo, __ATOMIC_RELAXED, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) return;
    }
}

static __device__ __forceinline__ float eff_mass(float inv_a, float inv_b, M3 iia, M3 iib, V3 ra, V3 rb, V3 dir) {
    V3 ca = cross3(ra, dir), cb = cross3(rb, dir);
    float k = ((inv_a + inv_b) + dot3(ca, mv(iia, ca))) + dot3(cb, mv(iib, cb));
    return k > 1.0e-9f ? 1.0f / k : 0.0f;
}
// world/src/world/geom.rs:30-38
static __device__ __forceinline__ void orthonormal_basis(V3 n, V3 &t1, V3 &t2) {
    float ax = fabsf(n.x), ay = fabsf(n.y), az = fabsf(n.z);
    V3 base = (ax <= ay && ax <= az) ? mk3(1, 0, 0) : (ay <= az ? mk3(0, 1, 0) : mk3(0, 0, 1));
    t1 = normalize_or_zero3(cross3(base, n));
    t2 = cross3(n, t1);
}

// solver.rs:494-590 constraint rows; islands.rs:58-66 unions; step.rs:160-163 constrained
// flags; step.rs:337-343 grounding flags; warm keys and anchors (solver.rs:720-731)
__global__ void k_prep(Dev d, TickParams tp) {
    if (aborted(d)) return;
    const uint32_t nct = min(d.cnt[CT_CONTACTS], d.cap_contacts);
    GRID_STRIDE(c, nct) {
        uint32_t ca = d.ct_a[c], cb = d.ct_b[c];
        bool terr = cb == RP_TERRAIN;
        uint32_t ai = d.col_body[ca], bi = terr ? RP_NONE : d.col_body[cb];
        uint32_t w = d.col_world[ca];
        uint32_t boff = d.w_body_off[w];
        float4 pn = d.ct_p[c];
        V3 point = L3(pn);
        float depth = pn.w;
        V3 n = L3(d.ct_n[c]);
        bool aw = d.awake[ai] != 0, bw = !terr && d.awake[bi] != 0;
        float inv_a = aw ? d.inv_mass[ai] : 0.0f, inv_b = bw ? d.inv_mass[bi] : 0.0f;
        M3 iia = inv_a > 0.0f ? load_m9(d.iiw + 9 * ai) : mzero();
        M3 iib = inv_b > 0.0f ? load_m9(d.iiw + 9 * bi) : mzero();
        uint32_t ma = d.col_mat[ca], mb = terr ? d.hf[w].mat : d.col_mat[cb];
        float4 pp = d.pair_tab[(ma & 15u) * 16u + (mb & 15u)];
        V3 t1, t2;
        orthonormal_basis(n, t1, t2);
        V3 pa = L3(d.pos[ai]);
        V3 ra = point - pa, rb = terr ? splat3(0.0f) : point - L3(d.pos[bi]);
        float mn = eff_mass(inv_a, inv_b, iia, iib, ra, rb, n);
        float m1 = eff_mass(inv_a, inv_b, iia, iib, ra, rb, t1);
        float m2 = eff_mass(inv_a, inv_b, iia, iib, ra, rb, t2);
        float pen = depth - 0.05f;
        float bias = pen > 0.0f ? fminf(fmaxf(pen - 0.005f, 0.0f) * (0.05f / tp.dt), 2.0f) : pen / tp.dt;
        V3 pva = L3(d.pre_lin[ai]) + cross3(L3(d.pre_ang[ai]), ra);
        V3 pvb = terr ? splat3(0.0f) : L3(d.pre_lin[bi]) + cross3(L3(d.pre_ang[bi]), rb);
        float v_pre = dot3(pvb - pva, n);
        V3 cva = L3(d.lin[ai]) + cross3(L3(d.ang[ai]), ra);
        V3 cvb = terr ? splat3(0.0f) : L3(d.lin[bi]) + cross3(L3(d.ang[bi]), rb);
        float approach = 0.5f * (v_pre + dot3(cvb - cva, n));
        // compliant pair (solver.rs:545-567): no Baumgarte bias or restitution; the implicit
        // Hunt-Crossley row's constants when its effective mass and spring are positive
        uint32_t cmf = 0;
        if (d.cm_on) {
            const float4 cm = d.pair_cm[(ma & 15u) * 16u + (mb & 15u)];
            float4 sp = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
            if (cm.x > 0.0f) {
                cmf = CM_SOFT;
                bias = 0.0f;
                const float vn0 = dot3(cvb - cva, n);
                const float pen_d = tp.damp_start ? pen : pen - (0.5f * tp.dt) * vn0;
                const float cdamp = (cm.x * fmaxf(cm.y, 0.0f)) * fmaxf(pen_d, 0.0f);
                const float denom = cdamp + tp.dt * cm.x;
                if (mn > 0.0f && denom > 0.0f) {
                    cmf |= CM_IMPLICIT;
                    sp.x = 1.0f / (tp.dt * denom);
                    sp.y = (((tp.dt * cm.x) / denom) * pen) / tp.dt;
                    sp.z = 1.0f / (1.0f / mn + sp.x);
                }
            }
            if (d.ct_patch[c] != RP_NONE) cmf |= CM_PATCH_PT;
            sp.w = __uint_as_float(cmf);
            d.cs_soft[c] = sp;
        }
        // restitution bias (solver.rs:569-588): a contact approaching faster than 1 m/s bounces
        // after the solve; for free bodies at its time of impact (W6), from the sides' normal
        // velocities at the step's start (last step's end linear velocity, the pre-step angular
        // one) and after this step's forces
        if (d.w6_on && approach < -1.0f && !(cmf & CM_SOFT)) {
            const float rbias = -pp.z * approach;
            if (!(rbias <= 0.0f)) {
                const bool toi = !(d.flags[ai] & FL_JOINTED) && (terr || !(d.flags[bi] & FL_JOINTED));
                float4 tv = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
                if (toi) {
                    tv.x = dot3(L3(d.last_lin[ai]) + cross3(L3(d.pre_ang[ai]), ra), n);
                    tv.y = dot3(cva, n);
                    if (!terr) {
                        tv.z = dot3(L3(d.last_lin[bi]) + cross3(L3(d.pre_ang[bi]), rb), n);
                        tv.w = dot3(cvb, n);
                    }
                }
                d.w6_toi[c] = make_float4(rbias, toi ? -pen : 0.0f, toi ? pp.z : 0.0f, toi ? 1.0f : 0.0f);
                d.w6_tv[c] = tv;
                if (!d.w_w6[w]) atomicOr(&d.w_w6[w], 1u);
            }
        }
        d.cs_ai[c] = ai; d.cs_bi[c] = bi;
        d.cs_inv[c] = make_float2(inv_a, inv_b);
        d.cs_ra[c] = S3(ra); d.cs_rb[c] = S3(rb);
        d.cs_n[c] = S3(n); d.cs_t1[c] = S3(t1); d.cs_t2[c] = S3(t2);
        d.cs_m[c] = make_float4(mn, m1, m2, bias);
        d.cs_mu[c] = make_float4(pp.x, pp.y, fmaxf(pp.w, 1.0e-4f), approach);
        uint32_t wa = inv_a > 0.0f ? ai : RP_NONE;
        uint32_t wb = (!terr && inv_b > 0.0f) ? bi : RP_NONE;
        if (d.nj && (wa != RP_NONE || wb != RP_NONE)) {
            // a body a joint writes (asleep, say) that this contact only reads: the contact waits
            // for that joint's row and joins its island
            if (wa == RP_NONE && (d.flags[ai] & FL_JW)) wa = ai;
            if (wb == RP_NONE && !terr && (d.flags[bi] & FL_JW)) wb = bi;
        }
        d.cs_wa[c] = wa; d.cs_wb[c] = wb;
        d.jn[c] = 0.0f; d.jn_pre[c] = 0.0f; d.jt[c] = make_float2(0.0f, 0.0f); d.ct_imp[c] = 0.0f;
        d.warm_idx[c] = -1;
        d.ct_key[c] = ((uint64_t)(ai - boff) << 32) | (terr ? 0xFFFFFFFFull : (uint64_t)(bi - boff));
        d.ct_anchor[c] = S3(qrot(qconj(LQ(d.rot[ai])), point - pa));
        // idempotent ORs: skipped once the bit is seen set (a world's contacts all mark one word)
        if (!(d.flags[ai] & FL_CONSTRAINED)) atomicOr(&d.flags[ai], FL_CONSTRAINED);
        if (!terr && !(d.flags[bi] & FL_CONSTRAINED)) atomicOr(&d.flags[bi], FL_CONSTRAINED);
        // grounding (step.rs:337-342): touching terrain contacts of rigid pairs
        if (terr && depth - 0.05f > -0.005f && !(cmf & CM_SOFT) && !(d.flags[ai] & FL_GROUND)) atomicOr(&d.flags[ai], FL_GROUND);
        if ((d.inv_mass[ai] > 0.0f || (!terr && d.inv_mass[bi] > 0.0f)) && !d.w_solve[w]) atomicOr(&d.w_solve[w], 1u);
        if (wa != RP_NONE && wb != RP_NONE) uf_union(d.root, wa, wb);
    }
}

// Ball joints before the contacts' prep: the constrained and jointed flags of both bodies
// (step.rs:164-173), and for an active joint (a dynamic body on either side, islands.rs:96-101) its
// rows from the tick-start poses (solver.rs:336-424: axis i of frame A, lever arms, 1 / K where
// K = wa + wb + (ra x u) Ia (ra x u) + (rb x u) Ib (rb x u) > 1e-9), the world's solve flag, the
// written-by-a-joint flags, and the union of its two bodies when both are dynamic.
// D6 joints (rp_d6.h, included below with the position solve)
static __device__ __forceinline__ void d6_store(const Dev &d, uint32_t q);
static __device__ __forceinline__ bool d6_is(const Dev &d, uint32_t q);
static __device__ __forceinline__ bool d6_motor(const Dev &d, uint32_t q);
static __device__ uint32_t d6_rows(const Dev &d, uint32_t q, float dt);
static __device__ __forceinline__ bool dist_is(const Dev &d, uint32_t q);
static __device__ __forceinline__ float dist_compliance(const Dev &d, uint32_t q);
__global__ void k_joint_prep(Dev d, TickParams tp) {
    if (aborted(d)) return;
    GRID_STRIDE(q, d.nj) {
        const uint32_t a = d.jt_a[q], b = d.jt_b[q];
        if ((d.flags[a] & (FL_CONSTRAINED | FL_JOINTED)) != (FL_CONSTRAINED | FL_JOINTED)) atomicOr(&d.flags[a], FL_CONSTRAINED | FL_JOINTED);
        if ((d.flags[b] & (FL_CONSTRAINED | FL_JOINTED)) != (FL_CONSTRAINED | FL_JOINTED)) atomicOr(&d.flags[b], FL_CONSTRAINED | FL_JOINTED);
        const float wa = d.inv_mass[a], wb = d.inv_mass[b];
        uint32_t m = 0;
        // A D6 or distance joint with a dynamic side is an island entry even without velocity rows
        // (its position row runs in its island's passes), joining its two bodies' islands.
        if (d6_is(d, q)) {  // a D6 joint: its rows (rp_d6.h); velocity motors mark their bodies
            if (d6_motor(d, q)) { atomicOr(&d.flags[a], FL_DRIVEN); atomicOr(&d.flags[b], FL_DRIVEN); }
            uint32_t n = 0;
            const bool act = wa > 0.0f || wb > 0.0f;
            if (act) {
                const uint32_t w = d.bworld[a];
                if (!d.w_solve[w]) atomicOr(&d.w_solve[w], 1u);
                n = d6_rows(d, q, tp.dt);
                if (n) {
                    if (wa > 0.0f && !(d.flags[a] & FL_JW)) atomicOr(&d.flags[a], FL_JW);
                    if (wb > 0.0f && !(d.flags[b] & FL_JW)) atomicOr(&d.flags[b], FL_JW);
                }
                if (wa > 0.0f && wb > 0.0f) uf_union(d.root, a, b);
            }
            d.d6_nrows[q] = n;
            d.jr_mask[q] = act ? 1u : 0u;
            continue;
        }
        if (dist_is(d, q)) {  // a distance joint: a rigid one's legacy row writes each side with mass
            const bool act = wa > 0.0f || wb > 0.0f;
            if (act) {
                const uint32_t w = d.bworld[a];
                if (!d.w_solve[w]) atomicOr(&d.w_solve[w], 1u);
                if (!(dist_compliance(d, q) > 0.0f)) {
                    if (wa > 0.0f && !(d.flags[a] & FL_JW)) atomicOr(&d.flags[a], FL_JW);
                    if (wb > 0.0f && !(d.flags[b] & FL_JW)) atomicOr(&d.flags[b], FL_JW);
                }
                if (wa > 0.0f && wb > 0.0f) uf_union(d.root, a, b);
            }
            d.d6_nrows[q] = 0;
            d.jr_mask[q] = act ? 1u : 0u;
            continue;
        }
        if (wa > 0.0f || wb > 0.0f) {
            const uint32_t w = d.bworld[a];
            if (!d.w_solve[w]) atomicOr(&d.w_solve[w], 1u);
            const V3 pa = L3(d.pos[a]), pb = L3(d.pos[b]);
            const Q4 qa = LQ(d.rot[a]), qb = LQ(d.rot[b]);
            const Q4 rot_a = qmul(qa, LQ(d.jt_far[q]));
            const V3 xa = pa + qrot(qa, L3(d.jt_fap[q])), xb = pb + qrot(qb, L3(d.jt_fbp[q]));
            const M3 ia = load_m9(d.iiw + 9 * a), ib = load_m9(d.iiw + 9 * b);
            const V3 ra = xa - pa, rb = xb - pb;
            for (uint32_t i = 0; i < 3; i++) {
                const V3 u = qrot(rot_a, unit_axis(i));
                const V3 ca = cross3(ra, u), cb = cross3(rb, u);
                const float k = ((wa + wb) + dot3(ca, mv(ia, ca))) + dot3(cb, mv(ib, cb));
                d.jr_acc[3 * q + i] = 0.0f;
                if (k <= 1.0e-9f) continue;
                d.jr_u[3 * q + i] = S3(u, 1.0f / k);
                m |= 1u << i;
            }
            d.jr_ra[q] = S3(ra); d.jr_rb[q] = S3(rb);
            if (m) {
                if (wa > 0.0f && !(d.flags[a] & FL_JW)) atomicOr(&d.flags[a], FL_JW);
                if (wb > 0.0f && !(d.flags[b] & FL_JW)) atomicOr(&d.flags[b], FL_JW);
                if (wa > 0.0f && wb > 0.0f) uf_union(d.root, a, b);
            }
        }
        d.jr_mask[q] = m;
    }
}

// Island key of each list entry (joints first, then contacts, then friction patches, so a stable
// sort keeps every joint ahead of the contacts of its island and every patch after them, as the
// reference runs joint batches, contact batches, then the patches in each iteration): the
// component root (lowest body index) of its written bodies.
__global__ void k_island_keys(Dev d) {
    if (aborted(d)) return;
    const uint32_t nrows = min(d.cnt[CT_ROWS], d.cap_contacts), nj = d.nj;
    GRID_STRIDE(r, nrows) {
        if (r < nj) {
            const uint32_t a = d.jt_a[r], b = d.jt_b[r];
            const uint32_t x = d.inv_mass[a] > 0.0f ? a : (d.inv_mass[b] > 0.0f ? b : RP_NONE);
            d.isl_k[r] = (d.jr_mask[r] && x != RP_NONE) ? uf_find(d.root, x) : d.nb;
            d.isl_v[r] = RP_JBIT | r;
            continue;
        }
        const uint32_t nct = min(d.cnt[CT_CONTACTS], d.cap_contacts);
        if (r - nj >= nct) {  // a friction patch: its first point's island, after its contacts
            const uint32_t p = r - nj - nct, c = d.pt_first[p];
            const uint32_t x = d.cs_wa[c] != RP_NONE ? d.cs_wa[c] : d.cs_wb[c];
            d.isl_k[r] = x == RP_NONE ? d.nb : uf_find(d.root, x);
            d.isl_v[r] = RP_PBIT | p;
            continue;
        }
        const uint32_t c = r - nj;
        uint32_t wa = d.cs_wa[c], wb = d.cs_wb[c];
        uint32_t x = wa != RP_NONE ? wa : wb;
        d.isl_k[r] = x == RP_NONE ? d.nb : uf_find(d.root, x);  // nb: no written body (sorts last)
        d.isl_v[r] = c;
    }
}

// Island segment starts over the sorted (key, contact) list; counts contacts with a writer.
__global__ void k_island_flags(Dev d, const uint32_t *keys) {
    if (aborted(d)) return;
    const uint32_t nct = min(d.cnt[CT_ROWS], d.cap_contacts);
    GRID_STRIDE(r, d.cap_contacts) {
        if (r >= nct) { d.isl_flag[r] = 0; continue; }
        uint32_t k = keys[r];
        d.isl_flag[r] = (k != d.nb && (r == 0 || keys[r - 1] != k)) ? 1u : 0u;
        if (k != d.nb && (r + 1 == nct || keys[r + 1] == d.nb)) d.cnt[CT_ISLC] = r + 1;
    }
}
__global__ void k_island_compact(Dev d, const uint32_t *keys) {
    if (aborted(d)) return;
    const uint32_t nct = min(d.cnt[CT_ROWS], d.cap_contacts);
    GRID_STRIDE(r, nct) {
        uint32_t k = keys[r];
        if (k != d.nb && (r == 0 || keys[r - 1] != k)) d.isl_start[d.isl_flag[r]] = r;
    }
}

// solver.rs:733-762 matching: per key group (contiguous, <= 4 contacts), the nearest unused
// entry within 1 cm, lowest index on ties.  Contacts writing no body get their friction
// accumulator here (their bodies' velocities are final); the others in the first sweep.
__global__ void k_warm_match(Dev d, TickParams tp) {
    if (aborted(d)) return;
    const uint32_t nct = min(d.cnt[CT_CONTACTS], d.cap_contacts);
    const uint32_t cur = tp.cur;
    GRID_STRIDE(c, nct) {
        uint64_t key = d.ct_key[c];
        uint32_t w = d.col_world[d.ct_a[c]];
        if (!d.w_solve[w]) continue;
        if (c > 0 && d.ct_key[c - 1] == key && d.col_world[d.ct_a[c - 1]] == w) continue;
        uint32_t e0 = d.wc_off[cur][w], e1 = d.wc_off[cur][w + 1];
        uint32_t lo = e0, hi = e1;
        while (lo < hi) { uint32_t mid = (lo + hi) >> 1; if (d.wk_key[cur][mid] < key) lo = mid + 1; else hi = mid; }
        uint32_t es = lo, ee = lo;
        while (ee < e1 && d.wk_key[cur][ee] == key) ee++;
        if (es == ee) continue;
        uint32_t taken = 0;  // entries per key stay well under 32
        for (uint32_t g = c; g < nct && d.ct_key[g] == key && d.col_world[d.ct_a[g]] == w; g++) {
            const uint32_t cmf = cm_flags(d, g);
            if (cmf & CM_SOFT) continue;  // compliant rows start cold (solver.rs:736)
            V3 an = L3(d.ct_anchor[g]);
            int best = -1; float bd = 0.0f;
            for (uint32_t e = es; e < ee && e - es < 32; e++) {
                if (taken & (1u << (e - es))) continue;
                float dd = len2(L3(d.wk_la[cur][e]) - an);
                if (dd <= 0.01f * 0.01f && (best < 0 || dd < bd)) { best = (int)e; bd = dd; }
            }
            if (best < 0) continue;
            taken |= 1u << (best - es);
            d.warm_idx[g] = best;
            // zero / learned warm start: solved rows start elsewhere (the match stays a feature)
            if (tp.ws_mode != 0 && (d.cs_wa[g] != RP_NONE || d.cs_wb[g] != RP_NONE)) continue;
            float ejn = d.wk_jn[cur][best];
            d.jn[g] = ejn;
            d.jn_pre[g] = ejn;
            if (d.cs_wa[g] == RP_NONE && d.cs_wb[g] == RP_NONE && !(cmf & CM_PATCH_PT)) {
                // slip speed from constant velocities; clamp as the iterations do
                uint32_t ai = d.cs_ai[g], bi = d.cs_bi[g];
                V3 ra = L3(d.cs_ra[g]), rb = L3(d.cs_rb[g]);
                V3 va = L3(d.lin[ai]) + cross3(L3(d.ang[ai]), ra);
                V3 vb = bi == RP_NONE ? splat3(0.0f) : L3(d.lin[bi]) + cross3(L3(d.ang[bi]), rb);
                V3 v = vb - va;
                V3 t1 = L3(d.cs_t1[g]), t2 = L3(d.cs_t2[g]);
                float vt1 = dot3(v, t1), vt2 = dot3(v, t2);
                float mu = mu_dyn(d.cs_mu[g], sqrtf(vt1 * vt1 + vt2 * vt2));
                V3 ejt = L3(d.wk_jt[cur][best]);
                V3 t = clamp_len_max(mk3(dot3(ejt, t1), dot3(ejt, t2), 0.0f), mu * ejn);
                d.jt[g] = make_float2(t.x, t.y);
            }
        }
    }
}

// ------------------------------------------------------------------ learned warm start
// The 12 per-contact features of the warm-start model, from the constraint as prepared and the
// velocities entering the solve: normal and tangential relative velocity, penetration, the
// sides' inverse masses, the angular terms (r x d) . I^-1 (r x d) for the normal (both sides) and
// tangent 1 (both sides), the previous frame's matched normal impulse (0 if none), and mu_s.
__global__ void k_features(Dev d, TickParams tp) {
    if (aborted(d)) return;
    const uint32_t nct = min(d.cnt[CT_CONTACTS], d.cap_contacts);
    GRID_STRIDE(c, nct) {
        const uint32_t ai = d.cs_ai[c], bi = d.cs_bi[c];
        const V3 ra = L3(d.cs_ra[c]), rb = L3(d.cs_rb[c]), n = L3(d.cs_n[c]), t1 = L3(d.cs_t1[c]), t2 = L3(d.cs_t2[c]);
        const float2 inv = d.cs_inv[c];
        V3 va = L3(d.lin[ai]) + cross3(L3(d.ang[ai]), ra);
        V3 vb = bi == RP_NONE ? splat3(0.0f) : L3(d.lin[bi]) + cross3(L3(d.ang[bi]), rb);
        V3 v = vb - va;
        M3 iia = inv.x > 0.0f ? load_m9(d.iiw + 9 * ai) : mzero();
        M3 iib = inv.y > 0.0f ? load_m9(d.iiw + 9 * bi) : mzero();
        V3 an = cross3(ra, n), bn = cross3(rb, n), at = cross3(ra, t1), bt = cross3(rb, t1);
        const int wi = d.warm_idx[c];
        float *f = d.feat + (size_t)c * 12;
        f[0] = dot3(v, n); f[1] = dot3(v, t1); f[2] = dot3(v, t2);
        f[3] = d.ct_p[c].w - 0.05f;
        f[4] = inv.x; f[5] = inv.y;
        f[6] = dot3(an, mv(iia, an)); f[7] = dot3(bn, mv(iib, bn)); f[8] = dot3(at, mv(iia, at)); f[9] = dot3(bt, mv(iib, bt));
        f[10] = wi >= 0 ? d.wk_jn[tp.cur][wi] : 0.0f;
        f[11] = d.cs_mu[c].x;
    }
}
// The previous-frame warm start of contact c in the contact's frame: (matched jn, matched jt
// projected on the current tangents), zeros when nothing matched (solver.rs:749-760 before the
// friction clamp).  The learned model predicts the converged impulses minus this.
static __device__ __forceinline__ V3 warm_base(const Dev &d, uint32_t c, uint32_t cur) {
    const int wi = d.warm_idx[c];
    if (wi < 0) return splat3(0.0f);
    const V3 ejt = L3(d.wk_jt[cur][wi]);
    return mk3(d.wk_jn[cur][wi], dot3(ejt, L3(d.cs_t1[c])), dot3(ejt, L3(d.cs_t2[c])));
}
// The contact entries of the island lists (joint entries are not logged), flagged for an
// exclusive scan that gives each its log row; newflag is free until the warm-cache store.
// The contacts the log takes: every solved contact, or in residual mode those of islands that
// reached the tolerance (an island stopped at the sweep limit holds no converged impulses).
__global__ void k_log_flags(Dev d, TickParams tp) {
    if (aborted(d)) return;
    const uint32_t nislc = min(d.cnt[CT_ISLC], d.cap_contacts);
    GRID_STRIDE(r, d.cap_contacts)
        d.newflag[r] = (r < nislc && !(d.isl_v2[r] & (RP_JBIT | RP_PBIT)) && (!tp.resid_on || d.rank_conv[r])) ? 1u : 0u;
}
// Appends the solved contacts of this tick (island-list order) with their converged impulses,
// the normal impulse before the shock pass and the friction accumulators (solver.rs:917-926),
// less the contact's previous-frame warm start.
__global__ void k_log_contacts(Dev d, TickParams tp) {
    if (aborted(d)) return;
    const uint32_t nislc = min(d.cnt[CT_ISLC], d.cap_contacts);
    const uint32_t base = d.log_n[0];
    const bool room = (uint64_t)base + d.cnt[CT_LOG_N] <= d.log_cap;
    GRID_STRIDE(r, nislc) {
        if (!room) break;
        const uint32_t c = d.isl_v2[r];
        if ((c & (RP_JBIT | RP_PBIT)) || (tp.resid_on && !d.rank_conv[r])) continue;
        const size_t o = (size_t)base + d.newflag[r];
        for (int k = 0; k < 12; k++) d.log_feat[o * 12 + k] = d.feat[(size_t)c * 12 + k];
        const float2 t = d.jt[c];
        const V3 b = warm_base(d, c, tp.cur);
        d.log_targ[o * 3] = d.jn_pre[c] - b.x; d.log_targ[o * 3 + 1] = t.x - b.y; d.log_targ[o * 3 + 2] = t.y - b.z;
    }
}
__global__ void k_log_advance(Dev d) {
    if (threadIdx.x != 0 || blockIdx.x != 0 || aborted(d)) return;
    const uint32_t nrow = d.cnt[CT_LOG_N];
    if ((uint64_t)d.log_n[0] + nrow <= d.log_cap) d.log_n[0] += nrow;
    else d.log_n[1] += nrow;
}

// ------------------------------------------------------------------ ordered sweeps
// Island lists: positions [0, nislc) of isl_v2 grouped by island (isl_start), contact index order
// within an island, which is the reference's constraint order.  An order of an island is a
// permutation of its segment; its chains give, for every rank and side, the previous rank of the
// segment that writes the same body (global ranks, RP_NONE when there is none).  A sweep then runs
// as dataflow: a rank executes once both predecessors have, which is serial Gauss-Seidel in the
// order (islands share no written body).
//
// Chains are built for all islands at once: the (body, rank, side) entries are radix-sorted by
// body, stably, so ranks stay increasing within a body, and each entry links to the one before.
__global__ void k_chain_count(Dev d) {
    if (threadIdx.x == 0 && blockIdx.x == 0) d.cnt[CT_CHAIN_N] = 2u * min(d.cnt[CT_ISLC], d.cap_contacts);
}
__global__ void k_chain_entries(Dev d, const uint32_t *ord, uint32_t *keys, uint32_t *vals) {
    if (aborted(d)) return;
    const uint32_t n2 = d.cnt[CT_CHAIN_N];
    GRID_STRIDE(e, n2) {
        uint32_t c = ord[e >> 1];
        uint32_t b;
        if (c & RP_JBIT) {  // a joint writes each of its dynamic bodies
            const uint32_t x = (e & 1u) ? d.jt_b[c & ~RP_JBIT] : d.jt_a[c & ~RP_JBIT];
            b = d.inv_mass[x] > 0.0f ? x : RP_NONE;
        } else if (c & RP_PBIT) {  // a patch writes its points' bodies
            const uint32_t f = d.pt_first[c & ~RP_PBIT];
            b = (e & 1u) ? d.cs_wb[f] : d.cs_wa[f];
        } else {
            b = (e & 1u) ? d.cs_wb[c] : d.cs_wa[c];
        }
        keys[e] = b == RP_NONE ? d.nb : b;
        vals[e] = e;
    }
}
// Writes each (rank, side)'s predecessor (pa / pb) and successor (sa / sb) on its body's chain:
// the neighbouring entries of the same body at a strictly different rank (a contact writing one
// body on both sides links past its own other side), so every wait is on an earlier rank.
__global__ void k_chain_link(Dev d, const uint32_t *keys, const uint32_t *vals, uint32_t *pa, uint32_t *pb, uint32_t *sa, uint32_t *sb) {
    if (aborted(d)) return;
    const uint32_t n2 = d.cnt[CT_CHAIN_N];
    GRID_STRIDE(i, n2) {
        uint32_t v = vals[i], k = keys[i], r = v >> 1;
        uint32_t pred = RP_NONE, succ = RP_NONE;
        if (k != d.nb) {
            uint32_t j = i;
            while (j > 0 && keys[j - 1] == k && (vals[j - 1] >> 1) == r) j--;
            if (j > 0 && keys[j - 1] == k) pred = vals[j - 1] >> 1;
            j = i + 1;
            while (j < n2 && keys[j] == k && (vals[j] >> 1) == r) j++;
            if (j < n2 && keys[j] == k) succ = vals[j] >> 1;
        }
        ((v & 1u) ? pb : pa)[r] = pred;
        ((v & 1u) ? sb : sa)[r] = succ;
    }
}

// Shock order (solver.rs:968-975): (sort_key6(lowest height along up), list position) within each
// island; terrain contacts count as -infinity.  Keys are clamped to +-2^43 units of 1e-6 m
// (heights beyond 8,796 km tie and keep list order) and offset to 44 unsigned bits, split into
// the low 32 and high 12 bits for the stable radix passes.
static __device__ __forceinline__ uint64_t shock_key44(float h) {
    int64_t k = sort_key6(h);
    const int64_t LIM = (int64_t)1 << 43;
    if (k < -LIM) k = -LIM;
    if (k > LIM - 1) k = LIM - 1;
    return (uint64_t)(k + LIM);
}
__global__ void k_shock_keys(Dev d, TickParams tp, uint32_t *klo, uint32_t *khi, uint32_t *vals) {
    if (aborted(d)) return;
    const uint32_t nislc = min(d.cnt[CT_ISLC], d.cap_contacts);
    const V3 up = mk3(tp.upx, tp.upy, tp.upz);
    GRID_STRIDE(r, nislc) {
        uint32_t c = d.isl_v2[r];
        uint64_t k;
        if (c & (RP_JBIT | RP_PBIT)) {
            k = ((uint64_t)1 << 44) - 1;  // joints and patches take no part in the shock passes: last, no-ops
        } else {
            uint32_t a = d.cs_ai[c], b = d.cs_bi[c];
            float h = b == RP_NONE ? -INFINITY : fminf(dot3(L3(d.pos[a]), up), dot3(L3(d.pos[b]), up));
            k = shock_key44(h);
        }
        klo[r] = (uint32_t)k;
        khi[r] = (uint32_t)(k >> 32);
        vals[r] = r;
    }
}
// keys_out[i] = src[perm[i]]
__global__ void k_gather_keys(Dev d, const uint32_t *src, const uint32_t *perm, uint32_t *keys_out) {
    if (aborted(d)) return;
    const uint32_t nislc = min(d.cnt[CT_ISLC], d.cap_contacts);
    GRID_STRIDE(i, nislc) keys_out[i] = src[perm[i]];
}
__global__ void k_shock_order(Dev d, const uint32_t *perm) {
    if (aborted(d)) return;
    const uint32_t nislc = min(d.cnt[CT_ISLC], d.cap_contacts);
    GRID_STRIDE(i, nislc) d.sh_ord[i] = d.isl_v2[perm[i]];
}

// Readiness flags of the island being swept.  Tier 0 (up to 4,096 ranks) and tier 1 (up to
// 131,072 ranks) keep one byte per rank in LDS; tier 2 uses a u32 slice of a global array owned
// by the island.  The flag of a rank holds the token of the last sweep that executed it; tokens
// increase through an island's sweeps, so a new chain set needs no reset.  LDS is reached through
// LDS-typed pointers only (ds_* instructions): flat accesses into the LDS aperture beyond its
// first 64 KiB fault on gfx950.
#define RP_LDS __attribute__((address_space(3)))
struct LdsFlags {
    RP_LDS uint8_t *f;
    static constexpr uint32_t max_token = 255u;
    __device__ __forceinline__ uint32_t load(uint32_t i) const { return __hip_atomic_load(&f[i], __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_WORKGROUP); }
    __device__ __forceinline__ void store(uint32_t i, uint32_t t) const {
        __hip_atomic_store(&f[i], (uint8_t)t, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_WORKGROUP);
    }
    __device__ __forceinline__ void clear(uint32_t i) const { f[i] = 0; }
    __device__ __forceinline__ void set(uint32_t i, uint32_t v) const { f[i] = (uint8_t)v; }
    // decrements byte i (never below zero: at most its value is subtracted), returns its old value
    __device__ __forceinline__ uint32_t dec(uint32_t i) const {
        RP_LDS uint32_t *w = (RP_LDS uint32_t *)f + (i >> 2);
        const uint32_t sh = 8u * (i & 3u);
        return (__hip_atomic_fetch_sub(w, 1u << sh, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP) >> sh) & 255u;
    }
};
struct GlbFlags {
    uint32_t *f;
    static constexpr uint32_t max_token = 0xFFFFFFFEu;
    __device__ __forceinline__ uint32_t load(uint32_t i) const { return __hip_atomic_load(&f[i], __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_WORKGROUP); }
    __device__ __forceinline__ void store(uint32_t i, uint32_t t) const { __hip_atomic_store(&f[i], t, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_WORKGROUP); }
    __device__ __forceinline__ void clear(uint32_t i) const { f[i] = 0; }
    __device__ __forceinline__ void set(uint32_t i, uint32_t v) const { f[i] = v; }
    __device__ __forceinline__ uint32_t dec(uint32_t i) const { return __hip_atomic_fetch_sub(&f[i], 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP); }
};
template <int TIER> struct Tier;
// Tiers 0 and 3 are one wave per island: their islands (towers, chains, piles of up to 4,096
// entries) are mostly deep chains a few to a few tens of rows wide, so a single wave keeps up with
// their levels, and a level's rows run in one pass of one wave rather than a pass in each of the
// waves they fall in.  Tiers 1 and 2 are the big islands (wide levels, split form, phase clocks).
template <> struct Tier<0> { static constexpr uint32_t threads = 64, flag_bytes = 256; };
template <> struct Tier<3> { static constexpr uint32_t threads = 64, flag_bytes = 4096; };
template <> struct Tier<1> { static constexpr uint32_t threads = 1024, flag_bytes = 131072; };
template <> struct Tier<2> { static constexpr uint32_t threads = 1024, flag_bytes = 0; };
// Tier 4: islands with a compliant row or a friction patch (any size), one workgroup each, flags
// in HBM, velocities in HBM
template <> struct Tier<4> { static constexpr uint32_t threads = 256, flag_bytes = 0; };
template <int TIER> constexpr bool BIG = TIER == 1 || TIER == 2;
// Island sizes up to tp.tier_max[0] go to tier 0, then up to tier_max[1] to tier 3, up to
// tier_max[2] to tier 1, larger ones to tier 2.  The host keeps each bound within its tier's flag
// storage (defaults: the storage itself), so lowering them only moves islands to wider tiers.
static __device__ __forceinline__ int tier_of(uint32_t n, const TickParams &tp) {
    return n <= tp.tier_max[0] ? 0 : (n <= tp.tier_max[1] ? 3 : (n <= tp.tier_max[2] ? 1 : 2));
}
// The islands of each sweep tier (by contact count; islands with a compliant row or a patch to
// tier 4), so each tier's launch walks only its own.
__global__ void k_tier_lists(Dev d, TickParams tp) {
    if (aborted(d)) return;
    const uint32_t nisl = d.cnt[CT_ISLANDS], nislc = min(d.cnt[CT_ISLC], d.cap_contacts);
    GRID_STRIDE(isl, nisl) {
        const uint32_t n = ((isl + 1 < nisl) ? d.isl_start[isl + 1] : nislc) - d.isl_start[isl];
        const int t = (d.cm_on && d.isl_cm[isl]) ? 4 : tier_of(n, tp);
        d.tier_list[t][atomicAdd(&d.cnt[CT_TIER_N + t], 1u)] = isl;
    }
}

// The island's sweep state: segment [s, s + n) and the token counter.
template <class FL> struct Island {
    FL fl;
    uint32_t s, n, token;
    __device__ void begin() {
        for (uint32_t i = threadIdx.x; i < n; i += blockDim.x) fl.clear(i);
        token = 0;
        __syncthreads();
    }
};

// An order of the island as execution slots j = 0..n-1 (global offset s + j): the contact at
// each slot and its predecessors' slots (RP_NONE: none).  Slot order is a topological order of
// the chains (list order, or level order), and readiness flags are per slot.
struct Slots { uint32_t *pa, *pb, *c; };

// Writes the slots of the order whose chains are (pa, pb) over island ranks, whose contacts are
// ordc[s + rank], with ex[s + slot] = rank (ex == nullptr: slot = rank).
template <class FL>
static __device__ void build_slots(const Dev &d, const Island<FL> &I, const uint32_t *ordc, const uint32_t *pa, const uint32_t *pb,
                                   const uint32_t *ex, Slots o) {
    const uint32_t s = I.s, n = I.n;
    if (ex) {
        for (uint32_t j = threadIdx.x; j < n; j += blockDim.x) d.slot_of[s + ex[s + j]] = j;
        __syncthreads();
    }
    for (uint32_t j = threadIdx.x; j < n; j += blockDim.x) {
        const uint32_t r = ex ? ex[s + j] : j, a = pa[s + r], b = pb[s + r];
        o.pa[s + j] = a == RP_NONE ? RP_NONE : (ex ? d.slot_of[a] : a - s);
        o.pb[s + j] = b == RP_NONE ? RP_NONE : (ex ? d.slot_of[b] : b - s);
        o.c[s + j] = ordc[s + r];
    }
    __syncthreads();
}

// One sweep over the slots: each lane walks its slots j = threadIdx.x + k * blockDim in
// increasing order; it loads the slot's constant row with prep(contact, slot) at once, waits until
// both predecessor slots carry this sweep's token, then runs run(contact, row) and marks the
// slot.  The lowest unfinished slot is always ready, so every lane finishes.  A wave none of whose
// lanes is ready sleeps briefly, leaving its SIMD to the waves that run rows.
template <class FL, class P, class R>
static __device__ void run_sweep(Island<FL> &I, const Slots &sl, P prep, R run) {
    if (I.token >= FL::max_token) I.begin();
    const uint32_t tok = ++I.token, s = I.s, n = I.n;
    uint32_t j = threadIdx.x, a = RP_NONE, b = RP_NONE, c = 0;
    decltype(prep(0u, 0u)) x{};
    bool fresh = true;
    while (j < n) {
        if (fresh) { a = sl.pa[s + j]; b = sl.pb[s + j]; c = sl.c[s + j]; x = prep(c, s + j); fresh = false; }
        bool ready = (a == RP_NONE || I.fl.load(a) >= tok) && (b == RP_NONE || I.fl.load(b) >= tok);
        if (!__any(ready)) __builtin_amdgcn_s_sleep(1);  // the whole wave waits: yield the SIMD
        if (ready) {
            run(c, x);
            I.fl.store(j, tok);
            j += blockDim.x;
            fresh = true;
        }
    }
    __syncthreads();
}
// One sweep over slots [j0, n) of an island whose slots below j0 have run already (this sweep's
// wide levels, launched across the GPU): predecessors below j0 count as done.
template <class FL, class P, class R>
static __device__ void run_sweep_from(Island<FL> &I, const Slots &sl, uint32_t j0, P prep, R run) {
    if (I.token >= FL::max_token) I.begin();
    const uint32_t tok = ++I.token, s = I.s, n = I.n;
    uint32_t j = j0 + threadIdx.x, a = RP_NONE, b = RP_NONE, c = 0;
    decltype(prep(0u, 0u)) x{};
    bool fresh = true;
    while (j < n) {
        if (fresh) { a = sl.pa[s + j]; b = sl.pb[s + j]; c = sl.c[s + j]; x = prep(c, s + j); fresh = false; }
        bool ready = (a == RP_NONE || a < j0 || I.fl.load(a) >= tok) && (b == RP_NONE || b < j0 || I.fl.load(b) >= tok);
        if (!__any(ready)) __builtin_amdgcn_s_sleep(1);  // the whole wave waits: yield the SIMD
        if (ready) {
            run(c, x);
            I.fl.store(j, tok);
            j += blockDim.x;
            fresh = true;
        }
    }
    __syncthreads();
}
struct NoRow {};
template <class FL, class F>
static __device__ void run_sweep(Island<FL> &I, const Slots &sl, F exec) {
    run_sweep(I, sl, [](uint32_t, uint32_t) { return NoRow{}; }, [&](uint32_t c, const NoRow &) { exec(c); });
}

// Wall-clock time per sweep phase of the large-island tiers (thread 0 of each workgroup adds
// its share), read by the profiled step.
struct PhaseClock {
    uint64_t t;
    __device__ PhaseClock() : t(wall_clock64()) {}
    __device__ void lap(const Dev &d, int k) {
        uint64_t now = wall_clock64();
        if (threadIdx.x == 0) atomicAdd(&d.ph[k], (unsigned long long)(now - t));
        t = now;
    }
};

// Level order of the island for the chains (pa, pb) with successors (sa, sb): ranks are grouped
// by level(r) = 1 + max(level of its predecessors), the reference's batch index
// (solver.rs:621-642) within the island.  Ranks of one level write no common body, so executing
// them in any order gives the same bits; the sweeps then keep every lane busy on independent
// ranks.  Built by frontier expansion: ranks without predecessors form level 1, and a rank joins
// the next level when its last predecessor leaves the current one.  In-degrees live in the
// island's flag storage (cleared again on return); ex[s + slot] receives local ranks.  Returns ex,
// or nullptr (list order, always valid) should the expansion not place every rank.
template <class FL>
static __device__ const uint32_t *level_order(Island<FL> &I, const uint32_t *pa, const uint32_t *pb, const uint32_t *sa, const uint32_t *sb,
                                              uint32_t *ex, unsigned long long *rounds, uint32_t wide_min, uint32_t *wide_off = nullptr,
                                              uint32_t *wide_k = nullptr) {
    __shared__ uint32_t tail;
    const uint32_t s = I.s, n = I.n;
    if (threadIdx.x == 0) tail = 0;
    for (uint32_t r = threadIdx.x; r < n; r += blockDim.x) I.fl.set(r, (pa[s + r] != RP_NONE ? 1u : 0u) + (pb[s + r] != RP_NONE ? 1u : 0u));
    __syncthreads();
    for (uint32_t r = threadIdx.x; r < n; r += blockDim.x)
        if (pa[s + r] == RP_NONE && pb[s + r] == RP_NONE) ex[s + atomicAdd(&tail, 1u)] = r;
    __syncthreads();
    uint32_t fs = 0, fe = tail, nr = 0, nw = 0;
    bool leading = wide_off != nullptr;
    __syncthreads();  // every thread has the frontier's end before any appends to the next
    while (fs < fe) {
        // the island's leading run of wide levels: their start slots, then the run's end
        if (leading && fe - fs >= wide_min && nw < RP_WIDE_MAX_LEVELS) {
            if (threadIdx.x == 0) { wide_off[nw] = fs; wide_off[nw + 1] = fe; }
            nw++;
        } else {
            leading = false;
        }
        nr++;
        for (uint32_t j = fs + threadIdx.x; j < fe; j += blockDim.x) {
            const uint32_t r = ex[s + j];
            const uint32_t qa = sa[s + r], qb = sb[s + r];
            if (qa != RP_NONE && I.fl.dec(qa - s) == 1u) ex[s + atomicAdd(&tail, 1u)] = qa - s;
            if (qb != RP_NONE && I.fl.dec(qb - s) == 1u) ex[s + atomicAdd(&tail, 1u)] = qb - s;
        }
        __syncthreads();
        fs = fe;
        fe = tail;
        __syncthreads();
    }
    I.begin();
    if (threadIdx.x == 0) atomicAdd(rounds, (unsigned long long)nr);
    if (wide_k) *wide_k = fe == n ? nw : 0u;  // every thread stores the same count
    return fe == n ? ex : nullptr;
}

// solver.rs:455-464 apply_contact_impulse (+ dynamics.rs:193-205)
static __device__ __forceinline__ void apply_cimp(const Dev &d, uint32_t ai, uint32_t bi, float inv_a, float inv_b, V3 ra, V3 rb, V3 j) {
    if (inv_a > 0.0f) {
        V3 nj = -j;
        float im = d.inv_mass[ai];
        if (im > 0.0f && fin3(nj)) d.lin[ai] = S3(fin_v(L3(d.lin[ai]) + nj * im));
        V3 tau = cross3(ra, nj);
        if (d.inv_mass[ai] > 0.0f && fin3(tau)) d.ang[ai] = S3(fin_v(L3(d.ang[ai]) + mv(load_m9(d.iiw + 9 * ai), tau)));
    }
    if (inv_b > 0.0f) {
        float im = d.inv_mass[bi];
        if (im > 0.0f && fin3(j)) d.lin[bi] = S3(fin_v(L3(d.lin[bi]) + j * im));
        V3 tau = cross3(rb, j);
        if (d.inv_mass[bi] > 0.0f && fin3(tau)) d.ang[bi] = S3(fin_v(L3(d.ang[bi]) + mv(load_m9(d.iiw + 9 * bi), tau)));
    }
}
static __device__ __forceinline__ V3 rel_vel(const Dev &d, uint32_t ai, uint32_t bi, V3 ra, V3 rb) {
    V3 va = L3(d.lin[ai]) + cross3(L3(d.ang[ai]), ra);
    V3 vb = bi == RP_NONE ? splat3(0.0f) : L3(d.lin[bi]) + cross3(L3(d.ang[bi]), rb);
    return vb - va;
}
// A contact row's constants during the velocity solve: everything but the bodies' velocities
// and the row's accumulators (world inverse inertia is fixed until integration).
struct Row {
    uint32_t ai, bi;
    uint32_t la, lb;  // LDS velocity slots of the bodies the row writes (RP_NONE: the body's own, in HBM)
    float inv_a, inv_b, ima, imb;
    V3 ra, rb, n, t1, t2;
    float4 m4, mu4;
    M3 iia, iib;
};
static __device__ __forceinline__ Row load_row(const Dev &d, uint32_t c) {
    Row x;
    x.ai = d.cs_ai[c]; x.bi = d.cs_bi[c];
    x.la = RP_NONE; x.lb = RP_NONE;
    float2 inv = d.cs_inv[c];
    x.inv_a = inv.x; x.inv_b = inv.y;
    x.ra = L3(d.cs_ra[c]); x.rb = L3(d.cs_rb[c]); x.n = L3(d.cs_n[c]); x.t1 = L3(d.cs_t1[c]); x.t2 = L3(d.cs_t2[c]);
    x.m4 = d.cs_m[c]; x.mu4 = d.cs_mu[c];
    x.ima = d.inv_mass[x.ai];
    x.iia = load_m9(d.iiw + 9 * x.ai);
    if (x.bi != RP_NONE) { x.imb = d.inv_mass[x.bi]; x.iib = load_m9(d.iiw + 9 * x.bi); }
    else { x.imb = 0.0f; x.iib = mzero(); }
    return x;
}
// Rows stored per execution slot, component-major (srow[k * cap + slot], k < RP_SROW), so the
// lanes of a wave load consecutive addresses and a row is one round of independent loads.
static __device__ __forceinline__ float4 f4u(uint32_t a, uint32_t b, float z, float w) { return make_float4(__uint_as_float(a), __uint_as_float(b), z, w); }
// o: the row's first component; C: the distance between components
static __device__ __forceinline__ void store_row_at(float4 *o, size_t C, const Row &x) {
    o[0] = S3(x.ra, x.m4.x); o[C] = S3(x.rb, x.m4.y); o[2 * C] = S3(x.n, x.m4.z); o[3 * C] = S3(x.t1, x.m4.w);
    o[4 * C] = S3(x.t2, x.mu4.x); o[5 * C] = make_float4(x.mu4.y, x.mu4.z, x.inv_a, x.inv_b); o[6 * C] = f4u(x.ai, x.bi, x.ima, x.imb);
    o[7 * C] = make_float4(x.iia.c0.x, x.iia.c0.y, x.iia.c0.z, x.iia.c1.x); o[8 * C] = make_float4(x.iia.c1.y, x.iia.c1.z, x.iia.c2.x, x.iia.c2.y);
    o[9 * C] = make_float4(x.iia.c2.z, x.iib.c0.x, x.iib.c0.y, x.iib.c0.z); o[10 * C] = make_float4(x.iib.c1.x, x.iib.c1.y, x.iib.c1.z, x.iib.c2.x);
    o[11 * C] = make_float4(x.iib.c2.y, x.iib.c2.z, __uint_as_float(x.la), __uint_as_float(x.lb));
}
static __device__ __forceinline__ Row fetch_row_at(const float4 *o, size_t C) {
    float4 f0 = o[0], f1 = o[C], f2 = o[2 * C], f3 = o[3 * C], f4 = o[4 * C], f5 = o[5 * C], f6 = o[6 * C];
    float4 f7 = o[7 * C], f8 = o[8 * C], f9 = o[9 * C], f10 = o[10 * C], f11 = o[11 * C];
    Row x;
    x.ra = L3(f0); x.rb = L3(f1); x.n = L3(f2); x.t1 = L3(f3); x.t2 = L3(f4);
    x.m4 = make_float4(f0.w, f1.w, f2.w, f3.w);
    x.mu4 = make_float4(f4.w, f5.x, f5.y, 0.0f);
    x.inv_a = f5.z; x.inv_b = f5.w;
    x.ai = __float_as_uint(f6.x); x.bi = __float_as_uint(f6.y); x.ima = f6.z; x.imb = f6.w;
    x.iia.c0 = mk3(f7.x, f7.y, f7.z); x.iia.c1 = mk3(f7.w, f8.x, f8.y); x.iia.c2 = mk3(f8.z, f8.w, f9.x);
    x.iib.c0 = mk3(f9.y, f9.z, f9.w); x.iib.c1 = mk3(f10.x, f10.y, f10.z); x.iib.c2 = mk3(f10.w, f11.x, f11.y);
    x.la = __float_as_uint(f11.z); x.lb = __float_as_uint(f11.w);
    return x;
}
static __device__ __forceinline__ void store_row(const Dev &d, uint32_t slot, const Row &x) { store_row_at(d.srow + slot, d.cap_contacts, x); }
static __device__ __forceinline__ Row fetch_row(const Dev &d, uint32_t slot) { return fetch_row_at(d.srow + slot, d.cap_contacts); }
static __device__ __forceinline__ bool same_bits(V3 a, V3 b) {
    return __float_as_uint(a.x) == __float_as_uint(b.x) && __float_as_uint(a.y) == __float_as_uint(b.y) && __float_as_uint(a.z) == __float_as_uint(b.z);
}
// Velocities of a row's two bodies held in registers across its normal and friction updates.
struct Vel2 { V3 la, aa, lb, ab; };
static __device__ __forceinline__ V3 rel_vel2(const Row &x, const Vel2 &v) {
    V3 va = v.la + cross3(v.aa, x.ra);
    V3 vb = x.bi == RP_NONE ? splat3(0.0f) : v.lb + cross3(v.ab, x.rb);
    return vb - va;
}
// apply_cimp on the register copies (same operations and the same sanitised values as stored)
static __device__ __forceinline__ V3 sel3(bool c, V3 a, V3 b) { return mk3(c ? a.x : b.x, c ? a.y : b.y, c ? a.z : b.z); }
// Every candidate is computed and the kept ones selected: the values of the conditional form.
static __device__ __forceinline__ void apply_row2(const Row &x, Vel2 &v, V3 j) {
    const V3 nj = -j;
    const bool wa = x.inv_a > 0.0f && x.ima > 0.0f, wb = x.inv_b > 0.0f && x.imb > 0.0f;
    const V3 ta = cross3(x.ra, nj), tb = cross3(x.rb, j);
    const V3 la = fin_v(v.la + nj * x.ima), aa = fin_v(v.aa + mv(x.iia, ta));
    const V3 lb = fin_v(v.lb + j * x.imb), ab = fin_v(v.ab + mv(x.iib, tb));
    v.la = sel3(wa && fin3(nj), la, v.la);
    v.aa = sel3(wa && fin3(ta), aa, v.aa);
    v.lb = sel3(wb && fin3(j), lb, v.lb);
    v.ab = sel3(wb && fin3(tb), ab, v.ab);
}
// apply_cimp with the row's prefetched masses and inertias (same operations)
static __device__ __forceinline__ void apply_row(const Dev &d, const Row &x, V3 j) {
    if (x.inv_a > 0.0f) {
        V3 nj = -j;
        if (x.ima > 0.0f && fin3(nj)) d.lin[x.ai] = S3(fin_v(L3(d.lin[x.ai]) + nj * x.ima));
        V3 tau = cross3(x.ra, nj);
        if (x.ima > 0.0f && fin3(tau)) d.ang[x.ai] = S3(fin_v(L3(d.ang[x.ai]) + mv(x.iia, tau)));
    }
    if (x.inv_b > 0.0f) {
        if (x.imb > 0.0f && fin3(j)) d.lin[x.bi] = S3(fin_v(L3(d.lin[x.bi]) + j * x.imb));
        V3 tau = cross3(x.rb, j);
        if (x.imb > 0.0f && fin3(tau)) d.ang[x.bi] = S3(fin_v(L3(d.ang[x.bi]) + mv(x.iib, tau)));
    }
}

// Islands of this tier, one per workgroup (grid-stride); I is the island's sweep state with its
// readiness flags, lst its contacts in list order.
#define ISLAND_LOOP_BEGIN ISLAND_LOOP_BEGIN_FROM(d.tier_list[TIER], d.cnt[CT_TIER_N + TIER])
#define ISLAND_LOOP_BEGIN_FROM(LIST, COUNT)                                                        \
    __shared__ __attribute__((aligned(4))) uint8_t flag_lds[Tier<TIER>::flag_bytes > 0 ? Tier<TIER>::flag_bytes : 4]; \
    const uint32_t nisl = d.cnt[CT_ISLANDS], nislc = min(d.cnt[CT_ISLC], d.cap_contacts);         \
    const uint32_t *const tlist = (LIST), ntier = (COUNT);                                         \
    for (uint32_t q = blockIdx.x; q < ntier; q += gridDim.x) {                                     \
        const uint32_t isl = tlist[q];                                                             \
        const uint32_t s = d.isl_start[isl], n = ((isl + 1 < nisl) ? d.isl_start[isl + 1] : nislc) - s; \
        const uint32_t *lst = d.isl_v2 + s;                                                        \
        auto I = make_island<TIER>(d, flag_lds, s, n);                                             \
        I.begin();
#define ISLAND_LOOP_END }

template <int TIER>
static __device__ __forceinline__ auto make_island(const Dev &d, uint8_t *flag_lds, uint32_t s, uint32_t n) {
    if constexpr (TIER != 2 && TIER != 4) return Island<LdsFlags>{LdsFlags{(RP_LDS uint8_t *)flag_lds}, s, n, 0};
    else return Island<GlbFlags>{GlbFlags{d.g_done + s}, s, n, 0};
}

// ------------------------------------------------------------------ ball joints
// A ball joint locks the three translational axes of frame A (articulation/src/bodies.rs:326-331).
// Velocity level: one equality row per axis, target 0, unbounded (solver.rs:336-424 d6_rows with
// no limits, drives, damping, friction or armature), warm-started from last tick's rows when the
// same rows exist (solver.rs:700-715) and stored after the sweeps (solver.rs:945-952).  Position
// level: solve_d6's translational rows with SHAKE directions (bodies.rs:616-739), one pass per
// projection iteration before the contacts (step.rs:302-306).  A joint's rows write both its
// bodies when their inverse mass is positive, asleep or not (joint_row_apply, solver.rs:298-307;
// apply_position_delta / apply_orientation_delta, dynamics.rs:71-89, 207-210).
static __device__ __forceinline__ float jfin(float x) { return isfinite(x) ? x : 0.0f; }
static __device__ __forceinline__ M3 inertia_world_of(M3 il, Q4 q) {
    M3 r = mquat(q);
    return mmul(mmul(r, il), mtrans(r));
}
// dynamics.rs:193-205 apply_impulse(b, j), then apply_angular_impulse(b, r x j)
static __device__ __forceinline__ void joint_impulse(const Dev &d, uint32_t b, V3 j, V3 r) {
    const float im = d.inv_mass[b];
    if (im > 0.0f && fin3(j)) d.lin[b] = S3(fin_v(L3(d.lin[b]) + j * im));
    const V3 tau = cross3(r, j);
    if (im > 0.0f && fin3(tau)) d.ang[b] = S3(fin_v(L3(d.ang[b]) + mv(load_m9(d.iiw + 9 * b), tau)));
}
// solver.rs:298-307 joint_row_apply for a translational row
static __device__ __forceinline__ void joint_apply(const Dev &d, uint32_t a, uint32_t b, V3 ra, V3 rb, V3 u, float lam) {
    const V3 p = u * lam;
    if (d.inv_mass[a] > 0.0f) joint_impulse(d, a, -p, ra);
    if (d.inv_mass[b] > 0.0f) joint_impulse(d, b, p, rb);
}
// solver.rs:291-295 joint_row_vel
static __device__ __forceinline__ float joint_vel(const Dev &d, uint32_t a, uint32_t b, V3 ra, V3 rb, V3 u) {
    const V3 va = L3(d.lin[a]) + cross3(L3(d.ang[a]), ra);
    const V3 vb = L3(d.lin[b]) + cross3(L3(d.ang[b]), rb);
    return dot3(vb - va, u);
}
// warm start of joint q's rows
static __device__ void joint_warm(const Dev &d, uint32_t q) {
    const uint32_t m = d.jr_mask[q], wm = d.wj_mask[q];
    if (!m || !(wm >> 31) || (wm & 7u) != m) return;
    const uint32_t a = d.jt_a[q], b = d.jt_b[q];
    const V3 ra = L3(d.jr_ra[q]), rb = L3(d.jr_rb[q]);
    for (uint32_t i = 0; i < 3; i++) {
        if (!((m >> i) & 1u)) continue;
        const float lam = d.wj_lam[3 * q + i];  // clamped to the row's (-inf, inf): unchanged
        d.jr_acc[3 * q + i] = lam;
        if (lam != 0.0f) joint_apply(d, a, b, ra, rb, L3(d.jr_u[3 * q + i]), lam);
    }
}
// Where a sweep gathers its largest row residual (residual mode): an LDS word, and whether to.
// The flag is explicit: an LDS variable can sit at address 0, which compares equal to a null
// address_space(3) pointer (a kernel whose only LDS word is this one lost its residuals that way).
struct ResidSink { RP_LDS uint32_t *p; bool on; };
// One PGS pass over joint q's rows (solver.rs:822-832).  Residual mode: returns the bits of the
// largest row residual |d lambda| K (non-negative floats order as their bits; NaN above all).
static __device__ uint32_t joint_pgs(const Dev &d, uint32_t q, bool resid) {
    const uint32_t m = d.jr_mask[q], a = d.jt_a[q], b = d.jt_b[q];
    const V3 ra = L3(d.jr_ra[q]), rb = L3(d.jr_rb[q]);
    uint32_t rbits = 0;
    for (uint32_t i = 0; i < 3; i++) {
        if (!((m >> i) & 1u)) continue;
        const float4 u4 = d.jr_u[3 * q + i];
        const V3 u = L3(u4);
        const float v = joint_vel(d, a, b, ra, rb, u);
        const float old = d.jr_acc[3 * q + i];
        const float nw = old + (0.0f - v) * u4.w;
        d.jr_acc[3 * q + i] = nw;
        if (nw != old) joint_apply(d, a, b, ra, rb, u, nw - old);
        if (resid) rbits = max(rbits, __float_as_uint(fabsf(nw - old) / u4.w));
    }
    return rbits;
}
// joint q's rows become its warm rows (a joint without rows keeps its old ones)
static __device__ __forceinline__ void joint_store(const Dev &d, uint32_t q) {
    if (dist_is(d, q)) return;  // a distance row is not carried
    if (d6_is(d, q)) { d6_store(d, q); return; }
    const uint32_t m = d.jr_mask[q];
    if (!m) return;
    d.wj_mask[q] = 0x80000000u | m;
    for (uint32_t i = 0; i < 3; i++) if ((m >> i) & 1u) d.wj_lam[3 * q + i] = d.jr_acc[3 * q + i];
}
static __device__ __forceinline__ V3 normalize_fin(V3 v) {  // bodies.rs:19-27
    v = fin_v(v);
    const float l2 = len2(v);
    return l2 > 1.0e-12f ? v * (1.0f / sqrtf(l2)) : splat3(0.0f);
}
// dynamics.rs:207-210
static __device__ __forceinline__ void joint_pos_delta(const Dev &d, uint32_t b, V3 dp) {
    if (d.inv_mass[b] > 0.0f && fin3(dp)) d.pos[b] = S3(fin_v(L3(d.pos[b]) + dp));
}
// dynamics.rs:71-89
static __device__ __forceinline__ void joint_orient_delta(const Dev &d, uint32_t b, V3 dth) {
    if (d.inv_mass[b] <= 0.0f || !fin3(dth)) return;
    const V3 x = fin_v(dth);
    if (len2(x) <= 1.0e-24f) return;
    const Q4 dq = qnormalize(mkq(x.x * 0.5f, x.y * 0.5f, x.z * 0.5f, 1.0f));
    d.rot[b] = SQ(fin_q(qnormalize(qmul(dq, LQ(d.rot[b])))));
}
#include "rp_d6.h"
// One position pass over joint q (solve_step_with, bodies.rs:422-460, one iteration): skipped when
// both bodies are static or asleep; SHAKE directions from the step's start poses (the poses before
// integration) unless a body turned by 0.5 rad or more since.  A D6 joint runs solve_d6 (rp_d6.h).
static __device__ void joint_pos(const Dev &d, uint32_t q, float dt) {
    const uint32_t a = d.jt_a[q], b = d.jt_b[q];
    const bool fa = !((d.bflags[a] & BF_DYN) && d.awake[a]), fb = !((d.bflags[b] & BF_DYN) && d.awake[b]);
    if (fa && fb) return;
    if (d6_is(d, q)) { d6_pos(d, q, dt); return; }
    if (dist_is(d, q)) { dist_pos(d, q, dt); return; }
    const float wa = jfin(d.inv_mass[a]), wb = jfin(d.inv_mass[b]);
    const Q4 sa = LQ(d.orot[a]), sb = LQ(d.orot[b]);
    const float cmin = __uint_as_float(RP_SHAKE_COS_BITS);
    const bool shake = fminf(fabsf(qmul(sa, qconj(LQ(d.rot[a]))).w), 1.0f) >= cmin && fminf(fabsf(qmul(sb, qconj(LQ(d.rot[b]))).w), 1.0f) >= cmin;
    const Q4 fra = LQ(d.jt_far[q]);
    const V3 fap = L3(d.jt_fap[q]), fbp = L3(d.jt_fbp[q]);
    V3 la0 = splat3(0.0f), lb0 = splat3(0.0f);
    M3 ia0 = mzero(), ib0 = mzero();
    Q4 ra0 = qid();
    if (shake) {
        ra0 = qmul(sa, fra);
        la0 = fin_v(qrot(sa, fap)); lb0 = fin_v(qrot(sb, fbp));
        // bodies.rs:569-573 inv_inertia_at (the local inverse inertia of a static body is zero)
        ia0 = inertia_world_of(d.inv_mass[a] == 0.0f ? mzero() : load_m9(d.iil + 9 * a), sa);
        ib0 = inertia_world_of(d.inv_mass[b] == 0.0f ? mzero() : load_m9(d.iil + 9 * b), sb);
    }
    for (uint32_t i = 0; i < 3; i++) {
        const V3 pa = L3(d.pos[a]), pb = L3(d.pos[b]);
        const Q4 qa = LQ(d.rot[a]), qb = LQ(d.rot[b]);
        const V3 u = normalize_fin(qrot(qmul(qa, fra), unit_axis(i)));
        if (zero3(u)) continue;
        const V3 xa = fin_v(pa + qrot(qa, fap)), xb = fin_v(pb + qrot(qb, fbp));
        const V3 dd = fin_v(xb - xa);
        V3 un, lva, lvb; M3 iaw, ibw;
        if (shake) { un = normalize_fin(qrot(ra0, unit_axis(i))); lva = la0; lvb = lb0; iaw = ia0; ibw = ib0; }
        else { un = u; lva = xa - pa; lvb = xb - pb; iaw = inv_inertia_world(d, a, qa); ibw = inv_inertia_world(d, b, qb); }
        float c = jfin(jfin(dot3(dd, u)) - 0.0f);
        const V3 rau = fin_v(cross3(lva, un)), rbu = fin_v(cross3(lvb, un));
        const float wra = jfin(dot3(rau, mv(iaw, rau))), wrb = jfin(dot3(rbu, mv(ibw, rbu)));
        const float wsum = jfin(((wa + wb) + wra) + wrb);
        const float denom = jfin(wsum + 0.0f);
        if (denom <= 0.0f) continue;
        c = jfin(c + 0.0f);
        const float dl = jfin(-(c + 0.0f) / denom);
        joint_pos_delta(d, a, fin_v((-un) * jfin(dl * wa)));
        joint_pos_delta(d, b, fin_v(un * jfin(dl * wb)));
        joint_orient_delta(d, a, fin_v((-mv(iaw, rau)) * dl));
        joint_orient_delta(d, b, fin_v(mv(ibw, rbu) * dl));
    }
}

// Warm-start application to contact c (solver.rs:749-760): the previous frame's impulses (mode 0,
// the reference), nothing (mode 1), or the previous frame's plus the model's correction where the
// contact passed the model's gate (mode 2; other contacts take the reference's warm start, and a
// zero correction leaves every operation of the reference's).
static __device__ __forceinline__ void warm_row(const Dev &d, const TickParams &tp, uint32_t c) {
    if (c & RP_JBIT) { joint_warm(d, c & ~RP_JBIT); return; }
    const uint32_t cur = tp.cur;
    // a patch's point takes its carried normal impulse only (its friction is the patch's,
    // solver.rs:752); compliant rows were matched to nothing.  Both start as the reference's in
    // the learned mode too.
    const uint32_t cmf = cm_flags(d, c);
    if (tp.ws_mode == 2 && !cmf) {
        const int wi = d.warm_idx[c];
        const float4 pr = d.pred[c];
        const bool use = pr.w != 0.0f;
        if (!use && wi < 0) return;
        uint32_t ai = d.cs_ai[c], bi = d.cs_bi[c];
        V3 ra = L3(d.cs_ra[c]), rb = L3(d.cs_rb[c]), n3 = L3(d.cs_n[c]), t1 = L3(d.cs_t1[c]), t2 = L3(d.cs_t2[c]);
        float2 inv = d.cs_inv[c];
        float jn = 0.0f, x1 = 0.0f, x2 = 0.0f;
        if (wi >= 0) {
            jn = d.wk_jn[cur][wi];
            V3 ejt = L3(d.wk_jt[cur][wi]);
            x1 = dot3(ejt, t1); x2 = dot3(ejt, t2);
        }
        if (use) {
            if (pr.x != 0.0f) jn = fmaxf(jn + pr.x, 0.0f);
            if (pr.y != 0.0f) x1 += pr.y;
            if (pr.z != 0.0f) x2 += pr.z;
        }
        V3 v = rel_vel(d, ai, bi, ra, rb);
        float vt1 = dot3(v, t1), vt2 = dot3(v, t2);
        float mu = mu_dyn(d.cs_mu[c], sqrtf(vt1 * vt1 + vt2 * vt2));
        V3 t = clamp_len_max(mk3(x1, x2, 0.0f), mu * jn);
        d.jn[c] = jn;
        d.jt[c] = make_float2(t.x, t.y);
        if (wi < 0 && jn == 0.0f && t.x == 0.0f && t.y == 0.0f) return;
        V3 imp = n3 * jn;
        imp = imp + (t1 * t.x + t2 * t.y);
        apply_cimp(d, ai, bi, inv.x, inv.y, ra, rb, imp);
    } else if (tp.ws_mode != 1) {
        int wi = d.warm_idx[c];
        if (wi < 0) return;
        uint32_t ai = d.cs_ai[c], bi = d.cs_bi[c];
        V3 ra = L3(d.cs_ra[c]), rb = L3(d.cs_rb[c]), n3 = L3(d.cs_n[c]), t1 = L3(d.cs_t1[c]), t2 = L3(d.cs_t2[c]);
        float2 inv = d.cs_inv[c];
        float ejn = d.wk_jn[cur][wi];
        V3 imp = n3 * ejn;
        if (cmf & CM_PATCH_PT) {
            d.jn[c] = ejn;
            apply_cimp(d, ai, bi, inv.x, inv.y, ra, rb, imp);
            return;
        }
        V3 v = rel_vel(d, ai, bi, ra, rb);
        float vt1 = dot3(v, t1), vt2 = dot3(v, t2);
        float mu = mu_dyn(d.cs_mu[c], sqrtf(vt1 * vt1 + vt2 * vt2));
        V3 ejt = L3(d.wk_jt[cur][wi]);
        V3 t = clamp_len_max(mk3(dot3(ejt, t1), dot3(ejt, t2), 0.0f), mu * ejn);
        d.jt[c] = make_float2(t.x, t.y);
        imp = imp + (t1 * t.x + t2 * t.y);
        apply_cimp(d, ai, bi, inv.x, inv.y, ra, rb, imp);
    }
}

// A ball joint's rows in the layout of a contact row: the lever arms in ra, rb; axes 0, 1, 2 in n,
// t1, t2 with their 1/K in m4.x, m4.y, m4.z; the row mask (bits) in m4.w; the bodies, inverse
// masses and world inverse inertias where the contact keeps its own (inv_a, inv_b = the inverse
// masses: a joint writes a body exactly when its inverse mass is positive, joint_row_apply,
// solver.rs:298-307).  apply_row2 with the impulse u d lambda is joint_row_apply's pair of
// apply_impulse and apply_angular_impulse, the same operations.
static __device__ __forceinline__ Row load_jrow(const Dev &d, uint32_t q) {
    Row x;
    x.ai = d.jt_a[q]; x.bi = d.jt_b[q];
    x.la = RP_NONE; x.lb = RP_NONE;
    x.ima = d.inv_mass[x.ai]; x.imb = d.inv_mass[x.bi];
    x.inv_a = x.ima; x.inv_b = x.imb;
    x.ra = L3(d.jr_ra[q]); x.rb = L3(d.jr_rb[q]);
    const uint32_t m = d.jr_mask[q];
    const float4 u0 = (m & 1u) ? d.jr_u[3 * q] : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    const float4 u1 = (m & 2u) ? d.jr_u[3 * q + 1] : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    const float4 u2 = (m & 4u) ? d.jr_u[3 * q + 2] : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    x.n = L3(u0); x.t1 = L3(u1); x.t2 = L3(u2);
    x.m4 = make_float4(u0.w, u1.w, u2.w, __uint_as_float(m));
    x.mu4 = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    x.iia = load_m9(d.iiw + 9 * x.ai); x.iib = load_m9(d.iiw + 9 * x.bi);
    return x;
}
// ------------------------------------------------------------------ island velocities in LDS
// The one-wave and four-wave tiers keep the velocities of the bodies their island writes in LDS
// during the PGS iterations (up to VB bodies; any beyond stay in HBM): a row's wait and its
// velocity loads are then LDS round trips.  The warm start and the shock pass work in HBM, before
// the velocities are loaded and after they are written back.
template <uint32_t VB> struct LdsVel {
    RP_LDS float *lin, *ang;  // 4 floats per slot: the HBM row's bits
    RP_LDS uint32_t *body;
};
static __device__ __forceinline__ V3 lds_v3(const RP_LDS float *p) { return mk3(p[0], p[1], p[2]); }
static __device__ __forceinline__ void lds_put4(RP_LDS float *p, float4 v) { p[0] = v.x; p[1] = v.y; p[2] = v.z; p[3] = v.w; }
static __device__ __forceinline__ float4 lds_get4(const RP_LDS float *p) { return make_float4(p[0], p[1], p[2], p[3]); }
// The body island-list entry pos writes on side k (the chains' key, k_chain_entries).
static __device__ __forceinline__ uint32_t entry_body(const Dev &d, uint32_t pos, uint32_t k) {
    const uint32_t c = d.isl_v2[pos];
    if (c & RP_JBIT) {
        const uint32_t b = k ? d.jt_b[c & ~RP_JBIT] : d.jt_a[c & ~RP_JBIT];
        return d.inv_mass[b] > 0.0f ? b : RP_NONE;
    }
    return k ? d.cs_wb[c] : d.cs_wa[c];
}
// Slots for the island [s, s + n)'s written bodies: each chain head (an entry side with a written
// body and no predecessor) takes the next slot, loads its body's velocities and walks its body's
// chain, recording the slot at every entry side it reaches (lid[2 pos + k]; RP_NONE: HBM).
template <uint32_t VB>
static __device__ void lds_vel_begin(const Dev &d, uint32_t s, uint32_t n, const LdsVel<VB> &L, uint32_t *nloc) {
    if (threadIdx.x == 0) *nloc = 0;
    for (uint32_t e = threadIdx.x; e < 2 * n; e += blockDim.x) d.lid[2 * s + e] = RP_NONE;
    __syncthreads();
    for (uint32_t e = threadIdx.x; e < 2 * n; e += blockDim.x) {
        const uint32_t r = e >> 1, k = e & 1u;
        if ((k ? d.ch_pb : d.ch_pa)[s + r] != RP_NONE) continue;
        const uint32_t body = entry_body(d, s + r, k);
        if (body == RP_NONE) continue;
        const uint32_t id = atomicAdd(nloc, 1u), slot = id < VB ? id : RP_NONE;
        if (id < VB) { L.body[id] = body; lds_put4(L.lin + 4 * id, d.lin[body]); lds_put4(L.ang + 4 * id, d.ang[body]); }
        uint32_t cr = r, ck = k;
        while (true) {
            d.lid[2 * (s + cr) + ck] = slot;
            const uint32_t q = (ck ? d.ch_sb : d.ch_sa)[s + cr];
            if (q == RP_NONE) break;
            cr = q - s;
            ck = entry_body(d, q, 0) == body ? 0u : 1u;
        }
    }
    __syncthreads();
}
template <uint32_t VB>
static __device__ void lds_vel_end(const Dev &d, const LdsVel<VB> &L, uint32_t nloc) {
    __syncthreads();
    for (uint32_t i = threadIdx.x; i < min(nloc, VB); i += blockDim.x) {
        d.lin[L.body[i]] = lds_get4(L.lin + 4 * i);
        d.ang[L.body[i]] = lds_get4(L.ang + 4 * i);
    }
    __syncthreads();
}
template <uint32_t VB>
static __device__ __forceinline__ Vel2 load_vel2_l(const Dev &d, const Row &x, const LdsVel<VB> &L) {
    Vel2 v;
    if (x.la < VB) { v.la = lds_v3(L.lin + 4 * x.la); v.aa = lds_v3(L.ang + 4 * x.la); }
    else { v.la = L3(d.lin[x.ai]); v.aa = L3(d.ang[x.ai]); }
    if (x.bi == RP_NONE) { v.lb = splat3(0.0f); v.ab = splat3(0.0f); }
    else if (x.lb < VB) { v.lb = lds_v3(L.lin + 4 * x.lb); v.ab = lds_v3(L.ang + 4 * x.lb); }
    else { v.lb = L3(d.lin[x.bi]); v.ab = L3(d.ang[x.bi]); }
    return v;
}
template <uint32_t VB>
static __device__ __forceinline__ void store_vel2_l(const Dev &d, const Row &x, const Vel2 &v, const Vel2 &v0, const LdsVel<VB> &L) {
    if (!same_bits(v.la, v0.la)) { if (x.la < VB) lds_put4(L.lin + 4 * x.la, S3(v.la)); else d.lin[x.ai] = S3(v.la); }
    if (!same_bits(v.aa, v0.aa)) { if (x.la < VB) lds_put4(L.ang + 4 * x.la, S3(v.aa)); else d.ang[x.ai] = S3(v.aa); }
    if (x.bi == RP_NONE) return;
    if (!same_bits(v.lb, v0.lb)) { if (x.lb < VB) lds_put4(L.lin + 4 * x.lb, S3(v.lb)); else d.lin[x.bi] = S3(v.lb); }
    if (!same_bits(v.ab, v0.ab)) { if (x.lb < VB) lds_put4(L.ang + 4 * x.lb, S3(v.ab)); else d.ang[x.bi] = S3(v.ab); }
}
// PGS row of contact c (solver.rs:839-879): its constant row and accumulators are fetched, before
// the row's wait, by pgs_fetch; pgs_row keeps the two bodies' velocities in registers between the
// normal and the friction update and stores only the components that changed.  resid_bits (LDS)
// gathers the row residual of residual mode.
struct PgsIn { Row x; float jn; float2 jt; };
static __device__ __forceinline__ PgsIn pgs_fetch(const Dev &d, uint32_t c, uint32_t slot) {
    PgsIn in;
    if (c & RP_JBIT) { in.jn = 0.0f; in.jt = make_float2(0.0f, 0.0f); in.x.ai = RP_NONE; in.x.bi = RP_NONE; return in; }
    in.x = fetch_row(d, slot);
    in.jn = d.jn[c];
    in.jt = d.jt[c];
    return in;
}
static __device__ __forceinline__ Vel2 load_vel2(const Dev &d, const Row &x) {
    Vel2 v;
    v.la = L3(d.lin[x.ai]); v.aa = L3(d.ang[x.ai]);
    if (x.bi != RP_NONE) { v.lb = L3(d.lin[x.bi]); v.ab = L3(d.ang[x.bi]); } else { v.lb = splat3(0.0f); v.ab = splat3(0.0f); }
    return v;
}
// stores what changed (a body the row does not write keeps its bits)
static __device__ __forceinline__ void store_vel2(const Dev &d, const Row &x, const Vel2 &v, const Vel2 &v0) {
    if (!same_bits(v.la, v0.la)) d.lin[x.ai] = S3(v.la);
    if (!same_bits(v.aa, v0.aa)) d.ang[x.ai] = S3(v.aa);
    if (x.bi != RP_NONE) {
        if (!same_bits(v.lb, v0.lb)) d.lin[x.bi] = S3(v.lb);
        if (!same_bits(v.ab, v0.ab)) d.ang[x.bi] = S3(v.ab);
    }
}
// One entry's update on register copies of its bodies' velocities v and its accumulators acc: a
// contact (solver.rs:839-879; acc = jn, jt) or, with joint set, a ball joint's rows in a contact
// row's layout (solver.rs:822-832; acc = the three rows' lambdas).  The two run as three phases
// sharing their relative velocities and impulse applications: the normal row / axis 0, the
// friction pair / axis 1, axis 2 (joints only), each with its own operations.  Residual mode:
// returns the bits of the entry's residual, the oracle's operations (a contact's largest of its
// normal and friction residuals, a joint's largest row residual |d lambda| K).
static __device__ __forceinline__ uint32_t entry_core(const Row &x, Vel2 &v, float4 &acc, bool joint, bool resid) {
    const uint32_t m = joint ? __float_as_uint(x.m4.w) : 7u;
    // normal row: jn = max(jn + (bias - v.n) m, 0), applied; axis 0: lambda += (0 - v.u) K, applied if changed
    const float vn = dot3(rel_vel2(x, v), x.n);
    const float old0 = acc.x;
    float nw0 = old0 + ((joint ? 0.0f : x.m4.w) - vn) * x.m4.x;
    if (!joint) nw0 = fmaxf(nw0, 0.0f);
    if (!(m & 1u)) nw0 = old0;
    acc.x = nw0;
    if (!joint || nw0 != old0) apply_row2(x, v, x.n * (nw0 - old0));
    // friction: the tangent impulses clamped to mu jn together; axis 1
    const V3 vf = rel_vel2(x, v);
    const float vt1 = dot3(vf, x.t1);
    V3 imp;
    float c1, c2 = 0.0f;
    bool ap1;
    if (joint) {
        const float old1 = acc.y;
        float nw1 = old1 + (0.0f - vt1) * x.m4.y;
        if (!(m & 2u)) nw1 = old1;
        acc.y = nw1;
        c1 = nw1 - old1;
        imp = x.t1 * c1;
        ap1 = nw1 != old1;
    } else {
        const float vt2 = dot3(vf, x.t2);
        const float mu = mu_dyn(x.mu4, sqrtf(vt1 * vt1 + vt2 * vt2));
        const V3 ot = mk3(acc.y, acc.z, 0.0f);
        const V3 nt = clamp_len_max(ot - mk3(vt1 * x.m4.y, vt2 * x.m4.z, 0.0f), mu * nw0);
        acc.y = nt.x;
        acc.z = nt.y;
        const V3 dd = nt - ot;
        c1 = dd.x;
        c2 = dd.y;
        imp = x.t1 * dd.x + x.t2 * dd.y;
        ap1 = true;
    }
    if (ap1) apply_row2(x, v, imp);
    // axis 2
    if (joint && (m & 4u)) {
        const float old2 = acc.z;
        const float nw2 = old2 + (0.0f - dot3(rel_vel2(x, v), x.t2)) * x.m4.z;
        acc.z = nw2;
        c2 = nw2 - old2;
        if (nw2 != old2) apply_row2(x, v, x.t2 * c2);
    }
    if (!resid) return 0u;
    const float q0 = fabsf(nw0 - old0) / x.m4.x, q1 = fabsf(c1) / x.m4.y, q2 = fabsf(c2) / x.m4.z;
    if (joint)
        return max((m & 1u) ? __float_as_uint(q0) : 0u, max((m & 2u) ? __float_as_uint(q1) : 0u, (m & 4u) ? __float_as_uint(q2) : 0u));
    const float r0 = x.m4.x > 0.0f ? q0 / (1.0f + fabsf(x.m4.w)) : 0.0f;
    const float r1 = x.m4.y > 0.0f ? q1 : 0.0f;
    const float r2 = x.m4.z > 0.0f ? q2 : 0.0f;
    return __float_as_uint(fmaxf(r0, fmaxf(r1, r2)));
}
static __device__ __forceinline__ uint32_t pgs_core(const Row &x, Vel2 &v, float &jn, float2 &jt, bool resid) {
    float4 acc = make_float4(jn, jt.x, jt.y, 0.0f);
    const uint32_t r = entry_core(x, v, acc, false, resid);
    jn = acc.x;
    jt = make_float2(acc.y, acc.z);
    return r;
}
static __device__ __forceinline__ void pgs_row(const Dev &d, uint32_t c, const PgsIn &in, ResidSink resid_bits) {
    uint32_t r;
    if (c & RP_JBIT) {
        r = joint_pgs(d, c & ~RP_JBIT, resid_bits.on);
    } else {
        Vel2 v = load_vel2(d, in.x);
        const Vel2 v0 = v;
        float jn = in.jn;
        float2 jt = in.jt;
        r = pgs_core(in.x, v, jn, jt, resid_bits.on);
        d.jn[c] = jn;
        d.jt[c] = jt;
        store_vel2(d, in.x, v, v0);
    }
    if (resid_bits.on) __hip_atomic_fetch_max(resid_bits.p, r, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}

// A PGS entry of either kind with its accumulators (a contact's jn, jt; a joint's rows' lambdas).
struct PgsE { Row x; float4 acc; };
static __device__ __forceinline__ PgsE pgs_fetch_e(const Dev &d, uint32_t c, uint32_t slot) {
    PgsE e;
    e.x = fetch_row(d, slot);
    if (c & RP_JBIT) {
        const uint32_t q = c & ~RP_JBIT;
        e.acc = make_float4(d.jr_acc[3 * q], d.jr_acc[3 * q + 1], d.jr_acc[3 * q + 2], 0.0f);
    } else {
        const float2 jt = d.jt[c];
        e.acc = make_float4(d.jn[c], jt.x, jt.y, 0.0f);
    }
    return e;
}
template <uint32_t VB>
static __device__ __forceinline__ void pgs_row_l(const Dev &d, uint32_t c, const PgsE &in, const LdsVel<VB> &L, ResidSink resid_bits) {
    Vel2 v = load_vel2_l(d, in.x, L);
    const Vel2 v0 = v;
    float4 acc = in.acc;
    const bool joint = (c & RP_JBIT) != 0u;
    const uint32_t r = entry_core(in.x, v, acc, joint, resid_bits.on);
    if (joint) {
        const uint32_t q = c & ~RP_JBIT, m = __float_as_uint(in.x.m4.w);
        if (m & 1u) d.jr_acc[3 * q] = acc.x;
        if (m & 2u) d.jr_acc[3 * q + 1] = acc.y;
        if (m & 4u) d.jr_acc[3 * q + 2] = acc.z;
    } else {
        d.jn[c] = acc.x;
        d.jt[c] = make_float2(acc.y, acc.z);
    }
    store_vel2_l(d, in.x, v, v0, L);
    if (resid_bits.on) __hip_atomic_fetch_max(resid_bits.p, r, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}

// Velocity shock row of contact c (solver.rs:976-993): the lower body along up is held fixed.
static __device__ __forceinline__ void vshock_row(const Dev &d, const V3 up, uint32_t c) {
    if (c & (RP_JBIT | RP_PBIT)) return;
    if (cm_flags(d, c) & CM_SOFT) return;  // compliant rows have no shock pass (solver.rs:973)
    uint32_t ai = d.cs_ai[c], bi = d.cs_bi[c];
    V3 ra = L3(d.cs_ra[c]), rb = L3(d.cs_rb[c]), n3 = L3(d.cs_n[c]);
    float2 inv = d.cs_inv[c];
    float4 m4 = d.cs_m[c];
    float inv_a = inv.x, inv_b = inv.y, mn = m4.x;
    if (inv_a > 0.0f && inv_b > 0.0f) {
        M3 iia = load_m9(d.iiw + 9 * ai), iib = load_m9(d.iiw + 9 * bi);
        if (dot3(L3(d.pos[ai]), up) < dot3(L3(d.pos[bi]), up)) { inv_a = 0.0f; iia = mzero(); }
        else { inv_b = 0.0f; iib = mzero(); }
        mn = eff_mass(inv_a, inv_b, iia, iib, ra, rb, n3);
    }
    float v = dot3(rel_vel(d, ai, bi, ra, rb), n3);
    float old = d.jn[c];
    float nw = fmaxf(old + (m4.w - v) * mn, 0.0f);
    d.jn[c] = nw;
    apply_cimp(d, ai, bi, inv_a, inv_b, ra, rb, n3 * (nw - old));
}
// Shock order of island isl in slots ss, recording (tiers 1 and 2) its leading wide levels.
template <int TIER, class FL>
static __device__ void shock_slots(const Dev &d, const TickParams &tp, Island<FL> &I, uint32_t isl) {
    const uint32_t s = I.s, n = I.n;
    __shared__ uint32_t wide_ks;
    const uint32_t *exs = n > blockDim.x ? level_order(I, d.sh_pa, d.sh_pb, d.sh_sa, d.sh_sb, d.ex_shock, &d.ph[RP_PH_ROUNDS_SHOCK],
                                                       tp.wide_min, BIG<TIER> ? d.lvl_off_sh + s : nullptr, BIG<TIER> ? &wide_ks : nullptr)
                                          : nullptr;
    build_slots(d, I, d.sh_ord, d.sh_pa, d.sh_pb, exs, Slots{d.ss_pa, d.ss_pb, d.ss_c});
    if constexpr (BIG<TIER>) {
        const uint32_t K = exs ? wide_ks : 0u;
        if (threadIdx.x == 0) { d.isl_wsh[isl] = K; if (K) atomicMax(&d.cnt[CT_WSH_MAX], K); }
    }
}
// The pre-shock impulse snapshot (what the warm cache stores) and the velocity shock pass in
// shock order, for one island; with klsh > 0 (tiers 1 and 2) the pass itself follows as
// k_level_vshock + k_sweep_vshock_tail.
template <int TIER, class FL>
static __device__ void vel_shock(const Dev &d, const TickParams &tp, Island<FL> &I, const uint32_t *lst, uint32_t isl, PhaseClock &pc,
                                 uint32_t klsh) {
    const uint32_t n = I.n;
    const V3 up = mk3(tp.upx, tp.upy, tp.upz);
    for (uint32_t r = threadIdx.x; r < n; r += blockDim.x) {
        const uint32_t c = lst[r];
        if (c & RP_JBIT) joint_store(d, c & ~RP_JBIT);
        else if (!(c & RP_PBIT)) d.jn_pre[c] = d.jn[c];
    }
    __syncthreads();
    shock_slots<TIER>(d, tp, I, isl);
    if constexpr (BIG<TIER>) pc.lap(d, RP_PH_VEL_SHOCK_ORDER);
    if (BIG<TIER> && klsh) return;
    run_sweep(I, Slots{d.ss_pa, d.ss_pb, d.ss_c}, [&](uint32_t c) { vshock_row(d, up, c); });
    if constexpr (BIG<TIER>) pc.lap(d, RP_PH_VEL_SHOCK);
}

// Velocity solve of one island per workgroup: warm-start application and the PGS iterations in
// list order, the pre-shock impulse snapshot (what the warm cache stores), then the shock pass in
// shock order (built before the launch from the tick-start poses, which the solve leaves as is).
// order (tier 0 with the lane tier on): the tier's islands by decreasing size, of which this
// launch takes the first CT_WAVE_K.
template <int TIER>
__global__ void __launch_bounds__(Tier<TIER>::threads) k_sweep_vel(Dev d, TickParams tp, uint32_t split, const uint32_t *order) {
    if (aborted(d)) return;
    const uint32_t cur = tp.cur;
    const V3 up = mk3(tp.upx, tp.upy, tp.upz);
    ISLAND_LOOP_BEGIN_FROM(order ? order : d.tier_list[TIER], order ? d.cnt[CT_WAVE_K] : d.cnt[CT_TIER_N + TIER])
        if (threadIdx.x == 0 && TIER == 1) atomicAdd(&d.cnt[CT_ISLC_GLB], n);
        // islands larger than the workgroup sweep in level order (the ballistic and projection
        // sweeps reuse these slots); smaller ones give every rank its own lane in list order
        PhaseClock pc;
        __shared__ uint32_t wide_k;
        const uint32_t *exb = n > blockDim.x ? level_order(I, d.ch_pa, d.ch_pb, d.ch_sa, d.ch_sb, d.ex_base, &d.ph[RP_PH_ROUNDS_BASE],
                                                           tp.wide_min, BIG<TIER> ? d.lvl_off + s : nullptr, BIG<TIER> ? &wide_k : nullptr)
                                              : nullptr;
        const Slots sb{d.bs_pa, d.bs_pb, d.bs_c};
        build_slots(d, I, d.isl_v2, d.ch_pa, d.ch_pb, exb, sb);
        if constexpr (BIG<TIER>) {
            // the island's leading wide levels (sweeps may run them as GPU-wide launches)
            const uint32_t K = exb ? wide_k : 0u;
            if (threadIdx.x == 0) {
                d.big_list[atomicAdd(&d.cnt[CT_BIG_N], 1u)] = isl;
                d.isl_w[isl] = K;
                if (K) { d.wide_list[atomicAdd(&d.cnt[CT_WIDE_N], 1u)] = isl; atomicMax(&d.cnt[CT_WIDE_MAX], K); }
            }
            pc.lap(d, RP_PH_VEL_ORDER);
        }
        if (!(BIG<TIER> && split) && tp.ws_mode != 1) run_sweep(I, sb, [&](uint32_t c) { warm_row(d, tp, c); });
        if constexpr (BIG<TIER>) pc.lap(d, RP_PH_WARM);
        // PGS iterations (solver.rs:839-879).  The rows' constants are stored per slot once and
        // loaded, with the row's own accumulators, before its wait; the bodies' velocities are
        // loaded once and kept in registers between the normal and the friction update.
        if (BIG<TIER> && split) {
            // the row store (k_store_rows), warm start, sweeps and shock pass follow in launches
            // of their own
            if (threadIdx.x == 0) { atomicMax(&d.cnt[CT_ITERS_MAX], tp.iv); atomicAdd(&d.cnt[CT_ITERS_SUM], tp.iv * n); }
            continue;
        }
        // the rows per slot, with the LDS velocity slots of their bodies in the LDS tiers
        constexpr uint32_t VB = TIER == 0 ? 257u : (TIER == 3 ? 512u : 1u);
        __shared__ __attribute__((aligned(16))) float v_lin[4 * VB], v_ang[4 * VB];
        __shared__ uint32_t v_body[VB], v_n;
        const LdsVel<VB> L{(RP_LDS float *)v_lin, (RP_LDS float *)v_ang, (RP_LDS uint32_t *)v_body};
        if constexpr (!BIG<TIER>) lds_vel_begin<VB>(d, s, n, L, &v_n);
        for (uint32_t j = threadIdx.x; j < n; j += blockDim.x) {
            const uint32_t c = d.bs_c[s + j];
            if constexpr (BIG<TIER>) {
                if (!(c & RP_JBIT)) store_row(d, s + j, load_row(d, c));
            } else {
                const uint32_t r = exb ? exb[s + j] : j;
                Row x = (c & RP_JBIT) ? load_jrow(d, c & ~RP_JBIT) : load_row(d, c);
                x.la = d.lid[2 * (s + r)];
                x.lb = d.lid[2 * (s + r) + 1];
                store_row(d, s + j, x);
            }
        }
        __syncthreads();
        // residual mode: the sweep's largest row residual (non-negative floats order as their bits)
        __shared__ uint32_t resid_bits;
        const ResidSink rb_lds{(RP_LDS uint32_t *)&resid_bits, tp.resid_on != 0};
        uint32_t used = 0;
        bool conv = false;
        auto sweep = [&]() {
            if constexpr (BIG<TIER>)
                run_sweep(I, sb, [&](uint32_t c, uint32_t slot) { return pgs_fetch(d, c, slot); }, [&](uint32_t c, const PgsIn &in) { pgs_row(d, c, in, rb_lds); });
            else
                run_sweep(I, sb, [&](uint32_t c, uint32_t slot) { return pgs_fetch_e(d, c, slot); },
                          [&](uint32_t c, const PgsE &in) { pgs_row_l<VB>(d, c, in, L, rb_lds); });
        };
        if (!tp.resid_on) {
            for (uint32_t it = 0; it < tp.iv; it++) sweep();
            used = tp.iv;
        } else {
            for (uint32_t it = 0; it < tp.iv; it++) {
                if (threadIdx.x == 0) resid_bits = 0;
                __syncthreads();
                sweep();
                // every thread takes the sweep's residual before thread 0 may clear it for the
                // next sweep, so all threads leave the loop at the same sweep
                const float rmax = __uint_as_float(resid_bits);
                __syncthreads();
                used++;
                if (rmax <= tp.resid_tol) { conv = true; break; }
            }
        }
        if (threadIdx.x == 0) {
            atomicMax(&d.cnt[CT_ITERS_MAX], used); atomicAdd(&d.cnt[CT_ITERS_SUM], used * n);
            if (conv) { atomicAdd(&d.cnt[CT_ISL_CONV], 1u); atomicAdd(&d.cnt[CT_ISLC_CONV], n); atomicAdd(&d.cnt[CT_ITERS_SUM_CONV], used * n); }
        }
        if (tp.resid_on)  // which entries' impulses are converged ones (the contact log keeps only those)
            for (uint32_t r = threadIdx.x; r < n; r += blockDim.x) d.rank_conv[s + r] = conv ? 1u : 0u;
        if constexpr (BIG<TIER>) pc.lap(d, RP_PH_PGS);
        if constexpr (!BIG<TIER>) lds_vel_end<VB>(d, L, v_n);
        vel_shock<TIER>(d, tp, I, lst, isl, pc, 0u);
    ISLAND_LOOP_END
}

// ------------------------------------------------------------------ lane tier
// The velocity solve of the islands of up to tp.lane_max entries, one island per lane.  These are
// mostly chains a row or two wide (box towers, jointed capsule chains), whose rows a wave per
// island runs one at a time; a lane walking its island's entries in list order (a topological
// order of the chains, so the same bits as any other) keeps every lane of the wave busy instead.
// Waves take 64 islands of similar size (lane_list: the lane tier's islands by decreasing entry
// count) and keep their constant rows and accumulators in the lane layout: entry j, component k
// of lane l at lrow[base + (RP_LROW j + k) 64 + l], base = 64 RP_LROW x the largest island of
// each earlier wave, so a wave's load of one component of its lanes' j-th entries is 1 KiB
// contiguous.  With the islands sorted, the waves' spans sum to at most the island lists' length
// plus 64 x the largest island (cap_lane entries).
static __device__ __forceinline__ uint32_t island_len(const Dev &d, uint32_t isl, uint32_t nisl, uint32_t nislc) {
    return ((isl + 1 < nisl) ? d.isl_start[isl + 1] : nislc) - d.isl_start[isl];
}
static __device__ __forceinline__ uint32_t wave_sum(uint32_t v) {
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}
// Sort keys of tier 0's islands: decreasing size (tier 0 holds islands of at most 256 entries).
__global__ void k_lane_keys(Dev d, TickParams tp) {
    if (aborted(d)) return;
    const uint32_t m = d.cnt[CT_TIER_N], nisl = d.cnt[CT_ISLANDS], nislc = min(d.cnt[CT_ISLC], d.cap_contacts);
    GRID_STRIDE(q, m) {
        const uint32_t isl = d.tier_list[0][q];
        d.sk0[q] = tp.tier_max[0] - island_len(d, isl, nisl, nislc);
        d.sv0[q] = isl;
    }
}
// The split of tier 0's sorted islands: [0, CT_WAVE_K) (more than tp.quad_max entries, and at
// least the tp.wave_min largest: while a wave per island still fits the chip at once, its
// islands' levels run in parallel) solved a wave each, [CT_WAVE_K, CT_GRP_K) (more than
// tp.pair_max) by groups of four lanes, [CT_GRP_K, CT_LANE_K) (more than tp.lane_max) by lane
// pairs, the rest (CT_LANE_N of them) one per lane.
static __device__ uint32_t first_at_most(const Dev &d, const uint32_t *sorted, uint32_t m, uint32_t bound, uint32_t nisl, uint32_t nislc) {
    uint32_t lo = 0, hi = m;
    while (lo < hi) {
        const uint32_t mid = lo + (hi - lo) / 2;
        if (island_len(d, sorted[mid], nisl, nislc) > bound) lo = mid + 1;
        else hi = mid;
    }
    return lo;
}
__global__ void k_lane_split(Dev d, TickParams tp, const uint32_t *sorted) {
    if (threadIdx.x != 0 || blockIdx.x != 0 || aborted(d)) return;
    const uint32_t m = d.cnt[CT_TIER_N], nisl = d.cnt[CT_ISLANDS], nislc = min(d.cnt[CT_ISLC], d.cap_contacts);
    const uint32_t kw = max(first_at_most(d, sorted, m, tp.quad_max, nisl, nislc), min(m, tp.wave_min));
    const uint32_t k4 = max(first_at_most(d, sorted, m, tp.pair_max, nisl, nislc), kw);
    const uint32_t k2 = max(first_at_most(d, sorted, m, tp.lane_max, nisl, nislc), k4);
    d.cnt[CT_WAVE_K] = kw;
    d.cnt[CT_GRP_K] = k4;
    d.cnt[CT_LANE_K] = k2;
    d.cnt[CT_LANE_N] = m - k2;
}
// One lane-tier entry in registers: its row, accumulators (and entry), its bodies' velocities v
// and what HBM holds of them, m (a body the next entry also writes passes to it in registers and
// reaches HBM after the last consecutive entry that uses it).
struct LaneEntry { Row x; float4 acc; Vel2 v, m; };
// In the lane layout a row's la, lb (LDS slots elsewhere) hold the next entry's bodies, so the
// next entry's velocities load alongside its row instead of after it.
static __device__ __forceinline__ Vel2 load_vel2_ab(const Dev &d, uint32_t ai, uint32_t bi) {
    Vel2 v;
    v.la = L3(d.lin[ai]); v.aa = L3(d.ang[ai]);
    if (bi != RP_NONE) { v.lb = L3(d.lin[bi]); v.ab = L3(d.ang[bi]); } else { v.lb = splat3(0.0f); v.ab = splat3(0.0f); }
    return v;
}
static __device__ __forceinline__ void lane_fetch_row(const float4 *R, uint32_t j, LaneEntry &e) {
    const float4 *o = R + (size_t)j * (64u * RP_LROW);
    e.acc = o[64u * RP_SROW];
    e.x = fetch_row_at(o, 64u);
}
// entry j (bodies ai, bi) after the entry in cur: the bodies cur also has come from cur's registers
static __device__ __forceinline__ void lane_fetch(const Dev &d, const float4 *R, uint32_t j, LaneEntry &e, uint32_t ai, uint32_t bi,
                                                  const LaneEntry &cur) {
    lane_fetch_row(R, j, e);
    if (ai != cur.x.ai && ai != cur.x.bi) { e.v.la = L3(d.lin[ai]); e.v.aa = L3(d.ang[ai]); }
    if (bi == RP_NONE) { e.v.lb = splat3(0.0f); e.v.ab = splat3(0.0f); }
    else if (bi != cur.x.ai && bi != cur.x.bi) { e.v.lb = L3(d.lin[bi]); e.v.ab = L3(d.ang[bi]); }
    e.m = e.v;
}
// cur's bodies after its entry ran: to the next entry (bodies na, nb) in registers, or to HBM
static __device__ __forceinline__ void lane_retire(const Dev &d, const LaneEntry &cur, LaneEntry &nxt, bool more, uint32_t na, uint32_t nb) {
    const uint32_t a = cur.x.ai, b = cur.x.bi;
    if (more && a == na) { nxt.v.la = cur.v.la; nxt.v.aa = cur.v.aa; nxt.m.la = cur.m.la; nxt.m.aa = cur.m.aa; }
    else if (more && a == nb) { nxt.v.lb = cur.v.la; nxt.v.ab = cur.v.aa; nxt.m.lb = cur.m.la; nxt.m.ab = cur.m.aa; }
    else {
        if (!same_bits(cur.v.la, cur.m.la)) d.lin[a] = S3(cur.v.la);
        if (!same_bits(cur.v.aa, cur.m.aa)) d.ang[a] = S3(cur.v.aa);
    }
    if (b == RP_NONE) return;
    if (more && b == na) { nxt.v.la = cur.v.lb; nxt.v.aa = cur.v.ab; nxt.m.la = cur.m.lb; nxt.m.aa = cur.m.ab; }
    else if (more && b == nb) { nxt.v.lb = cur.v.lb; nxt.v.ab = cur.v.ab; nxt.m.lb = cur.m.lb; nxt.m.ab = cur.m.ab; }
    else {
        if (!same_bits(cur.v.lb, cur.m.lb)) d.lin[b] = S3(cur.v.lb);
        if (!same_bits(cur.v.ab, cur.m.ab)) d.ang[b] = S3(cur.v.ab);
    }
}
// Entry j of n (in cur): fetches entry j + 1 into nxt, runs entry j, stores its accumulators and
// hands its bodies' velocities on (lane_retire).  Returns the residual bits.
static __device__ __forceinline__ uint32_t lane_step(const Dev &d, float4 *R, uint32_t j, uint32_t n, LaneEntry &cur, LaneEntry &nxt,
                                                     bool resid) {
    const bool more = j + 1 < n;
    const uint32_t na = cur.x.la, nb = cur.x.lb;  // the next entry's bodies
    if (more) lane_fetch(d, R, j + 1, nxt, na, nb, cur);
    const uint32_t r = entry_core(cur.x, cur.v, cur.acc, (__float_as_uint(cur.acc.w) & RP_JBIT) != 0u, resid);
    R[(size_t)j * (64u * RP_LROW) + 64u * RP_SROW] = cur.acc;
    lane_retire(d, cur, nxt, more, na, nb);
    return r;
}
// the lane tier's waves b0, b0 + nb, ...
static __device__ void lane_vel(const Dev &d, const TickParams &tp, const uint32_t *sorted, uint32_t b0, uint32_t nb) {
    const uint32_t m = d.cnt[CT_LANE_N], nisl = d.cnt[CT_ISLANDS], nislc = min(d.cnt[CT_ISLC], d.cap_contacts);
    // a lane wave's islands are the velocity solve's longest serial chains: its instructions go
    // first on a shared SIMD
    __builtin_amdgcn_s_setprio(2);
    const uint32_t *const lane_list = sorted + d.cnt[CT_LANE_K];
    const uint32_t nwaves = (m + 63u) / 64u, l = threadIdx.x;
    const V3 up = mk3(tp.upx, tp.upy, tp.upz);
    for (uint32_t w = b0; w < nwaves; w += nb) {
        uint32_t before = 0;
        for (uint32_t v = l; v < w; v += 64u) before += island_len(d, lane_list[64u * v], nisl, nislc);
        before = wave_sum(before);
        const uint32_t nmax = island_len(d, lane_list[64u * w], nisl, nislc);
        if ((size_t)(before + nmax) * 64u > d.cap_lane) {  // cannot happen with sorted islands; replay with room
            if (l == 0) { d.cnt[CT_NEED_CONTACTS] = max(d.cnt[CT_NEED_CONTACTS], nislc + 64u * nmax); d.cnt[CT_ABORT] = 1u; }
            return;
        }
        float4 *R = d.lrow + (size_t)before * 64u * RP_LROW + l;
        const uint32_t q = 64u * w + l;
        const uint32_t isl = q < m ? lane_list[q] : 0u;
        const uint32_t s = q < m ? d.isl_start[isl] : 0u, n = q < m ? island_len(d, isl, nisl, nislc) : 0u;
        const uint32_t *lst = d.isl_v2 + s;
        if (!tp.lane_pos)  // the list order's slots, for the one-wave ballistic and projection sweeps
            for (uint32_t j = 0; j < n; j++) {
                const uint32_t a = d.ch_pa[s + j], b = d.ch_pb[s + j];
                d.bs_pa[s + j] = a == RP_NONE ? RP_NONE : a - s;
                d.bs_pb[s + j] = b == RP_NONE ? RP_NONE : b - s;
                d.bs_c[s + j] = lst[j];
            }
        if (tp.ws_mode != 1)
            for (uint32_t j = 0; j < n; j++) warm_row(d, tp, lst[j]);
        for (uint32_t j = 0; j < n; j++) {
            const uint32_t c = lst[j];
            float4 *o = R + (size_t)j * (64u * RP_LROW);
            uint32_t na = RP_NONE, nbd = RP_NONE;  // the next entry's bodies
            if (j + 1 < n) {
                const uint32_t cn = lst[j + 1];
                if (cn & RP_JBIT) { na = d.jt_a[cn & ~RP_JBIT]; nbd = d.jt_b[cn & ~RP_JBIT]; }
                else { na = d.cs_ai[cn]; nbd = d.cs_bi[cn]; }
            }
            if (c & RP_JBIT) {
                const uint32_t q = c & ~RP_JBIT;
                Row x = load_jrow(d, q);
                x.la = na; x.lb = nbd;
                store_row_at(o, 64u, x);
                o[64u * RP_SROW] = make_float4(d.jr_acc[3 * q], d.jr_acc[3 * q + 1], d.jr_acc[3 * q + 2], __uint_as_float(c));
                continue;
            }
            Row x = load_row(d, c);
            x.la = na; x.lb = nbd;
            store_row_at(o, 64u, x);
            const float2 jt = d.jt[c];
            o[64u * RP_SROW] = make_float4(d.jn[c], jt.x, jt.y, __uint_as_float(c));
        }
        // PGS iterations (solver.rs:839-879); residual mode: a lane stops at its island's tolerance
        const bool resid = tp.resid_on != 0;
        bool act = n > 0, conv = false;
        uint32_t used = 0;
        for (uint32_t it = 0; it < tp.iv; it++) {
            if (!__any(act)) break;
            if (!act) continue;
            // entry j + 1's row, accumulators and bodies' velocities load while entry j runs (two
            // buffers in turn); the velocities of bodies entry j wrote are taken from its registers
            LaneEntry A, B;
            lane_fetch_row(R, 0, A);
            A.v = load_vel2(d, A.x);
            A.m = A.v;
            uint32_t rbits = 0;
            for (uint32_t j = 0; j < n; j += 2) {
                rbits = max(rbits, lane_step(d, R, j, n, A, B, resid));
                if (j + 1 < n) rbits = max(rbits, lane_step(d, R, j + 1, n, B, A, resid));
            }
            used++;
            if (resid && __uint_as_float(rbits) <= tp.resid_tol) { conv = true; act = false; }
        }
        if (n) {
            atomicMax(&d.cnt[CT_ITERS_MAX], used); atomicAdd(&d.cnt[CT_ITERS_SUM], used * n);
            if (conv) { atomicAdd(&d.cnt[CT_ISL_CONV], 1u); atomicAdd(&d.cnt[CT_ISLC_CONV], n); atomicAdd(&d.cnt[CT_ITERS_SUM_CONV], used * n); }
        }
        // the accumulators, the pre-shock impulse snapshot (what the warm cache stores) and, in
        // residual mode, which entries' impulses are converged ones
        for (uint32_t j = 0; j < n; j++) {
            const float4 acc = R[(size_t)j * (64u * RP_LROW) + 64u * RP_SROW];
            const uint32_t c = __float_as_uint(acc.w);
            if (c & RP_JBIT) {
                const uint32_t q = c & ~RP_JBIT, m = d.jr_mask[q];
                if (m & 1u) d.jr_acc[3 * q] = acc.x;
                if (m & 2u) d.jr_acc[3 * q + 1] = acc.y;
                if (m & 4u) d.jr_acc[3 * q + 2] = acc.z;
                joint_store(d, q);
            } else {
                d.jn[c] = acc.x;
                d.jt[c] = make_float2(acc.y, acc.z);
                d.jn_pre[c] = acc.x;
            }
            if (resid) d.rank_conv[s + j] = conv ? 1u : 0u;
        }
        // the velocity shock pass in shock order (solver.rs:964-993)
        for (uint32_t j = 0; j < n; j++) vshock_row(d, up, d.sh_ord[s + j]);
    }
}

// ------------------------------------------------------------------ lane groups
// The velocity solve of tier 0's mid-sized islands by groups of G lanes, 64 / G islands per wave:
// the dataflow sweep of one wave per island (slots in list order, lane li of the group takes
// slots li, li + G, ...; a slot runs once both its predecessor slots carry the sweep's token),
// within each group.  These islands are a few rows wide, so G lanes keep up with their levels
// while a wave per island would run 64-wide instructions for those few rows.  A group's lanes are
// one wave's: their memory operations are ordered by the wave's program order, and the readiness
// flags (LDS, one byte per slot) need only wavefront scope.  Rows and accumulators stay in the
// slot layout of the one-wave sweeps (srow, jn / jt, jr_acc), each slot's by its own lane.
template <int G> struct LaneGroup;
template <> struct LaneGroup<2> { static constexpr uint32_t slots = 128; };  // islands of up to 128 entries
template <> struct LaneGroup<4> { static constexpr uint32_t slots = 256; };  // (tier 0's bound)
static __device__ __forceinline__ uint32_t gflag_load(const RP_LDS uint8_t *f, uint32_t i) {
    return __hip_atomic_load(&f[i], __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_WAVEFRONT);
}
static __device__ __forceinline__ void gflag_store(RP_LDS uint8_t *f, uint32_t i, uint32_t t) {
    __hip_atomic_store(&f[i], (uint8_t)t, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_WAVEFRONT);
}
// One sweep of an island of n slots by lane li of its group; pred(j, a, b) gives slot j's
// predecessor slots (RP_NONE: none, always lower slots), prep(j) its row, run(j, row) runs it.  A
// lane's next slot is often ready as soon as this one is done, so its row loads before this one
// runs.  The group's lowest unfinished slot is always ready.
template <int G, class PRED, class PRE, class RUN>
static __device__ __forceinline__ void group_sweep(RP_LDS uint8_t *fl, uint32_t tok, uint32_t n, uint32_t li, PRED pred, PRE prep, RUN run) {
    uint32_t j = li, a = RP_NONE, b = RP_NONE;
    decltype(prep(0u)) x{};
    if (j < n) { pred(j, a, b); x = prep(j); }
    while (j < n) {
        const bool ready = (a == RP_NONE || gflag_load(fl, a) >= tok) && (b == RP_NONE || gflag_load(fl, b) >= tok);
        if (ready) {
            const uint32_t jn = j + G;
            uint32_t an = RP_NONE, bn = RP_NONE;
            decltype(prep(0u)) xn{};
            if (jn < n) { pred(jn, an, bn); xn = prep(jn); }
            run(j, x);
            gflag_store(fl, j, tok);
            j = jn; a = an; b = bn; x = xn;
        }
    }
}
// the largest residual bits over the G lanes of each group
template <int G>
static __device__ __forceinline__ uint32_t group_max(uint32_t v) {
    for (int o = G / 2; o > 0; o >>= 1) v = max(v, (uint32_t)__shfl_xor((int)v, o, 64));
    return v;
}
// the waves b0, b0 + nb, ... of the islands solved by groups of G lanes
template <int G>
static __device__ void group_vel(const Dev &d, const TickParams &tp, const uint32_t *order, uint32_t b0, uint32_t nb) {
    constexpr uint32_t NG = 64u / G, FC = LaneGroup<G>::slots;
    __shared__ __attribute__((aligned(4))) uint8_t flags[NG * FC];
    const uint32_t lo = G == 4 ? d.cnt[CT_WAVE_K] : d.cnt[CT_GRP_K], hi = G == 4 ? d.cnt[CT_GRP_K] : d.cnt[CT_LANE_K];
    const uint32_t nisl = d.cnt[CT_ISLANDS], nislc = min(d.cnt[CT_ISLC], d.cap_contacts);
    const uint32_t gi = threadIdx.x / G, li = threadIdx.x % G;
    RP_LDS uint8_t *fl = (RP_LDS uint8_t *)flags + gi * FC;
    const V3 up = mk3(tp.upx, tp.upy, tp.upz);
    const bool resid = tp.resid_on != 0;
    const uint32_t nwaves = (hi - lo + NG - 1u) / NG;
    for (uint32_t w = b0; w < nwaves; w += nb) {
        const uint32_t q = lo + w * NG + gi;
        const uint32_t isl = q < hi ? order[q] : 0u;
        const uint32_t s = q < hi ? d.isl_start[isl] : 0u, n = q < hi ? min(island_len(d, isl, nisl, nislc), FC) : 0u;
        const uint32_t *lst = d.isl_v2 + s;
        auto list_pred = [&](uint32_t j, uint32_t &a, uint32_t &b) {
            const uint32_t pa = d.ch_pa[s + j], pb = d.ch_pb[s + j];
            a = pa == RP_NONE ? RP_NONE : pa - s;
            b = pb == RP_NONE ? RP_NONE : pb - s;
        };
        uint32_t tok = 0;
        auto next_token = [&]() {  // the flags hold byte tokens: cleared when they run out
            if (tok == 255u) {
                for (uint32_t j = li; j < n; j += G) fl[j] = 0;
                __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");  // the clears precede every later flag access
                tok = 0;
            }
            return ++tok;
        };
        for (uint32_t j = li; j < n; j += G) {
            fl[j] = 0;
            // the list order's slots, for the one-wave ballistic and projection sweeps
            const uint32_t a = d.ch_pa[s + j], b = d.ch_pb[s + j];
            d.bs_pa[s + j] = a == RP_NONE ? RP_NONE : a - s;
            d.bs_pb[s + j] = b == RP_NONE ? RP_NONE : b - s;
            d.bs_c[s + j] = lst[j];
        }
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
        // warm start in list order (solver.rs:749-760)
        if (tp.ws_mode != 1)
            group_sweep<G>(fl, next_token(), n, li, list_pred, [](uint32_t) { return NoRow{}; }, [&](uint32_t j, const NoRow &) { warm_row(d, tp, lst[j]); });
        for (uint32_t j = li; j < n; j += G) {
            const uint32_t c = lst[j];
            store_row(d, s + j, (c & RP_JBIT) ? load_jrow(d, c & ~RP_JBIT) : load_row(d, c));
        }
        // PGS iterations (solver.rs:839-879); residual mode: a group stops at its island's tolerance
        bool act = n > 0, conv = false;
        uint32_t used = 0;
        for (uint32_t it = 0; it < tp.iv; it++) {
            if (!__any(act)) break;
            const uint32_t t = next_token();
            uint32_t rbits = 0;
            if (act)
                group_sweep<G>(fl, t, n, li, list_pred,
                               [&](uint32_t j) {
                                   LaneEntry e;
                                   const uint32_t c = lst[j];
                                   e.x = fetch_row(d, s + j);
                                   if (c & RP_JBIT) {
                                       const uint32_t jq = c & ~RP_JBIT;
                                       e.acc = make_float4(d.jr_acc[3 * jq], d.jr_acc[3 * jq + 1], d.jr_acc[3 * jq + 2], __uint_as_float(c));
                                   } else {
                                       const float2 jt = d.jt[c];
                                       e.acc = make_float4(d.jn[c], jt.x, jt.y, __uint_as_float(c));
                                   }
                                   return e;
                               },
                               [&](uint32_t, LaneEntry &e) {
                                   const uint32_t c = __float_as_uint(e.acc.w);
                                   Vel2 v = load_vel2(d, e.x);
                                   const Vel2 v0 = v;
                                   rbits = max(rbits, entry_core(e.x, v, e.acc, (c & RP_JBIT) != 0u, resid));
                                   if (c & RP_JBIT) {
                                       const uint32_t jq = c & ~RP_JBIT, m = __float_as_uint(e.x.m4.w);
                                       if (m & 1u) d.jr_acc[3 * jq] = e.acc.x;
                                       if (m & 2u) d.jr_acc[3 * jq + 1] = e.acc.y;
                                       if (m & 4u) d.jr_acc[3 * jq + 2] = e.acc.z;
                                   } else {
                                       d.jn[c] = e.acc.x;
                                       d.jt[c] = make_float2(e.acc.y, e.acc.z);
                                   }
                                   store_vel2(d, e.x, v, v0);
                               });
            const uint32_t rg = group_max<G>(rbits);
            if (act) {
                used++;
                if (resid && __uint_as_float(rg) <= tp.resid_tol) { conv = true; act = false; }
            }
        }
        if (n && li == 0) {
            atomicMax(&d.cnt[CT_ITERS_MAX], used); atomicAdd(&d.cnt[CT_ITERS_SUM], used * n);
            if (conv) { atomicAdd(&d.cnt[CT_ISL_CONV], 1u); atomicAdd(&d.cnt[CT_ISLC_CONV], n); atomicAdd(&d.cnt[CT_ITERS_SUM_CONV], used * n); }
        }
        // the pre-shock impulse snapshot (what the warm cache stores), converged entries
        for (uint32_t j = li; j < n; j += G) {
            const uint32_t c = lst[j];
            if (c & RP_JBIT) joint_store(d, c & ~RP_JBIT);
            else d.jn_pre[c] = d.jn[c];
            if (resid) d.rank_conv[s + j] = conv ? 1u : 0u;
        }
        // the velocity shock pass in shock order (solver.rs:964-993): slots are shock ranks
        group_sweep<G>(fl, next_token(), n, li,
                       [&](uint32_t j, uint32_t &a, uint32_t &b) {
                           const uint32_t pa = d.sh_pa[s + j], pb = d.sh_pb[s + j];
                           a = pa == RP_NONE ? RP_NONE : pa - s;
                           b = pb == RP_NONE ? RP_NONE : pb - s;
                       },
                       [](uint32_t) { return NoRow{}; }, [&](uint32_t j, const NoRow &) { vshock_row(d, up, d.sh_ord[s + j]); });
    }
}

// Tier 0's velocity solve with the lane tier on, but for its largest islands (k_sweep_vel<0>): the
// lane tier, and in a launch of their own (when their bounds leave them any islands) the groups of
// four lanes in the first ng workgroups, the lane pairs in the rest.  The launches run alongside
// one another; keeping the lane tier's code apart keeps its loop in the instruction cache.
__global__ void __launch_bounds__(64) k_lane_vel(Dev d, TickParams tp, const uint32_t *sorted) {
    if (aborted(d)) return;
    lane_vel(d, tp, sorted, blockIdx.x, gridDim.x);
}
__global__ void __launch_bounds__(64) k_group_vel(Dev d, TickParams tp, const uint32_t *sorted, uint32_t ng) {
    if (aborted(d)) return;
    const uint32_t b = blockIdx.x;
    if (b < ng) group_vel<4>(d, tp, sorted, b, ng);
    else group_vel<2>(d, tp, sorted, b - ng, gridDim.x - ng);
}

// Every slot of wide level L of the islands in list[0..count) whose leading wide levels
// (kw[isl] of them, starts at off[s + ...]) include L, spread over the launch's lanes: the rows of
// one level write pairwise distinct bodies, and the launch boundaries order the levels (the
// reference's level batches, solver.rs:621-642, across the GPU instead of one workgroup).
template <class F>
static __device__ __forceinline__ void for_wide_level(const Dev &d, const uint32_t *list, uint32_t count, const uint32_t *kw, const uint32_t *off,
                                                      uint32_t L, F row) {
    for (uint32_t i = 0; i < count; i++) {
        const uint32_t isl = list[i];
        if (L >= kw[isl]) continue;
        const uint32_t s = d.isl_start[isl], lo = off[s + L], hi = off[s + L + 1];
        for (uint32_t j = lo + blockIdx.x * blockDim.x + threadIdx.x; j < hi; j += gridDim.x * blockDim.x) row(s + j);
    }
}
// Every slot of every big island (the 1024-thread tiers) this tick, across the launch's lanes.
template <class F>
static __device__ __forceinline__ void for_big_island_slots(const Dev &d, F f) {
    const uint32_t nbig = d.cnt[CT_BIG_N], nisl = d.cnt[CT_ISLANDS], nislc = min(d.cnt[CT_ISLC], d.cap_contacts);
    for (uint32_t i = 0; i < nbig; i++) {
        const uint32_t isl = d.big_list[i], s = d.isl_start[isl];
        const uint32_t n = (isl + 1 < nisl ? d.isl_start[isl + 1] : nislc) - s;
        for (uint32_t j = blockIdx.x * blockDim.x + threadIdx.x; j < n; j += gridDim.x * blockDim.x) f(s + j);
    }
}

// The split form of the big-island velocity solve, after k_sweep_vel (split): one PGS sweep over
// the slots after the wide levels that k_level_vel ran (kl of them launched)...
template <int TIER>
__global__ void __launch_bounds__(Tier<TIER>::threads) k_sweep_vel_tail(Dev d, TickParams tp, uint32_t kl) {
    if (aborted(d)) return;
    ISLAND_LOOP_BEGIN
        const uint32_t k = min(d.isl_w[isl], kl);
        run_sweep_from(I, Slots{d.bs_pa, d.bs_pb, d.bs_c}, k ? d.lvl_off[s + k] : 0u,
                       [&](uint32_t c, uint32_t slot) { return pgs_fetch(d, c, slot); }, [&](uint32_t c, const PgsIn &in) { pgs_row(d, c, in, ResidSink{nullptr, false}); });
    ISLAND_LOOP_END
}
// ... and, after the sweeps, the pre-shock snapshot and the shock pass.
template <int TIER>
__global__ void __launch_bounds__(Tier<TIER>::threads) k_sweep_vel_post(Dev d, TickParams tp, uint32_t klsh) {
    if (aborted(d)) return;
    ISLAND_LOOP_BEGIN
        PhaseClock pc;
        vel_shock<TIER>(d, tp, I, lst, isl, pc, klsh);
    ISLAND_LOOP_END
}
// Wide level L of the velocity shock pass across the GPU ...
__global__ void __launch_bounds__(256) k_level_vshock(Dev d, TickParams tp, uint32_t L) {
    if (aborted(d)) return;
    const V3 up = mk3(tp.upx, tp.upy, tp.upz);
    for_wide_level(d, d.big_list, d.cnt[CT_BIG_N], d.isl_wsh, d.lvl_off_sh, L, [&](uint32_t slot) { vshock_row(d, up, d.ss_c[slot]); });
}
// ... and the pass's tail in one workgroup.
template <int TIER>
__global__ void __launch_bounds__(Tier<TIER>::threads) k_sweep_vshock_tail(Dev d, TickParams tp, uint32_t klsh) {
    if (aborted(d)) return;
    const V3 up = mk3(tp.upx, tp.upy, tp.upz);
    ISLAND_LOOP_BEGIN
        const uint32_t k = min(d.isl_wsh[isl], klsh);
        run_sweep_from(I, Slots{d.ss_pa, d.ss_pb, d.ss_c}, k ? d.lvl_off_sh[s + k] : 0u, [](uint32_t, uint32_t) { return NoRow{}; },
                       [&](uint32_t c, const NoRow &) { vshock_row(d, up, c); });
    ISLAND_LOOP_END
}
// Wide level L of the PGS sweep (split form).
__global__ void __launch_bounds__(256) k_level_vel(Dev d, uint32_t L) {
    if (aborted(d)) return;
    for_wide_level(d, d.wide_list, d.cnt[CT_WIDE_N], d.isl_w, d.lvl_off, L, [&](uint32_t slot) {
        const uint32_t c = d.bs_c[slot];
        pgs_row(d, c, pgs_fetch(d, c, slot), ResidSink{nullptr, false});
    });
}
// The constant rows of every big island's slots (split form), across the GPU.
__global__ void __launch_bounds__(256) k_store_rows(Dev d) {
    if (aborted(d)) return;
    for_big_island_slots(d, [&](uint32_t slot) {
        const uint32_t c = d.bs_c[slot];
        if (!(c & RP_JBIT)) store_row(d, slot, load_row(d, c));
    });
}
// Warm start of the split form: wide level L across the GPU, then the tail in one workgroup.
__global__ void __launch_bounds__(256) k_level_warm(Dev d, TickParams tp, uint32_t L) {
    if (aborted(d)) return;
    for_wide_level(d, d.wide_list, d.cnt[CT_WIDE_N], d.isl_w, d.lvl_off, L, [&](uint32_t slot) { warm_row(d, tp, d.bs_c[slot]); });
}
template <int TIER>
__global__ void __launch_bounds__(Tier<TIER>::threads) k_sweep_warm_tail(Dev d, TickParams tp, uint32_t kl) {
    if (aborted(d)) return;
    ISLAND_LOOP_BEGIN
        const uint32_t k = min(d.isl_w[isl], kl);
        run_sweep_from(I, Slots{d.bs_pa, d.bs_pb, d.bs_c}, k ? d.lvl_off[s + k] : 0u, [](uint32_t, uint32_t) { return NoRow{}; },
                       [&](uint32_t c, const NoRow &) { warm_row(d, tp, c); });
    ISLAND_LOOP_END
}

// solver.rs:1247-1270 loaded flags and impulse write-back
__global__ void k_loaded(Dev d) {
    if (aborted(d)) return;
    const uint32_t nct = min(d.cnt[CT_CONTACTS], d.cap_contacts);
    GRID_STRIDE(c, nct) {
        uint32_t w = d.col_world[d.ct_a[c]];
        if (!d.w_solve[w]) { d.ct_imp[c] = 0.0f; continue; }
        float2 t = d.jt[c];
        float jn = d.jn[c];
        if (jn > 0.0f || t.x != 0.0f || t.y != 0.0f) {
            atomicOr(&d.flags[d.cs_ai[c]], FL_LOADED);
            if (d.cs_bi[c] != RP_NONE) atomicOr(&d.flags[d.cs_bi[c]], FL_LOADED);
        }
        d.ct_imp[c] = jn;
    }
}

// ------------------------------------------------------------------ warm cache store
// solver.rs:917-944: the new cache of a world is its contacts' nonzero impulses (contact order is
// key order) merged with the old entries of idle pairs whose key has no new entry.
static __device__ __forceinline__ bool idle_body(const Dev &d, uint32_t w, uint32_t local) {
    if (local == 0xFFFFFFFFu) return true;
    uint32_t b = d.w_body_off[w] + local;
    if (b >= d.w_body_off[w + 1] || !(d.bflags[b] & BF_ALIVE)) return false;
    return !(d.bflags[b] & BF_DYN) || !d.awake[b];
}
static __device__ __forceinline__ bool new_entry(const Dev &d, uint32_t c) {
    if (cm_flags(d, c) & CM_SOFT) return false;  // compliant rows are not carried (solver.rs:922)
    float2 t = d.jt[c];
    V3 jt = L3(d.cs_t1[c]) * t.x + L3(d.cs_t2[c]) * t.y;
    return !(d.jn_pre[c] == 0.0f && zero3(jt));
}
static __device__ __forceinline__ void world_contacts(const Dev &d, uint32_t w, uint32_t &c0, uint32_t &c1) {
    c0 = min(d.c_off[d.w_col_off[w]], d.cap_contacts);
    c1 = min(d.c_off[d.w_col_off[w + 1]], d.cap_contacts);
}
static __device__ __forceinline__ uint32_t lower_bound_key(const uint64_t *a, uint32_t lo, uint32_t hi, uint64_t k) {
    while (lo < hi) { uint32_t mid = (lo + hi) >> 1; if (a[mid] < k) lo = mid + 1; else hi = mid; }
    return lo;
}

__global__ void k_warm_newflags(Dev d) {
    if (aborted(d)) return;
    const uint32_t nct = min(d.cnt[CT_CONTACTS], d.cap_contacts);
    GRID_STRIDE(c, d.cap_contacts) {
        uint32_t f = 0;
        if (c < nct) f = (d.w_solve[d.col_world[d.ct_a[c]]] && new_entry(d, c)) ? 1u : 0u;
        d.newflag[c] = f;
    }
}
__global__ void k_warm_keepflags(Dev d, TickParams tp) {
    if (aborted(d)) return;
    const uint32_t cur = tp.cur, tot = d.wc_off[cur][d.nw];
    GRID_STRIDE(e, d.cap_warm) {
        uint32_t f = 0;
        if (e < tot) {
            uint32_t w = upper_bound_u32(d.wc_off[cur], d.nw, e) - 1;
            if (!d.w_solve[w]) f = 1;
            else {
                uint64_t k = d.wk_key[cur][e];
                if (idle_body(d, w, (uint32_t)(k >> 32)) && idle_body(d, w, (uint32_t)k)) {
                    uint32_t c0, c1;
                    world_contacts(d, w, c0, c1);
                    bool has_new = false;
                    for (uint32_t c = lower_bound_key(d.ct_key, c0, c1, k); c < c1 && d.ct_key[c] == k; c++) has_new |= d.newflag[c + 1] != d.newflag[c];
                    f = has_new ? 0u : 1u;
                }
            }
        }
        d.keepflag[e] = f;
    }
}
// After both flag arrays are scanned: positions by rank in own list + smaller keys in the other.
__global__ void k_warm_merge(Dev d, TickParams tp) {
    if (aborted(d)) return;
    const uint32_t cur = tp.cur, nxt = cur ^ 1u;
    const uint32_t nct = min(d.cnt[CT_CONTACTS], d.cap_contacts), tot = d.wc_off[cur][d.nw];
    GRID_STRIDE(c, nct) {
        if (d.newflag[c + 1] == d.newflag[c]) continue;  // not a new entry
        uint32_t w = d.col_world[d.ct_a[c]];
        uint64_t k = d.ct_key[c];
        uint32_t e = lower_bound_key(d.wk_key[cur], d.wc_off[cur][w], d.wc_off[cur][w + 1], k);
        uint32_t o = d.newflag[c] + d.keepflag[e];
        if (o >= d.cap_warm) continue;
        float2 t = d.jt[c];
        d.wk_key[nxt][o] = k; d.wk_la[nxt][o] = d.ct_anchor[c]; d.wk_jn[nxt][o] = d.jn_pre[c];
        d.wk_jt[nxt][o] = S3(L3(d.cs_t1[c]) * t.x + L3(d.cs_t2[c]) * t.y);
    }
    GRID_STRIDE(e, tot) {
        if (d.keepflag[e + 1] == d.keepflag[e]) continue;
        uint32_t w = upper_bound_u32(d.wc_off[cur], d.nw, e) - 1;
        uint64_t k = d.wk_key[cur][e];
        uint32_t c0, c1;
        world_contacts(d, w, c0, c1);
        uint32_t c = lower_bound_key(d.ct_key, c0, c1, k);
        uint32_t o = d.keepflag[e] + d.newflag[c];
        if (o >= d.cap_warm) continue;
        d.wk_key[nxt][o] = k; d.wk_la[nxt][o] = d.wk_la[cur][e];
        d.wk_jn[nxt][o] = d.wk_jn[cur][e]; d.wk_jt[nxt][o] = d.wk_jt[cur][e];
    }
    GRID_STRIDE(w, d.nw + 1) {
        uint32_t c0 = w < d.nw ? min(d.c_off[d.w_col_off[w]], d.cap_contacts) : nct;
        uint32_t e0 = d.wc_off[cur][w];
        d.wc_off[nxt][w] = d.newflag[c0] + d.keepflag[e0];
    }
}

#include "rp_cm.h"
#include "rp_bounce.h"
#include "rp_query.h"
#include "rp_player.h"


// ------------------------------------------------------------------ S4
__global__ void k_ball_init(Dev d) {
    if (aborted(d)) return;
    GRID_STRIDE(b, d.nb) {
        uint32_t f = d.flags[b];
        d.ballf[b] = (!(f & FL_CONSTRAINED) || (!(f & FL_JOINTED) && !(f & FL_LOADED))) ? 1u : 0u;
    }
}

// step.rs:180-206 in contact order over each island; only awake dynamic bodies' flags can change
// an outcome, and those are chained.
// Ballistic classification of contact c (step.rs:154-203): bodies still approaching lose the flag.
static __device__ __forceinline__ void ball_row(const Dev &d, const TickParams &tp, uint32_t c) {
    if (c & (RP_JBIT | RP_PBIT)) return;
    uint32_t a = d.cs_ai[c], b = d.cs_bi[c];
    bool hb = b != RP_NONE;
    bool ba = d.ballf[a] != 0, bb = hb && d.ballf[b] != 0;
    if (!ba && !bb) return;
    V3 point = L3(d.ct_p[c]);
    V3 nrm = L3(d.cs_n[c]);
    V3 la = ba ? (L3(d.last_lin[a]) + L3(d.lin[a])) * 0.5f : L3(d.lin[a]);
    V3 wa = L3(d.ang[a]);
    V3 ra = point - L3(d.pos[a]);
    V3 lb = splat3(0.0f), wb = splat3(0.0f), rb = splat3(0.0f), vbl = splat3(0.0f);
    if (hb) { lb = bb ? (L3(d.last_lin[b]) + L3(d.lin[b])) * 0.5f : L3(d.lin[b]); wb = L3(d.ang[b]); rb = point - L3(d.pos[b]); vbl = L3(d.lin[b]); }
    float approach = dot3((lb + cross3(wb, rb)) - (la + cross3(wa, ra)), nrm);
    float vend = dot3((vbl + cross3(wb, rb)) - (L3(d.lin[a]) + cross3(wa, ra)), nrm);
    float s0 = -(d.ct_p[c].w - 0.05f);
    if (s0 + (2.0f * vend) * tp.dt < 0.0f || s0 + approach * tp.dt < 0.0f) {
        if (d.cs_wa[c] != RP_NONE) d.ballf[a] = 0;
        if (d.cs_wb[c] != RP_NONE) d.ballf[b] = 0;
    }
}
// order (tier 0 with the lane tier's projection on): the tier's islands by decreasing size, of
// which this launch takes the first CT_LANE_K (the rest are the lane tier's, k_lane_ball).
template <int TIER>
__global__ void __launch_bounds__(Tier<TIER>::threads) k_sweep_ball(Dev d, TickParams tp, const uint32_t *order) {
    if (aborted(d)) return;
    ISLAND_LOOP_BEGIN_FROM(order ? order : d.tier_list[TIER], order ? d.cnt[CT_LANE_K] : d.cnt[CT_TIER_N + TIER])
        PhaseClock pc;
        run_sweep(I, Slots{d.bs_pa, d.bs_pb, d.bs_c}, [&](uint32_t c) { ball_row(d, tp, c); });
        if constexpr (BIG<TIER>) pc.lap(d, RP_PH_BALL);
    ISLAND_LOOP_END
}

static __device__ Q4 from_axis_angle(V3 axis, float angle) {
    float s, c;
    rp_sincosf(angle * 0.5f, &s, &c);
    V3 v = axis * s;
    return mkq(v.x, v.y, v.z, c);
}
static __device__ Q4 turn(Q4 q, V3 w, float h) {  // step.rs:230-233
    if (len2(w) <= 1.0e-18f) return q;
    return qnormalize(qmul(from_axis_angle(normalize3(w), len3(w) * h), q));
}

// step.rs:208-280
__global__ void k_integrate(Dev d, TickParams tp) {
    if (aborted(d)) return;
    GRID_STRIDE(i, d.nb) {
        float4 p4 = d.pos[i], q4 = d.rot[i];
        d.opos[i] = p4; d.orot[i] = q4;
        d.ppos[i] = make_float4(0, 0, 0, 0); d.prot[i] = make_float4(0, 0, 0, 1);
        if (!(d.bflags[i] & BF_DYN) || !d.awake[i]) continue;
        const float dt = tp.dt;
        V3 p = L3(p4); Q4 q = LQ(q4);
        V3 vl = L3(d.lin[i]), va = L3(d.ang[i]);
        p = p + vl * dt;
        if (d.ballf[i]) p = p + (L3(d.last_lin[i]) - vl) * (0.5f * dt);
        if (len2(va) > 0.0f) {
            M3 l = load_m9(d.iiw + 9 * i);
            float det = mdet(l);
            bool has_m = isfinite(det) && fabsf(det) > 1.0e-20f;
            V3 mom = has_m ? mv(minverse(l), va) : splat3(0.0f);
            Q4 start = q;
            q = turn(start, va, dt);
            if (has_m) {
                M3 il = d.inv_mass[i] == 0.0f ? mzero() : load_m9(d.iil + 9 * i);
                float mean = ((il.c0.x + il.c1.y) + il.c2.z) / 3.0f;
                M3 dev = msub(il, mdiag(splat3(mean)));
                bool principal = len3(cross3(mom, va)) < (1.0e-5f * len3(mom)) * len3(va);
                float gain = principal ? 0.0f : (dt * fmaxf(fmaxf(len3(dev.c0), len3(dev.c1)), len3(dev.c2))) * len3(mom);
                float gq = ceilf(gain / 0.5f);
                uint32_t n = !(gq >= 1.0f) ? 1u : (gq > 256.0f ? 256u : (uint32_t)gq);
                float h = dt / (float)n;
                Q4 qq = start; V3 wv = va;
                for (uint32_t s = 0; s < n; s++) {
                    Q4 q0 = qq;
                    qq = turn(q0, wv, h);
                    M3 r = mquat(qq);
                    V3 wend = mv(mmul(mmul(r, il), mtrans(r)), mom);
                    for (int t = 0; t < 2; t++) {
                        qq = turn(q0, (wv + wend) * 0.5f, h);
                        r = mquat(qq);
                        wend = mv(mmul(mmul(r, il), mtrans(r)), mom);
                    }
                    wv = wend;
                }
                q = qq;
                set_vel(d, i, vl, wv);
            }
        }
        // a bounced free body's pseudo-velocity moves it and never enters its velocity (step.rs:271-277)
        if (d.w6_on && d.w_w6[d.nw + d.bworld[i]]) {
            p = p + L3(d.w6_pl[i]) * dt;
            const V3 pa = L3(d.w6_pa[i]);
            if (len2(pa) > 0.0f) q = qnormalize(qmul(from_axis_angle(normalize3(pa), len3(pa) * dt), q));
        }
        set_pose(d, i, p, q);
        d.ppos[i] = d.pos[i]; d.prot[i] = d.rot[i];
    }
}

// ------------------------------------------------------------------ S5
// step.rs:423-443 anchors at the step-start poses (pose before integration for every body)
__global__ void k_posc_prep(Dev d) {
    if (aborted(d)) return;
    const uint32_t nislc = d.cnt[CT_ISLC];
    GRID_STRIDE(r, nislc) {
        uint32_t c = d.isl_v2[r];
        if (c & (RP_JBIT | RP_PBIT)) continue;
        uint32_t a = d.cs_ai[c], b = d.cs_bi[c];
        V3 point = L3(d.ct_p[c]);
        V3 la = qrot(qconj(LQ(d.orot[a])), point - L3(d.opos[a]));
        V3 lb = b == RP_NONE ? splat3(0.0f) : qrot(qconj(LQ(d.orot[b])), point - L3(d.opos[b]));
        // a compliant pair is left out of the projection (step.rs:433): its depth never exceeds
        // the slop, so its item does nothing
        const float pen0 = (cm_flags(d, c) & CM_SOFT) ? -INFINITY : d.ct_p[c].w - 0.05f;
        d.pc_la[c] = S3(la, pen0);
        d.pc_lb[c] = S3(lb);
    }
}

// step.rs:447-497 one projection pass item, split into its constants (loaded before the item's
// wait) and the pose-dependent part
struct PRow {
    uint32_t a, b;
    bool am, bm;
    float4 la4;
    V3 lb, n, pt;
    float ima, imb;
    M3 ila, ilb;  // local inverse inertia
};
static __device__ __forceinline__ PRow load_prow(const Dev &d, uint32_t c) {
    PRow x;
    if (c & (RP_JBIT | RP_PBIT)) {  // a joint's position pass (a patch has none): only the entry travels
        x.a = c; x.b = RP_NONE; x.am = x.bm = false; x.la4 = make_float4(0, 0, 0, 0); x.lb = x.n = x.pt = splat3(0.0f);
        x.ima = x.imb = 0.0f; x.ila = x.ilb = mzero();
        return x;
    }
    x.a = d.cs_ai[c]; x.b = d.cs_bi[c];
    const bool hb = x.b != RP_NONE;
    x.ima = d.inv_mass[x.a];
    x.imb = hb ? d.inv_mass[x.b] : 0.0f;
    x.am = x.ima > 0.0f && d.awake[x.a];
    x.bm = hb && x.imb > 0.0f && d.awake[x.b];
    x.la4 = d.pc_la[c];
    x.n = L3(d.cs_n[c]);
    x.lb = hb ? L3(d.pc_lb[c]) : splat3(0.0f);
    x.pt = hb ? splat3(0.0f) : L3(d.ct_p[c]);
    x.ila = load_m9(d.iil + 9 * x.a);
    x.ilb = hb ? load_m9(d.iil + 9 * x.b) : mzero();
    return x;
}
static __device__ __forceinline__ void store_prow(const Dev &d, uint32_t slot, const PRow &x) {
    float4 *o = d.srow + slot;
    const size_t C = d.cap_contacts;
    o[0] = f4u(x.a, x.b, x.ima, x.imb);
    o[C] = make_float4(__uint_as_float((x.am ? 1u : 0u) | (x.bm ? 2u : 0u)), x.lb.x, x.lb.y, x.lb.z);
    o[2 * C] = x.la4; o[3 * C] = S3(x.n, x.pt.x); o[4 * C] = make_float4(x.pt.y, x.pt.z, x.ila.c0.x, x.ila.c0.y);
    o[5 * C] = make_float4(x.ila.c0.z, x.ila.c1.x, x.ila.c1.y, x.ila.c1.z); o[6 * C] = make_float4(x.ila.c2.x, x.ila.c2.y, x.ila.c2.z, x.ilb.c0.x);
    o[7 * C] = make_float4(x.ilb.c0.y, x.ilb.c0.z, x.ilb.c1.x, x.ilb.c1.y); o[8 * C] = make_float4(x.ilb.c1.z, x.ilb.c2.x, x.ilb.c2.y, x.ilb.c2.z);
}
static __device__ __forceinline__ PRow fetch_prow(const Dev &d, uint32_t slot) {
    const float4 *o = d.srow + slot;
    const size_t C = d.cap_contacts;
    float4 f0 = o[0], f1 = o[C], f2 = o[2 * C], f3 = o[3 * C], f4 = o[4 * C], f5 = o[5 * C], f6 = o[6 * C], f7 = o[7 * C], f8 = o[8 * C];
    PRow x;
    x.a = __float_as_uint(f0.x); x.b = __float_as_uint(f0.y); x.ima = f0.z; x.imb = f0.w;
    const uint32_t fl = __float_as_uint(f1.x);
    x.am = (fl & 1u) != 0; x.bm = (fl & 2u) != 0;
    x.lb = mk3(f1.y, f1.z, f1.w); x.la4 = f2; x.n = L3(f3); x.pt = mk3(f3.w, f4.x, f4.y);
    x.ila.c0 = mk3(f4.z, f4.w, f5.x); x.ila.c1 = mk3(f5.y, f5.z, f5.w); x.ila.c2 = mk3(f6.x, f6.y, f6.z);
    x.ilb.c0 = mk3(f6.w, f7.x, f7.y); x.ilb.c1 = mk3(f7.z, f7.w, f8.x); x.ilb.c2 = mk3(f8.y, f8.z, f8.w);
    return x;
}
static __device__ void project_row(const Dev &d, const PRow &x, bool shock, V3 up, float dt) {
    if (x.a & RP_JBIT) { if (!shock) joint_pos(d, x.a & ~RP_JBIT, dt); return; }
    if (x.a & RP_PBIT) return;
    const uint32_t a = x.a, b = x.b;
    const bool hb = b != RP_NONE;
    bool am = x.am, bm = x.bm;
    if (shock && am && bm) {
        float ha = dot3(L3(d.pos[a]), up), hbb = dot3(L3(d.pos[b]), up);
        if (ha < hbb) am = false; else bm = false;
    }
    const float4 la4 = x.la4;
    const V3 n = x.n;
    V3 pa = L3(d.pos[a]); Q4 qa = LQ(d.rot[a]);
    V3 wa = pa + qrot(qa, L3(la4));
    V3 pb = splat3(0.0f), wb; Q4 qb = qid();
    if (hb) { pb = L3(d.pos[b]); qb = LQ(d.rot[b]); wb = pb + qrot(qb, x.lb); }
    else wb = x.pt;
    float pen = la4.w - dot3(wb - wa, n);
    float err = pen - 0.005f;
    if (err <= 0.0f) return;
    V3 ra = wa - pa, rb = hb ? wb - pb : splat3(0.0f);
    float ima = 0.0f, imb = 0.0f; M3 iia = mzero(), iib = mzero();
    if (am) { ima = x.ima; iia = inertia_world_of(x.ila, qa); }
    if (hb && bm) { imb = x.imb; iib = inertia_world_of(x.ilb, qb); }
    V3 ca = cross3(ra, n), cb = cross3(rb, n);
    float wsum = ((ima + imb) + dot3(ca, mv(iia, ca))) + dot3(cb, mv(iib, cb));
    if (wsum <= 1.0e-9f) return;
    float lambda = err / wsum;
    for (int side = 0; side < 2; side++) {
        uint32_t id; V3 p; Q4 q; float im; M3 ii; V3 r; V3 j;
        if (side == 0) { if (!am) continue; id = a; p = pa; q = qa; im = ima; ii = iia; r = ra; j = (-n) * lambda; }
        else { if (!(hb && bm)) continue; id = b; p = pb; q = qb; im = imb; ii = iib; r = rb; j = n * lambda; }
        V3 dth = mv(ii, cross3(r, j));
        d.slin[id] = S3(L3(d.slin[id]) + j * im);
        d.sang[id] = S3(L3(d.sang[id]) + dth);
        Q4 dq = qmul(mkq(dth.x, dth.y, dth.z, 0.0f), q);
        dq = mkq(dq.x * 0.5f, dq.y * 0.5f, dq.z * 0.5f, dq.w * 0.5f);
        Q4 nr = qnormalize(mkq(q.x + dq.x, q.y + dq.y, q.z + dq.z, q.w + dq.w));
        set_pose(d, id, p + j * im, nr);
    }
}

// Position projection iterations in list order (step.rs:297-310); the shock pass follows in its
// own launch once the shock order of the projected poses is built.
template <int TIER>
__global__ void __launch_bounds__(Tier<TIER>::threads) k_sweep_proj(Dev d, TickParams tp, uint32_t split, const uint32_t *order) {
    if (aborted(d)) return;
    const V3 up = mk3(tp.upx, tp.upy, tp.upz);
    ISLAND_LOOP_BEGIN_FROM(order ? order : d.tier_list[TIER], order ? d.cnt[CT_LANE_K] : d.cnt[CT_TIER_N + TIER])
        PhaseClock pc;
        const Slots sb{d.bs_pa, d.bs_pb, d.bs_c};
        if (BIG<TIER> && split) continue;  // k_store_prows, then k_level_proj + k_sweep_proj_tail
        for (uint32_t j = threadIdx.x; j < n; j += blockDim.x) store_prow(d, s + j, load_prow(d, d.bs_c[s + j]));
        __syncthreads();
        for (uint32_t it = 0; it < tp.ip; it++)
            run_sweep(I, sb, [&](uint32_t, uint32_t slot) { return fetch_prow(d, slot); }, [&](uint32_t, const PRow &x) { project_row(d, x, false, up, tp.dt); });
        if constexpr (BIG<TIER>) pc.lap(d, RP_PH_PROJ);
    ISLAND_LOOP_END
}
// Split form: one projection sweep over the slots after the wide levels k_level_proj ran.
template <int TIER>
__global__ void __launch_bounds__(Tier<TIER>::threads) k_sweep_proj_tail(Dev d, TickParams tp, uint32_t kl) {
    if (aborted(d)) return;
    const V3 up = mk3(tp.upx, tp.upy, tp.upz);
    ISLAND_LOOP_BEGIN
        PhaseClock pc;
        const uint32_t k = min(d.isl_w[isl], kl);
        run_sweep_from(I, Slots{d.bs_pa, d.bs_pb, d.bs_c}, k ? d.lvl_off[s + k] : 0u,
                       [&](uint32_t, uint32_t slot) { return fetch_prow(d, slot); }, [&](uint32_t, const PRow &x) { project_row(d, x, false, up, tp.dt); });
        if constexpr (BIG<TIER>) pc.lap(d, RP_PH_PROJ);
    ISLAND_LOOP_END
}
// The projection rows of every big island's slots (split form), across the GPU.
__global__ void __launch_bounds__(256) k_store_prows(Dev d) {
    if (aborted(d)) return;
    for_big_island_slots(d, [&](uint32_t slot) { store_prow(d, slot, load_prow(d, d.bs_c[slot])); });
}
// Wide level L of the projection sweep, across the GPU (as k_level_vel).
__global__ void __launch_bounds__(256) k_level_proj(Dev d, TickParams tp, uint32_t L) {
    if (aborted(d)) return;
    const V3 up = mk3(tp.upx, tp.upy, tp.upz);
    for_wide_level(d, d.wide_list, d.cnt[CT_WIDE_N], d.isl_w, d.lvl_off, L, [&](uint32_t slot) { project_row(d, fetch_prow(d, slot), false, up, tp.dt); });
}
// Projection shock pass (step.rs:299-301, 447-497 with shock = true) in shock order.
template <int TIER>
__global__ void __launch_bounds__(Tier<TIER>::threads) k_sweep_proj_shock(Dev d, TickParams tp, uint32_t klsh, const uint32_t *order) {
    if (aborted(d)) return;
    const V3 up = mk3(tp.upx, tp.upy, tp.upz);
    ISLAND_LOOP_BEGIN_FROM(order ? order : d.tier_list[TIER], order ? d.cnt[CT_LANE_K] : d.cnt[CT_TIER_N + TIER])
        PhaseClock pc;
        shock_slots<TIER>(d, tp, I, isl);
        const Slots ss{d.ss_pa, d.ss_pb, d.ss_c};
        if constexpr (BIG<TIER>) pc.lap(d, RP_PH_PROJ_SHOCK_ORDER);
        if (BIG<TIER> && klsh) continue;  // k_level_pshock + k_sweep_pshock_tail follow
        run_sweep(I, ss, [&](uint32_t c, uint32_t) { return load_prow(d, c); }, [&](uint32_t, const PRow &x) { project_row(d, x, true, up, tp.dt); });
        if constexpr (BIG<TIER>) pc.lap(d, RP_PH_PROJ_SHOCK);
    ISLAND_LOOP_END
}

// The lane tier's ballistic pass, projection iterations and projection shock pass (when the
// projection iterations are few, tp.lane_pos): one island per thread, entries in list order (shock
// order for the shock pass), the same rows as the sweeps.
#define LANE_ISLANDS_BEGIN                                                                           \
    const uint32_t nisl = d.cnt[CT_ISLANDS], nislc = min(d.cnt[CT_ISLC], d.cap_contacts);            \
    const uint32_t *const lane_list = order + d.cnt[CT_LANE_K];                                      \
    GRID_STRIDE(q, d.cnt[CT_LANE_N]) {                                                               \
        const uint32_t isl = lane_list[q], s = d.isl_start[isl], n = island_len(d, isl, nisl, nislc);
__global__ void __launch_bounds__(64) k_lane_ball(Dev d, TickParams tp, const uint32_t *order) {
    if (aborted(d)) return;
    LANE_ISLANDS_BEGIN
        for (uint32_t j = 0; j < n; j++) ball_row(d, tp, d.isl_v2[s + j]);
    }
}
__global__ void __launch_bounds__(64) k_lane_proj(Dev d, TickParams tp, const uint32_t *order) {
    if (aborted(d)) return;
    const V3 up = mk3(tp.upx, tp.upy, tp.upz);
    LANE_ISLANDS_BEGIN
        for (uint32_t it = 0; it < tp.ip; it++)
            for (uint32_t j = 0; j < n; j++) project_row(d, load_prow(d, d.isl_v2[s + j]), false, up, tp.dt);
    }
}
__global__ void __launch_bounds__(64) k_lane_pshock(Dev d, TickParams tp, const uint32_t *order) {
    if (aborted(d)) return;
    const V3 up = mk3(tp.upx, tp.upy, tp.upz);
    LANE_ISLANDS_BEGIN
        for (uint32_t j = 0; j < n; j++) project_row(d, load_prow(d, d.sh_ord[s + j]), true, up, tp.dt);
    }
}

// Split form of the projection shock pass: wide level L across the GPU, then the tail.
__global__ void __launch_bounds__(256) k_level_pshock(Dev d, TickParams tp, uint32_t L) {
    if (aborted(d)) return;
    const V3 up = mk3(tp.upx, tp.upy, tp.upz);
    for_wide_level(d, d.big_list, d.cnt[CT_BIG_N], d.isl_wsh, d.lvl_off_sh, L,
                   [&](uint32_t slot) { project_row(d, load_prow(d, d.ss_c[slot]), true, up, tp.dt); });
}
template <int TIER>
__global__ void __launch_bounds__(Tier<TIER>::threads) k_sweep_pshock_tail(Dev d, TickParams tp, uint32_t klsh) {
    if (aborted(d)) return;
    const V3 up = mk3(tp.upx, tp.upy, tp.upz);
    ISLAND_LOOP_BEGIN
        const uint32_t k = min(d.isl_wsh[isl], klsh);
        run_sweep_from(I, Slots{d.ss_pa, d.ss_pb, d.ss_c}, k ? d.lvl_off_sh[s + k] : 0u, [&](uint32_t c, uint32_t) { return load_prow(d, c); },
                       [&](uint32_t, const PRow &x) { project_row(d, x, true, up, tp.dt); });
    ISLAND_LOOP_END
}

// ------------------------------------------------------------------ S6 / S7
__global__ void k_s6(Dev d, TickParams tp) {
    if (aborted(d)) return;
    GRID_STRIDE(i, d.nb) {
        if ((d.bflags[i] & BF_DYN) && d.awake[i]) {
            V3 vl = L3(d.lin[i]), va = L3(d.ang[i]);
            V3 np = L3(d.pos[i]); Q4 nq = LQ(d.rot[i]);
            Q4 pq = LQ(d.prot[i]);
            // no split impulse on a jointed body (step.rs:322-327)
            const bool jb = (d.flags[i] & FL_JOINTED) != 0;
            V3 dl = (np - L3(d.ppos[i])) - (jb ? splat3(0.0f) : L3(d.slin[i]));
            if (!zero3(dl)) vl = vl + dl / tp.dt;
            if ((d.flags[i] & FL_GROUND) && vl.y < 0.0f) {
                vl.y = 0.0f;
                np.y = d.opos[i].y;
                set_pose(d, i, np, nq);
            }
            if (!(d.flags[i] & FL_DRIVEN) && !eqq(nq, pq)) {
                Q4 dq = qmul(nq, qconj(pq));
                V3 da = mk3(dq.x, dq.y, dq.z) * 2.0f;
                if (dq.w < 0.0f) da = -da;
                va = va + (da - (jb ? splat3(0.0f) : L3(d.sang[i]))) / tp.dt;
            }
            set_vel(d, i, vl, va);
        }
        d.last_lin[i] = d.lin[i];
        // sleep.rs:59-80
        uint32_t bf = d.bflags[i];
        if ((bf & BF_SLEEP_OK) && d.awake[i] && (bf & BF_DYN)) {
            if (len2(L3(d.lin[i])) < 0.01f && len2(L3(d.ang[i])) < 0.01f) {
                uint32_t t = (d.sticks[i] + 1u) & 0xFFFFu;
                d.sticks[i] = t;
                if (t > 30u) d.awake[i] = 0;
            } else d.sticks[i] = 0;
        }
    }
}

// Overflow: restore the pre-tick state, keep the failed tick from repeating its mutations.
__global__ void k_restore_backup(Dev d) {
    GRID_STRIDE(b, d.nb) {
        d.pos[b] = d.bk_pos[b]; d.rot[b] = d.bk_rot[b]; d.lin[b] = d.bk_lin[b]; d.ang[b] = d.bk_ang[b];
        d.last_lin[b] = d.bk_last[b]; d.awake[b] = d.bk_awake[b]; d.sticks[b] = d.bk_sticks[b];
    }
    GRID_STRIDE(q, d.nj) {
        d.wj_mask[q] = d.bk_wj_mask[q];
        for (uint32_t i = 0; i < 3; i++) d.wj_lam[3 * q + i] = d.bk_wj_lam[3 * q + i];
        if (d.d6_on) {
            d.d6_wn[q] = d.bk_d6_wn[q];
            for (uint32_t r = 0; r < 32; r++) { d.d6_wtag[32 * q + r] = d.bk_d6_wtag[32 * q + r]; d.d6_wlam[32 * q + r] = d.bk_d6_wlam[32 * q + r]; }
        }
    }
    GRID_STRIDE(p, d.np) { d.pl_state[p] = d.bk_pl_state[p]; d.pl_grounded[p] = d.bk_pl_grounded[p]; }
}

// Capacity checks after each sizing scan; sets the abort flag so later kernels return at once.
// Once a tick is aborted the later sizing scans run over stale counts, so their checks record
// nothing: only the first overflow's needs are grown before the replay.
__global__ void k_check(Dev d, uint32_t which) {
    if (threadIdx.x != 0 || blockIdx.x != 0 || d.cnt[CT_ABORT]) return;
    if (which == 0) {
        if (d.cnt[CT_PAIRS] > d.cap_pairs) { d.cnt[CT_NEED_PAIRS] = d.cnt[CT_PAIRS]; d.cnt[CT_ABORT] = 1; }
        if (d.cnt[CT_CANDS] > d.cap_cands) { d.cnt[CT_NEED_CANDS] = d.cnt[CT_CANDS]; d.cnt[CT_ABORT] = 1; }
    } else if (which == 1) {
        // the island lists hold the joints too, and the friction patches (at most one per two
        // contacts; their count is added once they are found)
        d.cnt[CT_ROWS] = d.cnt[CT_CONTACTS] + d.nj;
        const uint32_t need = d.cnt[CT_ROWS] + (d.cm_on ? d.cnt[CT_CONTACTS] / 2u : 0u);
        if (need > d.cap_contacts) { d.cnt[CT_NEED_CONTACTS] = need; d.cnt[CT_ABORT] = 1; }
    } else if (which == 2) {
        uint32_t wt = d.cnt[CT_NEW_TOTAL] + d.cnt[CT_KEEP_TOTAL];
        if (d.cm_on) wt = max(wt, d.cnt[CT_WP_NEW] + d.cnt[CT_WP_KEEP]);  // the patch cache shares the capacity
        if (wt > d.cap_warm) { d.cnt[CT_NEED_WARM] = wt; d.cnt[CT_ABORT] = 1; }
    }
}
__global__ void k_zero_u32(uint32_t *p, uint32_t n) { GRID_STRIDE(i, n) p[i] = 0; }
__global__ void k_tick_done(Dev d) {
    if (threadIdx.x != 0 || blockIdx.x != 0 || d.cnt[CT_ABORT]) return;
    uint32_t *c = d.cnt;
    c[CT_DONE] += 1;
    c[CT_SUM_PAIRS] += c[CT_PAIRS]; c[CT_SUM_CANDS] += c[CT_CANDS]; c[CT_SUM_CONTACTS] += c[CT_CONTACTS];
    c[CT_SUM_ISLC] += c[CT_ISLC]; c[CT_SUM_ISLANDS] += c[CT_ISLANDS]; c[CT_SUM_ISLC_GLB] += c[CT_ISLC_GLB];
    c[CT_SUM_ITERS_MAX] += c[CT_ITERS_MAX];
    const uint64_t isum = ((uint64_t)c[CT_SUM_ITERS_SUM_HI] << 32 | c[CT_SUM_ITERS_SUM]) + c[CT_ITERS_SUM];
    c[CT_SUM_ITERS_SUM] = (uint32_t)isum; c[CT_SUM_ITERS_SUM_HI] = (uint32_t)(isum >> 32);
    const uint64_t csum = ((uint64_t)c[CT_SUM_ITERS_CONV_HI] << 32 | c[CT_SUM_ITERS_CONV]) + c[CT_ITERS_SUM_CONV];
    c[CT_SUM_ITERS_CONV] = (uint32_t)csum; c[CT_SUM_ITERS_CONV_HI] = (uint32_t)(csum >> 32);
    c[CT_SUM_WIDE] += c[CT_WIDE_MAX]; c[CT_SUM_WIDE_SH] += c[CT_WSH_MAX];
    c[CT_SUM_ISL_CONV] += c[CT_ISL_CONV]; c[CT_SUM_ISLC_CONV] += c[CT_ISLC_CONV]; c[CT_SUM_WS_LEARNED] += c[CT_WS_LEARNED];
}
__global__ void k_set_counter(uint32_t *cnt, uint32_t idx, uint32_t v) { if (threadIdx.x == 0 && blockIdx.x == 0) cnt[idx] = v; }
__global__ void k_copy_u32(const uint32_t *src, uint32_t *dst, uint32_t n) { GRID_STRIDE(i, n) dst[i] = src[i]; }
__global__ void k_copy_counted(const uint32_t *src, uint32_t *dst, const uint32_t *n) { GRID_STRIDE(i, *n) dst[i] = src[i]; }

// ------------------------------------------------------------------ host-side launch sequence
extern "C" void rp_launch_mlp_infer(const float *p, const __bf16 *w1p, const __bf16 *w2p, const __bf16 *w3p, const float *feat, const uint32_t *n_ptr,
                                    uint32_t cap, float *out, uint32_t *n_trusted, hipStream_t st);
static inline uint32_t grid_for(uint32_t n, uint32_t block, uint32_t max_grid = 8192) {
    uint32_t g = (n + block - 1) / block;
    if (g < 1) g = 1;
    if (g > max_grid) g = max_grid;
    return g;
}

// Exclusive scan of n u32 in place; the total goes to *total (device).  n + 1 slots are written
// when with_total_slot (the scan of counts gives offsets with the end in slot n).
static void scan_u32(uint32_t *a, uint32_t n, uint32_t *tmp, uint32_t *total, hipStream_t st) {
    uint32_t nb = (n + RP_SCAN_BLOCK - 1) / RP_SCAN_BLOCK;
    if (nb < 1) nb = 1;
    hipLaunchKernelGGL(k_scan_reduce, dim3(nb), dim3(256), 0, st, a, n, tmp);
    hipLaunchKernelGGL(k_scan_sums, dim3(1), dim3(256), 0, st, tmp, nb, total);
    hipLaunchKernelGGL(k_scan_apply, dim3(nb), dim3(256), 0, st, a, n, tmp, a);
}

__global__ void k_store_total(uint32_t *a, uint32_t n, const uint32_t *total) { if (threadIdx.x == 0 && blockIdx.x == 0) a[n] = *total; }

// Stable LSD radix sort of (keys, vals) by the low `bits` of the keys; n read from device.
// Result ends in (k_out, v_out); ping-pong buffers given.
static void radix_sort(const Dev &d, uint32_t *k0, uint32_t *v0, uint32_t *k1, uint32_t *v1, const uint32_t *n_dev, uint32_t cap,
                       uint32_t bits, hipStream_t st, uint32_t **k_out, uint32_t **v_out) {
    uint32_t nblocks = (cap + RP_RS_TILE - 1) / RP_RS_TILE;
    if (nblocks < 1) nblocks = 1;
    uint32_t *ka = k0, *va = v0, *kb = k1, *vb = v1;
    for (uint32_t shift = 0; shift < bits; shift += 8) {
        hipLaunchKernelGGL(k_rs_hist, dim3(nblocks), dim3(256), 0, st, ka, n_dev, cap, shift, nblocks, d.rs_hist);
        scan_u32(d.rs_hist, 256 * nblocks, d.scan_tmp, d.cnt + CT_SCRATCH, st);
        hipLaunchKernelGGL(k_rs_scatter, dim3(nblocks), dim3(256), 0, st, ka, va, n_dev, cap, shift, nblocks, d.rs_hist, kb, vb);
        uint32_t *t = ka; ka = kb; kb = t;
        t = va; va = vb; vb = t;
    }
    *k_out = ka; *v_out = va;
}

// Radix bits covering keys 0..maxkey, in whole 8-bit passes.
static uint32_t key_bits(uint32_t maxkey) {
    uint32_t b = 8;
    while (b < 32 && (uint64_t)maxkey >= (1ull << b)) b += 8;
    return b;
}

// Chains of the order `ord` (contact at each island-list rank) into (pa, pb).
static void build_chains(const Dev &d, const uint32_t *ord, uint32_t *pa, uint32_t *pb, uint32_t *sa, uint32_t *sb, uint32_t body_bits,
                         hipStream_t st) {
    const uint32_t B = 256, cap2 = 2u * d.cap_contacts;
    hipLaunchKernelGGL(k_chain_entries, dim3(grid_for(cap2, B)), dim3(B), 0, st, d, ord, d.sk0, d.sv0);
    uint32_t *ks, *vs;
    radix_sort(d, d.sk0, d.sv0, d.sk1, d.sv1, d.cnt + CT_CHAIN_N, cap2, body_bits, st, &ks, &vs);
    hipLaunchKernelGGL(k_chain_link, dim3(grid_for(cap2, B)), dim3(B), 0, st, d, ks, vs, pa, pb, sa, sb);
}

// Shock order of every island from the current poses into sh_ord, with its chains: stable radix
// passes over the key's low 32 bits, then its high 12 bits, then the island (root body) key,
// starting from list order, so equal keys keep list order.
static void build_shock_order(const Dev &d, const TickParams &tp, uint32_t body_bits, hipStream_t st) {
    const uint32_t B = 256, C = d.cap_contacts;
    const uint32_t *n_dev = d.cnt + CT_ISLC;
    hipLaunchKernelGGL(k_shock_keys, dim3(grid_for(C, B)), dim3(B), 0, st, d, tp, d.sk0, d.sh_hi, d.sv0);
    uint32_t *ka, *va, *kb, *vb, *kc, *vc;
    radix_sort(d, d.sk0, d.sv0, d.sk1, d.sv1, n_dev, C, 32, st, &ka, &va);
    uint32_t *kf = ka == d.sk0 ? d.sk1 : d.sk0, *vf = va == d.sv0 ? d.sv1 : d.sv0;
    hipLaunchKernelGGL(k_gather_keys, dim3(grid_for(C, B)), dim3(B), 0, st, d, d.sh_hi, va, kf);
    radix_sort(d, kf, va, ka, vf, n_dev, C, 16, st, &kb, &vb);
    kf = kb == d.sk0 ? d.sk1 : d.sk0;
    vf = vb == d.sv0 ? d.sv1 : d.sv0;
    hipLaunchKernelGGL(k_gather_keys, dim3(grid_for(C, B)), dim3(B), 0, st, d, d.isl_k2, vb, kf);
    radix_sort(d, kf, vb, kb, vf, n_dev, C, body_bits, st, &kc, &vc);
    hipLaunchKernelGGL(k_shock_order, dim3(grid_for(C, B)), dim3(B), 0, st, d, vc);
    build_chains(d, d.sh_ord, d.sh_pa, d.sh_pb, d.sh_sa, d.sh_sb, body_bits, st);
}

// Enqueues one tick; returns the launch status (hipGetLastError after the last launch).  When ev
// is given, ev[k] is recorded on st at stage boundary k (RP_NMARKS events; stage k spans
// ev[k]..ev[k+1]).
// grids: workgroups of the island sweeps per tier (<= 4,096 contacts, <= 131,072, larger); any
// positive counts give the same results (islands are taken grid-stride).
extern "C" hipError_t rp_launch_tick(const Dev &d, const TickParams &tp, const uint32_t *grids, hipStream_t st, hipEvent_t *ev,
                                     const TierStreams *ts) {
    const uint32_t B = 256;
    auto mark = [&](int k) { if (ev) (void)hipEventRecord(ev[k], st); };
    // the four tiers of one sweep stage: tier 0 on side[0], tier 3 then tier 2 on side[1], tier 1 on
    // the tick's stream, concurrently (all solve disjoint islands); in order without side streams
    auto tiers = [&](auto k0, auto k3, auto k1, auto k2) {
        if (!ts) { k0(st); k3(st); k1(st); k2(st); return; }
        (void)hipEventRecord(ts->ev[0], st);
        (void)hipStreamWaitEvent(ts->side[0], ts->ev[0], 0);
        (void)hipStreamWaitEvent(ts->side[1], ts->ev[0], 0);
        k0(ts->side[0]);
        k1(st);
        k3(ts->side[1]);
        k2(ts->side[1]);
        (void)hipEventRecord(ts->ev[1], ts->side[0]);
        (void)hipEventRecord(ts->ev[2], ts->side[1]);
        (void)hipStreamWaitEvent(st, ts->ev[1], 0);
        (void)hipStreamWaitEvent(st, ts->ev[2], 0);
    };
    mark(RP_ST_PRE);
    uint32_t *scratch_total = d.cnt + CT_SCRATCH;
    if (d.np) {  // World::update_players runs first (step.rs:51)
        hipLaunchKernelGGL(k_player_backup, dim3(grid_for(d.nb, B)), dim3(B), 0, st, d);
        hipLaunchKernelGGL(k_players, dim3(d.nw), dim3(RPP_T), 0, st, d, tp);
    }
    hipLaunchKernelGGL(k_bodies_pre, dim3(grid_for(d.nb > d.nj ? d.nb : d.nj, B)), dim3(B), 0, st, d, tp);
    hipLaunchKernelGGL(k_col_aabb, dim3(grid_for(d.nc, B)), dim3(B), 0, st, d, tp);
    mark(RP_ST_BROAD);
    if (!d.bp_radix) {
        hipLaunchKernelGGL(k_bp_sort, dim3(d.nw), dim3(1024), 0, st, d);
    } else {
        hipLaunchKernelGGL(k_zero_u32, dim3(grid_for(d.nw, B)), dim3(B), 0, st, d.w_nsap, d.nw);
        hipLaunchKernelGGL(k_zero_u32, dim3(grid_for(d.nw, B)), dim3(B), 0, st, d.w_nover, d.nw);
        hipLaunchKernelGGL(k_set_counter, dim3(1), dim3(1), 0, st, d.cnt, (uint32_t)CT_NCOL, d.nc);
        hipLaunchKernelGGL(k_bp_keys, dim3(grid_for(d.nc, B)), dim3(B), 0, st, d, d.bp_k0, d.bp_v0);
        uint32_t *ka, *va, *kb, *vb;
        radix_sort(d, d.bp_k0, d.bp_v0, d.bp_k1, d.bp_v1, d.cnt + CT_NCOL, d.nc, 32, st, &ka, &va);
        uint32_t *kf = ka == d.bp_k0 ? d.bp_k1 : d.bp_k0, *vf = va == d.bp_v0 ? d.bp_v1 : d.bp_v0;
        hipLaunchKernelGGL(k_bp_world_keys, dim3(grid_for(d.nc, B)), dim3(B), 0, st, d, va, kf);
        radix_sort(d, kf, va, ka, vf, d.cnt + CT_NCOL, d.nc, key_bits(d.nw - 1u), st, &kb, &vb);
        hipLaunchKernelGGL(k_bp_ranks, dim3(grid_for(d.nc, B)), dim3(B), 0, st, d, vb);
    }
    hipLaunchKernelGGL(k_zero_u32, dim3(grid_for(d.nc, B)), dim3(B), 0, st, d.out_off, d.nc);
    hipLaunchKernelGGL(k_bp_enum<0>, dim3(grid_for(d.nc * RP_BP_G, B)), dim3(B), 0, st, d);
    scan_u32(d.out_off, d.nc, d.scan_tmp, d.cnt + CT_PAIRS, st);
    hipLaunchKernelGGL(k_store_total, dim3(1), dim3(1), 0, st, d.out_off, d.nc, d.cnt + CT_PAIRS);
    scan_u32(d.cand_off, d.nc, d.scan_tmp, d.cnt + CT_CANDS, st);
    hipLaunchKernelGGL(k_store_total, dim3(1), dim3(1), 0, st, d.cand_off, d.nc, d.cnt + CT_CANDS);
    hipLaunchKernelGGL(k_check, dim3(1), dim3(64), 0, st, d, 0u);
    hipLaunchKernelGGL(k_zero_u32, dim3(grid_for(d.nc, B)), dim3(B), 0, st, d.bp_fill, d.nc);
    hipLaunchKernelGGL(k_bp_enum<1>, dim3(grid_for(d.nc * RP_BP_G, B)), dim3(B), 0, st, d);
    hipLaunchKernelGGL(k_bp_sort_lists, dim3(grid_for(d.nc, B)), dim3(B), 0, st, d);
    mark(RP_ST_NARROW_PAIRS);
    hipLaunchKernelGGL(k_narrow_pairs, dim3(grid_for(d.cap_pairs, B)), dim3(B), 0, st, d);
    mark(RP_ST_NARROW_TERRAIN);
    hipLaunchKernelGGL(k_narrow_terrain, dim3(grid_for(d.cap_cands, B, 16384)), dim3(B), 0, st, d);
    mark(RP_ST_CONTACTS);
    hipLaunchKernelGGL(k_terrain_cull, dim3(grid_for(d.nc, B)), dim3(B), 0, st, d);
    hipLaunchKernelGGL(k_contact_count, dim3(grid_for(d.nc, B)), dim3(B), 0, st, d);
    scan_u32(d.c_off, d.nc, d.scan_tmp, d.cnt + CT_CONTACTS, st);
    hipLaunchKernelGGL(k_store_total, dim3(1), dim3(1), 0, st, d.c_off, d.nc, d.cnt + CT_CONTACTS);
    hipLaunchKernelGGL(k_check, dim3(1), dim3(64), 0, st, d, 1u);
    hipLaunchKernelGGL(k_contact_write, dim3(grid_for(d.nc, B)), dim3(B), 0, st, d);
    if (d.cm_on) {  // friction patches: runs of a patch-friction pair's contacts
        hipLaunchKernelGGL(k_patch_heads, dim3(grid_for(d.cap_contacts, B)), dim3(B), 0, st, d);
        scan_u32(d.pt_head, d.cap_contacts, d.scan_tmp, d.cnt + CT_NPATCH, st);
        hipLaunchKernelGGL(k_store_total, dim3(1), dim3(1), 0, st, d.pt_head, d.cap_contacts, d.cnt + CT_NPATCH);
        hipLaunchKernelGGL(k_patch_fill, dim3(grid_for(d.cap_contacts, B)), dim3(B), 0, st, d);
        hipLaunchKernelGGL(k_rows_patches, dim3(1), dim3(1), 0, st, d);
    }
    mark(RP_ST_PREP);
    hipLaunchKernelGGL(k_zero_u32, dim3(grid_for(d.nw, B)), dim3(B), 0, st, d.w_solve, d.nw);
    if (d.w6_on) hipLaunchKernelGGL(k_zero_u32, dim3(grid_for(2 * d.nw, B)), dim3(B), 0, st, d.w_w6, 2 * d.nw);
    if (d.nj) hipLaunchKernelGGL(k_joint_prep, dim3(grid_for(d.nj, B)), dim3(B), 0, st, d, tp);
    hipLaunchKernelGGL(k_prep, dim3(grid_for(d.cap_contacts, B)), dim3(B), 0, st, d, tp);
    if (d.cm_on) hipLaunchKernelGGL(k_patch_prep, dim3(grid_for(d.cap_patch, B)), dim3(B), 0, st, d);
    hipLaunchKernelGGL(k_island_keys, dim3(grid_for(d.cap_contacts, B)), dim3(B), 0, st, d);
    const uint32_t body_bits = key_bits(d.nb);  // body ids and the nb sentinel
    uint32_t *ks, *vs;
    radix_sort(d, d.isl_k, d.isl_v, d.isl_k2, d.isl_v2, d.cnt + CT_ROWS, d.cap_contacts, body_bits, st, &ks, &vs);
    if (vs != d.isl_v2) {  // the sweeps read (isl_k2, isl_v2)
        hipLaunchKernelGGL(k_copy_u32, dim3(grid_for(d.cap_contacts, B)), dim3(B), 0, st, vs, d.isl_v2, d.cap_contacts);
        hipLaunchKernelGGL(k_copy_u32, dim3(grid_for(d.cap_contacts, B)), dim3(B), 0, st, ks, d.isl_k2, d.cap_contacts);
    }
    hipLaunchKernelGGL(k_set_counter, dim3(1), dim3(1), 0, st, d.cnt, (uint32_t)CT_ISLC, 0u);
    hipLaunchKernelGGL(k_island_flags, dim3(grid_for(d.cap_contacts, B)), dim3(B), 0, st, d, d.isl_k2);
    scan_u32(d.isl_flag, d.cap_contacts, d.scan_tmp, d.cnt + CT_ISLANDS, st);
    hipLaunchKernelGGL(k_island_compact, dim3(grid_for(d.cap_contacts, B)), dim3(B), 0, st, d, d.isl_k2);
    for (uint32_t t = 0; t < 5; t++) hipLaunchKernelGGL(k_set_counter, dim3(1), dim3(1), 0, st, d.cnt, (uint32_t)CT_TIER_N + t, 0u);
    if (d.cm_on) {  // islands with a compliant row or a patch go to tier 4
        hipLaunchKernelGGL(k_zero_u32, dim3(grid_for(d.cap_contacts + 1, B)), dim3(B), 0, st, d.isl_cm, d.cap_contacts + 1);
        hipLaunchKernelGGL(k_isl_cm, dim3(grid_for(d.cap_contacts, B)), dim3(B), 0, st, d, d.isl_k2);
    }
    hipLaunchKernelGGL(k_tier_lists, dim3(grid_for(d.cap_contacts, B)), dim3(B), 0, st, d, tp);
    hipLaunchKernelGGL(k_warm_match, dim3(grid_for(d.cap_contacts, B)), dim3(B), 0, st, d, tp);
    hipLaunchKernelGGL(k_set_counter, dim3(1), dim3(1), 0, st, d.cnt, (uint32_t)CT_ISLC_GLB, 0u);
    hipLaunchKernelGGL(k_set_counter, dim3(1), dim3(1), 0, st, d.cnt, (uint32_t)CT_ITERS_MAX, 0u);
    hipLaunchKernelGGL(k_set_counter, dim3(1), dim3(1), 0, st, d.cnt, (uint32_t)CT_ITERS_SUM, 0u);
    hipLaunchKernelGGL(k_set_counter, dim3(1), dim3(1), 0, st, d.cnt, (uint32_t)CT_ISL_CONV, 0u);
    hipLaunchKernelGGL(k_set_counter, dim3(1), dim3(1), 0, st, d.cnt, (uint32_t)CT_ISLC_CONV, 0u);
    hipLaunchKernelGGL(k_set_counter, dim3(1), dim3(1), 0, st, d.cnt, (uint32_t)CT_ITERS_SUM_CONV, 0u);
    hipLaunchKernelGGL(k_set_counter, dim3(1), dim3(1), 0, st, d.cnt, (uint32_t)CT_WS_LEARNED, 0u);
    hipLaunchKernelGGL(k_set_counter, dim3(1), dim3(1), 0, st, d.cnt, (uint32_t)CT_WIDE_N, 0u);
    hipLaunchKernelGGL(k_set_counter, dim3(1), dim3(1), 0, st, d.cnt, (uint32_t)CT_WIDE_MAX, 0u);
    hipLaunchKernelGGL(k_set_counter, dim3(1), dim3(1), 0, st, d.cnt, (uint32_t)CT_BIG_N, 0u);
    hipLaunchKernelGGL(k_set_counter, dim3(1), dim3(1), 0, st, d.cnt, (uint32_t)CT_WSH_MAX, 0u);
    hipLaunchKernelGGL(k_chain_count, dim3(1), dim3(1), 0, st, d);
    if (tp.ws_mode == 2 || tp.log_on) hipLaunchKernelGGL(k_features, dim3(grid_for(d.cap_contacts, B)), dim3(B), 0, st, d, tp);
    if (tp.ws_mode == 2)
        rp_launch_mlp_infer(d.mlp.p, d.mlp.w1p, d.mlp.w2p, d.mlp.w3p, d.feat, d.cnt + CT_CONTACTS, d.cap_contacts, (float *)d.pred,
                            d.cnt + CT_WS_LEARNED, st);
    mark(RP_ST_ORDERS);
    build_chains(d, d.isl_v2, d.ch_pa, d.ch_pb, d.ch_sa, d.ch_sb, body_bits, st);
    build_shock_order(d, tp, body_bits, st);
    // the lane tier's islands by decreasing size (the radix buffers are free until the
    // projection's shock order)
    const uint32_t *sorted0 = nullptr;
    if (tp.lane_max) {
        hipLaunchKernelGGL(k_lane_keys, dim3(grid_for(d.cap_contacts, B)), dim3(B), 0, st, d, tp);
        uint32_t *lk, *lv;
        radix_sort(d, d.sk0, d.sv0, d.sk1, d.sv1, d.cnt + CT_TIER_N, d.cap_contacts, key_bits(tp.tier_max[0] - 1u), st, &lk, &lv);
        // kept for the projection passes (the radix buffers serve the projection's shock order)
        hipLaunchKernelGGL(k_copy_counted, dim3(grid_for(d.cap_contacts, B, 4096)), dim3(B), 0, st, lv, d.lane_order, d.cnt + CT_TIER_N);
        sorted0 = d.lane_order;
        hipLaunchKernelGGL(k_lane_split, dim3(1), dim3(64), 0, st, d, tp, sorted0);
    }
    const uint32_t g_lane = (d.cap_contacts + 63u) / 64u < 8192u ? (d.cap_contacts + 63u) / 64u : 8192u;
    const uint32_t g0 = grids[0], g1 = grids[1], g2 = grids[2];
    // wide levels: kl level launches per sweep (fixed iteration counts only); 0 keeps the
    // one-kernel sweeps.  Any kl gives the same results: levels beyond it join the tail sweep.
    const uint32_t kl = grids[3], split = (kl > 0 && !tp.resid_on) ? 1u : 0u, GL = 64;
    // shock passes in the split form: klsh wide-level launches of the shock order per pass
    const uint32_t klsh = split ? grids[4] : 0u;
    // tier 0 on side[0] and tier 3 on side[1] alongside `big` on the tick's stream, joined back
    auto fork0 = [&](auto k0, auto k3, auto big) {
        if (!ts) { k0(st); k3(st); big(st); return; }
        (void)hipEventRecord(ts->ev[0], st);
        (void)hipStreamWaitEvent(ts->side[0], ts->ev[0], 0);
        (void)hipStreamWaitEvent(ts->side[1], ts->ev[0], 0);
        k0(ts->side[0]);
        k3(ts->side[1]);
        big(st);
        (void)hipEventRecord(ts->ev[1], ts->side[0]);
        (void)hipEventRecord(ts->ev[2], ts->side[1]);
        (void)hipStreamWaitEvent(st, ts->ev[1], 0);
        (void)hipStreamWaitEvent(st, ts->ev[2], 0);
    };
    auto big_vel = [&](hipStream_t s) {
        hipLaunchKernelGGL(k_sweep_vel<1>, dim3(g1), dim3(Tier<1>::threads), 0, s, d, tp, split, (const uint32_t *)nullptr);
        hipLaunchKernelGGL(k_sweep_vel<2>, dim3(g2), dim3(Tier<2>::threads), 0, s, d, tp, split, (const uint32_t *)nullptr);
        if (!split) return;
        hipLaunchKernelGGL(k_store_rows, dim3(1024), dim3(256), 0, s, d);
        if (tp.ws_mode != 1) {
            for (uint32_t L = 0; L < kl; L++) hipLaunchKernelGGL(k_level_warm, dim3(GL), dim3(256), 0, s, d, tp, L);
            hipLaunchKernelGGL(k_sweep_warm_tail<1>, dim3(g1), dim3(Tier<1>::threads), 0, s, d, tp, kl);
            hipLaunchKernelGGL(k_sweep_warm_tail<2>, dim3(g2), dim3(Tier<2>::threads), 0, s, d, tp, kl);
        }
        for (uint32_t it = 0; it < tp.iv; it++) {
            for (uint32_t L = 0; L < kl; L++) hipLaunchKernelGGL(k_level_vel, dim3(GL), dim3(256), 0, s, d, L);
            hipLaunchKernelGGL(k_sweep_vel_tail<1>, dim3(g1), dim3(Tier<1>::threads), 0, s, d, tp, kl);
            hipLaunchKernelGGL(k_sweep_vel_tail<2>, dim3(g2), dim3(Tier<2>::threads), 0, s, d, tp, kl);
        }
        hipLaunchKernelGGL(k_sweep_vel_post<1>, dim3(g1), dim3(Tier<1>::threads), 0, s, d, tp, klsh);
        hipLaunchKernelGGL(k_sweep_vel_post<2>, dim3(g2), dim3(Tier<2>::threads), 0, s, d, tp, klsh);
        if (!klsh) return;
        for (uint32_t L = 0; L < klsh; L++) hipLaunchKernelGGL(k_level_vshock, dim3(GL), dim3(256), 0, s, d, tp, L);
        hipLaunchKernelGGL(k_sweep_vshock_tail<1>, dim3(g1), dim3(Tier<1>::threads), 0, s, d, tp, klsh);
        hipLaunchKernelGGL(k_sweep_vshock_tail<2>, dim3(g2), dim3(Tier<2>::threads), 0, s, d, tp, klsh);
    };
    auto big_proj = [&](hipStream_t s) {
        hipLaunchKernelGGL(k_sweep_proj<1>, dim3(g1), dim3(Tier<1>::threads), 0, s, d, tp, split, (const uint32_t *)nullptr);
        hipLaunchKernelGGL(k_sweep_proj<2>, dim3(g2), dim3(Tier<2>::threads), 0, s, d, tp, split, (const uint32_t *)nullptr);
        if (!split) return;
        hipLaunchKernelGGL(k_store_prows, dim3(1024), dim3(256), 0, s, d);
        for (uint32_t it = 0; it < tp.ip; it++) {
            for (uint32_t L = 0; L < kl; L++) hipLaunchKernelGGL(k_level_proj, dim3(GL), dim3(256), 0, s, d, tp, L);
            hipLaunchKernelGGL(k_sweep_proj_tail<1>, dim3(g1), dim3(Tier<1>::threads), 0, s, d, tp, kl);
            hipLaunchKernelGGL(k_sweep_proj_tail<2>, dim3(g2), dim3(Tier<2>::threads), 0, s, d, tp, kl);
        }
    };
    // the small islands: tier 0 and tier 3 (mid-sized), one wave each
    const uint32_t g3 = g0 > 1024u ? g0 : 1024u;
    // tier 0: with the lane tier on, the lanes' and lane groups' islands on side[0] and the largest
    // islands (a wave each) ahead of tier 1 on the tick's stream
    auto vel0 = [&](hipStream_t s) {
        if (tp.lane_max) hipLaunchKernelGGL(k_lane_vel, dim3(g_lane), dim3(64), 0, s, d, tp, sorted0);
        else hipLaunchKernelGGL(k_sweep_vel<0>, dim3(g0), dim3(Tier<0>::threads), 0, s, d, tp, 0u, (const uint32_t *)nullptr);
    };
    auto vel0w = [&](hipStream_t s) {
        if (!tp.lane_max) return;
        if (tp.quad_max > tp.lane_max) hipLaunchKernelGGL(k_group_vel, dim3(2u * g_lane), dim3(64), 0, s, d, tp, sorted0, g_lane);
        hipLaunchKernelGGL(k_sweep_vel<0>, dim3(g0), dim3(Tier<0>::threads), 0, s, d, tp, 0u, sorted0);
    };
    auto vel3 = [&](hipStream_t s) { hipLaunchKernelGGL(k_sweep_vel<3>, dim3(g3), dim3(Tier<3>::threads), 0, s, d, tp, 0u, (const uint32_t *)nullptr); };
    // profiled with side streams: tiers 0 and 3 run alongside on side[0] and side[1], timed as one
    // stage by events ev[RP_NMARKS + 2 * which] and ev[RP_NMARKS + 2 * which + 1] on side[0]; the
    // larger tiers take the stages of the tick's stream
    auto small_timed = [&](auto k0, auto k3, int which) {
        (void)hipEventRecord(ts->ev[0], st);
        (void)hipStreamWaitEvent(ts->side[0], ts->ev[0], 0);
        (void)hipStreamWaitEvent(ts->side[1], ts->ev[0], 0);
        (void)hipEventRecord(ev[RP_NMARKS + 2 * which], ts->side[0]);
        k0(ts->side[0]);
        k3(ts->side[1]);
        (void)hipEventRecord(ts->ev[2], ts->side[1]);
        (void)hipStreamWaitEvent(ts->side[0], ts->ev[2], 0);
        (void)hipEventRecord(ev[RP_NMARKS + 2 * which + 1], ts->side[0]);
        (void)hipEventRecord(ts->ev[1], ts->side[0]);
    };
    if (ev) {  // profiled: a stage per tier (the split big-island solve counts as large)
        mark(RP_ST_VEL_SMALL);
        if (ts) small_timed(vel0, [&](hipStream_t s) { vel0w(s); vel3(s); }, 0);
        else { vel0(st); vel0w(st); vel3(st); }
        mark(RP_ST_VEL_LARGE);
        if (split) big_vel(st);
        else hipLaunchKernelGGL(k_sweep_vel<1>, dim3(g1), dim3(Tier<1>::threads), 0, st, d, tp, 0u, (const uint32_t *)nullptr);
        mark(RP_ST_VEL_HUGE);
        if (!split) hipLaunchKernelGGL(k_sweep_vel<2>, dim3(g2), dim3(Tier<2>::threads), 0, st, d, tp, 0u, (const uint32_t *)nullptr);
        if (ts) (void)hipStreamWaitEvent(st, ts->ev[1], 0);
    } else if (split) {
        fork0(vel0, vel3, [&](hipStream_t s) { vel0w(s); big_vel(s); });
    } else {
        tiers(vel0, vel3, [&](hipStream_t s) {
                  vel0w(s);
                  hipLaunchKernelGGL(k_sweep_vel<1>, dim3(g1), dim3(Tier<1>::threads), 0, s, d, tp, 0u, (const uint32_t *)nullptr);
              },
              [&](hipStream_t s) { hipLaunchKernelGGL(k_sweep_vel<2>, dim3(g2), dim3(Tier<2>::threads), 0, s, d, tp, 0u, (const uint32_t *)nullptr); });
    }
    const uint32_t g4 = 512;
    if (d.cm_on) {  // islands with compliant rows or patches, then the patches in no island
        hipLaunchKernelGGL(k_sweep_cm, dim3(g4), dim3(Tier<4>::threads), 0, st, d, tp);
        hipLaunchKernelGGL(k_free_patches, dim3(grid_for(d.cap_patch, B)), dim3(B), 0, st, d, tp);
    }
    if (tp.log_on) {
        hipLaunchKernelGGL(k_log_flags, dim3(grid_for(d.cap_contacts, B)), dim3(B), 0, st, d, tp);
        scan_u32(d.newflag, d.cap_contacts, d.scan_tmp, d.cnt + CT_LOG_N, st);
        hipLaunchKernelGGL(k_log_contacts, dim3(grid_for(d.cap_contacts, B)), dim3(B), 0, st, d, tp);
        hipLaunchKernelGGL(k_log_advance, dim3(1), dim3(1), 0, st, d);
    }
    mark(RP_ST_WARM_STORE);
    // warm cache for the next tick (into buffer cur ^ 1)
    hipLaunchKernelGGL(k_warm_newflags, dim3(grid_for(d.cap_contacts, B)), dim3(B), 0, st, d);
    scan_u32(d.newflag, d.cap_contacts, d.scan_tmp, d.cnt + CT_NEW_TOTAL, st);
    hipLaunchKernelGGL(k_store_total, dim3(1), dim3(1), 0, st, d.newflag, d.cap_contacts, d.cnt + CT_NEW_TOTAL);
    hipLaunchKernelGGL(k_warm_keepflags, dim3(grid_for(d.cap_warm, B)), dim3(B), 0, st, d, tp);
    scan_u32(d.keepflag, d.cap_warm, d.scan_tmp, d.cnt + CT_KEEP_TOTAL, st);
    hipLaunchKernelGGL(k_store_total, dim3(1), dim3(1), 0, st, d.keepflag, d.cap_warm, d.cnt + CT_KEEP_TOTAL);
    if (d.cm_on) {  // the carried patch impulses (into buffer cur ^ 1)
        hipLaunchKernelGGL(k_wp_newflags, dim3(grid_for(d.cap_patch, B)), dim3(B), 0, st, d);
        scan_u32(d.wp_new, d.cap_patch, d.scan_tmp, d.cnt + CT_WP_NEW, st);
        hipLaunchKernelGGL(k_store_total, dim3(1), dim3(1), 0, st, d.wp_new, d.cap_patch, d.cnt + CT_WP_NEW);
        hipLaunchKernelGGL(k_wp_keepflags, dim3(grid_for(d.cap_warm, B)), dim3(B), 0, st, d, tp);
        scan_u32(d.wp_keep, d.cap_warm, d.scan_tmp, d.cnt + CT_WP_KEEP, st);
        hipLaunchKernelGGL(k_store_total, dim3(1), dim3(1), 0, st, d.wp_keep, d.cap_warm, d.cnt + CT_WP_KEEP);
    }
    hipLaunchKernelGGL(k_check, dim3(1), dim3(64), 0, st, d, 2u);
    hipLaunchKernelGGL(k_warm_merge, dim3(grid_for(d.cap_contacts > d.cap_warm ? d.cap_contacts : d.cap_warm, B)), dim3(B), 0, st, d, tp);
    if (d.cm_on) hipLaunchKernelGGL(k_wp_merge, dim3(grid_for(d.cap_warm > d.cap_patch ? d.cap_warm : d.cap_patch, B)), dim3(B), 0, st, d, tp);
    if (d.w6_on) hipLaunchKernelGGL(k_w6, dim3(d.w6_slots), dim3(64), 0, st, d, tp);
    hipLaunchKernelGGL(k_loaded, dim3(grid_for(d.cap_contacts, B)), dim3(B), 0, st, d);
    if (d.cm_on) hipLaunchKernelGGL(k_patch_loaded, dim3(grid_for(d.cap_patch, B)), dim3(B), 0, st, d);
    mark(RP_ST_BALLISTIC);
    hipLaunchKernelGGL(k_ball_init, dim3(grid_for(d.nb, B)), dim3(B), 0, st, d);
    // tier 0's share of each pass below: with few projection iterations (tp.lane_pos), the lane
    // tier's islands one per thread (k_lane_*) and the rest a wave each ahead of tier 1
    const uint32_t *const NO = nullptr, *const lp = tp.lane_pos ? sorted0 : nullptr;
    auto ball0 = [&](hipStream_t s) {
        if (lp) hipLaunchKernelGGL(k_lane_ball, dim3(g_lane), dim3(64), 0, s, d, tp, lp);
        else hipLaunchKernelGGL(k_sweep_ball<0>, dim3(g0), dim3(Tier<0>::threads), 0, s, d, tp, NO);
    };
    tiers(ball0, [&](hipStream_t s) { hipLaunchKernelGGL(k_sweep_ball<3>, dim3(g3), dim3(Tier<3>::threads), 0, s, d, tp, NO); },
          [&](hipStream_t s) {
              if (lp) hipLaunchKernelGGL(k_sweep_ball<0>, dim3(g0), dim3(Tier<0>::threads), 0, s, d, tp, lp);
              hipLaunchKernelGGL(k_sweep_ball<1>, dim3(g1), dim3(Tier<1>::threads), 0, s, d, tp, NO);
          },
          [&](hipStream_t s) { hipLaunchKernelGGL(k_sweep_ball<2>, dim3(g2), dim3(Tier<2>::threads), 0, s, d, tp, NO); });
    if (d.cm_on) hipLaunchKernelGGL(k_sweep_ball<4>, dim3(g4), dim3(Tier<4>::threads), 0, st, d, tp, NO);
    mark(RP_ST_INTEGRATE);
    hipLaunchKernelGGL(k_integrate, dim3(grid_for(d.nb, B)), dim3(B), 0, st, d, tp);
    mark(RP_ST_POS_PREP);
    hipLaunchKernelGGL(k_posc_prep, dim3(grid_for(d.cap_contacts, B)), dim3(B), 0, st, d);
    if (d.d6_on) hipLaunchKernelGGL(k_d6_begin, dim3(grid_for(6 * d.nj, B)), dim3(B), 0, st, d);
    auto proj0 = [&](hipStream_t s) {
        if (lp) hipLaunchKernelGGL(k_lane_proj, dim3(g_lane), dim3(64), 0, s, d, tp, lp);
        else hipLaunchKernelGGL(k_sweep_proj<0>, dim3(g0), dim3(Tier<0>::threads), 0, s, d, tp, 0u, NO);
    };
    auto proj0w = [&](hipStream_t s) { if (lp) hipLaunchKernelGGL(k_sweep_proj<0>, dim3(g0), dim3(Tier<0>::threads), 0, s, d, tp, 0u, lp); };
    auto proj3 = [&](hipStream_t s) { hipLaunchKernelGGL(k_sweep_proj<3>, dim3(g3), dim3(Tier<3>::threads), 0, s, d, tp, 0u, NO); };
    if (ev) {
        mark(RP_ST_PROJ_SMALL);
        if (ts) small_timed(proj0, [&](hipStream_t s) { proj0w(s); proj3(s); }, 1);
        else { proj0(st); proj0w(st); proj3(st); }
        mark(RP_ST_PROJ_LARGE);
        if (split) big_proj(st);
        else hipLaunchKernelGGL(k_sweep_proj<1>, dim3(g1), dim3(Tier<1>::threads), 0, st, d, tp, 0u, NO);
        mark(RP_ST_PROJ_HUGE);
        if (!split) hipLaunchKernelGGL(k_sweep_proj<2>, dim3(g2), dim3(Tier<2>::threads), 0, st, d, tp, 0u, NO);
        if (ts) (void)hipStreamWaitEvent(st, ts->ev[1], 0);
    } else if (split) {
        fork0(proj0, proj3, [&](hipStream_t s) { proj0w(s); big_proj(s); });
    } else {
        tiers(proj0, proj3, [&](hipStream_t s) {
                  proj0w(s);
                  hipLaunchKernelGGL(k_sweep_proj<1>, dim3(g1), dim3(Tier<1>::threads), 0, s, d, tp, 0u, NO);
              },
              [&](hipStream_t s) { hipLaunchKernelGGL(k_sweep_proj<2>, dim3(g2), dim3(Tier<2>::threads), 0, s, d, tp, 0u, NO); });
    }
    if (d.cm_on) hipLaunchKernelGGL(k_sweep_proj<4>, dim3(g4), dim3(Tier<4>::threads), 0, st, d, tp, 0u, NO);
    mark(RP_ST_PROJ_ORDER);
    build_shock_order(d, tp, body_bits, st);  // from the projected poses
    mark(RP_ST_PROJ_SHOCK);
    auto pshock0 = [&](hipStream_t s) {
        if (lp) hipLaunchKernelGGL(k_lane_pshock, dim3(g_lane), dim3(64), 0, s, d, tp, lp);
        else hipLaunchKernelGGL(k_sweep_proj_shock<0>, dim3(g0), dim3(Tier<0>::threads), 0, s, d, tp, 0u, NO);
    };
    auto pshock0w = [&](hipStream_t s) { if (lp) hipLaunchKernelGGL(k_sweep_proj_shock<0>, dim3(g0), dim3(Tier<0>::threads), 0, s, d, tp, 0u, lp); };
    auto pshock3 = [&](hipStream_t s) { hipLaunchKernelGGL(k_sweep_proj_shock<3>, dim3(g3), dim3(Tier<3>::threads), 0, s, d, tp, 0u, NO); };
    auto big_pshock = [&](hipStream_t s) {
        hipLaunchKernelGGL(k_sweep_proj_shock<1>, dim3(g1), dim3(Tier<1>::threads), 0, s, d, tp, klsh, NO);
        hipLaunchKernelGGL(k_sweep_proj_shock<2>, dim3(g2), dim3(Tier<2>::threads), 0, s, d, tp, klsh, NO);
        if (!klsh) return;
        for (uint32_t L = 0; L < klsh; L++) hipLaunchKernelGGL(k_level_pshock, dim3(GL), dim3(256), 0, s, d, tp, L);
        hipLaunchKernelGGL(k_sweep_pshock_tail<1>, dim3(g1), dim3(Tier<1>::threads), 0, s, d, tp, klsh);
        hipLaunchKernelGGL(k_sweep_pshock_tail<2>, dim3(g2), dim3(Tier<2>::threads), 0, s, d, tp, klsh);
    };
    if (klsh) fork0(pshock0, pshock3, [&](hipStream_t s) { pshock0w(s); big_pshock(s); });
    else tiers(pshock0, pshock3, [&](hipStream_t s) {
                   pshock0w(s);
                   hipLaunchKernelGGL(k_sweep_proj_shock<1>, dim3(g1), dim3(Tier<1>::threads), 0, s, d, tp, 0u, NO);
               },
               [&](hipStream_t s) { hipLaunchKernelGGL(k_sweep_proj_shock<2>, dim3(g2), dim3(Tier<2>::threads), 0, s, d, tp, 0u, NO); });
    if (d.cm_on) hipLaunchKernelGGL(k_sweep_proj_shock<4>, dim3(g4), dim3(Tier<4>::threads), 0, st, d, tp, 0u, NO);
    mark(RP_ST_S6);
    hipLaunchKernelGGL(k_s6, dim3(grid_for(d.nb, B)), dim3(B), 0, st, d, tp);
    hipLaunchKernelGGL(k_tick_done, dim3(1), dim3(64), 0, st, d);
    mark(RP_ST_END);
    (void)scratch_total;
    return hipGetLastError();
}

// Plain copy of n float4 (bandwidth reference for the roofline and for calibrating HBM counters).
__global__ void __launch_bounds__(256) k_copy_probe(const float4 *__restrict__ src, float4 *__restrict__ dst, size_t n) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) dst[i] = src[i];
}
extern "C" void rp_launch_copy_probe(const float4 *src, float4 *dst, size_t n, hipStream_t st) {
    hipLaunchKernelGGL(k_copy_probe, dim3(4096), dim3(256), 0, st, src, dst, n);
}

extern "C" size_t rp_w6_ws_bytes() { return w6_ws_bytes(); }
extern "C" hipError_t rp_launch_query(const Dev &d, const RpQuery *qs, RpQueryHit *out, uint32_t n, hipStream_t st) {
    hipLaunchKernelGGL(k_query, dim3(n), dim3(RPQ_T), 0, st, d, qs, out);
    return hipGetLastError();
}
extern "C" void rp_launch_restore_backup(const Dev &d, hipStream_t st) {
    hipLaunchKernelGGL(k_restore_backup, dim3(grid_for(d.nb > d.nj ? d.nb : d.nj, 256)), dim3(256), 0, st, d);
}
