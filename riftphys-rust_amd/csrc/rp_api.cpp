// C ABI of the step library: the rphys_* entry points of crates/riftphys-ffi/src/riftphys_c.h
// over device-resident worlds.  Host code keeps a mirror of each world (bodies, heightfield,
// settings, carried state); the device holds the copy the step runs on.  Host reads and writes
// pull the device copy first; a step pushes host edits first.  Kernels live in rp_step.hip and
// rp_hash.hip.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <atomic>
#include <cmath>
#include <chrono>
#include <cstdio>
#include <cstring>
#include <memory>
#include <mutex>
#include <new>
#include <string>
#include <vector>

#include "../../include/riftphys_c.h"
#include "rp_dev.h"

struct HashDev {
    uint32_t *keyflag, *w_chunk, *w_len;
    uint64_t *w_epoch, *w_rng;
    uint8_t *stream;
    uint32_t *cv, *cv2;
    uint8_t *digest;
    uint32_t cap_chunks;
    uint32_t *w_jbytes;
    uint32_t *j_off;
};
extern "C" hipError_t rp_launch_tick(const Dev &d, const TickParams &tp, const uint32_t *grids, hipStream_t st, hipEvent_t *ev,
                                     const TierStreams *ts);
extern "C" void rp_launch_restore_backup(const Dev &d, hipStream_t st);
extern "C" size_t rp_w6_ws_bytes();
extern "C" hipError_t rp_launch_query(const Dev &d, const RpQuery *qs, RpQueryHit *out, uint32_t n, hipStream_t st);
extern "C" void rp_launch_copy_probe(const float4 *src, float4 *dst, size_t n, hipStream_t st);
extern "C" void rp_launch_mlp_prep(const float *p, __bf16 *w1p, __bf16 *w2p, __bf16 *w2t, __bf16 *w3p, __bf16 *w3t, hipStream_t st);
extern "C" void rp_launch_mlp_infer(const float *p, const __bf16 *w1p, const __bf16 *w2p, const __bf16 *w3p, const float *feat, const uint32_t *n_ptr,
                                    uint32_t cap, float *out, uint32_t *n_trusted, hipStream_t st);
extern "C" void rp_launch_mlp_train(const TrainArgs &a, hipStream_t st);
extern "C" void rp_launch_hash(const Dev &d, const HashDev &h, uint32_t cur, uint32_t max_extra_slots, hipStream_t st);

// ------------------------------------------------------------------ errors (lib.rs:56-82)
static thread_local std::string g_last_error;
static void set_err(const std::string &s) { g_last_error = s; }
extern "C" void rp_set_last_error(const char *msg) { g_last_error = msg ? msg : ""; }
static void clear_err() { g_last_error.clear(); }

struct F4 { float x, y, z, w; };
static inline F4 f4(float x, float y, float z, float w = 0.0f) { F4 r = {x, y, z, w}; return r; }

// ------------------------------------------------------------------ host-side arithmetic used at setup
// (the same operation order as the device; compiled with -ffp-contract=off)
struct HV { float x, y, z; };
static inline HV hv(float x, float y, float z) { HV r = {x, y, z}; return r; }
static inline HV hcross(HV a, HV b) { return hv(a.y * b.z - b.y * a.z, a.z * b.x - b.z * a.x, a.x * b.y - b.x * a.y); }
static inline float hdot(HV a, HV b) { return (a.x * b.x + a.y * b.y) + a.z * b.z; }
struct HM { HV c0, c1, c2; };
static HM hinv(HM m) {  // Mat3A::inverse
    HV t0 = hcross(m.c1, m.c2), t1 = hcross(m.c2, m.c0), t2 = hcross(m.c0, m.c1);
    float id = 1.0f / hdot(m.c2, t2);
    HV a = hv(t0.x * id, t0.y * id, t0.z * id), b = hv(t1.x * id, t1.y * id, t1.z * id), c = hv(t2.x * id, t2.y * id, t2.z * id);
    HM r; r.c0 = hv(a.x, b.x, c.x); r.c1 = hv(a.y, b.y, c.y); r.c2 = hv(a.z, b.z, c.z);
    return r;
}
static inline bool hfin(float x) { return std::isfinite(x); }
static inline uint32_t f2u(float x) { uint32_t u; std::memcpy(&u, &x, 4); return u; }
static inline float u2f(uint32_t u) { float x; std::memcpy(&x, &u, 4); return x; }
// D6 axis bits (rp_d6.h D6F_*)
enum { D6F_ENABLED_BIT = 1u, D6F_DRIVE_BIT = 2u, D6F_VEL_BIT = 4u, D6F_LIMIT_BIT = 8u };
// SSE2 Vec4 normalize (division by sqrt of the pairwise dot4)
static F4 qnorm(F4 q) {
    float l = std::sqrt((q.x * q.x + q.z * q.z) + (q.y * q.y + q.w * q.w));
    return f4(q.x / l, q.y / l, q.z / l, q.w / l);
}
static F4 fin_q(F4 q) {  // dynamics.rs:16-22
    if (!(hfin(q.x) && hfin(q.y) && hfin(q.z) && hfin(q.w))) return f4(0, 0, 0, 1);
    F4 n = qnorm(q);
    if (!(hfin(n.x) && hfin(n.y) && hfin(n.z) && hfin(n.w))) return f4(0, 0, 0, 1);
    if (n.w < 0.0f) n = f4(-n.x, -n.y, -n.z, -n.w);
    return n;
}
static float q6div(float x) { return std::isfinite(x) ? std::round(x / 1.0e-6f) * 1.0e-6f : 0.0f; }
static float q6mul(float x) { return std::round(x * 1.0e6f) * 1.0e-6f; }

// materials.rs:93-304: the core table entries the FFI can reach (Default 0, Grit 4 and the
// first sixteen ids), geometric-mean friction, max restitution, quantised, and the curated
// overrides.
struct MatP { float mu_s, mu_k, rest, vs; };
static const MatP MATS[16] = {
    {0.60f, 0.50f, 0.10f, 0.08f}, {0.05f, 0.03f, 0.02f, 0.02f}, {1.10f, 0.95f, 0.80f, 0.15f},
    {0.70f, 0.55f, 0.50f, 0.05f}, {1.30f, 1.05f, 0.00f, 0.12f}, {0.55f, 0.45f, 0.08f, 0.10f},
    {0.80f, 0.70f, 0.00f, 0.10f}, {0.90f, 0.80f, 0.00f, 0.12f}, {0.80f, 0.60f, 0.20f, 0.10f},
    {0.55f, 0.45f, 0.10f, 0.06f}, {1.10f, 0.95f, 0.00f, 0.18f}, {0.40f, 0.30f, 0.65f, 0.04f},
    {0.90f, 0.60f, 0.35f, 0.06f}, {0.61f, 0.47f, 0.40f, 0.05f}, {0.50f, 0.40f, 0.12f, 0.08f},
    {0.20f, 0.15f, 0.85f, 0.04f}};
static MatP table_pair(uint32_t a, uint32_t b) {
    MatP pa = MATS[a & 15], pb = MATS[b & 15], r;
    r.mu_s = q6mul(std::sqrt(std::fabs(pa.mu_s * pb.mu_s)));
    r.mu_k = q6mul(std::sqrt(std::fabs(pa.mu_k * pb.mu_k)));
    r.rest = q6mul(std::fmax(pa.rest, pb.rest));
    r.vs = q6mul(0.5f * (pa.vs + pb.vs));
    uint32_t x = std::min(a, b), y = std::max(a, b);
    MatP o = r; bool ov = false;
    if (x == 1 && y == 2) { ov = true; o = {0.20f, 0.10f, 0.05f, 0.08f}; }
    if (x == 2 && y == 7) { ov = true; o = {1.10f, 0.95f, 0.00f, 0.18f}; }
    if (x == 2 && y == 6) { ov = true; o = {1.00f, 0.85f, 0.00f, 0.16f}; }
    if (x == 3 && y == 3) { ov = true; o = {0.58f, 0.50f, 0.08f, 0.04f}; }
    if (ov) { o.mu_s = q6mul(o.mu_s); o.mu_k = q6mul(o.mu_k); o.rest = q6mul(o.rest); o.vs = q6mul(o.vs); return o; }
    return r;
}

// ------------------------------------------------------------------ host world mirror
// a pair's properties: friction and restitution, and its contact model (materials.rs:75-90:
// stiffness > 0 compliant with that dissipation; patch friction)
struct PairOverride { uint32_t a, b; MatP p; float stiff = 0.0f, diss = 0.0f; uint32_t patch = 0; };
struct HostWorld {
    uint32_t max_bodies = 0, max_colliders = 0, bodies_created = 0;
    // bodies
    std::vector<F4> pos, rot, lin, ang, last_lin, dims;
    std::vector<float> inv_mass, iil;
    std::vector<uint32_t> bflags, awake, sticks, btype;
    std::vector<uint32_t> col_body, col_mat;
    // sleep slots beyond the bodies (capacity): awake 1, ticks 0 unless changed
    // heightfield (terrain.rs:24-52 applied at set time)
    bool has_hf = false;
    uint32_t hnx = 0, hnz = 0, hmat = 4;
    float hsx = 1, hsz = 1, hox = 0, hoz = 0, hmin = 0, hmax = 0;
    std::vector<float> heights;
    // settings
    float gx = 0.0f, gy = -9.81f, gz = 0.0f;
    uint32_t iv = 20, ip = 20;
    uint32_t resid_on = 0, resid_max = 100;  // residual mode (build addition)
    float resid_tol = 1.0e-4f;
    uint32_t ws_mode = 0;                    // warm start: 0 previous frame, 1 zero, 2 learned
    std::vector<float> mlp;                  // the learned model (RP_MLP_NPARAM floats) for mode 2
    std::vector<PairOverride> ovr;
    uint32_t damp_start = 0;  // compliant rows' damping at the step-start penetration (world.rs:992)
    uint64_t epoch = 0, rng = 0xC0FFEEull | 1ull, tick = 0;
    // carried warm cache
    std::vector<uint64_t> wkey;
    std::vector<F4> wla, wjt;
    std::vector<float> wjn;
    // carried friction-patch impulses (world.rs:249 warm_patches): keys, (jt, tw)
    std::vector<uint64_t> pkey;
    std::vector<F4> pjt;
    // ball joints (world.rs:718 add_ball_joint): body indices, frames (position, rotation) in each
    // body's space, and the carried warm rows (mask: bit 31 stored, bits 0-2 the rows' axes)
    std::vector<uint32_t> jt_a, jt_b;
    std::vector<F4> jt_fap, jt_far, jt_fbp, jt_fbr;
    std::vector<uint32_t> wj_mask;
    std::vector<float> wj_lam;  // 3 per joint
    // D6 joints (bodies.rs:107-136): kind (0 ball, RP_JK_D6), stored hinge axis (RP_NONE: none),
    // 18 axis words (rp_d6.h), and the carried rows (count word, 32 tags and impulses per joint)
    std::vector<uint32_t> jt_kind, jt_hinge, d6_wn, d6_wtag;
    std::vector<F4> jt_ax;
    std::vector<float> d6_wlam;
    // kinematic players (world.rs:44-69): body index, radius, height, speed, vertical velocity,
    // grounded, input direction
    struct Player { uint32_t body; float radius, height, speed, vvel; uint32_t grounded; float input[3]; };
    std::vector<Player> players;
    uint32_t nb() const { return (uint32_t)pos.size(); }
    uint32_t nc() const { return (uint32_t)col_body.size(); }
    uint32_t nj() const { return (uint32_t)jt_a.size(); }
};

// The override of pair (a, b), made from the pair's table properties (rigid) when absent.
static PairOverride &pair_override(HostWorld &w, uint32_t a, uint32_t b) {
    const uint32_t x = std::min(a, b), y = std::max(a, b);
    for (auto &o : w.ovr) if (o.a == x && o.b == y) return o;
    w.ovr.push_back(PairOverride{x, y, table_pair(x, y)});
    return w.ovr.back();
}

static void drop_warm_of(HostWorld &w, uint32_t b) {  // world.rs:818-820
    size_t k = 0;
    for (size_t i = 0; i < w.wkey.size(); i++) {
        uint32_t a = (uint32_t)(w.wkey[i] >> 32), c = (uint32_t)w.wkey[i];
        if (a == b || c == b) continue;
        w.wkey[k] = w.wkey[i]; w.wla[k] = w.wla[i]; w.wjn[k] = w.wjn[i]; w.wjt[k] = w.wjt[i]; k++;
    }
    w.wkey.resize(k); w.wla.resize(k); w.wjn.resize(k); w.wjt.resize(k);
    k = 0;
    for (size_t i = 0; i < w.pkey.size(); i++) {  // its patches' (world.rs:820)
        uint32_t a = (uint32_t)(w.pkey[i] >> 32), c = (uint32_t)w.pkey[i];
        if (a == b || c == b) continue;
        w.pkey[k] = w.pkey[i]; w.pjt[k] = w.pjt[i]; k++;
    }
    w.pkey.resize(k); w.pjt.resize(k);
    for (uint32_t q = 0; q < w.nj(); q++)  // and its joints' warm rows (world.rs:822-826)
        if (w.jt_a[q] == b || w.jt_b[q] == b) { w.wj_mask[q] = 0; w.d6_wn[q] = 0; }
}

// ------------------------------------------------------------------ device scene (batch of worlds)
// Device allocations still to fail (rphys_debug_fail_device_allocs): forces the device-failure
// path from the host.
static std::atomic<uint32_t> g_fail_allocs{0};
static bool take_forced_failure() {
    uint32_t v = g_fail_allocs.load();
    while (v && !g_fail_allocs.compare_exchange_weak(v, v - 1)) {}
    return v != 0;
}
struct DeviceBuffers {
    std::vector<void *> ptrs;
    ~DeviceBuffers() { for (void *p : ptrs) (void)hipFree(p); }
    template <class T> bool alloc(T *&p, size_t n) {
        void *q = nullptr;
        if (take_forced_failure() || hipMalloc(&q, std::max<size_t>(n, 1) * sizeof(T)) != hipSuccess) { p = nullptr; return false; }
        ptrs.push_back(q);
        p = (T *)q;
        return true;
    }
};

struct Slot {  // device-side snapshot of a scene
    bool valid = false;
    DeviceBuffers buf;
    float4 *pos, *rot, *lin, *ang, *last;
    uint32_t *awake, *sticks;
    uint64_t *wkey; float4 *wla; float *wjn; float4 *wjt; uint32_t *woff;
    uint32_t *wj_mask; float *wj_lam;
    float4 *pl_state, *pl_input; uint32_t *pl_grounded;
    uint64_t *pkey; float4 *pjt; uint32_t *poff;  // carried patch impulses (contact-model scenes)
    uint32_t *d6_wn, *d6_wtag; float *d6_wlam;    // carried D6 rows (scenes with D6 joints)
    uint32_t cur; uint32_t nwarm, npatch;
    size_t cap_warm = 0;  // warm-cache entries the slot's buffers hold
    std::vector<uint64_t> ticks;
};

// the island sweeps' launch shape before any tuning: workgroups of tier 0 (tier 3 takes a quarter),
// tier 1 and tier 2
static constexpr uint32_t kDefaultGrids[3] = {8192, 256, 64};

struct Scene {
    Dev d{};
    HashDev h{};
    DeviceBuffers buf;
    hipStream_t st = nullptr;
    uint32_t cur = 0;
    // island-sweep workgroups per tier (rphys_batch_set_tuning), then the wide-level launches per
    // sweep (set from the previous tick's widest island; 0: one-kernel sweeps)
    uint32_t grids[5] = {kDefaultGrids[0], kDefaultGrids[1], kDefaultGrids[2], 0, 0};  // ... and the same for the shock order
    uint32_t graph_grids[5] = {0, 0, 0, 0, 0};  // the shape the graphs were captured with
    uint32_t max_extra = 0;
    std::vector<uint32_t> body_off, col_off;
    std::vector<uint32_t> col_body;  // scene body of each collider, as built (contacts name colliders)
    hipGraphExec_t graph[2] = {nullptr, nullptr};
    TickParams graph_tp{};  // parameters the graphs were captured with (cur aside)
    std::vector<std::unique_ptr<Slot>> slots;
    TierStreams ts{{nullptr, nullptr}, {nullptr, nullptr, nullptr}};  // island-sweep tiers run concurrently
    RpQuery *q_buf = nullptr;  // spatial queries in flight and their answers (q_cap of each)
    RpQueryHit *h_buf = nullptr;
    uint32_t q_cap = 0;
    ~Scene() {
        for (int i = 0; i < 2; i++) if (graph[i]) (void)hipGraphExecDestroy(graph[i]);
        for (int i = 0; i < 3; i++) if (ts.ev[i]) (void)hipEventDestroy(ts.ev[i]);
        for (int i = 0; i < 2; i++) if (ts.side[i]) (void)hipStreamDestroy(ts.side[i]);
        if (st) (void)hipStreamDestroy(st);
    }
};

static bool hip_ok(hipError_t e, const char *what) {
    if (e == hipSuccess) return true;
    set_err(std::string(what) + ": " + hipGetErrorString(e));
    return false;
}

struct Caps { uint32_t pairs, cands, contacts, warm; };

static uint64_t next_pow2(uint64_t v) { uint64_t m = 1; while (m < v) m <<= 1; return m; }

// Copies a learned warm-start model into the scene and prepares its bf16 tiles (stream-ordered).
static bool upload_model(Scene &S, const float *params) {
    const MlpWeights &m = S.d.mlp;
    if (!hip_ok(hipMemcpyAsync((void *)m.p, params, RP_MLP_NPARAM * sizeof(float), hipMemcpyHostToDevice, S.st), "model upload")) return false;
    rp_launch_mlp_prep(m.p, m.w1p, m.w2p, m.w2t, m.w3p, m.w3t, S.st);
    return hip_ok(hipGetLastError(), "model prep");
}

// Whether a contact of these worlds can carry a restitution bias: some pair of the materials in
// use (colliders with colliders, colliders with terrain) has restitution > 0 in the scene's pair
// table (the first world's overrides, as the uploaded table).  Without one the W6 passes never
// run and their buffers are not allocated.
static bool scene_bounces(const std::vector<HostWorld *> &ws) {
    bool used[16] = {}, env[16] = {};
    for (auto *w : ws) {
        for (uint32_t m : w->col_mat) used[m & 15u] = true;
        if (w->has_hf) env[w->hmat & 15u] = true;
    }
    for (uint32_t a = 0; a < 16; a++) {
        if (!used[a]) continue;
        for (uint32_t b = 0; b < 16; b++) {
            if (!used[b] && !env[b]) continue;
            MatP p = table_pair(a, b);
            uint32_t x = std::min(a, b), y = std::max(a, b);
            for (auto &o : ws[0]->ovr) if (o.a == x && o.b == y) p = o.p;
            if (p.rest > 0.0f) return true;
        }
    }
    return false;
}

// Whether the scene needs the contact models: a pair override of the first world (the uploaded
// table) that is compliant or patch friction, or carried patch impulses in any world.
static bool scene_cm(const std::vector<HostWorld *> &ws) {
    for (auto &o : ws[0]->ovr) if (o.stiff > 0.0f || o.patch) return true;
    for (auto *w : ws) if (!w->pkey.empty()) return true;
    return false;
}

// Builds a device scene for the given worlds (host mirrors) with the given capacities and
// uploads everything.  Returns nullptr (with the error set) on failure.
static Scene *build_scene(const std::vector<HostWorld *> &ws, Caps caps, uint32_t cur) {
    std::unique_ptr<Scene> S(new (std::nothrow) Scene());
    if (!S) { set_err("out of host memory"); return nullptr; }
    Dev &d = S->d;
    uint32_t nw = (uint32_t)ws.size(), nb = 0, nc = 0, nh = 0, nwarm = 0, nj = 0;
    S->body_off.push_back(0); S->col_off.push_back(0);
    for (auto *w : ws) nj += w->nj();
    for (auto *w : ws) {
        if (w->nc() > RP_MAX_WORLD_COLLIDERS) d.bp_radix = 1;  // the broadphase orders this scene by radix sorts
        nb += w->nb(); nc += w->nc(); nh += w->has_hf ? (uint32_t)w->heights.size() : 0u; nwarm += (uint32_t)w->wkey.size();
        S->body_off.push_back(nb); S->col_off.push_back(nc);
        uint32_t cap = std::max(w->max_bodies, w->nb());
        S->max_extra = std::max(S->max_extra, cap - w->nb());
    }
    uint32_t npk = 0;
    for (auto *w : ws) npk += (uint32_t)w->pkey.size();
    caps.warm = std::max(caps.warm, std::max(nwarm, npk) + 1024u);
    bool any_d6 = false;
    for (auto *w : ws) for (uint32_t k : w->jt_kind) any_d6 |= k != 0;  // D6 or distance joints
    d.d6_on = any_d6 ? 1u : 0u;
    d.cm_on = (scene_cm(ws) || any_d6) ? 1u : 0u;  // D6 joints are solved by the contact-model tier
    // the broadphase runs 32 lanes per collider over a 32-bit index space
    if (nc > (1u << 27)) { set_err("scene exceeds 134,217,728 colliders"); return nullptr; }
    d.nb = nb; d.nc = nc; d.nw = nw; d.nj = nj;
    d.cap_pairs = caps.pairs; d.cap_cands = caps.cands; d.cap_contacts = caps.contacts; d.cap_warm = caps.warm;
    DeviceBuffers &B = S->buf;
    bool ok = true;
    size_t NB = nb + 1, NC = nc + 1;
    ok &= B.alloc(d.pos, NB) && B.alloc(d.rot, NB) && B.alloc(d.lin, NB) && B.alloc(d.ang, NB) && B.alloc(d.last_lin, NB);
    ok &= B.alloc(d.inv_mass, NB) && B.alloc(d.bflags, NB) && B.alloc(d.iil, 9 * NB) && B.alloc(d.awake, NB) && B.alloc(d.sticks, NB);
    ok &= B.alloc(d.btype, NB) && B.alloc(d.bdims, NB) && B.alloc(d.bworld, NB);
    ok &= B.alloc(d.col_body, NC) && B.alloc(d.col_mat, NC) && B.alloc(d.col_world, NC);
    ok &= B.alloc(d.w_body_off, nw + 1) && B.alloc(d.w_col_off, nw + 1) && B.alloc(d.w_cap, nw) && B.alloc(d.hf, nw);
    ok &= B.alloc(d.heights, nh + 1) && B.alloc(d.pair_tab, 256);
    ok &= B.alloc(d.bk_pos, NB) && B.alloc(d.bk_rot, NB) && B.alloc(d.bk_lin, NB) && B.alloc(d.bk_ang, NB) && B.alloc(d.bk_last, NB);
    ok &= B.alloc(d.bk_awake, NB) && B.alloc(d.bk_sticks, NB);
    ok &= B.alloc(d.pre_lin, NB) && B.alloc(d.pre_ang, NB) && B.alloc(d.iiw, 9 * NB) && B.alloc(d.flags, NB) && B.alloc(d.ballf, NB) && B.alloc(d.root, NB);
    ok &= B.alloc(d.opos, NB) && B.alloc(d.orot, NB) && B.alloc(d.ppos, NB) && B.alloc(d.prot, NB) && B.alloc(d.slin, NB) && B.alloc(d.sang, NB);
    ok &= B.alloc(d.amin, NC) && B.alloc(d.amax, NC) && B.alloc(d.smin, NC) && B.alloc(d.smax, NC) && B.alloc(d.sweep, NC);
    ok &= B.alloc(d.ccls, NC) && B.alloc(d.bp_ord, NC) && B.alloc(d.bp_rank, NC);
    ok &= B.alloc(d.w_nsap, nw) && B.alloc(d.w_nover, nw) && B.alloc(d.w_solve, nw) && B.alloc(d.w_maxext, nw);
    ok &= B.alloc(d.out_off, NC) && B.alloc(d.cand_off, NC) && B.alloc(d.c_off, NC);
    size_t P = caps.pairs, T = caps.cands, C = caps.contacts, W = caps.warm;
    ok &= B.alloc(d.pair_i, P) && B.alloc(d.pair_j, P) && B.alloc(d.pair_n, P) && B.alloc(d.pc_p, 4 * P) && B.alloc(d.pc_n, 4 * P);
    ok &= B.alloc(d.tc_n, T) && B.alloc(d.tc_p, 4 * T) && B.alloc(d.tc_nf, 4 * T) && B.alloc(d.tc_cell, T);
    ok &= B.alloc(d.tcul_p, 4 * NC) && B.alloc(d.tcul_n, 4 * NC) && B.alloc(d.tcul_cell, 4 * NC) && B.alloc(d.tcul_cnt, NC);
    ok &= B.alloc(d.ct_a, C) && B.alloc(d.ct_b, C) && B.alloc(d.ct_cell, C) && B.alloc(d.ct_p, C) && B.alloc(d.ct_n, C) && B.alloc(d.ct_imp, C);
    ok &= B.alloc(d.cs_ai, C) && B.alloc(d.cs_bi, C) && B.alloc(d.cs_wa, C) && B.alloc(d.cs_wb, C) && B.alloc(d.cs_inv, C);
    ok &= B.alloc(d.cs_ra, C) && B.alloc(d.cs_rb, C) && B.alloc(d.cs_n, C) && B.alloc(d.cs_t1, C) && B.alloc(d.cs_t2, C) && B.alloc(d.cs_m, C) && B.alloc(d.cs_mu, C);
    ok &= B.alloc(d.jn, C) && B.alloc(d.jn_pre, C) && B.alloc(d.jt, C) && B.alloc(d.ct_key, C) && B.alloc(d.ct_anchor, C) && B.alloc(d.warm_idx, C);
    ok &= B.alloc(d.rank_conv, C);
    ok &= B.alloc(d.isl_k, C) && B.alloc(d.isl_v, C) && B.alloc(d.isl_k2, C) && B.alloc(d.isl_v2, C) && B.alloc(d.isl_start, C) && B.alloc(d.isl_flag, C + 1);
    ok &= B.alloc(d.ch_pa, C) && B.alloc(d.ch_pb, C) && B.alloc(d.sh_ord, C) && B.alloc(d.sh_pa, C) && B.alloc(d.sh_pb, C) && B.alloc(d.g_done, C);
    ok &= B.alloc(d.ch_sa, C) && B.alloc(d.ch_sb, C) && B.alloc(d.sh_sa, C) && B.alloc(d.sh_sb, C) && B.alloc(d.ex_base, C) && B.alloc(d.ex_shock, C);
    ok &= B.alloc(d.bs_pa, C) && B.alloc(d.bs_pb, C) && B.alloc(d.bs_c, C) && B.alloc(d.ss_pa, C) && B.alloc(d.ss_pb, C) && B.alloc(d.ss_c, C);
    ok &= B.alloc(d.slot_of, C) && B.alloc(d.lvl_off, C) && B.alloc(d.isl_w, C + 1) && B.alloc(d.wide_list, C + 1) && B.alloc(d.big_list, C + 1) && B.alloc(d.lvl_off_sh, C) && B.alloc(d.isl_wsh, C + 1) && B.alloc(d.bp_fill, NC) && B.alloc(d.tier_list[0], C + 1) && B.alloc(d.tier_list[1], C + 1) && B.alloc(d.tier_list[2], C + 1) && B.alloc(d.tier_list[3], C + 1) && B.alloc(d.ph, RP_PH_COUNT) && B.alloc(d.srow, (size_t)RP_SROW * C);
    d.cap_lane = C + RP_LANE_PAD;
    ok &= B.alloc(d.lrow, (size_t)RP_LROW * d.cap_lane) && B.alloc(d.lane_order, C + 1) && B.alloc(d.lid, 2 * C);
    ok &= B.alloc(d.feat, 12 * C) && B.alloc(d.pred, C);
    {
        float *mp = nullptr;
        ok &= B.alloc(mp, RP_MLP_NPARAM) && B.alloc(d.mlp.w1p, RP_MLP_W1P) && B.alloc(d.mlp.w2p, RP_MLP_W2P) && B.alloc(d.mlp.w2t, RP_MLP_W2P) &&
              B.alloc(d.mlp.w3p, RP_MLP_W3P) && B.alloc(d.mlp.w3t, RP_MLP_W3T);
        d.mlp.p = mp;
    }
    ok &= B.alloc(d.sk0, 2 * C) && B.alloc(d.sv0, 2 * C) && B.alloc(d.sk1, 2 * C) && B.alloc(d.sv1, 2 * C) && B.alloc(d.sh_hi, C);
    d.rs_blocks = (uint32_t)((std::max<size_t>(2 * C, NC) + 1023) / 1024);
    if (d.bp_radix) ok &= B.alloc(d.bp_k0, NC) && B.alloc(d.bp_v0, NC) && B.alloc(d.bp_k1, NC) && B.alloc(d.bp_v1, NC);
    ok &= B.alloc(d.rs_hist, 256 * (size_t)d.rs_blocks + 1);
    ok &= B.alloc(d.pc_la, C) && B.alloc(d.pc_lb, C);
    uint32_t np = 0;
    for (auto *w : ws) np += (uint32_t)w->players.size();
    d.np = np;
    ok &= B.alloc(d.w_player_off, nw + 1);
    if (np) {
        ok &= B.alloc(d.pl_body, np) && B.alloc(d.pl_grounded, np) && B.alloc(d.bk_pl_grounded, np) && B.alloc(d.pl_of_body, NB) &&
              B.alloc(d.pl_state, np) && B.alloc(d.bk_pl_state, np) && B.alloc(d.pl_input, np) && B.alloc(d.pl_bmin, np) && B.alloc(d.pl_bmax, np);
    }
    if (scene_bounces(ws)) {
        d.w6_on = 1;
        d.w6_slots = std::min<uint32_t>(nw, 32u);
        ok &= B.alloc(d.w_w6, 2 * (size_t)nw) && B.alloc(d.w6_toi, C) && B.alloc(d.w6_tv, C) && B.alloc(d.w6_mark, C) &&
              B.alloc(d.w6_rl, C) && B.alloc(d.w6_rc, C) && B.alloc(d.w6_pc, C) && B.alloc(d.w6_grp, C) && B.alloc(d.w6_comp, C) &&
              B.alloc(d.w6_cst, C + nw + 1) && B.alloc(d.w6_rt, C) && B.alloc(d.w6_racc, C) && B.alloc(d.w6_pu, C) &&
              B.alloc(d.w6_b, 2 * NB) && B.alloc(d.w6_par, NB) && B.alloc(d.w6_pl, NB) && B.alloc(d.w6_pa, NB) && B.alloc(d.w6_acc, 6 * NB);
        char *wsp = nullptr;
        ok &= B.alloc(wsp, rp_w6_ws_bytes() * d.w6_slots);
        d.w6_ws = wsp;
    }
    for (int k = 0; k < 2; k++) ok &= B.alloc(d.wk_key[k], W) && B.alloc(d.wk_la[k], W) && B.alloc(d.wk_jn[k], W) && B.alloc(d.wk_jt[k], W) && B.alloc(d.wc_off[k], nw + 1);
    ok &= B.alloc(d.tier_list[4], d.cm_on ? C + 1 : 1);
    if (d.cm_on) {  // contact models: per pair, per contact, per patch (one per two contacts), the patch cache
        d.cap_patch = (uint32_t)(C / 2 + 1);
        const size_t PT = d.cap_patch + 1;
        ok &= B.alloc(d.pair_cm, 256) && B.alloc(d.cs_soft, C) && B.alloc(d.ct_patch, C) && B.alloc(d.pt_head, C + 1) && B.alloc(d.isl_cm, C + 1);
        ok &= B.alloc(d.pt_first, PT) && B.alloc(d.pt_n, PT) && B.alloc(d.pt_ai, PT) && B.alloc(d.pt_bi, PT) && B.alloc(d.pt_inv, PT) &&
              B.alloc(d.pt_ra, PT) && B.alloc(d.pt_rb, PT) && B.alloc(d.pt_nn, PT) && B.alloc(d.pt_t1, PT) && B.alloc(d.pt_t2, PT) &&
              B.alloc(d.pt_m, PT) && B.alloc(d.pt_mu, PT) && B.alloc(d.pt_acc, PT) && B.alloc(d.pt_key, PT) && B.alloc(d.pt_soft, PT);
        for (int k = 0; k < 2; k++) ok &= B.alloc(d.wp_key[k], W) && B.alloc(d.wp_jt[k], W) && B.alloc(d.wpc_off[k], nw + 1);
        ok &= B.alloc(d.wp_new, PT) && B.alloc(d.wp_keep, W + 1);
    }
    ok &= B.alloc(d.newflag, C + 1) && B.alloc(d.keepflag, W + 1);
    size_t scan_elems = std::max<size_t>({NB, NC, C + 1, W + 1, 256 * (size_t)d.rs_blocks + 1, (size_t)nw + 8});
    ok &= B.alloc(d.scan_tmp, scan_elems / 1024 + 2 * nw + 64);
    ok &= B.alloc(d.cnt, CT_COUNT);
    {
        const size_t NJ = nj + 1;
        ok &= B.alloc(d.jt_a, NJ) && B.alloc(d.jt_b, NJ) && B.alloc(d.jt_fap, NJ) && B.alloc(d.jt_far, NJ) && B.alloc(d.jt_fbp, NJ) &&
              B.alloc(d.jt_fbr, NJ) && B.alloc(d.jr_u, 3 * NJ) && B.alloc(d.jr_ra, NJ) && B.alloc(d.jr_rb, NJ) && B.alloc(d.jr_mask, NJ) &&
              B.alloc(d.bk_wj_mask, NJ) && B.alloc(d.bk_wj_lam, 3 * NJ) &&
              B.alloc(d.jr_acc, 3 * NJ) && B.alloc(d.wj_mask, NJ) && B.alloc(d.wj_lam, 3 * NJ) && B.alloc(d.jadj_off, NB + 1) &&
              B.alloc(d.jadj, 2 * NJ) && B.alloc(d.w_joint_off, nw + 1);
        ok &= B.alloc(d.jt_kind, NJ);
        if (d.d6_on)
            ok &= B.alloc(d.jt_hinge, NJ) && B.alloc(d.jt_ax, 18 * NJ) && B.alloc(d.d6_row, 4 * 32 * NJ) && B.alloc(d.d6_nrows, NJ) &&
                  B.alloc(d.d6_racc, 32 * NJ) && B.alloc(d.d6_wn, NJ) && B.alloc(d.d6_wtag, 32 * NJ) && B.alloc(d.d6_wlam, 32 * NJ) &&
                  B.alloc(d.bk_d6_wn, NJ) && B.alloc(d.bk_d6_wtag, 32 * NJ) && B.alloc(d.bk_d6_wlam, 32 * NJ) && B.alloc(d.d6_plam, 6 * NJ);
    }
    // hash buffers
    HashDev &h = S->h;
    uint64_t bytes = 0;
    for (auto *w : ws) {
        uint64_t n = w->nb(), cap = std::max(w->max_bodies, w->nb());
        bytes += 48 + 56 * n + 8 + 24 + 8 + 12 * n + 8 + cap + 8 + 2 * cap + 1024 + (d.d6_on ? 168ull : 23ull) * w->nj();
    }
    bytes += 48ull * W;
    if (d.cm_on) bytes += 32ull * W;  // carried patch entries
    h.cap_chunks = (uint32_t)(bytes / 1024 + nw + 1);
    ok &= B.alloc(h.w_jbytes, nw + 1) && B.alloc(h.j_off, nj + 1);
    ok &= B.alloc(h.keyflag, W + 1) && B.alloc(h.w_chunk, nw + 1) && B.alloc(h.w_len, nw) && B.alloc(h.w_epoch, nw) && B.alloc(h.w_rng, nw);
    ok &= B.alloc(h.stream, (size_t)h.cap_chunks * 1024) && B.alloc(h.cv, 8 * (size_t)h.cap_chunks) && B.alloc(h.cv2, 8 * (size_t)h.cap_chunks) && B.alloc(h.digest, 32 * (size_t)nw);
    if (!ok) { set_err("device allocation failed (scene too large for the configured capacities)"); return nullptr; }
    if (!hip_ok(hipStreamCreateWithFlags(&S->st, hipStreamNonBlocking), "hipStreamCreate")) return nullptr;
    for (int i = 0; i < 2; i++)
        if (!hip_ok(hipStreamCreateWithFlags(&S->ts.side[i], hipStreamNonBlocking), "hipStreamCreate")) return nullptr;
    for (int i = 0; i < 3; i++)
        if (!hip_ok(hipEventCreateWithFlags(&S->ts.ev[i], hipEventDisableTiming), "hipEventCreate")) return nullptr;
    // ---- upload
    std::vector<F4> pos, rot, lin, ang, last, dims;
    std::vector<float> invm, iil, heights;
    std::vector<uint32_t> bfl, awk, stk, bty, bwd, cbody, cmat, cwld, wcap;
    std::vector<HfDesc> hfs;
    std::vector<uint64_t> wkey, ep, rn;
    std::vector<F4> wla, wjt;
    std::vector<float> wjn;
    std::vector<uint32_t> woff, poff_w;
    std::vector<uint64_t> pkey;
    std::vector<F4> pjt;
    woff.push_back(0); poff_w.push_back(0);
    for (uint32_t wi = 0; wi < nw; wi++) {
        HostWorld *w = ws[wi];
        uint32_t bo = S->body_off[wi];
        pos.insert(pos.end(), w->pos.begin(), w->pos.end()); rot.insert(rot.end(), w->rot.begin(), w->rot.end());
        lin.insert(lin.end(), w->lin.begin(), w->lin.end()); ang.insert(ang.end(), w->ang.begin(), w->ang.end());
        last.insert(last.end(), w->last_lin.begin(), w->last_lin.end()); dims.insert(dims.end(), w->dims.begin(), w->dims.end());
        invm.insert(invm.end(), w->inv_mass.begin(), w->inv_mass.end()); iil.insert(iil.end(), w->iil.begin(), w->iil.end());
        bfl.insert(bfl.end(), w->bflags.begin(), w->bflags.end()); awk.insert(awk.end(), w->awake.begin(), w->awake.end());
        stk.insert(stk.end(), w->sticks.begin(), w->sticks.end()); bty.insert(bty.end(), w->btype.begin(), w->btype.end());
        for (uint32_t i = 0; i < w->nb(); i++) bwd.push_back(wi);
        for (uint32_t c = 0; c < w->nc(); c++) { cbody.push_back(w->col_body[c] + bo); cmat.push_back(w->col_mat[c]); cwld.push_back(wi); }
        wcap.push_back(std::max(w->max_bodies, w->nb()));
        HfDesc hd{};
        if (w->has_hf) {
            hd.nx = w->hnx; hd.nz = w->hnz; hd.off = (uint32_t)heights.size(); hd.mat = w->hmat;
            hd.sx = w->hsx; hd.sz = w->hsz; hd.ox = w->hox; hd.oz = w->hoz; hd.hmin = w->hmin; hd.hmax = w->hmax; hd.valid = 1;
            heights.insert(heights.end(), w->heights.begin(), w->heights.end());
        }
        hfs.push_back(hd);
        wkey.insert(wkey.end(), w->wkey.begin(), w->wkey.end()); wla.insert(wla.end(), w->wla.begin(), w->wla.end());
        wjn.insert(wjn.end(), w->wjn.begin(), w->wjn.end()); wjt.insert(wjt.end(), w->wjt.begin(), w->wjt.end());
        woff.push_back((uint32_t)wkey.size());
        pkey.insert(pkey.end(), w->pkey.begin(), w->pkey.end()); pjt.insert(pjt.end(), w->pjt.begin(), w->pjt.end());
        poff_w.push_back((uint32_t)pkey.size());
        ep.push_back(w->epoch); rn.push_back(w->rng);
    }
    // joints (global body indices), their warm rows, each body's jointed partners, world offsets
    std::vector<uint32_t> ja, jb, jmask, jadj_off(nb + 1, 0), jadj, jwoff, jkind, jhinge, jwn, jwtag;
    std::vector<F4> jax;
    std::vector<float> jwlam;
    std::vector<F4> jfap, jfar, jfbp, jfbr;
    std::vector<float> jlam;
    jwoff.push_back(0);
    for (uint32_t wi = 0; wi < nw; wi++) {
        HostWorld *w = ws[wi];
        uint32_t bo = S->body_off[wi];
        for (uint32_t q = 0; q < w->nj(); q++) {
            ja.push_back(w->jt_a[q] + bo); jb.push_back(w->jt_b[q] + bo);
            jadj_off[w->jt_a[q] + bo + 1]++; jadj_off[w->jt_b[q] + bo + 1]++;
        }
        jfap.insert(jfap.end(), w->jt_fap.begin(), w->jt_fap.end()); jfar.insert(jfar.end(), w->jt_far.begin(), w->jt_far.end());
        jfbp.insert(jfbp.end(), w->jt_fbp.begin(), w->jt_fbp.end()); jfbr.insert(jfbr.end(), w->jt_fbr.begin(), w->jt_fbr.end());
        jmask.insert(jmask.end(), w->wj_mask.begin(), w->wj_mask.end()); jlam.insert(jlam.end(), w->wj_lam.begin(), w->wj_lam.end());
        jkind.insert(jkind.end(), w->jt_kind.begin(), w->jt_kind.end()); jhinge.insert(jhinge.end(), w->jt_hinge.begin(), w->jt_hinge.end());
        jax.insert(jax.end(), w->jt_ax.begin(), w->jt_ax.end()); jwn.insert(jwn.end(), w->d6_wn.begin(), w->d6_wn.end());
        jwtag.insert(jwtag.end(), w->d6_wtag.begin(), w->d6_wtag.end()); jwlam.insert(jwlam.end(), w->d6_wlam.begin(), w->d6_wlam.end());
        jwoff.push_back((uint32_t)ja.size());
    }
    for (uint32_t b = 0; b < nb; b++) jadj_off[b + 1] += jadj_off[b];
    jadj.resize(jadj_off[nb]);
    {
        std::vector<uint32_t> fill(jadj_off.begin(), jadj_off.end() - 1);
        for (size_t q = 0; q < ja.size(); q++) { jadj[fill[ja[q]]++] = jb[q]; jadj[fill[jb[q]]++] = ja[q]; }
    }
    std::vector<F4> ptab(256), pcm(256, f4(0, 0, 0, 0));
    {
        HostWorld *w0 = ws[0];
        for (uint32_t a = 0; a < 16; a++)
            for (uint32_t b = 0; b < 16; b++) {
                MatP p = table_pair(a, b);
                uint32_t x = std::min(a, b), y = std::max(a, b);
                for (auto &o : w0->ovr)
                    if (o.a == x && o.b == y) { p = o.p; pcm[a * 16 + b] = f4(o.stiff, o.diss, o.patch ? 1.0f : 0.0f); }
                ptab[a * 16 + b] = f4(p.mu_s, p.mu_k, p.rest, p.vs);
            }
    }
    auto up = [&](void *dst, const void *src, size_t bytes) { return bytes == 0 || hip_ok(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, S->st), "upload"); };
    ok = up(d.pos, pos.data(), pos.size() * 16) && up(d.rot, rot.data(), rot.size() * 16) && up(d.lin, lin.data(), lin.size() * 16) &&
         up(d.ang, ang.data(), ang.size() * 16) && up(d.last_lin, last.data(), last.size() * 16) && up(d.bdims, dims.data(), dims.size() * 16) &&
         up(d.inv_mass, invm.data(), invm.size() * 4) && up(d.iil, iil.data(), iil.size() * 4) && up(d.bflags, bfl.data(), bfl.size() * 4) &&
         up(d.awake, awk.data(), awk.size() * 4) && up(d.sticks, stk.data(), stk.size() * 4) && up(d.btype, bty.data(), bty.size() * 4) &&
         up(d.bworld, bwd.data(), bwd.size() * 4) && up(d.col_body, cbody.data(), cbody.size() * 4) && up(d.col_mat, cmat.data(), cmat.size() * 4) &&
         up(d.col_world, cwld.data(), cwld.size() * 4) && up(d.w_body_off, S->body_off.data(), S->body_off.size() * 4) &&
         up(d.w_col_off, S->col_off.data(), S->col_off.size() * 4) && up(d.w_cap, wcap.data(), wcap.size() * 4) &&
         up(d.hf, hfs.data(), hfs.size() * sizeof(HfDesc)) && up(d.heights, heights.data(), heights.size() * 4) &&
         up(d.pair_tab, ptab.data(), ptab.size() * 16) && up(d.wk_key[cur], wkey.data(), wkey.size() * 8) &&
         up(d.wk_la[cur], wla.data(), wla.size() * 16) && up(d.wk_jn[cur], wjn.data(), wjn.size() * 4) && up(d.wk_jt[cur], wjt.data(), wjt.size() * 16) &&
         up(d.wc_off[cur], woff.data(), woff.size() * 4) && up(h.w_epoch, ep.data(), ep.size() * 8) && up(h.w_rng, rn.data(), rn.size() * 8) &&
         up(d.jt_a, ja.data(), ja.size() * 4) && up(d.jt_b, jb.data(), jb.size() * 4) && up(d.jt_fap, jfap.data(), jfap.size() * 16) &&
         up(d.jt_far, jfar.data(), jfar.size() * 16) && up(d.jt_fbp, jfbp.data(), jfbp.size() * 16) && up(d.jt_fbr, jfbr.data(), jfbr.size() * 16) &&
         up(d.wj_mask, jmask.data(), jmask.size() * 4) && up(d.wj_lam, jlam.data(), jlam.size() * 4) &&
         up(d.jadj_off, jadj_off.data(), jadj_off.size() * 4) && up(d.jadj, jadj.data(), jadj.size() * 4) &&
         up(d.w_joint_off, jwoff.data(), jwoff.size() * 4) && up(d.jt_kind, jkind.data(), jkind.size() * 4);
    if (ok && d.d6_on)
        ok = up(d.jt_hinge, jhinge.data(), jhinge.size() * 4) && up(d.jt_ax, jax.data(), jax.size() * 16) && up(d.d6_wn, jwn.data(), jwn.size() * 4) &&
             up(d.d6_wtag, jwtag.data(), jwtag.size() * 4) && up(d.d6_wlam, jwlam.data(), jwlam.size() * 4);
    {
        std::vector<uint32_t> poff{0}, pbody, pgnd, pof(nb + 1, RP_NONE);
        std::vector<F4> pst, pin;
        for (size_t wi = 0; wi < ws.size(); wi++) {
            for (auto &pl : ws[wi]->players) {
                const uint32_t gb = S->body_off[wi] + pl.body;
                pof[gb] = (uint32_t)pbody.size();
                pbody.push_back(gb); pgnd.push_back(pl.grounded);
                pst.push_back(f4(pl.radius, pl.height, pl.speed, pl.vvel));
                pin.push_back(f4(pl.input[0], pl.input[1], pl.input[2]));
            }
            poff.push_back((uint32_t)pbody.size());
        }
        ok = ok && up(d.w_player_off, poff.data(), poff.size() * 4);
        if (np) ok = ok && up(d.pl_body, pbody.data(), np * 4) && up(d.pl_grounded, pgnd.data(), np * 4) && up(d.pl_of_body, pof.data(), pof.size() * 4) &&
                     up(d.pl_state, pst.data(), np * 16) && up(d.pl_input, pin.data(), np * 16);
    }
    if (ok && d.cm_on)
        ok = up(d.pair_cm, pcm.data(), pcm.size() * 16) && up(d.wp_key[cur], pkey.data(), pkey.size() * 8) &&
             up(d.wp_jt[cur], pjt.data(), pjt.size() * 16) && up(d.wpc_off[cur], poff_w.data(), poff_w.size() * 4);
    if (ok) ok = hip_ok(hipMemsetAsync(d.cnt, 0, CT_COUNT * 4, S->st), "memset");
    if (ok && ws[0]->mlp.size() == RP_MLP_NPARAM) ok = upload_model(*S, ws[0]->mlp.data());
    S->col_body = cbody;
    if (ok) ok = hip_ok(hipStreamSynchronize(S->st), "upload sync");
    if (!ok) return nullptr;
    S->cur = cur;
    return S.release();
}

// Downloads the mutable state of world wi of the scene into the host mirror.
static bool pull_world(Scene &S, uint32_t wi, HostWorld &w) {
    const Dev &d = S.d;
    uint32_t bo = S.body_off[wi], n = S.body_off[wi + 1] - bo;
    auto dn = [&](void *dst, const void *src, size_t bytes) { return bytes == 0 || hip_ok(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, S.st), "download"); };
    bool ok = dn(w.pos.data(), d.pos + bo, n * 16) && dn(w.rot.data(), d.rot + bo, n * 16) && dn(w.lin.data(), d.lin + bo, n * 16) &&
              dn(w.ang.data(), d.ang + bo, n * 16) && dn(w.last_lin.data(), d.last_lin + bo, n * 16) &&
              dn(w.awake.data(), d.awake + bo, n * 4) && dn(w.sticks.data(), d.sticks + bo, n * 4);
    uint32_t off[2];
    ok = ok && dn(off, d.wc_off[S.cur] + wi, 8) && hip_ok(hipStreamSynchronize(S.st), "download sync");
    if (!ok) return false;
    uint32_t ne = off[1] - off[0];
    w.wkey.resize(ne); w.wla.resize(ne); w.wjn.resize(ne); w.wjt.resize(ne);
    ok = dn(w.wkey.data(), d.wk_key[S.cur] + off[0], ne * 8) && dn(w.wla.data(), d.wk_la[S.cur] + off[0], ne * 16) &&
         dn(w.wjn.data(), d.wk_jn[S.cur] + off[0], ne * 4) && dn(w.wjt.data(), d.wk_jt[S.cur] + off[0], ne * 16);
    if (ok && w.nj()) {
        uint32_t joff[2];
        ok = dn(joff, d.w_joint_off + wi, 8) && hip_ok(hipStreamSynchronize(S.st), "download sync") &&
             dn(w.wj_mask.data(), d.wj_mask + joff[0], w.nj() * 4) && dn(w.wj_lam.data(), d.wj_lam + 3 * (size_t)joff[0], 3 * (size_t)w.nj() * 4);
        if (ok && d.d6_on)
            ok = dn(w.d6_wn.data(), d.d6_wn + joff[0], w.nj() * 4) && dn(w.d6_wtag.data(), d.d6_wtag + 32 * (size_t)joff[0], 32 * (size_t)w.nj() * 4) &&
                 dn(w.d6_wlam.data(), d.d6_wlam + 32 * (size_t)joff[0], 32 * (size_t)w.nj() * 4);
    }
    if (ok && d.cm_on) {
        uint32_t po[2];
        ok = dn(po, d.wpc_off[S.cur] + wi, 8) && hip_ok(hipStreamSynchronize(S.st), "download sync");
        if (ok) {
            const uint32_t np = po[1] - po[0];
            w.pkey.resize(np); w.pjt.resize(np);
            ok = dn(w.pkey.data(), d.wp_key[S.cur] + po[0], (size_t)np * 8) && dn(w.pjt.data(), d.wp_jt[S.cur] + po[0], (size_t)np * 16);
        }
    }
    if (ok && !w.players.empty() && d.np) {
        uint32_t poff = 0;
        const size_t np = w.players.size();
        std::vector<F4> st(np);
        std::vector<uint32_t> gnd(np);
        ok = dn(&poff, d.w_player_off + wi, 4) && hip_ok(hipStreamSynchronize(S.st), "download sync") &&
             dn(st.data(), d.pl_state + poff, np * 16) && dn(gnd.data(), d.pl_grounded + poff, np * 4) &&
             hip_ok(hipStreamSynchronize(S.st), "download sync");
        for (size_t k = 0; ok && k < np; k++) { w.players[k].vvel = st[k].w; w.players[k].grounded = gnd[k]; }
    }
    return ok && hip_ok(hipStreamSynchronize(S.st), "download sync");
}

// Test hooks, process-wide: the sweep tiers' island-size bounds and the wide-level minimum (so
// small scenes can be solved by every tier and by the split form), and the initial capacities
// of new scenes (so the overflow-and-replay path runs on small scenes).  Zero: the default.
static std::atomic<uint32_t> g_tier_max[3] = {{0}, {0}, {0}}, g_wide_min{0}, g_lane_max{0}, g_pair_max{0}, g_quad_max{0}, g_wave_min{0};
static std::atomic<uint32_t> g_init_caps[4] = {{0}, {0}, {0}, {0}};

static TickParams tick_params(const HostWorld &w, float dt, uint32_t iv) {
    TickParams tp{};
    tp.dt = dt;
    // tier bounds: each within its tier's flag storage and not below the previous tier's
    const uint32_t storage[3] = {256u, 4096u, 131072u};
    uint32_t lo = 0;
    for (int k = 0; k < 3; k++) {
        uint32_t v = g_tier_max[k].load();
        v = v ? std::min(v, storage[k]) : storage[k];
        tp.tier_max[k] = lo = std::max(v, lo);
    }
    tp.wide_min = g_wide_min.load() ? g_wide_min.load() : RP_WIDE_MIN;
    // tier 0 with the lane tier on: lanes up to lane_max entries, lane pairs up to pair_max (at most
    // 128), groups of four lanes up to quad_max, a wave each above (every bound within tier 0's)
    const uint32_t lm = g_lane_max.load(), pm = g_pair_max.load(), qm = g_quad_max.load();
    tp.lane_max = lm == RPHYS_LANE_NONE ? 0u : std::min(lm ? lm : RP_LANE_MAX, tp.tier_max[0]);
    tp.pair_max = std::min(std::max(tp.lane_max, pm ? pm : RP_PAIR_MAX), tp.tier_max[0]);
    if (tp.pair_max > 128u && tp.pair_max > tp.lane_max) tp.pair_max = std::max(128u, tp.lane_max);
    tp.quad_max = std::min(std::max(tp.pair_max, qm ? qm : RP_QUAD_MAX), tp.tier_max[0]);
    const uint32_t wm = g_wave_min.load();
    tp.wave_min = wm == 0 ? RP_WAVE_MIN : (wm == RPHYS_LANE_NONE ? 0u : wm);
    // gravity.rs:85-87 q6 of the uniform field
    tp.gx = std::isfinite(w.gx) ? q6mul(w.gx) : 0.0f;
    tp.gy = std::isfinite(w.gy) ? q6mul(w.gy) : 0.0f;
    tp.gz = std::isfinite(w.gz) ? q6mul(w.gz) : 0.0f;
    // step.rs:448 / solver.rs:968 up = -normalize(gravity) (Vec3::normalize = v * (1/|v|))
    float l2 = (w.gx * w.gx + w.gy * w.gy) + w.gz * w.gz;
    if (l2 > 0.0f) {
        float r = 1.0f / std::sqrt(l2);
        tp.upx = -(w.gx * r); tp.upy = -(w.gy * r); tp.upz = -(w.gz * r);
    } else { tp.upx = 0.0f; tp.upy = 1.0f; tp.upz = 0.0f; }
    tp.iv = iv; tp.ip = w.ip;
    if (w.resid_on) { tp.resid_on = 1; tp.resid_tol = w.resid_tol; tp.iv = w.resid_max; }
    tp.ws_mode = w.ws_mode;
    tp.damp_start = w.damp_start;
    // one island per thread serialises its rows over every projection iteration: worth it for a
    // few iterations only
    tp.lane_pos = (tp.lane_max && tp.ip <= RP_LANE_POS_ITERS) ? 1u : 0u;
    return tp;
}

// Runs n ticks on the scene.  On a capacity overflow the failed tick is rolled back, the scene is
// rebuilt with larger buffers (through the host mirrors) and stepping resumes.
struct ProfileAcc;
static bool run_ticks(Scene *&S, std::vector<HostWorld *> &ws, float dt, uint32_t iv, uint32_t n, bool use_graph, ProfileAcc *prof = nullptr);

static Caps initial_caps(const std::vector<HostWorld *> &ws) {
    uint64_t nc = 0;
    for (auto *w : ws) nc += w->nc();
    Caps c;
    c.pairs = (uint32_t)std::min<uint64_t>(16 * nc + 1024, 0x7FFFFFFF);
    c.cands = (uint32_t)std::min<uint64_t>(96 * nc + 4096, 0x7FFFFFFF);
    uint64_t nj = 0;
    for (auto *w : ws) nj += w->nj();
    c.contacts = (uint32_t)std::min<uint64_t>(12 * nc + nj + 1024, RP_PBIT - 1);  // the island lists hold joints too; entries stay below RP_PBIT
    c.warm = (uint32_t)std::min<uint64_t>(12 * nc + 1024, 0x7FFFFFFF);
    if (g_init_caps[0].load()) c.pairs = g_init_caps[0].load();
    if (g_init_caps[1].load()) c.cands = g_init_caps[1].load();
    if (g_init_caps[2].load()) c.contacts = g_init_caps[2].load();
    if (g_init_caps[3].load()) c.warm = g_init_caps[3].load();
    return c;
}

static bool capture_graph(Scene &S, const TickParams &tp0, uint32_t parity) {
    TickParams tp = tp0;
    tp.cur = parity;
    hipGraph_t g = nullptr;
    if (!hip_ok(hipStreamBeginCapture(S.st, hipStreamCaptureModeThreadLocal), "hipStreamBeginCapture")) return false;
    hipError_t le = rp_launch_tick(S.d, tp, S.grids, S.st, nullptr, &S.ts);
    if (!hip_ok(hipStreamEndCapture(S.st, &g), "hipStreamEndCapture")) return false;
    if (!hip_ok(le, "kernel launch (capture)")) { (void)hipGraphDestroy(g); return false; }
    bool ok = hip_ok(hipGraphInstantiate(&S.graph[parity], g, nullptr, nullptr, 0), "hipGraphInstantiate");
    (void)hipGraphDestroy(g);
    return ok;
}

static_assert(RP_NMARKS - 1 <= RPHYS_PROFILE_MAX_STAGES, "stage count exceeds RPhysProfile");
// Stage timing of the profiled step: events at the stage boundaries of each tick, summed.
// Up to PROFILE_BATCH ticks are in flight, each with its own event set, so the stream never
// waits on the host between ticks.
struct ProfileAcc {
    static constexpr uint32_t PROFILE_BATCH = 64;
    // RP_NMARKS stage marks on the tick's stream, then the small tier's velocity and projection
    // kernels' start and end events on the side stream they run on
    static constexpr int NEV = RP_NMARKS + 4;
    hipEvent_t ev[PROFILE_BATCH][NEV] = {};
    RPhysProfile *out = nullptr;
    bool init() {
        for (uint32_t t = 0; t < PROFILE_BATCH; t++)
            for (int k = 0; k < NEV; k++)
                if (!hip_ok(hipEventCreate(&ev[t][k]), "hipEventCreate")) return false;
        return true;
    }
    ~ProfileAcc() {
        for (uint32_t t = 0; t < PROFILE_BATCH; t++)
            for (int k = 0; k < NEV; k++) if (ev[t][k]) (void)hipEventDestroy(ev[t][k]);
    }
    // Adds the stage times of the first `done` ticks of the batch and the counter sums.
    bool add(uint32_t done, const uint32_t *cnt) {
        for (uint32_t t = 0; t < done; t++)
            for (int k = 0; k + 1 < RP_NMARKS; k++) {
                // the small tier's stages run alongside the larger tiers: their own events
                const int a = k == RP_ST_VEL_SMALL ? RP_NMARKS : (k == RP_ST_PROJ_SMALL ? RP_NMARKS + 2 : k);
                const int b = a == k ? k + 1 : a + 1;
                float ms = 0.0f;
                if (!hip_ok(hipEventElapsedTime(&ms, ev[t][a], ev[t][b]), "hipEventElapsedTime")) return false;
                out->stage_ms[k] += ms;
            }
        out->ticks += done;
        out->pairs += cnt[CT_SUM_PAIRS]; out->terrain_candidates += cnt[CT_SUM_CANDS]; out->contacts += cnt[CT_SUM_CONTACTS];
        out->island_contacts += cnt[CT_SUM_ISLC]; out->islands += cnt[CT_SUM_ISLANDS]; out->island_contacts_global += cnt[CT_SUM_ISLC_GLB];
        out->pgs_iters_max += cnt[CT_SUM_ITERS_MAX]; out->pgs_iters_sum += cnt[CT_SUM_ITERS_SUM] + 4294967296.0 * cnt[CT_SUM_ITERS_SUM_HI];
        out->islands_converged += cnt[CT_SUM_ISL_CONV]; out->island_contacts_converged += cnt[CT_SUM_ISLC_CONV];
        out->warm_start_learned += cnt[CT_SUM_WS_LEARNED];
        out->pgs_iters_sum_converged += cnt[CT_SUM_ITERS_CONV] + 4294967296.0 * cnt[CT_SUM_ITERS_CONV_HI];
        out->wide_levels += cnt[CT_SUM_WIDE]; out->wide_levels_shock += cnt[CT_SUM_WIDE_SH];
        return true;
    }
};

static bool run_ticks(Scene *&S, std::vector<HostWorld *> &ws, float dt, uint32_t iv, uint32_t n, bool use_graph, ProfileAcc *prof) {
    uint32_t remaining = n;
    Caps caps{S->d.cap_pairs, S->d.cap_cands, S->d.cap_contacts, S->d.cap_warm};
    if (prof) use_graph = false;
    while (remaining > 0) {
        TickParams tp = tick_params(*ws[0], dt, iv);
        tp.log_on = S->d.log_cap > 0 ? 1u : 0u;
        if (!hip_ok(hipMemsetAsync(S->d.cnt, 0, CT_COUNT * 4, S->st), "memset")) return false;
        if (use_graph) {
            TickParams key = tp;
            key.cur = 0;
            if (!S->graph[0] || !S->graph[1] || std::memcmp(&key, &S->graph_tp, sizeof key) != 0 ||
                std::memcmp(S->grids, S->graph_grids, sizeof S->grids) != 0) {
                for (int i = 0; i < 2; i++) if (S->graph[i]) { (void)hipGraphExecDestroy(S->graph[i]); S->graph[i] = nullptr; }
                if (!capture_graph(*S, tp, 0) || !capture_graph(*S, tp, 1)) return false;
                S->graph_tp = key;
                std::memcpy(S->graph_grids, S->grids, sizeof S->grids);
            }
            for (uint32_t k = 0; k < remaining; k++)
                if (!hip_ok(hipGraphLaunch(S->graph[(S->cur + k) & 1u], S->st), "hipGraphLaunch")) return false;
        } else {
            uint32_t batch = prof ? std::min(remaining, ProfileAcc::PROFILE_BATCH) : remaining;
            for (uint32_t k = 0; k < batch; k++) {
                tp.cur = (S->cur + k) & 1u;
                if (!hip_ok(rp_launch_tick(S->d, tp, S->grids, S->st, prof ? prof->ev[k] : nullptr, &S->ts), "kernel launch"))
                    return false;
            }
        }
        uint32_t cnt[CT_COUNT];
        if (!hip_ok(hipMemcpyAsync(cnt, S->d.cnt, sizeof(cnt), hipMemcpyDeviceToHost, S->st), "counters")) return false;
        if (!hip_ok(hipStreamSynchronize(S->st), "step")) return false;
        uint32_t done = cnt[CT_DONE];
        // level launches for the next ticks: the last tick's most leading wide levels of one
        // island, rounded up to a multiple of 8 (fewer graph re-captures); any count is exact
        auto levels = [](uint32_t k) { return k ? std::min<uint32_t>(RP_WIDE_MAX_LEVELS, (k + 7u) / 8u * 8u) : 0u; };
        S->grids[3] = levels(cnt[CT_WIDE_MAX]);
        S->grids[4] = levels(cnt[CT_WSH_MAX]);
        S->cur ^= (done & 1u);
        for (auto *w : ws) w->tick += done;
        remaining -= done;
        if (!cnt[CT_ABORT]) {
            if (done == 0 && remaining > 0) { set_err("step made no progress"); return false; }
            if (prof && !prof->add(done, cnt)) return false;
            continue;
        }
        if (prof && !prof->add(done, cnt)) return false;
        // roll back the failed tick and grow
        rp_launch_restore_backup(S->d, S->st);
        if (!hip_ok(hipStreamSynchronize(S->st), "restore")) return false;
        auto grow = [](uint32_t cap, uint32_t need) -> uint32_t {
            uint64_t v = std::max<uint64_t>((uint64_t)need + need / 2 + 1024, (uint64_t)cap * 2);
            return (uint32_t)std::min<uint64_t>(v, RP_PBIT - 1);  // island entries stay below RP_PBIT
        };
        if (cnt[CT_NEED_PAIRS]) caps.pairs = grow(caps.pairs, cnt[CT_NEED_PAIRS]);
        if (cnt[CT_NEED_CANDS]) caps.cands = grow(caps.cands, cnt[CT_NEED_CANDS]);
        if (cnt[CT_NEED_CONTACTS]) caps.contacts = grow(caps.contacts, cnt[CT_NEED_CONTACTS]);
        if (cnt[CT_NEED_WARM]) caps.warm = grow(caps.warm, cnt[CT_NEED_WARM]);
        for (uint32_t i = 0; i < ws.size(); i++) if (!pull_world(*S, i, *ws[i])) return false;
        uint32_t cur = S->cur;
        Scene *N = build_scene(ws, caps, cur);
        if (!N) return false;
        std::memcpy(N->grids, S->grids, sizeof S->grids);
        N->slots = std::move(S->slots);  // rollback slots hold body state and a warm cache no larger than the new capacity
        N->d.log_feat = S->d.log_feat; N->d.log_targ = S->d.log_targ; N->d.log_n = S->d.log_n; N->d.log_cap = S->d.log_cap;
        delete S;
        S = N;
    }
    return true;
}

// ------------------------------------------------------------------ the world handle
struct RPhysWorld {
    std::mutex mu;
    HostWorld w;
    Scene *scene = nullptr;
    bool dev_newer = false;    // device state newer than the host mirror
    bool need_build = true;    // host structure or state changed since the last upload
    bool have_contacts = false;  // the scene's contact arrays hold the last tick's contacts
    ~RPhysWorld() { delete scene; }
};

static bool pull(RPhysWorld *W) {
    if (W->dev_newer && W->scene) {
        if (!pull_world(*W->scene, 0, W->w)) return false;
        W->dev_newer = false;
    }
    return true;
}
static void host_changed(RPhysWorld *W) { W->need_build = true; }
static bool push(RPhysWorld *W) {
    if (!W->need_build && W->scene) return true;
    Caps caps = W->scene ? Caps{W->scene->d.cap_pairs, W->scene->d.cap_cands, W->scene->d.cap_contacts, W->scene->d.cap_warm}
                         : initial_caps({&W->w});
    uint32_t cur = W->scene ? W->scene->cur : 0;
    std::vector<HostWorld *> ws{&W->w};
    Scene *S = build_scene(ws, caps, cur);
    if (!S) return false;
    delete W->scene;
    W->scene = S;
    W->need_build = false;
    W->have_contacts = false;
    return true;
}

#define GUARD_BEGIN try {
#define GUARD_END                                                                     \
    }                                                                                 \
    catch (const std::bad_alloc &) { set_err("out of host memory"); return RPHYS_ERR_INTERNAL; } \
    catch (...) { set_err("internal error crossed the C ABI"); return RPHYS_ERR_INTERNAL; }

static bool fin3(const RPhysVec3 &v) { return hfin(v.x) && hfin(v.y) && hfin(v.z); }
static bool finq(const RPhysQuat &q) { return hfin(q.x) && hfin(q.y) && hfin(q.z) && hfin(q.w); }
static bool fin_iso(const RPhysIsometry &p) { return fin3(p.pos) && finq(p.rot); }
static bool fin_vel(const RPhysVelocity &v) { return fin3(v.lin) && fin3(v.ang); }
static bool fin_mat(const RPhysMaterial &m) { return hfin(m.mu_static) && hfin(m.mu_dynamic) && hfin(m.restitution); }

static RPhysResult bad_arg(const char *ctx, const char *msg) { set_err(std::string(ctx) + ": " + msg); return RPHYS_ERR_BAD_ARG; }

// lib.rs:369-409 checked_live_body_id (1-based ids, 0 and the terrain sentinel rejected)
static RPhysResult live_id(const RPhysWorld *W, uint32_t id, const char *ctx, uint32_t &out) {
    if (id == RPHYS_INVALID_BODY_ID) { set_err(std::string(ctx) + ": body_id == 0 (invalid)"); return RPHYS_ERR_INVALID_ID; }
    if (id == RPHYS_TERRAIN_BODY_ID) { set_err(std::string(ctx) + ": body_id is terrain sentinel (not a real body)"); return RPHYS_ERR_INVALID_ID; }
    uint32_t raw = id - 1;
    if (raw >= W->w.bodies_created) { set_err(std::string(ctx) + ": body_id out of range"); return RPHYS_ERR_INVALID_ID; }
    if (!(W->w.bflags[raw] & BF_ALIVE)) { set_err(std::string(ctx) + ": body_id not found (removed)"); return RPHYS_ERR_INVALID_ID; }
    out = raw;
    return RPHYS_OK;
}

// mass.rs:92-165 density primitives + MassProps::finalize; world.rs:595-614 add_body
static void add_body_host(HostWorld &w, uint32_t shape, const float dims[3], const RPhysIsometry &pose, const RPhysVelocity &vel, bool dynamic, float density) {
    float rho = density > 0.0f ? density : 1500.0f;
    float m = 0.0f, ix = 0.0f, iy = 0.0f, iz = 0.0f;
    bool ok = false;
    const float PI = 3.14159265358979323846f;
    if (dynamic) {
        if (shape == RPHYS_SHAPE_SPHERE) {
            float r = std::fabs(dims[0]);
            if (hfin(r) && r > 0 && hfin(rho) && rho > 0) { float r2 = r * r; float vol = ((4.0f / 3.0f) * PI * r2) * r; m = rho * vol; ix = iy = iz = (0.4f * m) * r2; ok = true; }
        } else if (shape == RPHYS_SHAPE_BOX) {
            float hx = std::fabs(dims[0]), hy = std::fabs(dims[1]), hz = std::fabs(dims[2]);
            if (hfin(hx) && hfin(hy) && hfin(hz) && hfin(rho) && rho > 0) {
                float dx = hx * 2.0f, dy = hy * 2.0f, dz = hz * 2.0f;
                if (dx > 0 && dy > 0 && dz > 0) {
                    m = rho * ((dx * dy) * dz);
                    float x2 = dx * dx, y2 = dy * dy, z2 = dz * dz;
                    ix = ((1.0f / 12.0f) * m) * (y2 + z2); iy = ((1.0f / 12.0f) * m) * (x2 + z2); iz = ((1.0f / 12.0f) * m) * (x2 + y2);
                    ok = true;
                }
            }
        } else {
            float r = std::fabs(dims[0]), hh = std::fabs(dims[1]);
            if (hfin(r) && r > 0 && hfin(rho) && rho > 0) {
                float h = hh * 2.0f, r2 = r * r;
                float vc = (PI * r2) * h, vs = ((4.0f / 3.0f) * PI * r2) * r;
                m = rho * (vc + vs);
                ix = ((0.25f * m) * r2) + (((1.0f / 12.0f) * m) * h) * h; iy = (0.5f * m) * r2; iz = ix;
                ok = true;
            }
        }
        if (ok && (!(hfin(m) && m > 0.0f) || m < 1.0e-12f)) ok = false;
        if (ok && !(hfin(ix) && ix > 0 && hfin(iy) && iy > 0 && hfin(iz) && iz > 0)) ok = false;
    }
    float inv_mass = (dynamic && ok) ? 1.0f / m : 0.0f;
    HM I; I.c0 = hv(ok ? ix : 0, 0, 0); I.c1 = hv(0, ok ? iy : 0, 0); I.c2 = hv(0, 0, ok ? iz : 0);
    float det = hdot(I.c2, hcross(I.c0, I.c1));
    bool have_ii = dynamic && hfin(det) && det > 0.0f;
    HM inv = have_ii ? hinv(I) : HM{};
    bool is_dyn = dynamic && hfin(inv_mass) && inv_mass > 0.0f;
    float im = is_dyn ? inv_mass : 0.0f;
    w.pos.push_back(fin3(pose.pos) ? f4(pose.pos.x, pose.pos.y, pose.pos.z) : f4(0, 0, 0));
    w.rot.push_back(fin_q(f4(pose.rot.x, pose.rot.y, pose.rot.z, pose.rot.w)));
    w.lin.push_back(fin3(vel.lin) ? f4(vel.lin.x, vel.lin.y, vel.lin.z) : f4(0, 0, 0));
    w.ang.push_back(fin3(vel.ang) ? f4(vel.ang.x, vel.ang.y, vel.ang.z) : f4(0, 0, 0));
    w.last_lin.push_back(f4(vel.lin.x, vel.lin.y, vel.lin.z));
    w.inv_mass.push_back(im);
    w.bflags.push_back((is_dyn ? BF_DYN : 0u) | BF_ALIVE | BF_SLEEP_OK);
    float m9[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    if (is_dyn) {
        bool inv_fin = hfin(inv.c0.x) && hfin(inv.c0.y) && hfin(inv.c0.z) && hfin(inv.c1.x) && hfin(inv.c1.y) && hfin(inv.c1.z) &&
                       hfin(inv.c2.x) && hfin(inv.c2.y) && hfin(inv.c2.z);
        if (have_ii && inv_fin) { float t[9] = {inv.c0.x, inv.c0.y, inv.c0.z, inv.c1.x, inv.c1.y, inv.c1.z, inv.c2.x, inv.c2.y, inv.c2.z}; std::memcpy(m9, t, sizeof m9); }
        else { m9[0] = im; m9[4] = im; m9[8] = im; }
    }
    w.iil.insert(w.iil.end(), m9, m9 + 9);
    w.awake.push_back(1); w.sticks.push_back(0);
    w.btype.push_back(shape);
    w.dims.push_back(f4(dims[0], shape == RPHYS_SHAPE_SPHERE ? 0.0f : dims[1], shape == RPHYS_SHAPE_BOX ? dims[2] : 0.0f));
    w.col_body.push_back(w.nb() - 1);
    w.col_mat.push_back(0);  // MaterialId::Default (lib.rs:543-557 keeps the id)
    w.bodies_created++;
}

extern "C" {

uint32_t rphys_get_abi_version(void) { return RPHYS_ABI_VERSION; }
const char *rphys_get_build_id(void) { return "riftphys-mi355x gfx950 step library"; }
const char *rphys_last_error_message(void) { return g_last_error.empty() ? nullptr : g_last_error.c_str(); }

RPhysResult rphys_last_error_copy(char *dst, uint32_t cap, uint32_t *out_len) {
    uint32_t len = (uint32_t)g_last_error.size();
    if (out_len) *out_len = len;
    if (cap == 0) return RPHYS_OK;
    if (!dst) return RPHYS_ERR_NULL;
    uint32_t n = std::min(len, cap - 1);
    std::memcpy(dst, g_last_error.data(), n);
    dst[n] = 0;
    return RPHYS_OK;
}

RPhysWorld *rphys_world_create(uint32_t max_bodies, uint32_t max_colliders) {
    clear_err();
    try {
        RPhysWorld *W = new RPhysWorld();
        W->w.max_bodies = max_bodies;
        W->w.max_colliders = max_colliders;
        return W;
    } catch (...) {
        set_err("panic in rphys_world_create");
        return nullptr;
    }
}
void rphys_world_destroy(RPhysWorld *world) { delete world; }

RPhysResult rphys_world_set_epoch(RPhysWorld *world, uint64_t epoch) {
    GUARD_BEGIN
    if (!world) { set_err("rphys_world_set_epoch: world is null"); return RPHYS_ERR_NULL; }
    std::lock_guard<std::mutex> lk(world->mu);
    if (!pull(world)) return RPHYS_ERR_INTERNAL;
    world->w.epoch = epoch;
    host_changed(world);
    return RPHYS_OK;
    GUARD_END
}
RPhysResult rphys_world_set_rng_seed(RPhysWorld *world, uint64_t seed) {
    GUARD_BEGIN
    if (!world) { set_err("rphys_world_set_rng_seed: world is null"); return RPHYS_ERR_NULL; }
    std::lock_guard<std::mutex> lk(world->mu);
    if (!pull(world)) return RPHYS_ERR_INTERNAL;
    world->w.rng = seed | 1ull;
    host_changed(world);
    return RPHYS_OK;
    GUARD_END
}

static RPhysResult add_common(RPhysWorld *world, const char *ctx, uint32_t shape, const float dims[3], const RPhysIsometry &pose,
                              const RPhysVelocity &vel, RPhysBodyType type, float mass, bool check_mass, float density, uint32_t *out) {
    if (!fin_iso(pose)) return bad_arg(ctx, "pose contains NaN/Inf");
    if (!fin_vel(vel)) return bad_arg(ctx, "vel contains NaN/Inf");
    if (type != RPHYS_BODY_STATIC && type != RPHYS_BODY_DYNAMIC) return bad_arg(ctx, "invalid body_type");
    if (check_mass && type == RPHYS_BODY_DYNAMIC && !(hfin(mass) && mass > 0.0f)) return bad_arg(ctx, "dynamic body mass must be finite and > 0");
    std::lock_guard<std::mutex> lk(world->mu);
    if (world->w.bodies_created + 1 > world->w.max_bodies || world->w.nc() + 1 > world->w.max_colliders) {
        set_err(std::string(ctx) + ": body capacity exceeded");
        return RPHYS_ERR_CAPACITY;
    }
    if (!pull(world)) return RPHYS_ERR_INTERNAL;
    add_body_host(world->w, shape, dims, pose, vel, type == RPHYS_BODY_DYNAMIC, density);
    host_changed(world);
    *out = world->w.bodies_created;  // 1-based
    return RPHYS_OK;
}

RPhysResult rphys_add_capsule_body(RPhysWorld *world, const RPhysCapsuleBodyDesc *desc, uint32_t *out_body_id) {
    GUARD_BEGIN
    if (out_body_id) *out_body_id = RPHYS_INVALID_BODY_ID;
    if (!world || !desc || !out_body_id) { set_err("rphys_add_capsule_body: null arg"); return RPHYS_ERR_NULL; }
    const auto &d = *desc;
    if (!fin_mat(d.material)) return bad_arg("rphys_add_capsule_body", "material contains NaN/Inf");
    if (!(hfin(d.radius) && d.radius > 0.0f)) return bad_arg("rphys_add_capsule_body", "radius must be finite and > 0");
    if (!(hfin(d.half_height) && d.half_height >= 0.0f)) return bad_arg("rphys_add_capsule_body", "half_height must be finite and >= 0");
    float dims[3] = {d.radius, d.half_height, 0.0f};
    return add_common(world, "rphys_add_capsule_body", RPHYS_SHAPE_CAPSULE, dims, d.pose, d.vel, d.body_type, d.mass, true, 0.0f, out_body_id);
    GUARD_END
}
RPhysResult rphys_add_box_body(RPhysWorld *world, const RPhysBoxBodyDesc *desc, uint32_t *out_body_id) {
    GUARD_BEGIN
    if (out_body_id) *out_body_id = RPHYS_INVALID_BODY_ID;
    if (!world || !desc || !out_body_id) { set_err("rphys_add_box_body: null arg"); return RPHYS_ERR_NULL; }
    const auto &d = *desc;
    if (!fin_mat(d.material)) return bad_arg("rphys_add_box_body", "material contains NaN/Inf");
    if (!(hfin(d.hx) && d.hx > 0.0f)) return bad_arg("rphys_add_box_body", "hx must be finite and > 0");
    if (!(hfin(d.hy) && d.hy > 0.0f)) return bad_arg("rphys_add_box_body", "hy must be finite and > 0");
    if (!(hfin(d.hz) && d.hz > 0.0f)) return bad_arg("rphys_add_box_body", "hz must be finite and > 0");
    float dims[3] = {d.hx, d.hy, d.hz};
    return add_common(world, "rphys_add_box_body", RPHYS_SHAPE_BOX, dims, d.pose, d.vel, d.body_type, d.mass, true, 0.0f, out_body_id);
    GUARD_END
}
RPhysResult rphys_add_sphere_body(RPhysWorld *world, const RPhysSphereBodyDesc *desc, uint32_t *out_body_id) {
    GUARD_BEGIN
    if (out_body_id) *out_body_id = RPHYS_INVALID_BODY_ID;
    if (!world || !desc || !out_body_id) { set_err("rphys_add_sphere_body: null arg"); return RPHYS_ERR_NULL; }
    const auto &d = *desc;
    if (!fin_mat(d.material)) return bad_arg("rphys_add_sphere_body", "material contains NaN/Inf");
    if (!(hfin(d.radius) && d.radius > 0.0f)) return bad_arg("rphys_add_sphere_body", "radius must be finite and > 0");
    float dims[3] = {d.radius, 0.0f, 0.0f};
    return add_common(world, "rphys_add_sphere_body", RPHYS_SHAPE_SPHERE, dims, d.pose, d.vel, d.body_type, d.mass, true, 0.0f, out_body_id);
    GUARD_END
}
RPhysResult rphys_add_body_density(RPhysWorld *world, RPhysShape shape, const float dims[3], const RPhysIsometry *pose, const RPhysVelocity *vel,
                                   RPhysBodyType body_type, float density, uint32_t *out_body_id) {
    GUARD_BEGIN
    if (out_body_id) *out_body_id = RPHYS_INVALID_BODY_ID;
    if (!world || !dims || !pose || !vel || !out_body_id) { set_err("rphys_add_body_density: null arg"); return RPHYS_ERR_NULL; }
    if (shape > RPHYS_SHAPE_CAPSULE) return bad_arg("rphys_add_body_density", "unknown shape");
    if (!hfin(density)) return bad_arg("rphys_add_body_density", "density must be finite");
    for (int i = 0; i < (shape == RPHYS_SHAPE_BOX ? 3 : (shape == RPHYS_SHAPE_CAPSULE ? 2 : 1)); i++) {
        bool okd = hfin(dims[i]) && (dims[i] > 0.0f || (shape == RPHYS_SHAPE_CAPSULE && i == 1 && dims[i] >= 0.0f));
        if (!okd) return bad_arg("rphys_add_body_density", "dimensions must be finite and > 0");
    }
    return add_common(world, "rphys_add_body_density", shape, dims, *pose, *vel, body_type, 1.0f, false, density, out_body_id);
    GUARD_END
}

RPhysResult rphys_body_remove(RPhysWorld *world, uint32_t body_id) {
    GUARD_BEGIN
    if (!world) { set_err("rphys_body_remove: world is null"); return RPHYS_ERR_NULL; }
    std::lock_guard<std::mutex> lk(world->mu);
    uint32_t b;
    RPhysResult r = live_id(world, body_id, "rphys_body_remove", b);
    if (r != RPHYS_OK) return r;
    if (!pull(world)) return RPHYS_ERR_INTERNAL;
    HostWorld &w = world->w;
    // world.rs:696-713: dead, deactivated (dynamics.rs:211-219), collider gone, warm entries dropped
    w.bflags[b] &= ~(BF_ALIVE | BF_DYN);
    w.lin[b] = f4(0, 0, 0); w.ang[b] = f4(0, 0, 0); w.inv_mass[b] = 0.0f;
    for (int k = 0; k < 9; k++) w.iil[9 * b + k] = 0.0f;
    size_t k = 0;
    for (size_t c = 0; c < w.col_body.size(); c++) {
        if (w.col_body[c] == b) continue;
        w.col_body[k] = w.col_body[c]; w.col_mat[k] = w.col_mat[c]; k++;
    }
    w.col_body.resize(k); w.col_mat.resize(k);
    drop_warm_of(w, b);
    host_changed(world);
    return RPHYS_OK;
    GUARD_END
}

RPhysResult rphys_body_get_pose(const RPhysWorld *cworld, uint32_t body_id, RPhysIsometry *out_pose) {
    GUARD_BEGIN
    if (!out_pose) { set_err("rphys_body_get_pose: out_pose is null"); return RPHYS_ERR_NULL; }
    out_pose->pos = {0, 0, 0}; out_pose->rot = {0, 0, 0, 1};
    if (!cworld) { set_err("rphys_body_get_pose: world is null"); return RPHYS_ERR_NULL; }
    RPhysWorld *world = const_cast<RPhysWorld *>(cworld);
    std::lock_guard<std::mutex> lk(world->mu);
    uint32_t b;
    RPhysResult r = live_id(world, body_id, "rphys_body_get_pose", b);
    if (r != RPHYS_OK) return r;
    if (!pull(world)) return RPHYS_ERR_INTERNAL;
    F4 p = world->w.pos[b], q = world->w.rot[b];
    out_pose->pos = {p.x, p.y, p.z}; out_pose->rot = {q.x, q.y, q.z, q.w};
    return RPHYS_OK;
    GUARD_END
}
RPhysResult rphys_body_set_pose(RPhysWorld *world, uint32_t body_id, const RPhysIsometry *pose) {
    GUARD_BEGIN
    if (!world || !pose) { set_err("rphys_body_set_pose: null arg"); return RPHYS_ERR_NULL; }
    if (!fin_iso(*pose)) return bad_arg("rphys_body_set_pose", "pose contains NaN/Inf");
    std::lock_guard<std::mutex> lk(world->mu);
    uint32_t b;
    RPhysResult r = live_id(world, body_id, "rphys_body_set_pose", b);
    if (r != RPHYS_OK) return r;
    if (!pull(world)) return RPHYS_ERR_INTERNAL;
    drop_warm_of(world->w, b);  // world.rs:501-503
    world->w.pos[b] = f4(pose->pos.x, pose->pos.y, pose->pos.z);
    world->w.rot[b] = fin_q(f4(pose->rot.x, pose->rot.y, pose->rot.z, pose->rot.w));
    host_changed(world);
    return RPHYS_OK;
    GUARD_END
}
RPhysResult rphys_body_get_velocity(const RPhysWorld *cworld, uint32_t body_id, RPhysVelocity *out_vel) {
    GUARD_BEGIN
    if (!out_vel) { set_err("rphys_body_get_velocity: out_vel is null"); return RPHYS_ERR_NULL; }
    out_vel->lin = {0, 0, 0}; out_vel->ang = {0, 0, 0};
    if (!cworld) { set_err("rphys_body_get_velocity: world is null"); return RPHYS_ERR_NULL; }
    RPhysWorld *world = const_cast<RPhysWorld *>(cworld);
    std::lock_guard<std::mutex> lk(world->mu);
    uint32_t b;
    RPhysResult r = live_id(world, body_id, "rphys_body_get_velocity", b);
    if (r != RPHYS_OK) return r;
    if (!pull(world)) return RPHYS_ERR_INTERNAL;
    F4 l = world->w.lin[b], a = world->w.ang[b];
    out_vel->lin = {l.x, l.y, l.z}; out_vel->ang = {a.x, a.y, a.z};
    return RPHYS_OK;
    GUARD_END
}
RPhysResult rphys_body_set_velocity(RPhysWorld *world, uint32_t body_id, const RPhysVelocity *vel) {
    GUARD_BEGIN
    if (!world || !vel) { set_err("rphys_body_set_velocity: null arg"); return RPHYS_ERR_NULL; }
    if (!fin_vel(*vel)) return bad_arg("rphys_body_set_velocity", "velocity contains NaN/Inf");
    std::lock_guard<std::mutex> lk(world->mu);
    uint32_t b;
    RPhysResult r = live_id(world, body_id, "rphys_body_set_velocity", b);
    if (r != RPHYS_OK) return r;
    if (!pull(world)) return RPHYS_ERR_INTERNAL;
    // world.rs:674-676 (+ dynamics.rs:143-147)
    world->w.lin[b] = f4(vel->lin.x, vel->lin.y, vel->lin.z);
    world->w.ang[b] = f4(vel->ang.x, vel->ang.y, vel->ang.z);
    world->w.last_lin[b] = f4(vel->lin.x, vel->lin.y, vel->lin.z);
    host_changed(world);
    return RPHYS_OK;
    GUARD_END
}
RPhysResult rphys_body_apply_impulse(RPhysWorld *world, uint32_t body_id, RPhysVec3 impulse) {
    GUARD_BEGIN
    if (!world) { set_err("rphys_body_apply_impulse: world is null"); return RPHYS_ERR_NULL; }
    if (!fin3(impulse)) return bad_arg("rphys_body_apply_impulse", "impulse contains NaN/Inf");
    std::lock_guard<std::mutex> lk(world->mu);
    uint32_t b;
    RPhysResult r = live_id(world, body_id, "rphys_body_apply_impulse", b);
    if (r != RPHYS_OK) return r;
    if (!pull(world)) return RPHYS_ERR_INTERNAL;
    // world.rs:689-692: wake, then v += j * inv_mass
    HostWorld &w = world->w;
    w.awake[b] = 1; w.sticks[b] = 0;
    float im = w.inv_mass[b];
    if (im > 0.0f) {
        F4 l = w.lin[b];
        float x = l.x + impulse.x * im, y = l.y + impulse.y * im, z = l.z + impulse.z * im;
        w.lin[b] = f4(hfin(x) ? x : 0.0f, hfin(y) ? y : 0.0f, hfin(z) ? z : 0.0f);
    }
    host_changed(world);
    return RPHYS_OK;
    GUARD_END
}
RPhysResult rphys_body_set_sleep_allowed(RPhysWorld *world, uint32_t body_id, RPhysBool allowed) {
    GUARD_BEGIN
    if (!world) { set_err("rphys_body_set_sleep_allowed: world is null"); return RPHYS_ERR_NULL; }
    std::lock_guard<std::mutex> lk(world->mu);
    uint32_t b;
    RPhysResult r = live_id(world, body_id, "rphys_body_set_sleep_allowed", b);
    if (r != RPHYS_OK) return r;
    if (!pull(world)) return RPHYS_ERR_INTERNAL;
    if (allowed) world->w.bflags[b] |= BF_SLEEP_OK;
    else { world->w.bflags[b] &= ~BF_SLEEP_OK; world->w.awake[b] = 1; world->w.sticks[b] = 0; }
    host_changed(world);
    return RPHYS_OK;
    GUARD_END
}

RPhysResult rphys_body_set_material_id(RPhysWorld *world, uint32_t body_id, uint32_t material_id) {
    GUARD_BEGIN
    if (!world) { set_err("rphys_body_set_material_id: world is null"); return RPHYS_ERR_NULL; }
    if (material_id > 15) return bad_arg("rphys_body_set_material_id", "material ids above 15 are not reachable through this ABI");
    std::lock_guard<std::mutex> lk(world->mu);
    uint32_t b;
    RPhysResult r = live_id(world, body_id, "rphys_body_set_material_id", b);
    if (r != RPHYS_OK) return r;
    if (!pull(world)) return RPHYS_ERR_INTERNAL;
    HostWorld &w = world->w;
    for (uint32_t c = 0; c < w.nc(); c++) if (w.col_body[c] == b) w.col_mat[c] = material_id;
    host_changed(world);
    return RPHYS_OK;
    GUARD_END
}

static RPhysResult step_impl(RPhysWorld *world, float dt, uint32_t substeps, uint32_t iv_override, uint32_t repeat, bool graph) {
    std::lock_guard<std::mutex> lk(world->mu);
    float sub = dt / (float)substeps;
    if (!hfin(sub)) { set_err("rphys_world_step_ex: sub_dt is not finite"); return RPHYS_ERR_BAD_ARG; }
    if (world->w.nb() == 0) { world->w.tick += (uint64_t)substeps * repeat; return RPHYS_OK; }
    if (world->need_build) { if (!pull(world) || !push(world)) return RPHYS_ERR_INTERNAL; }
    std::vector<HostWorld *> ws{&world->w};
    uint32_t iv = iv_override ? iv_override : world->w.iv;
    bool ok = run_ticks(world->scene, ws, sub, iv, substeps * repeat, graph);
    world->dev_newer = true;
    world->have_contacts = ok;
    return ok ? RPHYS_OK : RPHYS_ERR_INTERNAL;
}

RPhysResult rphys_world_step_ex(RPhysWorld *world, const RPhysStepParams *p) {
    GUARD_BEGIN
    if (!world || !p) { set_err("rphys_world_step_ex: null arg"); return RPHYS_ERR_NULL; }
    if (p->size_bytes < sizeof(RPhysStepParams)) { set_err("rphys_world_step_ex: size_bytes too small"); return RPHYS_ERR_BAD_ARG; }
    if (!hfin(p->dt)) { set_err("rphys_world_step_ex: dt is not finite"); return RPHYS_ERR_BAD_ARG; }
    if (p->dt < 0.0f) { set_err("rphys_world_step_ex: dt is negative"); return RPHYS_ERR_BAD_ARG; }
    if (p->substeps == 0) { set_err("rphys_world_step_ex: substeps == 0"); return RPHYS_ERR_BAD_ARG; }
    uint32_t iv = ((p->flags & RPHYS_STEP_FLAG_SOLVER_ITERS) && p->solver_iters > 0) ? p->solver_iters : 0;
    return step_impl(world, p->dt, p->substeps, iv, 1, false);
    GUARD_END
}
RPhysResult rphys_world_step(RPhysWorld *world, float dt) {
    RPhysStepParams p{(uint32_t)sizeof(RPhysStepParams), 0, 1, 0, dt};
    return rphys_world_step_ex(world, &p);
}
RPhysResult rphys_world_step_n(RPhysWorld *world, float dt, uint32_t n_ticks) {
    GUARD_BEGIN
    if (!world) { set_err("rphys_world_step_n: world is null"); return RPHYS_ERR_NULL; }
    if (!hfin(dt) || dt < 0.0f) { set_err("rphys_world_step_n: dt must be finite and >= 0"); return RPHYS_ERR_BAD_ARG; }
    if (n_ticks == 0) return RPHYS_OK;
    return step_impl(world, dt, 1, 0, n_ticks, n_ticks > 1);
    GUARD_END
}

// Profiled ticks of a scene (one world or a batch): stage events, counter sums, sweep phases.
static bool profiled_ticks(Scene *&S, std::vector<HostWorld *> &ws, float dt, uint32_t iv, uint32_t n_ticks, RPhysProfile *out) {
    ProfileAcc acc;
    acc.out = out;
    if (!acc.init()) return false;
    if (!hip_ok(hipMemsetAsync(S->d.ph, 0, RP_PH_COUNT * sizeof(unsigned long long), S->st), "memset")) return false;
    bool ok = run_ticks(S, ws, dt, iv, n_ticks, false, &acc);
    if (ok) {  // sweep-phase clocks (a capacity rebuild during the call restarts them)
        unsigned long long ph[RP_PH_COUNT];
        int dev = 0, khz = 0;
        ok = hip_ok(hipMemcpy(ph, S->d.ph, sizeof ph, hipMemcpyDeviceToHost), "phase clocks") && hip_ok(hipGetDevice(&dev), "hipGetDevice") &&
             hip_ok(hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, dev), "wall clock rate");
        if (ok) {
            static_assert(RP_PH_COUNT == 11, "RPhysProfile.sweep_phase holds 11 entries");
            for (int k = 0; k < RP_PH_COUNT; k++)
                out->sweep_phase[k] = k < RP_PH_ROUNDS_BASE ? (khz > 0 ? (double)ph[k] / (double)khz : 0.0) : (double)ph[k];
        }
    }
    return ok;
}

RPhysResult rphys_world_step_profiled(RPhysWorld *world, float dt, uint32_t n_ticks, RPhysProfile *out) {
    GUARD_BEGIN
    if (!world || !out) { set_err("rphys_world_step_profiled: null arg"); return RPHYS_ERR_NULL; }
    std::memset(out, 0, sizeof(*out));
    out->n_stages = RP_NMARKS - 1;
    if (!hfin(dt) || dt < 0.0f) { set_err("rphys_world_step_profiled: dt must be finite and >= 0"); return RPHYS_ERR_BAD_ARG; }
    std::lock_guard<std::mutex> lk(world->mu);
    if (n_ticks == 0) return RPHYS_OK;
    if (world->w.nb() == 0) { world->w.tick += n_ticks; return RPHYS_OK; }
    if (world->need_build) { if (!pull(world) || !push(world)) return RPHYS_ERR_INTERNAL; }
    std::vector<HostWorld *> ws{&world->w};
    bool ok = profiled_ticks(world->scene, ws, dt, world->w.iv, n_ticks, out);
    world->dev_newer = true;
    world->have_contacts = ok;
    return ok ? RPHYS_OK : RPHYS_ERR_INTERNAL;
    GUARD_END
}

RPhysResult rphys_debug_copy_probe(uint64_t bytes, uint32_t reps, double *gbps) {
    GUARD_BEGIN
    if (!gbps) { set_err("rphys_debug_copy_probe: null arg"); return RPHYS_ERR_NULL; }
    *gbps = 0.0;
    if (bytes < 16 || reps == 0) { set_err("rphys_debug_copy_probe: bytes >= 16 and reps > 0 required"); return RPHYS_ERR_BAD_ARG; }
    const size_t n = (size_t)(bytes / 16);
    DeviceBuffers B;
    float4 *a = nullptr, *b = nullptr;
    if (!B.alloc(a, n) || !B.alloc(b, n)) { set_err("rphys_debug_copy_probe: device allocation failed"); return RPHYS_ERR_CAPACITY; }
    hipStream_t st = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    bool ok = hip_ok(hipStreamCreateWithFlags(&st, hipStreamNonBlocking), "hipStreamCreate") && hip_ok(hipEventCreate(&e0), "hipEventCreate") &&
              hip_ok(hipEventCreate(&e1), "hipEventCreate") && hip_ok(hipMemsetAsync(a, 0, n * 16, st), "memset");
    if (ok) {
        rp_launch_copy_probe(a, b, n, st);  // warm
        ok = hip_ok(hipEventRecord(e0, st), "hipEventRecord");
        for (uint32_t r = 0; ok && r < reps; r++) rp_launch_copy_probe(a, b, n, st);
        float ms = 0.0f;
        ok = ok && hip_ok(hipEventRecord(e1, st), "hipEventRecord") && hip_ok(hipEventSynchronize(e1), "hipEventSynchronize") &&
             hip_ok(hipEventElapsedTime(&ms, e0, e1), "hipEventElapsedTime");
        if (ok && ms > 0.0f) *gbps = 2.0 * (double)n * 16.0 * reps / (ms * 1e-3) / 1e9;
    }
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    if (st) (void)hipStreamDestroy(st);
    return ok ? RPHYS_OK : RPHYS_ERR_INTERNAL;
    GUARD_END
}

const char *rphys_profile_stage_name(uint32_t stage) {
    static const char *names[RP_NMARKS - 1] = {"pre", "broadphase", "narrow_pairs", "narrow_terrain", "contacts", "prep_islands",
                                               "orders", "vel_solve_small", "vel_solve_large", "vel_solve_huge", "warm_store",
                                               "ballistic", "integrate", "pos_prep", "projection_small", "projection_large",
                                               "projection_huge", "projection_order", "projection_shock", "s6_sleep"};
    return stage < RP_NMARKS - 1 ? names[stage] : nullptr;
}

RPhysResult rphys_world_tick_index(const RPhysWorld *cworld, uint64_t *out_tick) {
    GUARD_BEGIN
    if (!out_tick) { set_err("rphys_world_tick_index: out_tick is null"); return RPHYS_ERR_NULL; }
    *out_tick = 0;
    if (!cworld) { set_err("rphys_world_tick_index: world is null"); return RPHYS_ERR_NULL; }
    RPhysWorld *world = const_cast<RPhysWorld *>(cworld);
    std::lock_guard<std::mutex> lk(world->mu);
    *out_tick = world->w.tick;
    return RPHYS_OK;
    GUARD_END
}
RPhysResult rphys_world_num_bodies(const RPhysWorld *cworld, uint32_t *out_num) {
    GUARD_BEGIN
    if (!out_num) { set_err("rphys_world_num_bodies: out_num is null"); return RPHYS_ERR_NULL; }
    *out_num = 0;
    if (!cworld) { set_err("rphys_world_num_bodies: world is null"); return RPHYS_ERR_NULL; }
    RPhysWorld *world = const_cast<RPhysWorld *>(cworld);
    std::lock_guard<std::mutex> lk(world->mu);
    *out_num = world->w.nb();
    return RPHYS_OK;
    GUARD_END
}

static bool device_hash(Scene &S, uint8_t *out, uint32_t nw) {
    rp_launch_hash(S.d, S.h, S.cur, S.max_extra, S.st);
    return hip_ok(hipMemcpyAsync(out, S.h.digest, 32 * (size_t)nw, hipMemcpyDeviceToHost, S.st), "hash copy") &&
           hip_ok(hipStreamSynchronize(S.st), "hash");
}

RPhysResult rphys_world_get_step_hash(const RPhysWorld *cworld, RPhysHash32 *out_hash) {
    GUARD_BEGIN
    if (!out_hash) { set_err("rphys_world_get_step_hash: out_hash is null"); return RPHYS_ERR_NULL; }
    std::memset(out_hash, 0, sizeof(*out_hash));
    if (!cworld) { set_err("rphys_world_get_step_hash: world is null"); return RPHYS_ERR_NULL; }
    RPhysWorld *world = const_cast<RPhysWorld *>(cworld);
    std::lock_guard<std::mutex> lk(world->mu);
    if (world->need_build) { if (!pull(world) || !push(world)) return RPHYS_ERR_INTERNAL; }
    if (world->w.nb() == 0 && !world->scene) {
        // an empty world still hashes its header and sleep slots: build a scene for it
        if (!push(world)) return RPHYS_ERR_INTERNAL;
    }
    if (!device_hash(*world->scene, out_hash->bytes, 1)) return RPHYS_ERR_INTERNAL;
    return RPHYS_OK;
    GUARD_END
}
RPhysResult rphys_debug_get_hash_input(const RPhysWorld *cworld, uint8_t *out, uint64_t cap, uint64_t *out_len) {
    GUARD_BEGIN
    if (!cworld || !out_len) { set_err("rphys_debug_get_hash_input: null arg"); return RPHYS_ERR_NULL; }
    *out_len = 0;
    RPhysWorld *world = const_cast<RPhysWorld *>(cworld);
    std::lock_guard<std::mutex> lk(world->mu);
    if (world->need_build) { if (!pull(world) || !push(world)) return RPHYS_ERR_INTERNAL; }
    if (!world->scene && !push(world)) return RPHYS_ERR_INTERNAL;
    Scene &S = *world->scene;
    rp_launch_hash(S.d, S.h, S.cur, S.max_extra, S.st);
    uint32_t chunk = 0, len = 0;
    if (!hip_ok(hipMemcpyAsync(&chunk, S.h.w_chunk, 4, hipMemcpyDeviceToHost, S.st), "hash input") ||
        !hip_ok(hipMemcpyAsync(&len, S.h.w_len, 4, hipMemcpyDeviceToHost, S.st), "hash input") || !hip_ok(hipStreamSynchronize(S.st), "hash input"))
        return RPHYS_ERR_INTERNAL;
    *out_len = len;
    if (out && cap && len) {
        const size_t n = std::min<uint64_t>(cap, len);
        if (!hip_ok(hipMemcpy(out, S.h.stream + (size_t)chunk * 1024u, n, hipMemcpyDeviceToHost), "hash input")) return RPHYS_ERR_INTERNAL;
    }
    return RPHYS_OK;
    GUARD_END
}
RPhysResult rphys_world_step_and_hash(RPhysWorld *world, float dt, RPhysHash32 *out_hash) {
    GUARD_BEGIN
    if (!out_hash) { set_err("rphys_world_step_and_hash: out_hash is null"); return RPHYS_ERR_NULL; }
    std::memset(out_hash, 0, sizeof(*out_hash));
    RPhysResult r = rphys_world_step(world, dt);
    if (r != RPHYS_OK) return r;
    return rphys_world_get_step_hash(world, out_hash);
    GUARD_END
}

// ---- spatial queries (ffi/src/lib.rs:1268-1410 over world/src/world/queries.rs): the device
// answers them from the world's current state (k_query, one workgroup per query)
static void ray_miss(RPhysRayHit *h) {  // lib.rs:470-479
    h->hit = RPHYS_FALSE; h->body_id = RPHYS_INVALID_BODY_ID; h->fraction = 1.0f;
    h->point = {0.0f, 0.0f, 0.0f}; h->normal = {0.0f, 1.0f, 0.0f};
}
static void sweep_miss(RPhysCapsuleSweepHit *h) {  // lib.rs:482-492
    h->hit = RPHYS_FALSE; h->body_id = RPHYS_INVALID_BODY_ID; h->fraction = 1.0f; h->started_overlapping = RPHYS_FALSE;
    h->point = {0.0f, 0.0f, 0.0f}; h->normal = {0.0f, 1.0f, 0.0f};
}
static bool fin3h(RPhysVec3 v) { return hfin(v.x) && hfin(v.y) && hfin(v.z); }
// checked_optional_live_body_id (lib.rs:412-421): 0 is none, else a live body
static RPhysResult ignore_index(RPhysWorld *world, uint32_t id, const char *ctx, uint32_t &out) {
    out = RP_NONE;
    if (id == 0) return RPHYS_OK;
    uint32_t b;
    RPhysResult r = live_id(world, id, ctx, b);
    if (r != RPHYS_OK) return r;
    out = b;
    return RPHYS_OK;
}
// Runs n queries on the device against the world's current state.
static bool run_queries(RPhysWorld *world, std::vector<RpQuery> &qs, std::vector<RpQueryHit> &hits) {
    if (world->need_build || !world->scene) { if (!pull(world) || !push(world)) return false; }
    Scene &S = *world->scene;
    const uint32_t n = (uint32_t)qs.size();
    if (n > S.q_cap) {
        uint32_t cap = std::max<uint32_t>(n, 64u);
        if (!S.buf.alloc(S.q_buf, cap) || !S.buf.alloc(S.h_buf, cap)) { set_err("query buffers: device allocation failed"); return false; }
        S.q_cap = cap;
    }
    hits.resize(n);
    return hip_ok(hipMemcpyAsync(S.q_buf, qs.data(), n * sizeof(RpQuery), hipMemcpyHostToDevice, S.st), "query upload") &&
           hip_ok(rp_launch_query(S.d, S.q_buf, S.h_buf, n, S.st), "query launch") &&
           hip_ok(hipMemcpyAsync(hits.data(), S.h_buf, n * sizeof(RpQueryHit), hipMemcpyDeviceToHost, S.st), "query download") &&
           hip_ok(hipStreamSynchronize(S.st), "query sync");
}
static uint32_t hit_body_id(const RpQueryHit &h) { return h.body == RP_TERRAIN ? RPHYS_TERRAIN_BODY_ID : h.body + 1; }

// The argument checks of rphys_world_raycast (lib.rs:1277-1323); returns OK with skip set when the
// query misses by definition (max_distance <= 0 or a zero direction).
static RPhysResult ray_query(RPhysWorld *world, const RPhysRayQuery &q, RpQuery &out, bool &skip) {
    skip = false;
    if (!fin3h(q.origin)) return bad_arg("rphys_world_raycast", "origin contains NaN/Inf");
    if (!fin3h(q.dir)) return bad_arg("rphys_world_raycast", "dir contains NaN/Inf");
    if (!hfin(q.max_distance)) return bad_arg("rphys_world_raycast", "max_distance is not finite");
    uint32_t ig;
    RPhysResult r = ignore_index(world, q.ignore_body, "rphys_world_raycast.ignore_body", ig);
    if (r != RPHYS_OK) return r;
    if (q.max_distance <= 0.0f) { skip = true; return RPHYS_OK; }
    const float len2 = (q.dir.x * q.dir.x + q.dir.y * q.dir.y) + q.dir.z * q.dir.z;
    if (!hfin(len2)) return bad_arg("rphys_world_raycast", "dir length overflow/invalid");
    if (len2 <= 1.0e-20f) { skip = true; return RPHYS_OK; }
    const float l = std::sqrt(len2);
    float d[3] = {q.dir.x / l, q.dir.y / l, q.dir.z / l};
    // World::raycast normalises again (queries.rs:55-57)
    const float len = std::sqrt((d[0] * d[0] + d[1] * d[1]) + d[2] * d[2]);
    if (!hfin(len) || len < 1.0e-12f) { skip = true; return RPHYS_OK; }
    out = RpQuery{};
    out.kind = 0; out.world = 0; out.ignore = ig;
    out.a[0] = q.origin.x; out.a[1] = q.origin.y; out.a[2] = q.origin.z;
    out.b[0] = d[0] / len; out.b[1] = d[1] / len; out.b[2] = d[2] / len;
    out.max_d = q.max_distance;
    return RPHYS_OK;
}
static RPhysResult sweep_query(RPhysWorld *world, const RPhysCapsuleSweepQuery &q, RpQuery &out) {
    if (!fin3h(q.from)) return bad_arg("rphys_world_capsule_sweep", "from contains NaN/Inf");
    if (!fin3h(q.to)) return bad_arg("rphys_world_capsule_sweep", "to contains NaN/Inf");
    if (!(hfin(q.radius) && q.radius > 0.0f)) return bad_arg("rphys_world_capsule_sweep", "radius must be finite and > 0");
    if (!(hfin(q.half_height) && q.half_height >= 0.0f)) return bad_arg("rphys_world_capsule_sweep", "half_height must be finite and >= 0");
    uint32_t ig;
    RPhysResult r = ignore_index(world, q.ignore_body, "rphys_world_capsule_sweep.ignore_body", ig);
    if (r != RPHYS_OK) return r;
    out = RpQuery{};
    out.kind = 1; out.world = 0; out.ignore = ig;
    out.a[0] = q.from.x; out.a[1] = q.from.y; out.a[2] = q.from.z;
    out.b[0] = q.to.x; out.b[1] = q.to.y; out.b[2] = q.to.z;
    out.radius = q.radius; out.half_height = q.half_height;
    return RPHYS_OK;
}

RPhysResult rphys_world_raycast_n(const RPhysWorld *cworld, const RPhysRayQuery *queries, uint32_t n, RPhysRayHit *out_hits) {
    GUARD_BEGIN
    if (out_hits) for (uint32_t i = 0; i < n; i++) ray_miss(&out_hits[i]);
    if (!cworld || (n && (!queries || !out_hits))) { set_err("rphys_world_raycast: null arg"); return RPHYS_ERR_NULL; }
    RPhysWorld *world = const_cast<RPhysWorld *>(cworld);
    std::lock_guard<std::mutex> lk(world->mu);
    std::vector<RpQuery> qs;
    std::vector<uint32_t> slot;
    for (uint32_t i = 0; i < n; i++) {
        RpQuery q;
        bool skip;
        RPhysResult r = ray_query(world, queries[i], q, skip);
        if (r != RPHYS_OK) return r;
        if (!skip) { qs.push_back(q); slot.push_back(i); }
    }
    if (qs.empty() || (world->w.nb() == 0 && !world->w.has_hf)) return RPHYS_OK;
    std::vector<RpQueryHit> hits;
    if (!run_queries(world, qs, hits)) return RPHYS_ERR_INTERNAL;
    for (size_t k = 0; k < qs.size(); k++) {
        const RpQueryHit &h = hits[k];
        if (!h.hit) continue;
        RPhysRayHit &o = out_hits[slot[k]];
        o.hit = RPHYS_TRUE; o.body_id = hit_body_id(h); o.fraction = std::min(std::max(h.fraction, 0.0f), 1.0f);
        o.point = {h.point[0], h.point[1], h.point[2]}; o.normal = {h.normal[0], h.normal[1], h.normal[2]};
    }
    return RPHYS_OK;
    GUARD_END
}
RPhysResult rphys_world_raycast(const RPhysWorld *world, const RPhysRayQuery *query, RPhysRayHit *out_hit) {
    if (out_hit) ray_miss(out_hit);
    if (!world || !query || !out_hit) { set_err("rphys_world_raycast: null arg"); return RPHYS_ERR_NULL; }
    return rphys_world_raycast_n(world, query, 1, out_hit);
}
RPhysResult rphys_world_capsule_sweep_n(const RPhysWorld *cworld, const RPhysCapsuleSweepQuery *queries, uint32_t n,
                                        RPhysCapsuleSweepHit *out_hits) {
    GUARD_BEGIN
    if (out_hits) for (uint32_t i = 0; i < n; i++) sweep_miss(&out_hits[i]);
    if (!cworld || (n && (!queries || !out_hits))) { set_err("rphys_world_capsule_sweep: null arg"); return RPHYS_ERR_NULL; }
    RPhysWorld *world = const_cast<RPhysWorld *>(cworld);
    std::lock_guard<std::mutex> lk(world->mu);
    std::vector<RpQuery> qs(n);
    for (uint32_t i = 0; i < n; i++) {
        RPhysResult r = sweep_query(world, queries[i], qs[i]);
        if (r != RPHYS_OK) return r;
    }
    if (qs.empty() || (world->w.nb() == 0 && !world->w.has_hf)) return RPHYS_OK;
    std::vector<RpQueryHit> hits;
    if (!run_queries(world, qs, hits)) return RPHYS_ERR_INTERNAL;
    for (uint32_t k = 0; k < n; k++) {
        const RpQueryHit &h = hits[k];
        if (!h.hit) continue;
        RPhysCapsuleSweepHit &o = out_hits[k];
        o.hit = RPHYS_TRUE; o.body_id = hit_body_id(h); o.fraction = std::min(std::max(h.fraction, 0.0f), 1.0f);
        o.started_overlapping = h.started ? RPHYS_TRUE : RPHYS_FALSE;
        o.point = {h.point[0], h.point[1], h.point[2]}; o.normal = {h.normal[0], h.normal[1], h.normal[2]};
    }
    return RPHYS_OK;
    GUARD_END
}
RPhysResult rphys_world_capsule_sweep(const RPhysWorld *world, const RPhysCapsuleSweepQuery *query, RPhysCapsuleSweepHit *out_hit) {
    if (out_hit) sweep_miss(out_hit);
    if (!world || !query || !out_hit) { set_err("rphys_world_capsule_sweep: null arg"); return RPHYS_ERR_NULL; }
    return rphys_world_capsule_sweep_n(world, query, 1, out_hit);
}

RPhysResult rphys_debug_set_sweep_tiers(uint32_t max_tier0, uint32_t max_tier3, uint32_t max_tier1, uint32_t wide_min) {
    g_tier_max[0] = max_tier0; g_tier_max[1] = max_tier3; g_tier_max[2] = max_tier1; g_wide_min = wide_min;
    return RPHYS_OK;
}
RPhysResult rphys_debug_set_lane_tier(uint32_t lane_max, uint32_t pair_max, uint32_t quad_max, uint32_t wave_min) {
    g_lane_max = lane_max;
    g_pair_max = pair_max;
    g_quad_max = quad_max;
    g_wave_min = wave_min;
    return RPHYS_OK;
}
RPhysResult rphys_debug_set_initial_capacity(uint32_t pairs, uint32_t cands, uint32_t contacts, uint32_t warm) {
    g_init_caps[0] = pairs; g_init_caps[1] = cands; g_init_caps[2] = contacts; g_init_caps[3] = warm;
    return RPHYS_OK;
}

RPhysResult rphys_debug_fail_device_allocs(uint32_t n, uint32_t *out_pending) {
    uint32_t prev = g_fail_allocs.exchange(n);
    if (out_pending) *out_pending = prev;
    return RPHYS_OK;
}

// Kinematic players (lib.rs:1580-1693).  add_player (world.rs:466-475): a static body with the
// controller's default capsule (r 0.4, half height 0.9, material Default) consuming one body and
// one collider; the controller then takes the described radius, height and speed.  The device
// moves players at the start of each tick (k_players); input changes go straight to the device.
RPhysResult rphys_add_player(RPhysWorld *world, const RPhysPlayerDesc *desc, uint32_t *out_player_idx) {
    GUARD_BEGIN
    if (out_player_idx) *out_player_idx = 0xFFFFFFFFu;
    if (!world || !desc || !out_player_idx) { set_err("rphys_add_player: null arg"); return RPHYS_ERR_NULL; }
    if (!fin_iso(desc->start_pose)) return bad_arg("rphys_add_player", "pose contains NaN/Inf");
    if (!(hfin(desc->radius) && desc->radius > 0.0f)) return bad_arg("rphys_add_player", "radius must be > 0");
    if (!(hfin(desc->height) && desc->height > 0.0f)) return bad_arg("rphys_add_player", "height must be > 0");
    std::lock_guard<std::mutex> lk(world->mu);
    if (world->w.bodies_created + 1 > world->w.max_bodies || world->w.nc() + 1 > world->w.max_colliders) {
        set_err("rphys_add_player: body capacity exceeded");
        return RPHYS_ERR_CAPACITY;
    }
    if (!pull(world)) return RPHYS_ERR_INTERNAL;
    HostWorld &w = world->w;
    const float dims[3] = {0.4f, 1.8f * 0.5f, 0.0f};
    const RPhysVelocity zero{{0.0f, 0.0f, 0.0f}, {0.0f, 0.0f, 0.0f}};
    add_body_host(w, 2u, dims, desc->start_pose, zero, false, 0.0f);
    HostWorld::Player pl{};
    pl.body = w.nb() - 1; pl.radius = desc->radius; pl.height = desc->height; pl.speed = desc->speed;
    w.players.push_back(pl);
    host_changed(world);
    *out_player_idx = (uint32_t)w.players.size() - 1;
    return RPHYS_OK;
    GUARD_END
}
RPhysResult rphys_player_set_input(RPhysWorld *world, uint32_t player_idx, RPhysVec3 move_dir) {
    GUARD_BEGIN
    if (!world) { set_err("rphys_player_set_input: null arg"); return RPHYS_ERR_NULL; }
    if (!fin3(move_dir)) return bad_arg("rphys_player_set_input", "move_dir NaN/Inf");
    std::lock_guard<std::mutex> lk(world->mu);
    HostWorld &w = world->w;
    if (player_idx >= w.players.size()) return bad_arg("rphys_player_set_input", "player_idx out of bounds");
    HostWorld::Player &pl = w.players[player_idx];
    pl.input[0] = move_dir.x; pl.input[1] = move_dir.y; pl.input[2] = move_dir.z;
    if (world->scene && !world->need_build && world->scene->d.np > player_idx) {
        const F4 in = f4(move_dir.x, move_dir.y, move_dir.z);
        Scene &S = *world->scene;
        if (!hip_ok(hipMemcpyAsync(S.d.pl_input + player_idx, &in, 16, hipMemcpyHostToDevice, S.st), "player input") ||
            !hip_ok(hipStreamSynchronize(S.st), "player input"))
            return RPHYS_ERR_INTERNAL;
    }
    return RPHYS_OK;
    GUARD_END
}
RPhysResult rphys_player_get_pose(const RPhysWorld *cworld, uint32_t player_idx, RPhysIsometry *out_pose, RPhysBool *out_grounded) {
    GUARD_BEGIN
    if (out_pose) *out_pose = RPhysIsometry{{0.0f, 0.0f, 0.0f}, {0.0f, 0.0f, 0.0f, 1.0f}};
    if (out_grounded) *out_grounded = RPHYS_FALSE;
    if (!cworld || !out_pose) { set_err("rphys_player_get_pose: null arg"); return RPHYS_ERR_NULL; }
    RPhysWorld *world = const_cast<RPhysWorld *>(cworld);
    std::lock_guard<std::mutex> lk(world->mu);
    const HostWorld &w = world->w;
    if (player_idx >= w.players.size()) return bad_arg("rphys_player_get_pose", "player_idx out of bounds");
    const uint32_t b = w.players[player_idx].body;
    F4 p = w.pos[b], q = w.rot[b];
    uint32_t gnd = w.players[player_idx].grounded;
    if (world->dev_newer && world->scene) {  // the device is ahead of the host mirror: read this player alone
        Scene &S = *world->scene;
        if (!hip_ok(hipMemcpyAsync(&p, S.d.pos + b, 16, hipMemcpyDeviceToHost, S.st), "player pose") ||
            !hip_ok(hipMemcpyAsync(&q, S.d.rot + b, 16, hipMemcpyDeviceToHost, S.st), "player pose") ||
            !hip_ok(hipMemcpyAsync(&gnd, S.d.pl_grounded + player_idx, 4, hipMemcpyDeviceToHost, S.st), "player pose") ||
            !hip_ok(hipStreamSynchronize(S.st), "player pose"))
            return RPHYS_ERR_INTERNAL;
    }
    *out_pose = RPhysIsometry{{p.x, p.y, p.z}, {q.x, q.y, q.z, q.w}};
    if (out_grounded) *out_grounded = gnd ? RPHYS_TRUE : RPHYS_FALSE;
    return RPHYS_OK;
    GUARD_END
}

// terrain.rs:24-52 from_heights, material Grit (lib.rs:1479,1536)
static void set_hf(HostWorld &w, uint32_t nx, uint32_t nz, float sx, float sz, float ox, float oz, std::vector<float> &&h) {
    w.hnx = nx; w.hnz = nz;
    w.hsx = (hfin(sx) && sx > 0.0f) ? q6div(sx) : 1.0f;
    w.hsz = (hfin(sz) && sz > 0.0f) ? q6div(sz) : 1.0f;
    w.hox = ox; w.hoz = oz;
    float mn = INFINITY, mx = -INFINITY;
    for (float &v : h) {
        float x = hfin(v) ? v : 0.0f;
        v = q6div(x);
        mn = std::fmin(mn, v); mx = std::fmax(mx, v);
    }
    w.hmin = q6div(mn); w.hmax = q6div(mx);
    w.heights = std::move(h);
    w.hmat = 4;
    w.has_hf = true;
}

RPhysResult rphys_world_set_heightfield_i16(RPhysWorld *world, const RPhysHeightfieldDesc *hf, float origin_x, float origin_z, float y_offset) {
    GUARD_BEGIN
    if (!world || !hf) { set_err("rphys_world_set_heightfield_i16: null arg"); return RPHYS_ERR_NULL; }
    const auto &d = *hf;
    if (d.size_bytes < sizeof(RPhysHeightfieldDesc)) { set_err("rphys_world_set_heightfield_i16: size_bytes too small"); return RPHYS_ERR_BAD_ARG; }
    if (d.width == 0 || d.height == 0) { set_err("rphys_world_set_heightfield_i16: width/height == 0"); return RPHYS_ERR_BAD_ARG; }
    if (d.row_stride != 0 && d.row_stride < d.width) return bad_arg("rphys_world_set_heightfield_i16", "row_stride must be >= width");
    if (!d.heights) { set_err("rphys_world_set_heightfield_i16: heights is null"); return RPHYS_ERR_NULL; }
    if (!(hfin(d.cell_size_x) && hfin(d.cell_size_z) && d.cell_size_x > 0.0f && d.cell_size_z > 0.0f)) { set_err("rphys_world_set_heightfield_i16: bad cell sizes"); return RPHYS_ERR_BAD_ARG; }
    uint32_t stride = d.row_stride == 0 ? d.width : d.row_stride;
    std::vector<float> h;
    h.reserve((size_t)d.width * d.height);
    for (uint32_t r = 0; r < d.height; r++)
        for (uint32_t c = 0; c < d.width; c++) h.push_back(((float)d.heights[(size_t)r * stride + c] * d.height_scale + d.height_offset) + y_offset);
    std::lock_guard<std::mutex> lk(world->mu);
    if (!pull(world)) return RPHYS_ERR_INTERNAL;
    set_hf(world->w, d.width, d.height, d.cell_size_x, d.cell_size_z, origin_x, origin_z, std::move(h));
    host_changed(world);
    return RPHYS_OK;
    GUARD_END
}
RPhysResult rphys_world_set_heightfield_raw32_square(RPhysWorld *world, const uint8_t *bytes, uint32_t bytes_len, float world_size_x,
                                                     float world_size_z, float origin_x, float origin_z, float y_offset) {
    GUARD_BEGIN
    if (!world || !bytes) { set_err("rphys_world_set_heightfield_raw32_square: null arg"); return RPHYS_ERR_NULL; }
    if (bytes_len < 4) return bad_arg("rphys_world_set_heightfield_raw32_square", "bytes_len too small to form a grid");
    size_t count = bytes_len / 4;
    size_t dim = (size_t)std::sqrt((double)count);
    if (dim * dim != count) return bad_arg("rphys_world_set_heightfield_raw32_square", "byte length is not a square number of f32s");
    float sx = world_size_x / std::fmax((float)dim - 1.0f, 1.0f);
    float sz = world_size_z / std::fmax((float)dim - 1.0f, 1.0f);
    std::vector<float> h(count);
    std::memcpy(h.data(), bytes, count * 4);
    for (float &v : h) v = v + y_offset;
    std::lock_guard<std::mutex> lk(world->mu);
    if (!pull(world)) return RPHYS_ERR_INTERNAL;
    set_hf(world->w, (uint32_t)dim, (uint32_t)dim, sx, sz, origin_x, origin_z, std::move(h));
    host_changed(world);
    return RPHYS_OK;
    GUARD_END
}
RPhysResult rphys_world_clear_heightfield(RPhysWorld *world) {
    GUARD_BEGIN
    if (!world) { set_err("rphys_world_clear_heightfield: world is null"); return RPHYS_ERR_NULL; }
    std::lock_guard<std::mutex> lk(world->mu);
    if (!pull(world)) return RPHYS_ERR_INTERNAL;
    world->w.has_hf = false;
    world->w.heights.clear();
    host_changed(world);
    return RPHYS_OK;
    GUARD_END
}

// world.rs:718 add_ball_joint (articulation/src/bodies.rs:326-331): an additive export, the
// reference's C ABI has no joints.  Joint indices are per world, in creation order, never reused.
RPhysResult rphys_world_add_ball_joint(RPhysWorld *world, uint32_t body_a, uint32_t body_b, const RPhysIsometry *frame_a,
                                       const RPhysIsometry *frame_b, uint32_t *out_joint_id) {
    GUARD_BEGIN
    if (out_joint_id) *out_joint_id = 0xFFFFFFFFu;
    if (!world || !frame_a || !frame_b || !out_joint_id) { set_err("rphys_world_add_ball_joint: null arg"); return RPHYS_ERR_NULL; }
    std::lock_guard<std::mutex> lk(world->mu);
    uint32_t a, b;
    RPhysResult r = live_id(world, body_a, "rphys_world_add_ball_joint", a);
    if (r != RPHYS_OK) return r;
    r = live_id(world, body_b, "rphys_world_add_ball_joint", b);
    if (r != RPHYS_OK) return r;
    if (a == b) { set_err("rphys_world_add_ball_joint: a body cannot be jointed to itself"); return RPHYS_ERR_BAD_ARG; }
    const float v[14] = {frame_a->pos.x, frame_a->pos.y, frame_a->pos.z, frame_a->rot.x, frame_a->rot.y, frame_a->rot.z, frame_a->rot.w,
                         frame_b->pos.x, frame_b->pos.y, frame_b->pos.z, frame_b->rot.x, frame_b->rot.y, frame_b->rot.z, frame_b->rot.w};
    for (float x : v) if (!std::isfinite(x)) { set_err("rphys_world_add_ball_joint: non-finite frame"); return RPHYS_ERR_BAD_ARG; }
    if (!pull(world)) return RPHYS_ERR_INTERNAL;
    HostWorld &w = world->w;
    *out_joint_id = w.nj();
    w.jt_a.push_back(a); w.jt_b.push_back(b);
    w.jt_fap.push_back(f4(v[0], v[1], v[2])); w.jt_far.push_back(f4(v[3], v[4], v[5], v[6]));
    w.jt_fbp.push_back(f4(v[7], v[8], v[9])); w.jt_fbr.push_back(f4(v[10], v[11], v[12], v[13]));
    w.wj_mask.push_back(0); w.wj_lam.insert(w.wj_lam.end(), 3, 0.0f);
    w.jt_kind.push_back(0); w.jt_hinge.push_back(RP_NONE); w.jt_ax.insert(w.jt_ax.end(), 18, f4(0, 0, 0));
    w.d6_wn.push_back(0); w.d6_wtag.insert(w.d6_wtag.end(), 32, 0u); w.d6_wlam.insert(w.d6_wlam.end(), 32, 0.0f);
    host_changed(world);
    return RPHYS_OK;
    GUARD_END
}
// bodies.rs:321-345 add_d6 / add_hinge (world.rs:719-720): a ball joint's frames and checks, then
// the constrained axes of t_mask / r_mask and the stored hinge axis
RPhysResult rphys_world_add_d6_joint(RPhysWorld *world, uint32_t body_a, uint32_t body_b, const RPhysIsometry *frame_a, const RPhysIsometry *frame_b,
                                     uint32_t t_mask, uint32_t r_mask, int32_t hinge_axis, uint32_t *out_joint_id) {
    GUARD_BEGIN
    if (out_joint_id) *out_joint_id = 0xFFFFFFFFu;
    if ((t_mask | r_mask) > 7u) return bad_arg("rphys_world_add_d6_joint", "axis masks hold bits 0..2");
    if (hinge_axis < -1 || hinge_axis > 2) return bad_arg("rphys_world_add_d6_joint", "hinge axis must be -1 (none), 0, 1 or 2");
    RPhysResult r = rphys_world_add_ball_joint(world, body_a, body_b, frame_a, frame_b, out_joint_id);
    if (r != RPHYS_OK) return r;
    std::lock_guard<std::mutex> lk(world->mu);
    HostWorld &w = world->w;
    const uint32_t q = *out_joint_id;
    w.jt_kind[q] = RP_JK_D6;
    w.jt_hinge[q] = hinge_axis < 0 ? RP_NONE : (uint32_t)hinge_axis;
    for (uint32_t k = 0; k < 6; k++) {
        const uint32_t en = ((k < 3 ? t_mask >> k : r_mask >> (k - 3)) & 1u) ? D6F_ENABLED_BIT : 0u;
        w.jt_ax[18 * q + 3 * k].w = u2f(en);
    }
    host_changed(world);
    return RPHYS_OK;
    GUARD_END
}
// world.rs:537-539 add_distance_joint -> bodies.rs:273-285: bodies' centres held at rest (rest =
// fin(rest), compliance = max(fin(compliance), 0)); frames are the bodies' own
RPhysResult rphys_world_add_distance_joint(RPhysWorld *world, uint32_t body_a, uint32_t body_b, float rest, float compliance, uint32_t *out_joint_id) {
    GUARD_BEGIN
    if (out_joint_id) *out_joint_id = 0xFFFFFFFFu;
    const RPhysIsometry id = {{0.0f, 0.0f, 0.0f}, {0.0f, 0.0f, 0.0f, 1.0f}};
    RPhysResult r = rphys_world_add_ball_joint(world, body_a, body_b, &id, &id, out_joint_id);
    if (r != RPHYS_OK) return r;
    std::lock_guard<std::mutex> lk(world->mu);
    HostWorld &w = world->w;
    const uint32_t q = *out_joint_id;
    w.jt_kind[q] = RP_JK_DIST;
    w.jt_ax[18 * q] = f4(hfin(rest) ? rest : 0.0f, std::fmax(hfin(compliance) ? compliance : 0.0f, 0.0f), 0.0f);
    host_changed(world);
    return RPHYS_OK;
    GUARD_END
}
RPhysResult rphys_world_add_hinge_joint(RPhysWorld *world, uint32_t body_a, uint32_t body_b, const RPhysIsometry *frame_a, const RPhysIsometry *frame_b,
                                        uint32_t axis, uint32_t *out_joint_id) {
    if (out_joint_id) *out_joint_id = 0xFFFFFFFFu;
    if (axis > 2) { set_err("rphys_world_add_hinge_joint: axis must be 0, 1 or 2"); return RPHYS_ERR_BAD_ARG; }
    return rphys_world_add_d6_joint(world, body_a, body_b, frame_a, frame_b, 7u, 7u & ~(1u << axis), (int32_t)axis, out_joint_id);
}
// The axis words of D6 joint `joint` (ERR_INVALID_ID for an unknown joint, ERR_BAD_ARG for a ball
// joint or an axis above 2); the world's lock is held.
static RPhysResult d6_axis_of(RPhysWorld *world, uint32_t joint, RPhysBool angular, uint32_t axis, const char *fn, F4 *&ax) {
    HostWorld &w = world->w;
    if (joint >= w.nj()) { set_err(std::string(fn) + ": unknown joint"); return RPHYS_ERR_INVALID_ID; }
    if (w.jt_kind[joint] != RP_JK_D6) return bad_arg(fn, "not a D6 joint");
    if (axis > 2) return bad_arg(fn, "axis must be 0, 1 or 2");
    ax = &w.jt_ax[18 * joint + 3 * (angular ? 3 + axis : axis)];
    return RPHYS_OK;
}
// a limit or drive on a rotational axis other than the stored hinge axis forgets it (bodies.rs:878-880)
static void d6_leave_hinge(HostWorld &w, uint32_t joint, RPhysBool angular, uint32_t axis) {
    if (angular && w.jt_hinge[joint] != RP_NONE && w.jt_hinge[joint] != axis) w.jt_hinge[joint] = RP_NONE;
}
// bodies.rs:361-390 set_d6_limit_t / set_d6_limit_r: on = 0 removes the limit; a limit enables its axis
RPhysResult rphys_world_set_joint_limit(RPhysWorld *world, uint32_t joint, RPhysBool angular, uint32_t axis, RPhysBool on, float lo, float hi) {
    GUARD_BEGIN
    if (!world) { set_err("rphys_world_set_joint_limit: world is null"); return RPHYS_ERR_NULL; }
    if (on && !(hfin(lo) && hfin(hi))) return bad_arg("rphys_world_set_joint_limit", "non-finite limit");
    std::lock_guard<std::mutex> lk(world->mu);
    F4 *ax;
    RPhysResult r = d6_axis_of(world, joint, angular, axis, "rphys_world_set_joint_limit", ax);
    if (r != RPHYS_OK) return r;
    if (!pull(world)) return RPHYS_ERR_INTERNAL;
    uint32_t f = f2u(ax[0].w);
    f = on ? (f | D6F_LIMIT_BIT | D6F_ENABLED_BIT) : (f & ~D6F_LIMIT_BIT);
    ax[0].w = u2f(f);
    ax[2] = f4(lo, hi, 0.0f);
    if (on) d6_leave_hinge(world->w, joint, angular, axis);
    host_changed(world);
    return RPHYS_OK;
    GUARD_END
}
// set_d6_drive_t / set_d6_drive_r: on = 0 removes the drive; a drive enables its axis
RPhysResult rphys_world_set_joint_drive(RPhysWorld *world, uint32_t joint, RPhysBool angular, uint32_t axis, RPhysBool on, float target, float kp,
                                        float kd, float max_impulse, RPhysBool velocity_mode) {
    GUARD_BEGIN
    if (!world) { set_err("rphys_world_set_joint_drive: world is null"); return RPHYS_ERR_NULL; }
    if (on && !(hfin(target) && hfin(kp) && hfin(kd) && hfin(max_impulse))) return bad_arg("rphys_world_set_joint_drive", "non-finite drive");
    std::lock_guard<std::mutex> lk(world->mu);
    F4 *ax;
    RPhysResult r = d6_axis_of(world, joint, angular, axis, "rphys_world_set_joint_drive", ax);
    if (r != RPHYS_OK) return r;
    if (!pull(world)) return RPHYS_ERR_INTERNAL;
    uint32_t f = f2u(ax[0].w) & ~(D6F_DRIVE_BIT | D6F_VEL_BIT);
    if (on) f |= D6F_DRIVE_BIT | D6F_ENABLED_BIT | (velocity_mode ? D6F_VEL_BIT : 0u);
    ax[0].w = u2f(f);
    ax[1] = f4(target, kp, kd, max_impulse);
    if (on) d6_leave_hinge(world->w, joint, angular, axis);
    host_changed(world);
    return RPHYS_OK;
    GUARD_END
}
// set_d6_damping (bodies.rs:370-378): negative or non-finite values become 0
RPhysResult rphys_world_set_joint_damping(RPhysWorld *world, uint32_t joint, RPhysBool angular, uint32_t axis, float damping, float friction) {
    GUARD_BEGIN
    if (!world) { set_err("rphys_world_set_joint_damping: world is null"); return RPHYS_ERR_NULL; }
    std::lock_guard<std::mutex> lk(world->mu);
    F4 *ax;
    RPhysResult r = d6_axis_of(world, joint, angular, axis, "rphys_world_set_joint_damping", ax);
    if (r != RPHYS_OK) return r;
    if (!pull(world)) return RPHYS_ERR_INTERNAL;
    ax[0].y = std::fmax(hfin(damping) ? damping : 0.0f, 0.0f);
    ax[0].z = std::fmax(hfin(friction) ? friction : 0.0f, 0.0f);
    host_changed(world);
    return RPHYS_OK;
    GUARD_END
}
// D6Axis::compliance of one axis (0 = rigid; m/N or rad/(N m))
RPhysResult rphys_world_set_joint_compliance(RPhysWorld *world, uint32_t joint, RPhysBool angular, uint32_t axis, float compliance) {
    GUARD_BEGIN
    if (!world) { set_err("rphys_world_set_joint_compliance: world is null"); return RPHYS_ERR_NULL; }
    if (!hfin(compliance)) return bad_arg("rphys_world_set_joint_compliance", "non-finite compliance");
    std::lock_guard<std::mutex> lk(world->mu);
    F4 *ax;
    RPhysResult r = d6_axis_of(world, joint, angular, axis, "rphys_world_set_joint_compliance", ax);
    if (r != RPHYS_OK) return r;
    if (!pull(world)) return RPHYS_ERR_INTERNAL;
    ax[0].x = compliance;
    host_changed(world);
    return RPHYS_OK;
    GUARD_END
}
uint32_t rphys_world_num_joints(const RPhysWorld *world) {
    if (!world) return 0;
    std::lock_guard<std::mutex> lk(const_cast<RPhysWorld *>(world)->mu);
    return world->w.nj();
}

RPhysResult rphys_world_set_pair_props(RPhysWorld *world, uint32_t a, uint32_t b, float mu_s, float mu_k, float rest, float vs) {
    GUARD_BEGIN
    if (!world) { set_err("rphys_world_set_pair_props: world is null"); return RPHYS_ERR_NULL; }
    if (a > 15 || b > 15) return bad_arg("rphys_world_set_pair_props", "material ids above 15 are not reachable through this ABI");
    if (!(hfin(mu_s) && hfin(mu_k) && hfin(rest) && hfin(vs))) return bad_arg("rphys_world_set_pair_props", "non-finite property");
    std::lock_guard<std::mutex> lk(world->mu);
    if (!pull(world)) return RPHYS_ERR_INTERNAL;
    pair_override(world->w, a, b).p = MatP{mu_s, mu_k, rest, vs};  // the pair's contact model stays
    host_changed(world);
    return RPHYS_OK;
    GUARD_END
}
RPhysResult rphys_world_set_pair_contact_model(RPhysWorld *world, uint32_t a, uint32_t b, float stiffness, float dissipation, RPhysBool patch_friction) {
    GUARD_BEGIN
    if (!world) { set_err("rphys_world_set_pair_contact_model: world is null"); return RPHYS_ERR_NULL; }
    if (a > 15 || b > 15) return bad_arg("rphys_world_set_pair_contact_model", "material ids above 15 are not reachable through this ABI");
    if (!(hfin(stiffness) && stiffness >= 0.0f)) return bad_arg("rphys_world_set_pair_contact_model", "stiffness must be finite and >= 0");
    if (!hfin(dissipation)) return bad_arg("rphys_world_set_pair_contact_model", "dissipation must be finite");
    std::lock_guard<std::mutex> lk(world->mu);
    if (!pull(world)) return RPHYS_ERR_INTERNAL;
    PairOverride &o = pair_override(world->w, a, b);  // its friction and restitution stay
    o.stiff = stiffness; o.diss = dissipation; o.patch = patch_friction ? 1u : 0u;
    host_changed(world);
    return RPHYS_OK;
    GUARD_END
}
RPhysResult rphys_world_set_compliant_damping_at_start(RPhysWorld *world, RPhysBool on) {
    GUARD_BEGIN
    if (!world) { set_err("rphys_world_set_compliant_damping_at_start: world is null"); return RPHYS_ERR_NULL; }
    std::lock_guard<std::mutex> lk(world->mu);
    world->w.damp_start = on ? 1u : 0u;
    return RPHYS_OK;
    GUARD_END
}
RPhysResult rphys_world_set_solver_iterations(RPhysWorld *world, uint32_t velocity, uint32_t position) {
    GUARD_BEGIN
    if (!world) { set_err("rphys_world_set_solver_iterations: world is null"); return RPHYS_ERR_NULL; }
    std::lock_guard<std::mutex> lk(world->mu);
    world->w.iv = std::max(velocity, 1u);
    world->w.ip = std::max(position, 1u);
    return RPHYS_OK;
    GUARD_END
}
RPhysResult rphys_world_set_residual_mode(RPhysWorld *world, uint32_t on, float tol, uint32_t max_iters) {
    GUARD_BEGIN
    if (!world) { set_err("rphys_world_set_residual_mode: world is null"); return RPHYS_ERR_NULL; }
    if (on && (!hfin(tol) || tol < 0.0f || max_iters == 0)) return bad_arg("rphys_world_set_residual_mode", "tol must be finite and >= 0, max_iters > 0");
    std::lock_guard<std::mutex> lk(world->mu);
    world->w.resid_on = on ? 1u : 0u;
    if (on) { world->w.resid_tol = tol; world->w.resid_max = max_iters; }
    return RPHYS_OK;
    GUARD_END
}
RPhysResult rphys_world_set_gravity(RPhysWorld *world, RPhysVec3 g) {
    GUARD_BEGIN
    if (!world) { set_err("rphys_world_set_gravity: world is null"); return RPHYS_ERR_NULL; }
    std::lock_guard<std::mutex> lk(world->mu);
    world->w.gx = g.x; world->w.gy = g.y; world->w.gz = g.z;
    return RPHYS_OK;
    GUARD_END
}

RPhysResult rphys_world_get_stats(const RPhysWorld *cworld, RPhysStepStats *out) {
    GUARD_BEGIN
    if (!cworld || !out) { set_err("rphys_world_get_stats: null arg"); return RPHYS_ERR_NULL; }
    std::memset(out, 0, sizeof(*out));
    RPhysWorld *world = const_cast<RPhysWorld *>(cworld);
    std::lock_guard<std::mutex> lk(world->mu);
    if (!world->scene) return RPHYS_OK;
    uint32_t c[CT_COUNT];
    Scene &S = *world->scene;
    if (!hip_ok(hipMemcpyAsync(c, S.d.cnt, sizeof c, hipMemcpyDeviceToHost, S.st), "stats") || !hip_ok(hipStreamSynchronize(S.st), "stats")) return RPHYS_ERR_INTERNAL;
    uint32_t wt = 0;
    if (!hip_ok(hipMemcpy(&wt, S.d.wc_off[S.cur] + S.d.nw, 4, hipMemcpyDeviceToHost), "stats")) return RPHYS_ERR_INTERNAL;
    out->pairs = c[CT_PAIRS]; out->contacts = c[CT_CONTACTS]; out->terrain_candidates = c[CT_CANDS];
    out->island_contacts = c[CT_ISLC]; out->islands = c[CT_ISLANDS]; out->warm_entries = wt;
    out->pgs_iters = c[CT_ITERS_MAX]; out->pgs_iters_sum = c[CT_ITERS_SUM];
    out->bounce_rows = c[CT_W6_TOI]; out->bounce_support_rows = c[CT_W6_SUPPORT];
    out->bounce_pseudo_rows = c[CT_W6_PSEUDO]; out->bounce_fallbacks = c[CT_W6_FALLBACK];
    out->relax_rows = c[CT_W6_RELAX];
    return RPHYS_OK;
    GUARD_END
}

struct ContactDump {
    std::vector<uint32_t> a, b, cell;
    std::vector<F4> p, n;
    std::vector<float> imp;
};
static bool dump_contacts(Scene &S, ContactDump &D) {
    uint32_t n = 0;
    if (!hip_ok(hipMemcpy(&n, S.d.cnt + CT_CONTACTS, 4, hipMemcpyDeviceToHost), "contacts")) return false;
    n = std::min(n, S.d.cap_contacts);
    D.a.resize(n); D.b.resize(n); D.cell.resize(n); D.p.resize(n); D.n.resize(n); D.imp.resize(n);
    if (n == 0) return true;
    return hip_ok(hipMemcpy(D.a.data(), S.d.ct_a, 4 * n, hipMemcpyDeviceToHost), "contacts") &&
           hip_ok(hipMemcpy(D.b.data(), S.d.ct_b, 4 * n, hipMemcpyDeviceToHost), "contacts") &&
           hip_ok(hipMemcpy(D.cell.data(), S.d.ct_cell, 4 * n, hipMemcpyDeviceToHost), "contacts") &&
           hip_ok(hipMemcpy(D.p.data(), S.d.ct_p, 16 * n, hipMemcpyDeviceToHost), "contacts") &&
           hip_ok(hipMemcpy(D.n.data(), S.d.ct_n, 16 * n, hipMemcpyDeviceToHost), "contacts") &&
           hip_ok(hipMemcpy(D.imp.data(), S.d.ct_imp, 4 * n, hipMemcpyDeviceToHost), "contacts");
}

RPhysResult rphys_world_get_contacts(const RPhysWorld *cworld, RPhysContactEvent *out_buffer, uint32_t buffer_capacity, uint32_t *out_count) {
    GUARD_BEGIN
    if (!cworld || !out_count) { set_err("rphys_world_get_contacts: null arg"); return RPHYS_ERR_NULL; }
    RPhysWorld *world = const_cast<RPhysWorld *>(cworld);
    std::lock_guard<std::mutex> lk(world->mu);
    *out_count = 0;
    if (!world->scene || !world->have_contacts) return RPHYS_OK;
    ContactDump D;
    if (!dump_contacts(*world->scene, D)) return RPHYS_ERR_INTERNAL;
    *out_count = (uint32_t)D.a.size();
    if (!out_buffer) return RPHYS_OK;
    uint32_t n = std::min(*out_count, buffer_capacity);
    // Colliders resolve to bodies through the layout the tick ran with (a removal since then
    // compacts the host's collider list); an entry whose body is gone is skipped, as lib.rs:1841-1854
    // skips ids it cannot resolve.  Like the reference, the count is that of the last tick's
    // contacts, and the events written are packed at the front of the buffer.
    const HostWorld &w = world->w;
    const std::vector<uint32_t> &cb = world->scene->col_body;
    auto body_of = [&](uint32_t c, uint32_t &id) {
        if (c >= cb.size()) return false;
        uint32_t b = cb[c];
        if (b >= w.nb() || !(w.bflags[b] & BF_ALIVE)) return false;
        id = b + 1;
        return true;
    };
    uint32_t k = 0;
    for (uint32_t i = 0; i < n; i++) {
        RPhysContactEvent e;
        if (!body_of(D.a[i], e.body_id_a)) continue;
        if (D.b[i] == RP_TERRAIN) e.body_id_b = RPHYS_TERRAIN_BODY_ID;
        else if (!body_of(D.b[i], e.body_id_b)) continue;
        e.position = {D.p[i].x, D.p[i].y, D.p[i].z};
        e.normal = {D.n[i].x, D.n[i].y, D.n[i].z};
        e.impulse = D.imp[i];
        out_buffer[k++] = e;
    }
    return RPHYS_OK;
    GUARD_END
}
RPhysResult rphys_debug_get_contacts_ex(const RPhysWorld *cworld, RPhysContactEx *out, uint32_t cap, uint32_t *out_count) {
    GUARD_BEGIN
    if (!cworld || !out_count) { set_err("rphys_debug_get_contacts_ex: null arg"); return RPHYS_ERR_NULL; }
    RPhysWorld *world = const_cast<RPhysWorld *>(cworld);
    std::lock_guard<std::mutex> lk(world->mu);
    *out_count = 0;
    if (!world->scene || !world->have_contacts) return RPHYS_OK;
    ContactDump D;
    if (!dump_contacts(*world->scene, D)) return RPHYS_ERR_INTERNAL;
    *out_count = (uint32_t)D.a.size();
    if (!out) return RPHYS_OK;
    uint32_t n = std::min(*out_count, cap);
    for (uint32_t i = 0; i < n; i++) {
        RPhysContactEx e;
        e.a_collider = D.a[i]; e.b_collider = D.b[i];
        e.point[0] = D.p[i].x; e.point[1] = D.p[i].y; e.point[2] = D.p[i].z; e.depth = D.p[i].w;
        e.normal[0] = D.n[i].x; e.normal[1] = D.n[i].y; e.normal[2] = D.n[i].z;
        uint32_t f; std::memcpy(&f, &D.n[i].w, 4);
        e.feature = f; e.cell = D.cell[i]; e.impulse = D.imp[i];
        out[i] = e;
    }
    return RPHYS_OK;
    GUARD_END
}
RPhysResult rphys_debug_get_pairs(const RPhysWorld *cworld, uint32_t *out_ij, uint32_t cap, uint32_t *out_count) {
    GUARD_BEGIN
    if (!cworld || !out_count) { set_err("rphys_debug_get_pairs: null arg"); return RPHYS_ERR_NULL; }
    RPhysWorld *world = const_cast<RPhysWorld *>(cworld);
    std::lock_guard<std::mutex> lk(world->mu);
    *out_count = 0;
    if (!world->scene || !world->have_contacts) return RPHYS_OK;
    Scene &S = *world->scene;
    uint32_t n = 0;
    if (!hip_ok(hipMemcpy(&n, S.d.cnt + CT_PAIRS, 4, hipMemcpyDeviceToHost), "pairs")) return RPHYS_ERR_INTERNAL;
    n = std::min(n, S.d.cap_pairs);
    *out_count = n;
    if (!out_ij) return RPHYS_OK;
    uint32_t m = std::min(n, cap);
    std::vector<uint32_t> a(m), b(m);
    if (m && (!hip_ok(hipMemcpy(a.data(), S.d.pair_i, 4 * m, hipMemcpyDeviceToHost), "pairs") ||
              !hip_ok(hipMemcpy(b.data(), S.d.pair_j, 4 * m, hipMemcpyDeviceToHost), "pairs"))) return RPHYS_ERR_INTERNAL;
    for (uint32_t i = 0; i < m; i++) { out_ij[2 * i] = a[i]; out_ij[2 * i + 1] = b[i]; }
    return RPHYS_OK;
    GUARD_END
}

RPhysResult rphys_debug_export_state(const RPhysWorld *cworld, float *pos3, float *rot4, float *lin3, float *ang3, float *last3, uint32_t *awake,
                                     uint32_t *sticks, uint64_t *wkeys, float *wla3, float *wjn, float *wjt3, uint32_t warm_cap, uint32_t *warm_count) {
    GUARD_BEGIN
    if (!cworld || !warm_count) { set_err("rphys_debug_export_state: null arg"); return RPHYS_ERR_NULL; }
    RPhysWorld *world = const_cast<RPhysWorld *>(cworld);
    std::lock_guard<std::mutex> lk(world->mu);
    if (!pull(world)) return RPHYS_ERR_INTERNAL;
    const HostWorld &w = world->w;
    for (uint32_t i = 0; i < w.nb(); i++) {
        if (pos3) { pos3[3 * i] = w.pos[i].x; pos3[3 * i + 1] = w.pos[i].y; pos3[3 * i + 2] = w.pos[i].z; }
        if (rot4) { rot4[4 * i] = w.rot[i].x; rot4[4 * i + 1] = w.rot[i].y; rot4[4 * i + 2] = w.rot[i].z; rot4[4 * i + 3] = w.rot[i].w; }
        if (lin3) { lin3[3 * i] = w.lin[i].x; lin3[3 * i + 1] = w.lin[i].y; lin3[3 * i + 2] = w.lin[i].z; }
        if (ang3) { ang3[3 * i] = w.ang[i].x; ang3[3 * i + 1] = w.ang[i].y; ang3[3 * i + 2] = w.ang[i].z; }
        if (last3) { last3[3 * i] = w.last_lin[i].x; last3[3 * i + 1] = w.last_lin[i].y; last3[3 * i + 2] = w.last_lin[i].z; }
        if (awake) awake[i] = w.awake[i];
        if (sticks) sticks[i] = w.sticks[i];
    }
    uint32_t n = (uint32_t)w.wkey.size();
    *warm_count = n;
    for (uint32_t k = 0; k < n && k < warm_cap; k++) {
        if (wkeys) wkeys[k] = w.wkey[k];
        if (wla3) { wla3[3 * k] = w.wla[k].x; wla3[3 * k + 1] = w.wla[k].y; wla3[3 * k + 2] = w.wla[k].z; }
        if (wjn) wjn[k] = w.wjn[k];
        if (wjt3) { wjt3[3 * k] = w.wjt[k].x; wjt3[3 * k + 1] = w.wjt[k].y; wjt3[3 * k + 2] = w.wjt[k].z; }
    }
    return RPHYS_OK;
    GUARD_END
}

RPhysResult rphys_debug_get_joint_rows(const RPhysWorld *cworld, uint32_t joint, uint8_t *tags, float *impulses, uint32_t cap, uint32_t *out_n) {
    GUARD_BEGIN
    if (!cworld || !out_n) { set_err("rphys_debug_get_joint_rows: null arg"); return RPHYS_ERR_NULL; }
    *out_n = 0;
    RPhysWorld *world = const_cast<RPhysWorld *>(cworld);
    std::lock_guard<std::mutex> lk(world->mu);
    if (joint >= world->w.nj()) { set_err("rphys_debug_get_joint_rows: unknown joint"); return RPHYS_ERR_INVALID_ID; }
    if (!pull(world)) return RPHYS_ERR_INTERNAL;
    const HostWorld &w = world->w;
    uint32_t n = 0;
    if (w.jt_kind[joint] == RP_JK_D6) {
        const uint32_t wn = w.d6_wn[joint];
        if (wn >> 31)
            for (uint32_t r = 0; r < (wn & 0xFFu); r++, n++)
                if (n < cap) { if (tags) tags[n] = (uint8_t)w.d6_wtag[32 * joint + r]; if (impulses) impulses[n] = w.d6_wlam[32 * joint + r]; }
    } else if (w.wj_mask[joint] >> 31) {  // a ball joint: its locked axes' rows, tags axis << 3
        for (uint32_t i = 0; i < 3; i++) {
            if (!((w.wj_mask[joint] >> i) & 1u)) continue;
            if (n < cap) { if (tags) tags[n] = (uint8_t)(i << 3); if (impulses) impulses[n] = w.wj_lam[3 * joint + i]; }
            n++;
        }
    }
    *out_n = n;
    return RPHYS_OK;
    GUARD_END
}

// ------------------------------------------------------------------ snapshots (lib.rs:1696-1800)
struct RPhysWorldSnapshot {
    uint64_t tick;
    std::vector<F4> pos, rot, lin, ang, last_lin;
    std::vector<uint32_t> awake, sticks;
    std::vector<uint64_t> wkey;
    std::vector<F4> wla, wjt;
    std::vector<float> wjn;
    std::vector<uint32_t> wj_mask;
    std::vector<float> wj_lam;
    std::vector<HostWorld::Player> players;
    std::vector<uint64_t> pkey;
    std::vector<F4> pjt;
    std::vector<uint32_t> d6_wn, d6_wtag;
    std::vector<float> d6_wlam;
};
RPhysWorldSnapshot *rphys_world_snapshot_create(const RPhysWorld *cworld) {
    clear_err();
    try {
        if (!cworld) { set_err("rphys_world_snapshot_create: world is null"); return nullptr; }
        RPhysWorld *world = const_cast<RPhysWorld *>(cworld);
        std::lock_guard<std::mutex> lk(world->mu);
        if (!pull(world)) return nullptr;
        const HostWorld &w = world->w;
        auto *s = new RPhysWorldSnapshot();
        s->tick = w.tick;
        s->pos = w.pos; s->rot = w.rot; s->lin = w.lin; s->ang = w.ang;
        for (uint32_t i = 0; i < w.nb(); i++)
            if (!(w.bflags[i] & BF_ALIVE)) { s->pos[i] = f4(0, 0, 0); s->rot[i] = f4(0, 0, 0, 1); s->lin[i] = f4(0, 0, 0); s->ang[i] = f4(0, 0, 0); }
        s->last_lin = w.last_lin; s->awake = w.awake; s->sticks = w.sticks;
        s->wkey = w.wkey; s->wla = w.wla; s->wjn = w.wjn; s->wjt = w.wjt;
        s->wj_mask = w.wj_mask; s->wj_lam = w.wj_lam;
        s->players = w.players;
        s->pkey = w.pkey; s->pjt = w.pjt;
        s->d6_wn = w.d6_wn; s->d6_wtag = w.d6_wtag; s->d6_wlam = w.d6_wlam;
        return s;
    } catch (...) {
        set_err("panic in rphys_world_snapshot_create");
        return nullptr;
    }
}
RPhysResult rphys_world_snapshot_restore(RPhysWorld *world, const RPhysWorldSnapshot *s) {
    GUARD_BEGIN
    if (!world || !s) { set_err("rphys_world_snapshot_restore: null arg"); return RPHYS_ERR_NULL; }
    std::lock_guard<std::mutex> lk(world->mu);
    if (!pull(world)) return RPHYS_ERR_INTERNAL;
    HostWorld &w = world->w;
    w.tick = s->tick;
    w.players = s->players;  // lib.rs:1761: the players as snapshot
    for (uint32_t i = 0; i < s->pos.size() && i < w.nb(); i++) {
        if (!(w.bflags[i] & BF_ALIVE)) continue;
        // set_body_pose (world.rs:501-506) then set_body_vel (world.rs:674-676)
        w.pos[i] = s->pos[i];
        w.rot[i] = fin_q(s->rot[i]);
        w.lin[i] = s->lin[i]; w.ang[i] = s->ang[i];
    }
    // restore_carried_state (world.rs:927-961): warm cache, last velocities, sleep, dead keys dropped
    w.wkey.clear(); w.wla.clear(); w.wjn.clear(); w.wjt.clear();
    auto alive = [&](uint32_t b) { return b == 0xFFFFFFFFu || (b < w.nb() && (w.bflags[b] & BF_ALIVE)); };
    for (size_t k = 0; k < s->wkey.size(); k++) {
        if (!alive((uint32_t)(s->wkey[k] >> 32)) || !alive((uint32_t)s->wkey[k])) continue;
        w.wkey.push_back(s->wkey[k]); w.wla.push_back(s->wla[k]); w.wjn.push_back(s->wjn[k]); w.wjt.push_back(s->wjt[k]);
    }
    w.pkey.clear(); w.pjt.clear();
    for (size_t k = 0; k < s->pkey.size(); k++) {
        if (!alive((uint32_t)(s->pkey[k] >> 32)) || !alive((uint32_t)s->pkey[k])) continue;
        w.pkey.push_back(s->pkey[k]); w.pjt.push_back(s->pjt[k]);
    }
    for (uint32_t i = 0; i < s->last_lin.size() && i < w.nb(); i++) w.last_lin[i] = s->last_lin[i];
    for (uint32_t i = 0; i < s->awake.size() && i < w.nb(); i++) { w.awake[i] = s->awake[i]; w.sticks[i] = s->sticks[i]; }
    // warm joint rows as snapshot (joints added since start cold); a dead body's joints drop theirs
    for (uint32_t q = 0; q < w.nj(); q++) {
        const bool had = q < s->wj_mask.size();
        w.wj_mask[q] = had ? s->wj_mask[q] : 0u;
        for (int i = 0; i < 3; i++) w.wj_lam[3 * q + i] = had ? s->wj_lam[3 * q + i] : 0.0f;
        w.d6_wn[q] = had ? s->d6_wn[q] : 0u;
        for (int i = 0; i < 32; i++) { w.d6_wtag[32 * q + i] = had ? s->d6_wtag[32 * q + i] : 0u; w.d6_wlam[32 * q + i] = had ? s->d6_wlam[32 * q + i] : 0.0f; }
        if (!alive(w.jt_a[q]) || !alive(w.jt_b[q])) { w.wj_mask[q] = 0; w.d6_wn[q] = 0; }
    }
    host_changed(world);
    return RPHYS_OK;
    GUARD_END
}
void rphys_world_snapshot_free(RPhysWorldSnapshot *s) { delete s; }

// ------------------------------------------------------------------ batches
struct RPhysBatch {
    std::vector<HostWorld> worlds;
    Scene *scene = nullptr;
    DeviceBuffers logbuf;  // contact log (outlives scene rebuilds)
    ~RPhysBatch() { delete scene; }
};

RPhysBatch *rphys_batch_create(RPhysWorld *const *worlds, uint32_t n_worlds) {
    clear_err();
    try {
        if (!worlds || n_worlds == 0) { set_err("rphys_batch_create: no worlds"); return nullptr; }
        auto *B = new RPhysBatch();
        B->worlds.reserve(n_worlds);
        for (uint32_t i = 0; i < n_worlds; i++) {
            if (!worlds[i]) { delete B; set_err("rphys_batch_create: null world"); return nullptr; }
            std::lock_guard<std::mutex> lk(worlds[i]->mu);
            if (!pull(worlds[i])) { delete B; return nullptr; }
            B->worlds.push_back(worlds[i]->w);
        }
        std::vector<HostWorld *> ws;
        for (auto &w : B->worlds) ws.push_back(&w);
        B->scene = build_scene(ws, initial_caps(ws), 0);
        if (!B->scene) { delete B; return nullptr; }
        return B;
    } catch (...) {
        set_err("panic in rphys_batch_create");
        return nullptr;
    }
}
void rphys_batch_destroy(RPhysBatch *b) { delete b; }

static RPhysResult batch_step(RPhysBatch *b, float dt, uint32_t n, bool graph) {
    if (!b) { set_err("rphys_batch_step: batch is null"); return RPHYS_ERR_NULL; }
    if (!hfin(dt) || dt < 0.0f) { set_err("rphys_batch_step: dt must be finite and >= 0"); return RPHYS_ERR_BAD_ARG; }
    std::vector<HostWorld *> ws;
    for (auto &w : b->worlds) ws.push_back(&w);
    return run_ticks(b->scene, ws, dt, b->worlds[0].iv, n, graph) ? RPHYS_OK : RPHYS_ERR_INTERNAL;
}
RPhysResult rphys_batch_step_n(RPhysBatch *b, float dt, uint32_t n) { GUARD_BEGIN return batch_step(b, dt, n, false); GUARD_END }
RPhysResult rphys_batch_step_graph(RPhysBatch *b, float dt, uint32_t n) { GUARD_BEGIN return batch_step(b, dt, n, true); GUARD_END }
RPhysResult rphys_batch_step_profiled(RPhysBatch *b, float dt, uint32_t n, RPhysProfile *out) {
    GUARD_BEGIN
    if (!b || !out) { set_err("rphys_batch_step_profiled: null arg"); return RPHYS_ERR_NULL; }
    std::memset(out, 0, sizeof(*out));
    out->n_stages = RP_NMARKS - 1;
    if (!hfin(dt) || dt < 0.0f) { set_err("rphys_batch_step_profiled: dt must be finite and >= 0"); return RPHYS_ERR_BAD_ARG; }
    if (n == 0) return RPHYS_OK;
    std::vector<HostWorld *> ws;
    for (auto &w : b->worlds) ws.push_back(&w);
    return profiled_ticks(b->scene, ws, dt, b->worlds[0].iv, n, out) ? RPHYS_OK : RPHYS_ERR_INTERNAL;
    GUARD_END
}

RPhysResult rphys_batch_get_hashes(const RPhysBatch *b, uint8_t *out) {
    GUARD_BEGIN
    if (!b || !out) { set_err("rphys_batch_get_hashes: null arg"); return RPHYS_ERR_NULL; }
    return device_hash(*b->scene, out, b->scene->d.nw) ? RPHYS_OK : RPHYS_ERR_INTERNAL;
    GUARD_END
}
RPhysResult rphys_batch_total_bodies(const RPhysBatch *b, uint32_t *out) {
    if (!b || !out) { set_err("rphys_batch_total_bodies: null arg"); return RPHYS_ERR_NULL; }
    *out = b->scene->d.nb;
    return RPHYS_OK;
}
RPhysResult rphys_batch_get_state(const RPhysBatch *cb, uint32_t wi, float *pos3, float *rot4, float *lin3, float *ang3) {
    GUARD_BEGIN
    if (!cb) { set_err("rphys_batch_get_state: null arg"); return RPHYS_ERR_NULL; }
    RPhysBatch *b = const_cast<RPhysBatch *>(cb);
    if (wi >= b->worlds.size()) { set_err("rphys_batch_get_state: world index out of range"); return RPHYS_ERR_BAD_ARG; }
    if (!pull_world(*b->scene, wi, b->worlds[wi])) return RPHYS_ERR_INTERNAL;
    const HostWorld &w = b->worlds[wi];
    for (uint32_t i = 0; i < w.nb(); i++) {
        if (pos3) { pos3[3 * i] = w.pos[i].x; pos3[3 * i + 1] = w.pos[i].y; pos3[3 * i + 2] = w.pos[i].z; }
        if (rot4) { rot4[4 * i] = w.rot[i].x; rot4[4 * i + 1] = w.rot[i].y; rot4[4 * i + 2] = w.rot[i].z; rot4[4 * i + 3] = w.rot[i].w; }
        if (lin3) { lin3[3 * i] = w.lin[i].x; lin3[3 * i + 1] = w.lin[i].y; lin3[3 * i + 2] = w.lin[i].z; }
        if (ang3) { ang3[3 * i] = w.ang[i].x; ang3[3 * i + 1] = w.ang[i].y; ang3[3 * i + 2] = w.ang[i].z; }
    }
    return RPHYS_OK;
    GUARD_END
}

RPhysResult rphys_batch_get_world(RPhysBatch *b, uint32_t wi, RPhysWorld *dst) {
    GUARD_BEGIN
    if (!b || !dst) { set_err("rphys_batch_get_world: null arg"); return RPHYS_ERR_NULL; }
    if (wi >= b->worlds.size()) return bad_arg("rphys_batch_get_world", "world index out of range");
    if (!pull_world(*b->scene, wi, b->worlds[wi])) return RPHYS_ERR_INTERNAL;
    std::lock_guard<std::mutex> lk(dst->mu);
    dst->w = b->worlds[wi];
    dst->dev_newer = false;
    dst->have_contacts = false;
    host_changed(dst);
    return RPHYS_OK;
    GUARD_END
}

// Device-side rollback slots: body state, sleep state and the warm cache of every world.  Slots
// 0..63 are the caller's; slot 64 is the autotuner's.
static const uint32_t TUNE_SLOT = 64;
static RPhysResult batch_snapshot(RPhysBatch *b, uint32_t slot) {
    Scene &S = *b->scene;
    if (S.slots.size() <= slot) S.slots.resize(slot + 1);
    const Dev &d = S.d;
    uint32_t nwarm = 0, npatch = 0;
    if (!hip_ok(hipMemcpyAsync(&nwarm, d.wc_off[S.cur] + d.nw, 4, hipMemcpyDeviceToHost, S.st), "snapshot") ||
        (d.cm_on && !hip_ok(hipMemcpyAsync(&npatch, d.wpc_off[S.cur] + d.nw, 4, hipMemcpyDeviceToHost, S.st), "snapshot")) ||
        !hip_ok(hipStreamSynchronize(S.st), "snapshot"))
        return RPHYS_ERR_INTERNAL;
    // A slot made before a capacity rebuild can be smaller than today's caches: replace it.
    if (S.slots[slot] && S.slots[slot]->cap_warm < std::max(nwarm, npatch)) S.slots[slot].reset();
    if (!S.slots[slot]) {
        std::unique_ptr<Slot> sl(new Slot());
        size_t NB = d.nb + 1, W = std::max<size_t>(d.cap_warm, std::max(nwarm, npatch));
        sl->cap_warm = W;
        bool ok = sl->buf.alloc(sl->pos, NB) && sl->buf.alloc(sl->rot, NB) && sl->buf.alloc(sl->lin, NB) && sl->buf.alloc(sl->ang, NB) &&
                  sl->buf.alloc(sl->last, NB) && sl->buf.alloc(sl->awake, NB) && sl->buf.alloc(sl->sticks, NB) &&
                  sl->buf.alloc(sl->wkey, W) && sl->buf.alloc(sl->wla, W) && sl->buf.alloc(sl->wjn, W) && sl->buf.alloc(sl->wjt, W) &&
                  sl->buf.alloc(sl->woff, d.nw + 1) && sl->buf.alloc(sl->wj_mask, d.nj + 1) && sl->buf.alloc(sl->wj_lam, 3 * ((size_t)d.nj + 1)) &&
                  sl->buf.alloc(sl->pl_state, d.np + 1) && sl->buf.alloc(sl->pl_input, d.np + 1) && sl->buf.alloc(sl->pl_grounded, d.np + 1) &&
                  sl->buf.alloc(sl->pkey, d.cm_on ? W : 1) && sl->buf.alloc(sl->pjt, d.cm_on ? W : 1) && sl->buf.alloc(sl->poff, d.nw + 1) &&
                  sl->buf.alloc(sl->d6_wn, d.d6_on ? d.nj : 1) && sl->buf.alloc(sl->d6_wtag, d.d6_on ? 32 * (size_t)d.nj : 1) &&
                  sl->buf.alloc(sl->d6_wlam, d.d6_on ? 32 * (size_t)d.nj : 1);
        if (!ok) { set_err("rphys_batch_snapshot: device allocation failed"); return RPHYS_ERR_INTERNAL; }
        S.slots[slot] = std::move(sl);
    }
    Slot &sl = *S.slots[slot];
    size_t n = d.nb;
    auto cp = [&](void *dst, const void *src, size_t bytes) { return bytes == 0 || hip_ok(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, S.st), "snapshot copy"); };
    bool ok = cp(sl.pos, d.pos, 16 * n) && cp(sl.rot, d.rot, 16 * n) && cp(sl.lin, d.lin, 16 * n) && cp(sl.ang, d.ang, 16 * n) &&
              cp(sl.last, d.last_lin, 16 * n) && cp(sl.awake, d.awake, 4 * n) && cp(sl.sticks, d.sticks, 4 * n) &&
              cp(sl.wkey, d.wk_key[S.cur], 8 * (size_t)nwarm) && cp(sl.wla, d.wk_la[S.cur], 16 * (size_t)nwarm) &&
              cp(sl.wjn, d.wk_jn[S.cur], 4 * (size_t)nwarm) && cp(sl.wjt, d.wk_jt[S.cur], 16 * (size_t)nwarm) &&
              cp(sl.woff, d.wc_off[S.cur], 4 * ((size_t)d.nw + 1)) && cp(sl.wj_mask, d.wj_mask, 4 * (size_t)d.nj) &&
              cp(sl.wj_lam, d.wj_lam, 12 * (size_t)d.nj) && cp(sl.pl_state, d.pl_state, 16 * (size_t)d.np) &&
              cp(sl.pl_input, d.pl_input, 16 * (size_t)d.np) && cp(sl.pl_grounded, d.pl_grounded, 4 * (size_t)d.np);
    if (ok && d.cm_on)
        ok = cp(sl.pkey, d.wp_key[S.cur], 8 * (size_t)npatch) && cp(sl.pjt, d.wp_jt[S.cur], 16 * (size_t)npatch) &&
             cp(sl.poff, d.wpc_off[S.cur], 4 * ((size_t)d.nw + 1));
    if (ok && d.d6_on)
        ok = cp(sl.d6_wn, d.d6_wn, 4 * (size_t)d.nj) && cp(sl.d6_wtag, d.d6_wtag, 128 * (size_t)d.nj) && cp(sl.d6_wlam, d.d6_wlam, 128 * (size_t)d.nj);
    if (!ok || !hip_ok(hipStreamSynchronize(S.st), "snapshot")) return RPHYS_ERR_INTERNAL;
    sl.cur = S.cur; sl.nwarm = nwarm; sl.npatch = npatch; sl.valid = true;
    sl.ticks.clear();
    for (auto &w : b->worlds) sl.ticks.push_back(w.tick);
    return RPHYS_OK;
}
RPhysResult rphys_batch_snapshot(RPhysBatch *b, uint32_t slot) {
    GUARD_BEGIN
    if (!b) { set_err("rphys_batch_snapshot: batch is null"); return RPHYS_ERR_NULL; }
    if (slot > 63) return bad_arg("rphys_batch_snapshot", "slot must be < 64");
    return batch_snapshot(b, slot);
    GUARD_END
}
static RPhysResult batch_restore(RPhysBatch *b, uint32_t slot) {
    Scene &S = *b->scene;
    if (slot >= S.slots.size() || !S.slots[slot] || !S.slots[slot]->valid) return bad_arg("rphys_batch_restore", "empty slot");
    Slot &sl = *S.slots[slot];
    const Dev &d = S.d;
    if (sl.nwarm > d.cap_warm || sl.npatch > d.cap_warm) { set_err("rphys_batch_restore: warm cache larger than the scene's capacity"); return RPHYS_ERR_INTERNAL; }
    size_t n = d.nb;
    auto cp = [&](void *dst, const void *src, size_t bytes) { return bytes == 0 || hip_ok(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, S.st), "restore copy"); };
    // Snapshot poses are already sanitised (set_pose's normalisation maps them to themselves for
    // every pose a step produces: |q|^2 rounds to 1 within 1 ulp); the batch copies them back.
    bool ok = cp(d.pos, sl.pos, 16 * n) && cp(d.rot, sl.rot, 16 * n) && cp(d.lin, sl.lin, 16 * n) && cp(d.ang, sl.ang, 16 * n) &&
              cp(d.last_lin, sl.last, 16 * n) && cp(d.awake, sl.awake, 4 * n) && cp(d.sticks, sl.sticks, 4 * n) &&
              cp(d.wk_key[sl.cur], sl.wkey, 8 * (size_t)sl.nwarm) && cp(d.wk_la[sl.cur], sl.wla, 16 * (size_t)sl.nwarm) &&
              cp(d.wk_jn[sl.cur], sl.wjn, 4 * (size_t)sl.nwarm) && cp(d.wk_jt[sl.cur], sl.wjt, 16 * (size_t)sl.nwarm) &&
              cp(d.wc_off[sl.cur], sl.woff, 4 * ((size_t)d.nw + 1)) && cp(d.wj_mask, sl.wj_mask, 4 * (size_t)d.nj) &&
              cp(d.wj_lam, sl.wj_lam, 12 * (size_t)d.nj) && cp(d.pl_state, sl.pl_state, 16 * (size_t)d.np) &&
              cp(d.pl_input, sl.pl_input, 16 * (size_t)d.np) && cp(d.pl_grounded, sl.pl_grounded, 4 * (size_t)d.np);
    if (ok && d.cm_on)
        ok = cp(d.wp_key[sl.cur], sl.pkey, 8 * (size_t)sl.npatch) && cp(d.wp_jt[sl.cur], sl.pjt, 16 * (size_t)sl.npatch) &&
             cp(d.wpc_off[sl.cur], sl.poff, 4 * ((size_t)d.nw + 1));
    if (ok && d.d6_on)
        ok = cp(d.d6_wn, sl.d6_wn, 4 * (size_t)d.nj) && cp(d.d6_wtag, sl.d6_wtag, 128 * (size_t)d.nj) && cp(d.d6_wlam, sl.d6_wlam, 128 * (size_t)d.nj);
    if (!ok || !hip_ok(hipStreamSynchronize(S.st), "restore")) return RPHYS_ERR_INTERNAL;
    S.cur = sl.cur;
    for (size_t i = 0; i < b->worlds.size() && i < sl.ticks.size(); i++) b->worlds[i].tick = sl.ticks[i];
    return RPHYS_OK;
}
RPhysResult rphys_batch_restore(RPhysBatch *b, uint32_t slot) {
    GUARD_BEGIN
    if (!b) { set_err("rphys_batch_restore: batch is null"); return RPHYS_ERR_NULL; }
    if (slot > 63) return bad_arg("rphys_batch_restore", "slot must be < 64");
    return batch_restore(b, slot);
    GUARD_END
}

// ------------------------------------------------------------------ launch-shape tuning
RPhysResult rphys_batch_set_tuning(RPhysBatch *b, const RPhysTuning *t) {
    GUARD_BEGIN
    if (!b || !t) { set_err("rphys_batch_set_tuning: null arg"); return RPHYS_ERR_NULL; }
    if (t->size_bytes < sizeof(RPhysTuning)) return bad_arg("rphys_batch_set_tuning", "size_bytes too small");
    const uint32_t g[3] = {t->grid_small, t->grid_large, t->grid_huge};
    for (uint32_t k = 0; k < 3; k++)
        if (g[k] == 0 || g[k] > 65535) return bad_arg("rphys_batch_set_tuning", "grids must be in 1..65535");
    std::memcpy(b->scene->grids, g, sizeof g);
    return RPHYS_OK;
    GUARD_END
}
RPhysResult rphys_batch_get_tuning(const RPhysBatch *b, RPhysTuning *t) {
    GUARD_BEGIN
    if (!b || !t) { set_err("rphys_batch_get_tuning: null arg"); return RPHYS_ERR_NULL; }
    t->size_bytes = sizeof(RPhysTuning);
    t->grid_small = b->scene->grids[0]; t->grid_large = b->scene->grids[1]; t->grid_huge = b->scene->grids[2];
    return RPHYS_OK;
    GUARD_END
}
// Grid search over the island-sweep grids: every candidate steps the same `ticks` graph-captured
// ticks from one device snapshot (once to capture, once timed); the fastest is installed and the
// batch is left at the snapshot.  The per-world hashes after the ticks must agree across all
// candidates, or the search fails.
RPhysResult rphys_batch_autotune(RPhysBatch *b, float dt, uint32_t ticks, RPhysTuning *best, double *best_ms_per_tick) {
    GUARD_BEGIN
    if (!b || !best || !best_ms_per_tick) { set_err("rphys_batch_autotune: null arg"); return RPHYS_ERR_NULL; }
    if (!hfin(dt) || dt < 0.0f) return bad_arg("rphys_batch_autotune", "dt must be finite and >= 0");
    if (ticks == 0 || ticks > 64) return bad_arg("rphys_batch_autotune", "ticks must be in 1..64");
    static const uint32_t G0[] = {1024, 2048, 4096, 8192, 16384, 32768}, G1[] = {128, 256, 512, 1024, 2048}, G2[] = {64, 128, 256};
    RPhysResult r = batch_snapshot(b, TUNE_SLOT);
    if (r != RPHYS_OK) return r;
    const size_t nw = b->worlds.size();
    std::vector<uint8_t> ref, got(32 * nw);
    // one timed run of `ticks` graph-captured ticks from the snapshot with shape g (captured by an
    // untimed run first when `capture`), its hashes checked against the first candidate's
    auto timed = [&](const std::array<uint32_t, 3> &g, bool capture, double &sec) -> RPhysResult {
        std::memcpy(b->scene->grids, g.data(), sizeof(uint32_t) * 3);
        RPhysResult rr;
        if (capture && ((rr = batch_restore(b, TUNE_SLOT)) != RPHYS_OK || (rr = batch_step(b, dt, ticks, true)) != RPHYS_OK)) return rr;
        if ((rr = batch_restore(b, TUNE_SLOT)) != RPHYS_OK) return rr;
        const auto t0 = std::chrono::steady_clock::now();
        if ((rr = batch_step(b, dt, ticks, true)) != RPHYS_OK) return rr;
        sec = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        if (!device_hash(*b->scene, got.data(), b->scene->d.nw)) return RPHYS_ERR_INTERNAL;
        if (ref.empty()) ref = got;
        else if (got != ref) { set_err("rphys_batch_autotune: a launch shape changed the results"); return RPHYS_ERR_INTERNAL; }
        return RPHYS_OK;
    };
    std::vector<std::pair<double, std::array<uint32_t, 3>>> res;
    for (uint32_t g0 : G0)
        for (uint32_t g1 : G1)
            for (uint32_t g2 : G2) {
                double sec = 0.0;
                if ((r = timed({g0, g1, g2}, true, sec)) != RPHYS_OK) return r;
                res.push_back({sec, {g0, g1, g2}});
            }
    // one run is noisy: the four fastest shapes and the default shape run three more times each,
    // and the best of each shape's runs decides
    std::sort(res.begin(), res.end(), [](const auto &x, const auto &y) { return x.first < y.first; });
    const std::array<uint32_t, 3> dflt = {kDefaultGrids[0], kDefaultGrids[1], kDefaultGrids[2]};
    std::vector<std::pair<double, std::array<uint32_t, 3>>> lead(res.begin(), res.begin() + std::min<size_t>(4, res.size()));
    for (const auto &x : res)
        if (x.second == dflt && std::none_of(lead.begin(), lead.end(), [&](const auto &l) { return l.second == dflt; })) lead.push_back(x);
    for (auto &l : lead) {
        for (int k = 0; k < 3; k++) {
            double sec = 0.0;
            if ((r = timed(l.second, k == 0, sec)) != RPHYS_OK) return r;
            l.first = std::min(l.first, sec);
        }
    }
    const auto win = *std::min_element(lead.begin(), lead.end(), [](const auto &x, const auto &y) { return x.first < y.first; });
    const double best_s = win.first;
    const uint32_t bg[3] = {win.second[0], win.second[1], win.second[2]};
    std::memcpy(b->scene->grids, bg, sizeof bg);
    if ((r = batch_restore(b, TUNE_SLOT)) != RPHYS_OK) return r;
    best->size_bytes = sizeof(RPhysTuning);
    best->grid_small = bg[0]; best->grid_large = bg[1]; best->grid_huge = bg[2];
    *best_ms_per_tick = 1000.0 * best_s / ticks;
    return RPHYS_OK;
    GUARD_END
}

// ------------------------------------------------------------------ learned warm start
uint32_t rphys_mlp_param_count(void) { return RP_MLP_NPARAM; }
uint32_t rphys_mlp_gradient_rows(void) { return RP_MLP_STEP_CONTACTS; }

static RPhysResult check_mode(const char *ctx, uint32_t mode, const float *params) {
    if (mode > 2) return bad_arg(ctx, "mode must be 0 (previous frame), 1 (zero) or 2 (learned)");
    if (mode == 2 && !params) { set_err(std::string(ctx) + ": the learned mode needs parameters"); return RPHYS_ERR_NULL; }
    if (mode == 2)
        for (uint32_t i = 0; i < RP_MLP_NPARAM; i++)
            if (!hfin(params[i])) return bad_arg(ctx, "parameters must be finite");
    if (mode == 2)
        for (uint32_t k = 0; k < 12; k++)
            if (!(params[RP_MLP_XSTD + k] > 0.0f)) return bad_arg(ctx, "x_std must be positive");
    return RPHYS_OK;
}

RPhysResult rphys_world_set_warm_start(RPhysWorld *world, uint32_t mode, const float *params) {
    GUARD_BEGIN
    if (!world) { set_err("rphys_world_set_warm_start: world is null"); return RPHYS_ERR_NULL; }
    RPhysResult r = check_mode("rphys_world_set_warm_start", mode, params);
    if (r != RPHYS_OK) return r;
    std::lock_guard<std::mutex> lk(world->mu);
    world->w.ws_mode = mode;
    if (mode == 2) {
        world->w.mlp.assign(params, params + RP_MLP_NPARAM);
        if (world->scene && !upload_model(*world->scene, params)) return RPHYS_ERR_INTERNAL;
    }
    return RPHYS_OK;
    GUARD_END
}

RPhysResult rphys_batch_set_warm_start(RPhysBatch *b, uint32_t mode, const float *params) {
    GUARD_BEGIN
    if (!b) { set_err("rphys_batch_set_warm_start: batch is null"); return RPHYS_ERR_NULL; }
    RPhysResult r = check_mode("rphys_batch_set_warm_start", mode, params);
    if (r != RPHYS_OK) return r;
    for (auto &w : b->worlds) {
        w.ws_mode = mode;
        if (mode == 2) w.mlp.assign(params, params + RP_MLP_NPARAM);
    }
    if (mode == 2 && !upload_model(*b->scene, params)) return RPHYS_ERR_INTERNAL;
    return RPHYS_OK;
    GUARD_END
}

RPhysResult rphys_batch_set_residual_mode(RPhysBatch *b, uint32_t on, float tol, uint32_t max_iters) {
    GUARD_BEGIN
    if (!b) { set_err("rphys_batch_set_residual_mode: batch is null"); return RPHYS_ERR_NULL; }
    if (on && (!hfin(tol) || tol < 0.0f || max_iters == 0)) return bad_arg("rphys_batch_set_residual_mode", "tol must be finite and >= 0, max_iters > 0");
    for (auto &w : b->worlds) {
        w.resid_on = on ? 1u : 0u;
        if (on) { w.resid_tol = tol; w.resid_max = max_iters; }
    }
    return RPHYS_OK;
    GUARD_END
}

RPhysResult rphys_batch_set_contact_log(RPhysBatch *b, uint32_t capacity_rows) {
    GUARD_BEGIN
    if (!b) { set_err("rphys_batch_set_contact_log: batch is null"); return RPHYS_ERR_NULL; }
    if (!hip_ok(hipStreamSynchronize(b->scene->st), "sync")) return RPHYS_ERR_INTERNAL;
    Dev &d = b->scene->d;
    d.log_feat = nullptr; d.log_targ = nullptr; d.log_n = nullptr; d.log_cap = 0;
    for (void *q : b->logbuf.ptrs) (void)hipFree(q);
    b->logbuf.ptrs.clear();
    if (capacity_rows == 0) return RPHYS_OK;
    float *f = nullptr, *t = nullptr;
    uint32_t *n = nullptr;
    if (!b->logbuf.alloc(f, 12 * (size_t)capacity_rows) || !b->logbuf.alloc(t, 3 * (size_t)capacity_rows) || !b->logbuf.alloc(n, 2)) {
        set_err("rphys_batch_set_contact_log: device allocation failed");
        return RPHYS_ERR_CAPACITY;
    }
    if (!hip_ok(hipMemset(n, 0, 8), "memset")) return RPHYS_ERR_INTERNAL;
    d.log_feat = f; d.log_targ = t; d.log_n = n; d.log_cap = capacity_rows;
    return RPHYS_OK;
    GUARD_END
}

RPhysResult rphys_batch_contact_log_rows(RPhysBatch *b, uint32_t *rows, uint32_t *dropped) {
    GUARD_BEGIN
    if (!b || !rows) { set_err("rphys_batch_contact_log_rows: null arg"); return RPHYS_ERR_NULL; }
    *rows = 0;
    if (dropped) *dropped = 0;
    if (!b->scene->d.log_n) return RPHYS_OK;
    uint32_t n[2];
    if (!hip_ok(hipStreamSynchronize(b->scene->st), "sync") || !hip_ok(hipMemcpy(n, b->scene->d.log_n, 8, hipMemcpyDeviceToHost), "log rows"))
        return RPHYS_ERR_INTERNAL;
    *rows = n[0];
    if (dropped) *dropped = n[1];
    return RPHYS_OK;
    GUARD_END
}

RPhysResult rphys_batch_read_contact_log(RPhysBatch *b, uint32_t first, uint32_t count, float *feat12, float *targ3) {
    GUARD_BEGIN
    if (!b || !feat12 || !targ3) { set_err("rphys_batch_read_contact_log: null arg"); return RPHYS_ERR_NULL; }
    uint32_t rows = 0;
    RPhysResult r = rphys_batch_contact_log_rows(b, &rows, nullptr);
    if (r != RPHYS_OK) return r;
    if ((uint64_t)first + count > rows) return bad_arg("rphys_batch_read_contact_log", "range beyond the logged rows");
    const Dev &d = b->scene->d;
    if (count && (!hip_ok(hipMemcpy(feat12, d.log_feat + 12 * (size_t)first, 48 * (size_t)count, hipMemcpyDeviceToHost), "log read") ||
                  !hip_ok(hipMemcpy(targ3, d.log_targ + 3 * (size_t)first, 12 * (size_t)count, hipMemcpyDeviceToHost), "log read")))
        return RPHYS_ERR_INTERNAL;
    return RPHYS_OK;
    GUARD_END
}

// Glorot-uniform weights from XorShift64(seed), zero biases; normalisation from the data.
static void init_params(uint64_t seed, const double *xm, const double *xs, const double *ys, float *p) {
    uint64_t st = seed ? seed : 0x9E3779B97F4A7C15ull;
    auto next = [&]() { st ^= st << 13; st ^= st >> 7; st ^= st << 17; return (double)(st >> 11) * (1.0 / 9007199254740992.0); };
    auto fill = [&](uint32_t off, uint32_t n, uint32_t fan_in, uint32_t fan_out) {
        double a = std::sqrt(6.0 / (double)(fan_in + fan_out));
        for (uint32_t i = 0; i < n; i++) p[off + i] = (float)((2.0 * next() - 1.0) * a);
    };
    std::memset(p, 0, RP_MLP_NPARAM * sizeof(float));
    fill(RP_MLP_W1, 64 * 12, 12, 64);
    fill(RP_MLP_W2, 64 * 64, 64, 64);
    // the output layer starts at zero: the untrained model is the previous-frame warm start
    for (int k = 0; k < 12; k++) { p[RP_MLP_XMEAN + k] = (float)xm[k]; p[RP_MLP_XSTD + k] = (float)xs[k]; }
    for (int k = 0; k < 3; k++) p[RP_MLP_YSCALE + k] = (float)ys[k];
}

static uint64_t gcd64(uint64_t a, uint64_t b) { while (b) { uint64_t t = a % b; a = b; b = t; } return a; }

// Runs the trainer on device data (feat [n][12], targ [n][3]); params in/out on the host.
static bool train_on_device(const float *feat, const float *targ, uint64_t n, uint32_t epochs, uint32_t steps_per_epoch, float lr, uint64_t seed,
                            float *params, float *epoch_loss, bool grad_only, float *grad_out) {
    DeviceBuffers B;
    float *p = nullptr, *m = nullptr, *v = nullptr, *loss = nullptr, *grad = nullptr;
    uint64_t *perm = nullptr;
    if (!B.alloc(p, RP_MLP_NPARAM) || !B.alloc(m, RP_MLP_NW) || !B.alloc(v, RP_MLP_NW) || !B.alloc(loss, epochs) || !B.alloc(grad, RP_MLP_NW) ||
        !B.alloc(perm, 2 * (size_t)epochs)) {
        set_err("training: device allocation failed");
        return false;
    }
    std::vector<uint64_t> pr(2 * (size_t)epochs);
    uint64_t st = seed ^ 0xD1B54A32D192ED03ull;
    auto nxt = [&]() { st ^= st << 13; st ^= st >> 7; st ^= st << 17; return st; };
    for (uint32_t e = 0; e < epochs; e++) {
        uint64_t a = 1, b = 0;
        if (!grad_only && n > 1) {
            do { a = 1 + nxt() % (n - 1); } while (gcd64(a, n) != 1);
            b = nxt() % n;
        }
        pr[2 * e] = a; pr[2 * e + 1] = b;
    }
    bool ok = hip_ok(hipMemcpy(p, params, RP_MLP_NPARAM * sizeof(float), hipMemcpyHostToDevice), "params") &&
              hip_ok(hipMemset(m, 0, RP_MLP_NW * sizeof(float)), "memset") && hip_ok(hipMemset(v, 0, RP_MLP_NW * sizeof(float)), "memset") &&
              hip_ok(hipMemcpy(perm, pr.data(), pr.size() * 8, hipMemcpyHostToDevice), "perm");
    if (!ok) return false;
    TrainArgs A{};
    A.p = p; A.m = m; A.v = v; A.feat = feat; A.targ = targ; A.n = n; A.perm = perm;
    A.epochs = epochs; A.steps_per_epoch = steps_per_epoch; A.lr = lr; A.beta1 = 0.9f; A.beta2 = 0.999f; A.eps = 1.0e-8f;
    A.loss = loss; A.grad = grad; A.grad_only = grad_only ? 1u : 0u;
    rp_launch_mlp_train(A, nullptr);
    ok = hip_ok(hipGetLastError(), "trainer launch") && hip_ok(hipDeviceSynchronize(), "training") &&
         hip_ok(hipMemcpy(params, p, RP_MLP_NPARAM * sizeof(float), hipMemcpyDeviceToHost), "params") &&
         hip_ok(hipMemcpy(epoch_loss, loss, epochs * sizeof(float), hipMemcpyDeviceToHost), "loss");
    if (ok && grad_only) ok = hip_ok(hipMemcpy(grad_out, grad, RP_MLP_NW * sizeof(float), hipMemcpyDeviceToHost), "grad");
    return ok;
}

RPhysResult rphys_batch_train_warm_start(RPhysBatch *b, const RPhysMlpTrainConfig *cfg, float *params_out, float *epoch_loss) {
    GUARD_BEGIN
    if (!b || !cfg || !params_out || !epoch_loss) { set_err("rphys_batch_train_warm_start: null arg"); return RPHYS_ERR_NULL; }
    if (cfg->size_bytes < sizeof(RPhysMlpTrainConfig)) return bad_arg("rphys_batch_train_warm_start", "size_bytes too small");
    if (cfg->epochs == 0 || !hfin(cfg->lr) || cfg->lr <= 0.0f) return bad_arg("rphys_batch_train_warm_start", "epochs > 0 and a positive lr required");
    uint32_t rows = 0;
    RPhysResult r = rphys_batch_contact_log_rows(b, &rows, nullptr);
    if (r != RPHYS_OK) return r;
    if (rows < RP_MLP_STEP_CONTACTS) return bad_arg("rphys_batch_train_warm_start", "the contact log holds fewer rows than one training step");
    // normalisation from the logged data (f64 sums)
    const Dev &d = b->scene->d;
    std::vector<float> f(12 * (size_t)rows), t(3 * (size_t)rows);
    if (!hip_ok(hipMemcpy(f.data(), d.log_feat, f.size() * 4, hipMemcpyDeviceToHost), "log") ||
        !hip_ok(hipMemcpy(t.data(), d.log_targ, t.size() * 4, hipMemcpyDeviceToHost), "log"))
        return RPHYS_ERR_INTERNAL;
    double xm[12] = {0}, xs[12] = {0}, ys[3] = {0};
    for (size_t i = 0; i < rows; i++)
        for (int k = 0; k < 12; k++) xm[k] += f[12 * i + k];
    for (int k = 0; k < 12; k++) xm[k] /= (double)rows;
    for (size_t i = 0; i < rows; i++)
        for (int k = 0; k < 12; k++) { double e = f[12 * i + k] - xm[k]; xs[k] += e * e; }
    for (size_t i = 0; i < rows; i++)
        for (int k = 0; k < 3; k++) ys[k] += (double)t[3 * i + k] * t[3 * i + k];
    // a feature (nearly) constant in the data keeps a small scale, so other values of it fall
    // outside the gate and take the previous-frame warm start
    for (int k = 0; k < 12; k++) xs[k] = std::max(std::sqrt(xs[k] / (double)rows), std::max(1.0e-3 * std::fabs(xm[k]), 1.0e-6));
    for (int k = 0; k < 3; k++) { ys[k] = std::sqrt(ys[k] / (double)rows); if (!(ys[k] > 1.0e-6)) ys[k] = 1.0; }
    init_params(cfg->seed, xm, xs, ys, params_out);
    uint32_t spe = rows / RP_MLP_STEP_CONTACTS;
    return train_on_device(d.log_feat, d.log_targ, rows, cfg->epochs, spe, cfg->lr, cfg->seed, params_out, epoch_loss, false, nullptr)
               ? RPHYS_OK : RPHYS_ERR_INTERNAL;
    GUARD_END
}

RPhysResult rphys_mlp_forward(const float *params, const float *feat12, uint32_t n, float *out3) {
    GUARD_BEGIN
    if (!params || (n && (!feat12 || !out3))) { set_err("rphys_mlp_forward: null arg"); return RPHYS_ERR_NULL; }
    RPhysResult r = check_mode("rphys_mlp_forward", 2, params);
    if (r != RPHYS_OK || n == 0) return r;
    DeviceBuffers B;
    float *p = nullptr, *f = nullptr, *o = nullptr;
    __bf16 *w1p, *w2p, *w2t, *w3p, *w3t;
    if (!B.alloc(p, RP_MLP_NPARAM) || !B.alloc(f, 12 * (size_t)n) || !B.alloc(o, 4 * (size_t)n) || !B.alloc(w1p, RP_MLP_W1P) || !B.alloc(w2p, RP_MLP_W2P) ||
        !B.alloc(w2t, RP_MLP_W2P) || !B.alloc(w3p, RP_MLP_W3P) || !B.alloc(w3t, RP_MLP_W3T)) {
        set_err("rphys_mlp_forward: device allocation failed");
        return RPHYS_ERR_CAPACITY;
    }
    std::vector<float> host(4 * (size_t)n);
    bool ok = hip_ok(hipMemcpy(p, params, RP_MLP_NPARAM * 4, hipMemcpyHostToDevice), "upload") &&
              hip_ok(hipMemcpy(f, feat12, 48 * (size_t)n, hipMemcpyHostToDevice), "upload");
    if (ok) {
        rp_launch_mlp_prep(p, w1p, w2p, w2t, w3p, w3t, nullptr);
        rp_launch_mlp_infer(p, w1p, w2p, w3p, f, nullptr, n, o, nullptr, nullptr);
        ok = hip_ok(hipGetLastError(), "launch") && hip_ok(hipDeviceSynchronize(), "forward") &&
             hip_ok(hipMemcpy(host.data(), o, host.size() * 4, hipMemcpyDeviceToHost), "download");
    }
    if (!ok) return RPHYS_ERR_INTERNAL;
    for (uint32_t i = 0; i < n; i++)
        for (int k = 0; k < 3; k++) out3[3 * (size_t)i + k] = host[4 * (size_t)i + k];
    return RPHYS_OK;
    GUARD_END
}

// Test plumbing for device-pointer entry points: plain device buffers and copies.
RPhysResult rphys_debug_device_alloc(uint64_t bytes, void **out) {
    if (!out) { set_err("rphys_debug_device_alloc: null arg"); return RPHYS_ERR_NULL; }
    *out = nullptr;
    return hip_ok(hipMalloc(out, bytes ? bytes : 1), "rphys_debug_device_alloc") ? RPHYS_OK : RPHYS_ERR_CAPACITY;
}
RPhysResult rphys_debug_device_free(void *p) { return hip_ok(hipFree(p), "rphys_debug_device_free") ? RPHYS_OK : RPHYS_ERR_INTERNAL; }
RPhysResult rphys_debug_device_copy(void *dst, const void *src, uint64_t bytes, uint32_t to_host) {
    if (!dst || !src) { set_err("rphys_debug_device_copy: null arg"); return RPHYS_ERR_NULL; }
    return hip_ok(hipMemcpy(dst, src, bytes, to_host ? hipMemcpyDeviceToHost : hipMemcpyHostToDevice), "rphys_debug_device_copy")
               ? RPHYS_OK : RPHYS_ERR_INTERNAL;
}

RPhysResult rphys_mlp_gradient(const float *params, const float *feat12, const float *targ3, float *loss, float *grad) {
    GUARD_BEGIN
    if (!params || !feat12 || !targ3 || !loss || !grad) { set_err("rphys_mlp_gradient: null arg"); return RPHYS_ERR_NULL; }
    RPhysResult r = check_mode("rphys_mlp_gradient", 2, params);
    if (r != RPHYS_OK) return r;
    const uint32_t n = RP_MLP_STEP_CONTACTS;
    DeviceBuffers B;
    float *f = nullptr, *t = nullptr;
    if (!B.alloc(f, 12 * (size_t)n) || !B.alloc(t, 3 * (size_t)n)) { set_err("rphys_mlp_gradient: device allocation failed"); return RPHYS_ERR_CAPACITY; }
    if (!hip_ok(hipMemcpy(f, feat12, 48 * (size_t)n, hipMemcpyHostToDevice), "upload") || !hip_ok(hipMemcpy(t, targ3, 12 * (size_t)n, hipMemcpyHostToDevice), "upload"))
        return RPHYS_ERR_INTERNAL;
    std::vector<float> p(params, params + RP_MLP_NPARAM);
    return train_on_device(f, t, n, 1, 1, 1.0e-3f, 0, p.data(), loss, true, grad) ? RPHYS_OK : RPHYS_ERR_INTERNAL;
    GUARD_END
}

}  // extern "C"
