/* Oracle (TEST INFRASTRUCTURE ONLY): C restatement of the RiftPhys reference step.
 *
 * Follows, in evaluation order:
 *   World::step                       crates/riftphys-world/src/world/step.rs:40-420
 *   S1 forces + gravity::eval Uniform step.rs:53-68, crates/riftphys-gravity/src/gravity.rs:79-88
 *   update_colliders_aabb / aabb_of   world.rs:793-798, crates/riftphys-geom/src/shape.rs:18-76,
 *                                     crates/riftphys-geom/src/aabb.rs:21-31
 *   sweep + swept AABB                step.rs:75-81, world.rs:97-100
 *   GridBroadphase::run               crates/riftphys-collision/src/broadphase.rs:29-99 (restated
 *                                     as its equivalent: the sorted set of inclusively overlapping
 *                                     (i<j) pairs of colliders with finite min.x, plus oversized)
 *   narrowphase (pairs, terrain)      step.rs:95-142, narrowphase.rs (or_narrow.h)
 *   HeightField::query_aabb           crates/riftphys-terrain/src/terrain.rs:300-359
 *   cull_and_orient_contacts          step.rs:499-542
 *   build_islands (non-empty test)    islands.rs:27-92
 *   jointed-pair exclusion            step.rs:85-93, 102
 *   solve_islands_parallel            solver.rs:477-1008, 1247-1271; ball-joint
 *                                     rows, warm start and store solver.rs:272-307, 336-424,
 *                                     644-715, 788-835, 945-952 (articulation bodies.rs:326-331)
 *   compliant (Hunt-Crossley) rows   solver.rs:545-567, 848-856; left out of the warm start and
 *                                     store, shock, W6 supports, projection and grounding
 *                                     (solver.rs:736, 922, 973, 1020, 1087, 1175; step.rs:340, 433)
 *   D6 joints (no armature)           solver.rs:262-307, 336-424 (d6_rows), 700-715, 822-832,
 *                                     945-952; bodies.rs:50-53, 321-390, 616-873 (solve_d6),
 *                                     895-922, 1010-1030; step.rs:353-365 (velocity motors)
 *   friction patches                  solver.rs:219-260 (patch_radius), 592-619, 763-782, 865,
 *                                     882-915, 927-944, 1257-1264; warm_patches in the hash and
 *                                     snapshots (world.rs:249, 818-821, 896-897, 935, 952)
 *   restitution relax + W6 bounce     solver.rs:569-588 (time-of-impact rows), 995-1245 (relax,
 *                                     velocity pass with the interior-point LCP, pseudo pass),
 *                                     solver.rs:121-212 and bounce_lcp.rs (or_bounce.h)
 *   S4 integrate + gyroscopic turn    step.rs:154-280 (with W6 pseudo-velocities, step.rs:271-277)
 *   S5 joints + contact projection    step.rs:284-310, 423-497; ball joints' solve_d6 translational
 *                                     rows with SHAKE directions bodies.rs:422-460, 562-573, 616-739,
 *                                     dynamics.rs:71-89, 207-210
 *   S6 velocity from corrections      step.rs:312-379
 *   S7 sleep + AABB refresh           sleep.rs:59-80, step.rs:381-384
 *   step_hash + hash_carried          provenance.rs:6-24, world.rs:882-909
 *   snapshots                         ffi/src/lib.rs:1696-1800, world.rs:912-961
 *
 * The PGS here runs serially in contact-index order; the reference's level schedule
 * (solver.rs:621-642) is proven by its own construction to give the same result (constraints
 * in one level share no written body), and the level count is computed for the stats.
 *
 * Parity pins (tests/test_oracle_*.py): the reference's golden bits G1 (six distance-joint state
 * hashes, distance_bits_test.rs:47-63) and G2/G3 (six W6 scene step hashes, restitution_toi_test.rs
 * :596-619 / docs/PARITY_SPEC.md:2120-2126) reproduced bit for bit; the W6 closed forms of
 * restitution_toi_test.rs (e^2 h apexes, the zero-g 1-D collision, a resting box not lifted, side-hit
 * momentum) and the LCP numerics against dense algebra and a brute-force LCP; BLAKE3 published digests; XorShift64 closed form;
 * broadphase vs brute-force O(N^2) overlap; narrowphase unit tests restated from
 * narrowphase.rs:735-776 and collision/tests/box_sphere_inside_test.rs; terrain query yields
 * (terrain.rs:360-394); closed-form free fall and resting heights (contact_resting_test.rs:30-45,
 * reality/src/analytic.rs free fall); invariants (manifold <= 4 points, n.(pb-pa) >= 0,
 * acc_jn >= 0, |jt| <= mu jn, snapshot/restore hash equality); ball joints (test_oracle_joints.py):
 * a rigid pendulum's large-amplitude half period and energy, anchor drift, a hanging body's joint
 * impulse m g dt, jointed pairs without contacts, the SHAKE test's acosf boundary; contact models
 * (test_oracle_contact_models.py): compliant_contact_test.rs's resting sinks m g / (4 k) and F / k,
 * bounces against Hunt-Crossley's ODE, Coulomb sliding and a slope hold on compliant rows;
 * patch_friction_test.rs's spin-down angles with the face's mean radius and the corners' radius,
 * Coulomb sliding, the patch radius against quadrature; D6 joints (test_oracle_d6.py):
 * hinge_angle_test.rs's stops at -1 and -2.2 rad, joint_damping_test.rs's discrete damper decay
 * rate and friction stop angle, joint_energy_test.rs's pendulum period (RK4) and bounded energy, a
 * velocity motor's impulse cap, a position drive's hold, a sprung hinge's static sag.  The hinge
 * twist's atan2 is the same double evaluation.  The patch point order's atan2 is a
 * once-rounded double evaluation (or_math.h), within 1 ulp of the platform's atan2f; only the order
 * of the angles' keys reaches the results.  glam's SSE2
 * operation order and the platform libm are third-party (glam 0.27.0, libm 0.2.15); G1-G3 pin them
 * on the paths they cover (box-box contacts on a static box, distance joints, W6).  No reference
 * golden holds a heightfield contact, a sphere or capsule contact or a ball joint, so bit-level
 * agreement with the Rust build there is "parity unpinned": those paths are pinned by the closed
 * forms and invariants above.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include "oracle.h"
#include "or_math.h"
#include "or_narrow.h"
#include "or_bounce.h"

#define TERRAIN_CI 0xFFFFFFFFu
static const float MARGIN = 0.05f;     /* step.rs:12 MELEE_CCD_MARGIN */
static const float POS_SLOP = 0.005f;  /* step.rs:16 */
/* step.rs:14 / solver.rs:18 PAR_MIN: below this many items the reference runs a batch on the
 * calling thread; at or above it rayon splits it.  The oracle's OpenMP regions (narrowphase pairs,
 * terrain colliders, the PGS level batches) use the same threshold; each writes only its own
 * items, so the results are the single-threaded ones bit for bit (or_set_threads). */
#define PAR_MIN 64u
static int g_threads = 1;
void or_set_threads(int n) { g_threads = n > 0 ? n : 1; }

/* ------------------------------------------------------------------ quantisers */
/* det.rs:8-13 q6: (x / 1e-6).round() * 1e-6, non-finite -> 0 */
static float q6(float x) { if (!isfinite(x)) return 0.0f; return roundf(x / 1.0e-6f) * 1.0e-6f; }
/* determinism.rs:57-65 qf32 with q = 1e-6 */
static float qf32(float x) { if (!isfinite(x)) return NAN; return roundf(x / 1.0e-6f) * 1.0e-6f; }
/* materials.rs:319-321 quantize: (x * 1e6).round() * 1e-6 */
static float mquant(float x) { return roundf(x * 1.0e6f) * 1.0e-6f; }
/* gravity.rs:43-49 */
static float gq6(float x) { return isfinite(x) ? roundf(x * 1.0e6f) * 1.0e-6f : 0.0f; }
static uint32_t canon_bits(float x) {
    if (!isfinite(x)) return 0x7fc00000u;
    uint32_t b; memcpy(&b, &x, 4);
    return b == 0x80000000u ? 0u : b;
}
static float fclampf(float x, float lo, float hi) { if (x < lo) x = lo; if (x > hi) x = hi; return x; }

/* ------------------------------------------------------------------ materials */
/* materials.rs:93-240: the core entries (ids 0..15) the FFI can reach: Default and Grit */
typedef struct { float mu_s, mu_k, rest, vs; } MatP;
static const MatP MATS[16] = {
    {0.60f, 0.50f, 0.10f, 0.08f}, {0.05f, 0.03f, 0.02f, 0.02f}, {1.10f, 0.95f, 0.80f, 0.15f},
    {0.70f, 0.55f, 0.50f, 0.05f}, {1.30f, 1.05f, 0.00f, 0.12f}, {0.55f, 0.45f, 0.08f, 0.10f},
    {0.80f, 0.70f, 0.00f, 0.10f}, {0.90f, 0.80f, 0.00f, 0.12f}, {0.80f, 0.60f, 0.20f, 0.10f},
    {0.55f, 0.45f, 0.10f, 0.06f}, {1.10f, 0.95f, 0.00f, 0.18f}, {0.40f, 0.30f, 0.65f, 0.04f},
    {0.90f, 0.60f, 0.35f, 0.06f}, {0.61f, 0.47f, 0.40f, 0.05f}, {0.50f, 0.40f, 0.12f, 0.08f},
    {0.20f, 0.15f, 0.85f, 0.04f}};
enum { MAT_DEFAULT = 0, MAT_ICE = 1, MAT_RUBBER = 2, MAT_STEEL = 3, MAT_GRIT = 4, MAT_CONCRETE = 6, MAT_ASPHALT = 7 };

/* The solver's pair properties (materials.rs:75-90 MatPairProps): friction, restitution, Stribeck
 * speed, and the contact model: stiffness k > 0 makes the pair compliant (Hunt-Crossley, with
 * dissipation d), patch solves a manifold's friction as one patch.  The table's pairs are rigid. */
typedef struct { float mu_s, mu_k, rest, vs, stiff, diss; int patch; } PairP;
/* materials.rs:278-304 mix_pair + pair_props + override_for */
static PairP table_pair(uint32_t a, uint32_t b) {
    MatP pa = MATS[a & 15], pb = MATS[b & 15];
    PairP r;
    r.stiff = 0.0f; r.diss = 0.0f; r.patch = 0;
    r.mu_s = mquant(sqrtf(fabsf(pa.mu_s * pb.mu_s)));
    r.mu_k = mquant(sqrtf(fabsf(pa.mu_k * pb.mu_k)));
    r.rest = mquant(fmaxf(pa.rest, pb.rest));
    r.vs = mquant(0.5f * (pa.vs + pb.vs));
    uint32_t x = a < b ? a : b, y = a < b ? b : a;
    int ov = 0; PairP o = r;
    if (x == MAT_ICE && y == MAT_RUBBER) { ov = 1; o.mu_s = 0.20f; o.mu_k = 0.10f; o.rest = 0.05f; o.vs = 0.08f; }
    if (x == MAT_RUBBER && y == MAT_ASPHALT) { ov = 1; o.mu_s = 1.10f; o.mu_k = 0.95f; o.rest = 0.00f; o.vs = 0.18f; }
    if (x == MAT_RUBBER && y == MAT_CONCRETE) { ov = 1; o.mu_s = 1.00f; o.mu_k = 0.85f; o.rest = 0.00f; o.vs = 0.16f; }
    if (x == MAT_STEEL && y == MAT_STEEL) { ov = 1; o.mu_s = 0.58f; o.mu_k = 0.50f; o.rest = 0.08f; o.vs = 0.04f; }
    if (ov) { o.mu_s = mquant(o.mu_s); o.mu_k = mquant(o.mu_k); o.rest = mquant(o.rest); o.vs = mquant(o.vs); return o; }
    return r;
}

/* ------------------------------------------------------------------ world */
typedef struct { V3 min, max; } Aabb;
typedef struct {
    uint32_t a, b;  /* collider indices; b = TERRAIN_CI for terrain */
    V3 point, normal;
    float depth, impulse;
    uint32_t env_mat, feat, cell;
} Contact;
typedef struct { uint32_t key[4]; V3 la; float jn; V3 jt; } Warm;
/* A friction patch's carried impulses (solver.rs WarmPatch): sliding jt (world frame), torsion tw */
typedef struct { uint32_t key[4]; V3 jt; float tw; } WarmP;
typedef struct { uint32_t a, b; PairP p; } PairOverride;
/* A ball joint (articulation/src/bodies.rs:326-331 add_ball: the three translational axes of frame
 * A locked, the rotational ones free) between bodies a and b, with body-local frames (pos, rot);
 * or a distance joint between the two centres of mass (bodies.rs:29-50, 273-285: no anchors and
 * no damping, the rows the reference keeps bit for bit), with rest length, compliance (m/N, 0 =
 * rigid) and the step's XPBD multiplier. */
enum { J_BALL = 0, J_DIST = 1, J_D6 = 2 };
/* One axis of a D6 joint (bodies.rs:52-104 D6Axis): enabled (constrained), compliance (XPBD, 0 =
 * rigid), an optional drive (target, kp, kd, max impulse, velocity mode) and limit (min, max),
 * damping and Coulomb friction, and the step's XPBD multiplier.  No armature. */
typedef struct {
    uint8_t enabled, has_drive, vel_mode, has_limit;
    float compliance, damping, friction;
    float d_target, d_kp, d_kd, d_max;
    float lim_min, lim_max;
    float lam;
} D6Ax;
/* A D6 joint (bodies.rs:107-136 Generic6Dof) keeps t[3] (translational axes of frame A) and r[3]
 * (rotational), and its stored hinge axis (-1: derived from the rows, d6_hinge_axis). */
typedef struct { uint32_t a, b; V3 fap, fbp; Q4 fra, frb; int kind; float rest, compliance, lam; D6Ax t[3], r[3]; int hinge; } Joint;
/* A joint's warm rows (world.rs:250 warm_joints): row tags and impulses of its last solve (a D6
 * joint has at most 5 rows per axis). */
#define OR_JROWS_MAX 32
typedef struct { uint8_t has, n, tag[OR_JROWS_MAX]; float lam[OR_JROWS_MAX]; } WarmJ;
typedef struct { uint32_t body; float radius, height, speed, vvel; int grounded; V3 input; } OrPlayer;

struct OrWorld {
    uint32_t cap, n;
    V3 *pos; Q4 *rot; V3 *lin, *ang; float *inv_mass; uint8_t *dyn; M3 *iil;
    uint8_t *alive; V3 *last_lin;
    uint8_t *awake; uint16_t *sticks; uint8_t *sallow;
    /* colliders (one per body, in body order, removed bodies drop out) */
    uint32_t nc; uint32_t *cbody; ShapeD *cshape; uint32_t *cmat; Aabb *caabb;
    /* heightfield */
    int has_hf; uint32_t hnx, hnz; float hsx, hsz, hox, hoz, hmin, hmax; float *hts; uint32_t hmat;
    V3 gravity; uint64_t epoch, rng, tick;
    uint32_t iv, ip;
    int resid_on; float resid_tol; uint32_t resid_max;
    PairOverride ovr[32]; uint32_t novr;
    int damp_at_start;  /* world.rs:992 set_compliant_damping_at_start */
    Warm *warm; uint32_t nwarm;
    WarmP *wpatch; uint32_t nwpatch;  /* world.rs:249 warm_patches, sorted by key */
    Joint *joints; WarmJ *wj; uint32_t nj, capj;
    /* per-step outputs */
    uint32_t *pairs; uint32_t npairs;
    Contact *contacts; uint32_t ncontacts;
    uint8_t *loaded; int loaded_valid;
    /* W6 pseudo-velocities of the last solve (linear, angular), used by S4 when has_pseudo */
    V3 *pl, *pa; int has_pseudo;
    /* kinematic players (world.rs:44-69 PlayerController) */
    OrPlayer *players; uint32_t np, capp;
    OrStats stats;
};

static void *xrealloc(void *p, size_t n) { void *q = realloc(p, n ? n : 1); if (!q) abort(); return q; }

OrWorld *or_world_create(uint32_t cap) {
    OrWorld *w = (OrWorld *)calloc(1, sizeof(OrWorld));
    w->cap = cap;
    w->pos = calloc(cap + 1, sizeof(V3)); w->rot = calloc(cap + 1, sizeof(Q4));
    w->lin = calloc(cap + 1, sizeof(V3)); w->ang = calloc(cap + 1, sizeof(V3));
    w->inv_mass = calloc(cap + 1, 4); w->dyn = calloc(cap + 1, 1); w->iil = calloc(cap + 1, sizeof(M3));
    w->alive = calloc(cap + 1, 1); w->last_lin = calloc(cap + 1, sizeof(V3));
    w->awake = calloc(cap + 1, 1); w->sticks = calloc(cap + 1, 2); w->sallow = calloc(cap + 1, 1);
    for (uint32_t i = 0; i < cap; i++) { w->awake[i] = 1; w->sallow[i] = 1; }
    w->cbody = calloc(cap + 1, 4); w->cshape = calloc(cap + 1, sizeof(ShapeD));
    w->cmat = calloc(cap + 1, 4); w->caabb = calloc(cap + 1, sizeof(Aabb));
    w->loaded = calloc(cap + 1, 1);
    w->pl = calloc(cap + 1, sizeof(V3)); w->pa = calloc(cap + 1, sizeof(V3));
    w->gravity = v3(0.0f, -9.81f, 0.0f);
    w->rng = 0xC0FFEEull | 1ull; /* world.rs:413 XorShift64::new(0xC0FFEE) */
    w->iv = 20; w->ip = 20;      /* world.rs:425 */
    return w;
}

void or_world_destroy(OrWorld *w) {
    if (!w) return;
    free(w->joints); free(w->wj);
    free(w->pos); free(w->rot); free(w->lin); free(w->ang); free(w->inv_mass); free(w->dyn); free(w->iil);
    free(w->alive); free(w->last_lin); free(w->awake); free(w->sticks); free(w->sallow); free(w->pl); free(w->pa); free(w->players);
    free(w->cbody); free(w->cshape); free(w->cmat); free(w->caabb); free(w->loaded);
    free(w->hts); free(w->warm); free(w->wpatch); free(w->pairs); free(w->contacts);
    free(w);
}

void or_set_epoch(OrWorld *w, uint64_t e) { w->epoch = e; }
void or_set_rng_seed(OrWorld *w, uint64_t s) { w->rng = s | 1ull; }
void or_set_gravity(OrWorld *w, float x, float y, float z) { w->gravity = v3(x, y, z); }
/* world.rs:994-996 */
void or_set_iterations(OrWorld *w, uint32_t v, uint32_t p) { w->iv = v ? v : 1; w->ip = p ? p : 1; }
void or_set_residual_mode(OrWorld *w, int on, float tol, uint32_t max_iters) { w->resid_on = on; w->resid_tol = tol; w->resid_max = max_iters; }
/* world.rs:975-986 */
static PairP pair_props(const OrWorld *w, uint32_t a, uint32_t b);
/* The pair's override, created from its current properties when absent (NULL when full). */
static PairP *pair_override(OrWorld *w, uint32_t a, uint32_t b) {
    uint32_t x = a < b ? a : b, y = a < b ? b : a;
    for (uint32_t i = 0; i < w->novr; i++) if (w->ovr[i].a == x && w->ovr[i].b == y) return &w->ovr[i].p;
    if (w->novr >= 32) return NULL;
    /* the reference keeps overrides in a Vec and finds the first match; order of distinct keys is irrelevant */
    PairOverride o; o.a = x; o.b = y; o.p = pair_props(w, a, b);
    w->ovr[w->novr++] = o;
    return &w->ovr[w->novr - 1].p;
}
/* friction and restitution of a pair (its contact model stays) */
void or_set_pair_props(OrWorld *w, uint32_t a, uint32_t b, float mu_s, float mu_k, float rest, float vs) {
    PairP *p = pair_override(w, a, b);
    if (!p) return;
    p->mu_s = mu_s; p->mu_k = mu_k; p->rest = rest; p->vs = vs;
}
/* the pair's contact model (materials.rs:75-90 stiffness, dissipation, patch_friction; its friction
 * and restitution stay) */
void or_set_pair_contact_model(OrWorld *w, uint32_t a, uint32_t b, float stiffness, float dissipation, int patch) {
    PairP *p = pair_override(w, a, b);
    if (!p) return;
    p->stiff = stiffness; p->diss = dissipation; p->patch = patch != 0;
}
void or_set_compliant_damping_at_start(OrWorld *w, int on) { w->damp_at_start = on != 0; }
static PairP pair_props(const OrWorld *w, uint32_t a, uint32_t b) {
    uint32_t x = a < b ? a : b, y = a < b ? b : a;
    for (uint32_t i = 0; i < w->novr; i++) if (w->ovr[i].a == x && w->ovr[i].b == y) return w->ovr[i].p;
    return table_pair(a, b);
}
static PairP contact_props(const OrWorld *w, const Contact *c) {
    return pair_props(w, w->cmat[c->a], c->b == TERRAIN_CI ? c->env_mat : w->cmat[c->b]);
}

/* dynamics.rs:14-25 */
static float dfin(float x) { return isfinite(x) ? x : 0.0f; }
static V3 dfin_v(V3 v) { return v3(dfin(v.x), dfin(v.y), dfin(v.z)); }
static Q4 fin_q(Q4 q) {
    if (!qfinite(q)) return qident();
    Q4 qq = qnormalize(q);
    if (!qfinite(qq)) return qident();
    if (qq.w < 0.0f) qq = qneg(qq);
    return qq;
}
/* dynamics.rs:130-147 */
static void b_set_pose(OrWorld *w, uint32_t i, V3 p, Q4 q) {
    if (vfinite(p)) w->pos[i] = dfin_v(p);
    if (qfinite(q)) w->rot[i] = fin_q(q);
}
static void b_set_vel(OrWorld *w, uint32_t i, V3 l, V3 a) {
    if (vfinite(l)) w->lin[i] = dfin_v(l);
    if (vfinite(a)) w->ang[i] = dfin_v(a);
}
/* dynamics.rs:158-165 */
static M3 inv_inertia_world(const OrWorld *w, uint32_t i) {
    if (w->inv_mass[i] == 0.0f) return m3zero();
    M3 r = m3quat(w->rot[i]);
    return m3mul(m3mul(r, w->iil[i]), m3t(r));
}
static M3 inv_inertia_local_of(const OrWorld *w, uint32_t i) {
    if (w->inv_mass[i] == 0.0f) return m3zero();
    return w->iil[i];
}
/* dynamics.rs:193-205 */
static void apply_impulse(OrWorld *w, uint32_t i, V3 j) {
    float im = w->inv_mass[i];
    if (im > 0.0f && vfinite(j)) w->lin[i] = dfin_v(vadd(w->lin[i], vmul(j, im)));
}
static void apply_ang_impulse(OrWorld *w, uint32_t i, V3 tau) {
    if (w->inv_mass[i] > 0.0f && vfinite(tau)) {
        V3 dw = m3v(inv_inertia_world(w, i), tau);
        w->ang[i] = dfin_v(vadd(w->ang[i], dw));
    }
}

/* shape.rs:18-76 + aabb.rs:21-31 */
static Aabb aabb_new(V3 a, V3 b) {
    V3 qa = v3(qf32(a.x), qf32(a.y), qf32(a.z)), qb = v3(qf32(b.x), qf32(b.y), qf32(b.z));
    Aabb r; r.min = vmin(qa, qb); r.max = vmax(qa, qb); return r;
}
static Aabb aabb_of(ShapeD s, V3 p, Q4 q) {
    V3 he;
    if (s.type == OR_SPHERE) he = v3s(fabsf(s.d[0]));
    else if (s.type == OR_BOX) {
        V3 e = v3(fabsf(s.d[0]), fabsf(s.d[1]), fabsf(s.d[2]));
        M3 r = m3quat(q);
        M3 ar = m3cols(vabs(r.c0), vabs(r.c1), vabs(r.c2));
        he = m3v(ar, e);
    } else {
        float rr = fabsf(s.d[0]), h = fabsf(s.d[1]);
        V3 axw = vmul(qrot(q, v3(0, 1, 0)), h);
        he = vadd(vabs(axw), v3s(rr));
    }
    return aabb_new(vsub(p, he), vadd(p, he));
}

/* mass.rs:92-165 (density-based primitives) */
static void mass_props(ShapeD s, float rho, float *mass, V3 *idiag, int *ok) {
    *ok = 0;
    if (s.type == OR_SPHERE) {
        float r = fabsf(s.d[0]);
        if (!(isfinite(r) && r > 0 && isfinite(rho) && rho > 0)) return;
        float r2 = r * r;
        float vol = ((4.0f / 3.0f) * 3.14159265358979323846f * r2) * r;
        float m = rho * vol;
        float ii = (0.4f * m) * r2;
        *mass = m; *idiag = v3s(ii);
    } else if (s.type == OR_BOX) {
        V3 he = v3(fabsf(s.d[0]), fabsf(s.d[1]), fabsf(s.d[2]));
        if (!(vfinite(he) && isfinite(rho) && rho > 0)) return;
        V3 d = vmul(he, 2.0f);
        if (!(d.x > 0 && d.y > 0 && d.z > 0)) return;
        float vol = (d.x * d.y) * d.z;
        float m = rho * vol;
        float x2 = d.x * d.x, y2 = d.y * d.y, z2 = d.z * d.z;
        *mass = m;
        *idiag = v3(((1.0f / 12.0f) * m) * (y2 + z2), ((1.0f / 12.0f) * m) * (x2 + z2), ((1.0f / 12.0f) * m) * (x2 + y2));
    } else {
        float r = fabsf(s.d[0]), hh = fabsf(s.d[1]);
        if (!(isfinite(r) && r > 0 && isfinite(rho) && rho > 0)) return;
        float h = hh * 2.0f, r2 = r * r;
        float pi = 3.14159265358979323846f;
        float vc = (pi * r2) * h;
        float vsph = ((4.0f / 3.0f) * pi * r2) * r;
        float m = rho * (vc + vsph);
        float ix = ((0.25f * m) * r2) + (((1.0f / 12.0f) * m) * h) * h;
        float iy = (0.5f * m) * r2;
        *mass = m; *idiag = v3(ix, iy, ix);
    }
    /* MassProps::finalize */
    if (!(isfinite(*mass) && *mass > 0.0f) || *mass < 1.0e-12f) return;
    if (!(isfinite(idiag->x) && idiag->x > 0 && isfinite(idiag->y) && idiag->y > 0 && isfinite(idiag->z) && idiag->z > 0)) return;
    *ok = 1;
}

/* ffi/src/lib.rs:677-902 add_*_body -> world.rs:595-622 add_body/add_collider */
static int32_t add_body(OrWorld *w, int shape, const float *dims, const float *pose, const float *vel, int dynamic, float density,
                        int with_collider) {
    if (w->n >= w->cap) return -1;
    uint32_t i = w->n++;
    ShapeD s; s.type = shape; s.d[0] = dims[0]; s.d[1] = dims[1]; s.d[2] = dims[2];
    float rho = density > 0.0f ? density : 1500.0f; /* materials.rs:362 MaterialId::Default density */
    float mass = 0; V3 idiag = v3s(0); int ok = 0;
    if (dynamic) mass_props(s, rho, &mass, &idiag, &ok);
    float inv_mass = (dynamic && ok) ? 1.0f / mass : 0.0f;
    M3 inertia = ok ? m3diag(idiag) : m3zero();
    float det = m3det(inertia);
    int have_ii = dynamic && isfinite(det) && det > 0.0f;
    M3 iil = have_ii ? m3inv(inertia) : m3zero();
    /* dynamics.rs:98-119 Bodies::add */
    int is_dyn = dynamic && isfinite(inv_mass) && inv_mass > 0.0f;
    float im = is_dyn ? dfin(inv_mass) : 0.0f;
    V3 p = v3(pose[0], pose[1], pose[2]);
    Q4 q = q4(pose[3], pose[4], pose[5], pose[6]);
    V3 l = v3(vel[0], vel[1], vel[2]), a = v3(vel[3], vel[4], vel[5]);
    w->pos[i] = vfinite(p) ? p : v3s(0);
    w->rot[i] = fin_q(q);
    w->lin[i] = vfinite(l) ? l : v3s(0);
    w->ang[i] = vfinite(a) ? a : v3s(0);
    w->inv_mass[i] = im;
    w->dyn[i] = (uint8_t)is_dyn;
    if (!is_dyn) w->iil[i] = m3zero();
    else if (have_ii && m3finite(iil)) w->iil[i] = iil;
    else w->iil[i] = m3diag(v3s(im));
    w->last_lin[i] = l;
    if (i >= w->cap) { /* capacity grows only through the FFI check; never here */ }
    w->awake[i] = 1; w->sticks[i] = 0;
    w->alive[i] = 1;
    if (!with_collider) return (int32_t)i;
    uint32_t c = w->nc++;
    w->cbody[c] = i; w->cshape[c] = s; w->cmat[c] = MAT_DEFAULT;
    w->caabb[c] = aabb_of(s, w->pos[i], w->rot[i]);
    return (int32_t)i;
}

int32_t or_add_body(OrWorld *w, int shape, const float *dims, const float *pose, const float *vel, int dynamic, float density) {
    return add_body(w, shape, dims, pose, vel, dynamic, density, 1);
}
/* world.rs:595-613 add_body alone: a body with the shape's mass properties and no collider. */
int32_t or_add_body_nc(OrWorld *w, int shape, const float *dims, const float *pose, const float *vel, int dynamic, float density) {
    return add_body(w, shape, dims, pose, vel, dynamic, density, 0);
}

static void drop_warm_of(OrWorld *w, uint32_t b) {
    uint32_t k = 0;
    for (uint32_t i = 0; i < w->nwarm; i++)
        if (w->warm[i].key[0] != b && w->warm[i].key[1] != b) w->warm[k++] = w->warm[i];
    w->nwarm = k;
    k = 0;
    for (uint32_t i = 0; i < w->nwpatch; i++)
        if (w->wpatch[i].key[0] != b && w->wpatch[i].key[1] != b) w->wpatch[k++] = w->wpatch[i];
    w->nwpatch = k;
    for (uint32_t q = 0; q < w->nj; q++)
        if (w->joints[q].a == b || w->joints[q].b == b) w->wj[q].has = 0;
}

/* world.rs:718 add_ball_joint: frames {px,py,pz,qx,qy,qz,qw} in the bodies' local spaces.
 * Returns the joint index (joint ids are indices, never reused), or -1. */
int32_t or_add_ball_joint(OrWorld *w, uint32_t a, uint32_t b, const float *fa, const float *fb) {
    if (a >= w->n || b >= w->n) return -1;
    if (w->nj == w->capj) {
        w->capj = w->capj ? 2 * w->capj : 64;
        w->joints = xrealloc(w->joints, sizeof(Joint) * w->capj);
        w->wj = xrealloc(w->wj, sizeof(WarmJ) * w->capj);
    }
    Joint j;
    memset(&j, 0, sizeof j);
    j.a = a; j.b = b; j.fap = v3(fa[0], fa[1], fa[2]); j.fbp = v3(fb[0], fb[1], fb[2]);
    j.fra = q4(fa[3], fa[4], fa[5], fa[6]); j.frb = q4(fb[3], fb[4], fb[5], fb[6]);
    j.kind = J_BALL; j.hinge = -1;
    w->joints[w->nj] = j;
    memset(&w->wj[w->nj], 0, sizeof(WarmJ));
    return (int32_t)w->nj++;
}
/* world.rs:537-539 add_distance_joint -> bodies.rs:273-285: rest = fin(rest),
 * compliance = max(fin(compliance), 0). */
int32_t or_add_distance_joint(OrWorld *w, uint32_t a, uint32_t b, float rest, float compliance) {
    float z[7] = {0, 0, 0, 0, 0, 0, 1};
    int32_t q = or_add_ball_joint(w, a, b, z, z);
    if (q < 0) return q;
    Joint *j = &w->joints[q];
    j->kind = J_DIST;
    j->rest = dfin(rest);
    j->compliance = fmaxf(dfin(compliance), 0.0f);
    return q;
}
uint32_t or_num_joints(const OrWorld *w) { return w->nj; }
/* bodies.rs:321-345 add_d6 / add_hinge: a D6 joint with the translational axes of t_mask and the
 * rotational axes of r_mask constrained (bit i: axis i of frame A), and its stored hinge axis
 * (0..2, or -1: none).  add_hinge(axis) is t_mask 7, r_mask 7 & ~(1 << axis), hinge axis. */
int32_t or_add_d6_joint(OrWorld *w, uint32_t a, uint32_t b, const float *fa, const float *fb, uint32_t t_mask, uint32_t r_mask, int32_t hinge) {
    int32_t q = or_add_ball_joint(w, a, b, fa, fb);
    if (q < 0) return q;
    Joint *j = &w->joints[q];
    j->kind = J_D6;
    j->hinge = hinge >= 0 && hinge < 3 ? hinge : -1;
    for (int i = 0; i < 3; i++) { j->t[i].enabled = (t_mask >> i) & 1u; j->r[i].enabled = (r_mask >> i) & 1u; }
    return q;
}
static D6Ax *d6_axis(OrWorld *w, uint32_t q, int angular, uint32_t axis) {
    if (q >= w->nj || axis > 2 || w->joints[q].kind != J_D6) return NULL;
    return angular ? &w->joints[q].r[axis] : &w->joints[q].t[axis];
}
/* leave_hinge (bodies.rs:878-880): a limit or drive on another rotational axis forgets the stored one */
static void d6_leave_hinge(Joint *j, uint32_t axis) { if (j->hinge >= 0 && (uint32_t)j->hinge != axis) j->hinge = -1; }
/* bodies.rs:347-390 set_d6_limit_t / set_d6_limit_r: a limit enables its axis */
void or_set_d6_limit(OrWorld *w, uint32_t q, int angular, uint32_t axis, int on, float lo, float hi) {
    D6Ax *x = d6_axis(w, q, angular, axis);
    if (!x) return;
    x->has_limit = on != 0; x->lim_min = lo; x->lim_max = hi;
    if (on) { x->enabled = 1; if (angular) d6_leave_hinge(&w->joints[q], axis); }
}
/* set_d6_drive_t / set_d6_drive_r: a drive enables its axis */
void or_set_d6_drive(OrWorld *w, uint32_t q, int angular, uint32_t axis, int on, float target, float kp, float kd, float max_impulse, int vel_mode) {
    D6Ax *x = d6_axis(w, q, angular, axis);
    if (!x) return;
    x->has_drive = on != 0; x->d_target = target; x->d_kp = kp; x->d_kd = kd; x->d_max = max_impulse; x->vel_mode = vel_mode != 0;
    if (on) { x->enabled = 1; if (angular) d6_leave_hinge(&w->joints[q], axis); }
}
/* set_d6_damping (bodies.rs:370-378): negative or non-finite values become 0 */
void or_set_d6_damping(OrWorld *w, uint32_t q, int angular, uint32_t axis, float damping, float friction) {
    D6Ax *x = d6_axis(w, q, angular, axis);
    if (!x) return;
    x->damping = fmaxf(dfin(damping), 0.0f); x->friction = fmaxf(dfin(friction), 0.0f);
}
/* an axis's compliance (D6Axis::compliance, m/N or rad/(N m); 0 = rigid) */
void or_set_d6_compliance(OrWorld *w, uint32_t q, int angular, uint32_t axis, float compliance) {
    D6Ax *x = d6_axis(w, q, angular, axis);
    if (!x) return;
    x->compliance = compliance;
}
/* Joint q's warm rows: returns their count (0: none stored), tags and impulses into the arrays (room for 32). */
uint32_t or_get_joint_warm(const OrWorld *w, uint32_t q, uint8_t *tags, float *lams) {
    if (q >= w->nj || !w->wj[q].has) return 0;
    for (uint32_t k = 0; k < w->wj[q].n; k++) { tags[k] = w->wj[q].tag[k]; lams[k] = w->wj[q].lam[k]; }
    return w->wj[q].n;
}
/* The SHAKE direction test of the position solve, turned(a, b) < 0.5 rad, for a given |w| of a * conj(b). */
int or_shake_ok(float qw) { return 2.0f * acosf(fminf(fabsf(qw), 1.0f)) < 0.5f; }

/* world.rs:696-713 remove_body */
int or_remove_body(OrWorld *w, uint32_t id) {
    if (id >= w->n || !w->alive[id]) return 0;
    w->alive[id] = 0;
    w->lin[id] = v3s(0); w->ang[id] = v3s(0); w->inv_mass[id] = 0; w->dyn[id] = 0; w->iil[id] = m3zero();
    uint32_t k = 0;
    for (uint32_t c = 0; c < w->nc; c++) {
        if (w->cbody[c] == id) continue;
        w->cbody[k] = w->cbody[c]; w->cshape[k] = w->cshape[c]; w->cmat[k] = w->cmat[c]; w->caabb[k] = w->caabb[c]; k++;
    }
    w->nc = k;
    drop_warm_of(w, id);
    return 1;
}

/* terrain.rs:24-52 from_heights (heights already include offsets); ffi/src/lib.rs:1479,1536 use Grit */
void or_set_heightfield(OrWorld *w, uint32_t nx, uint32_t nz, float sx, float sz, float ox, float oz, const float *h, uint32_t mat) {
    free(w->hts);
    w->hts = malloc(sizeof(float) * (size_t)nx * nz);
    w->hnx = nx; w->hnz = nz;
    w->hsx = (isfinite(sx) && sx > 0) ? q6(sx) : 1.0f;
    w->hsz = (isfinite(sz) && sz > 0) ? q6(sz) : 1.0f;
    w->hox = ox; w->hoz = oz;
    float mn = INFINITY, mx = -INFINITY;
    for (size_t i = 0; i < (size_t)nx * nz; i++) {
        float v = isfinite(h[i]) ? h[i] : 0.0f;
        v = q6(v);
        w->hts[i] = v;
        mn = fminf(mn, v); mx = fmaxf(mx, v);
    }
    w->hmin = q6(mn); w->hmax = q6(mx);
    w->hmat = mat;
    w->has_hf = 1;
}

/* world.rs:501-513 set_body_pose (drops the body's warm entries, refreshes its AABB) */
void or_set_body_pose(OrWorld *w, uint32_t id, const float *pose) {
    drop_warm_of(w, id);
    b_set_pose(w, id, v3(pose[0], pose[1], pose[2]), q4(pose[3], pose[4], pose[5], pose[6]));
    for (uint32_t c = 0; c < w->nc; c++)
        if (w->cbody[c] == id) w->caabb[c] = aabb_of(w->cshape[c], v3(pose[0], pose[1], pose[2]), q4(pose[3], pose[4], pose[5], pose[6]));
}
/* world.rs:674-676 set_body_vel */
void or_set_body_vel(OrWorld *w, uint32_t id, const float *vel) {
    b_set_vel(w, id, v3(vel[0], vel[1], vel[2]), v3(vel[3], vel[4], vel[5]));
    w->last_lin[id] = v3(vel[0], vel[1], vel[2]);
}
/* world.rs:614-622 add_collider's material: the body's collider takes material id `mat`
 * (materials.rs core ids 0..15) */
void or_set_body_material(OrWorld *w, uint32_t id, uint32_t mat) {
    for (uint32_t c = 0; c < w->nc; c++) if (w->cbody[c] == id) w->cmat[c] = mat;
}
void or_set_sleep_allowed(OrWorld *w, uint32_t id, int allowed) {
    w->sallow[id] = (uint8_t)(allowed != 0);
    if (!allowed) { w->awake[id] = 1; w->sticks[id] = 0; }
}

/* ------------------------------------------------------------------ broadphase */
static int aabb_overlaps(const Aabb *a, const Aabb *b) {
    return !(a->max.x < b->min.x || a->min.x > b->max.x || a->max.y < b->min.y || a->min.y > b->max.y ||
             a->max.z < b->min.z || a->min.z > b->max.z);
}
static int cmp_u64(const void *x, const void *y) {
    uint64_t a = *(const uint64_t *)x, b = *(const uint64_t *)y;
    return a < b ? -1 : a > b;
}
/* broadphase.rs:29-99: grid cells of 2.5 m; a shared cell is implied by an inclusive overlap
 * (floor(x/2.5) is monotone), so the sorted, deduplicated result is every overlapping (i<j) pair
 * of colliders with finite min.x, plus oversized (> 12.5 m) colliders against everything. */
static uint32_t broadphase(const Aabb *sw, uint32_t n, uint32_t **out) {
    const float cell = 2.5f;
    size_t cap = 1024, cnt = 0;
    uint64_t *buf = malloc(cap * 8);
    uint8_t *valid = malloc(n + 1), *over = malloc(n + 1);
    /* sweep along x for the regular colliders */
    uint32_t *ord = malloc(sizeof(uint32_t) * (n + 1));
    uint32_t nv = 0;
    for (uint32_t i = 0; i < n; i++) {
        valid[i] = isfinite(sw[i].min.x) ? 1 : 0;
        float dx = sw[i].max.x - sw[i].min.x, dy = sw[i].max.y - sw[i].min.y, dz = sw[i].max.z - sw[i].min.z;
        over[i] = valid[i] && (dx > cell * 5.0f || dy > cell * 5.0f || dz > cell * 5.0f);
        if (valid[i] && !over[i]) ord[nv++] = i;
    }
    /* insertion-free O(N^2)-bounded sweep: sort by min.x then scan */
    uint64_t *keys = malloc(8 * (nv + 1));
    for (uint32_t k = 0; k < nv; k++) {
        float x = sw[ord[k]].min.x; uint32_t b; memcpy(&b, &x, 4);
        b = (b & 0x80000000u) ? ~b : (b | 0x80000000u);
        keys[k] = ((uint64_t)b << 32) | ord[k];
    }
    qsort(keys, nv, 8, cmp_u64);
    for (uint32_t k = 0; k < nv; k++) {
        uint32_t i = (uint32_t)keys[k];
        for (uint32_t m = k + 1; m < nv; m++) {
            uint32_t j = (uint32_t)keys[m];
            if (sw[j].min.x > sw[i].max.x) break;
            if (aabb_overlaps(&sw[i], &sw[j])) {
                if (cnt == cap) { cap *= 2; buf = xrealloc(buf, cap * 8); }
                uint32_t a = i < j ? i : j, b = i < j ? j : i;
                buf[cnt++] = ((uint64_t)a << 32) | b;
            }
        }
    }
    for (uint32_t i = 0; i < n; i++) {
        if (!over[i]) continue;
        for (uint32_t j = 0; j < n; j++) {
            if (j == i || !aabb_overlaps(&sw[i], &sw[j])) continue;
            if (cnt == cap) { cap *= 2; buf = xrealloc(buf, cap * 8); }
            uint32_t a = i < j ? i : j, b = i < j ? j : i;
            buf[cnt++] = ((uint64_t)a << 32) | b;
        }
    }
    qsort(buf, cnt, 8, cmp_u64);
    size_t u = 0;
    for (size_t k = 0; k < cnt; k++) if (u == 0 || buf[k] != buf[u - 1]) buf[u++] = buf[k];
    uint32_t *pairs = malloc(sizeof(uint32_t) * 2 * (u + 1));
    for (size_t k = 0; k < u; k++) { pairs[2 * k] = (uint32_t)(buf[k] >> 32); pairs[2 * k + 1] = (uint32_t)buf[k]; }
    free(buf); free(valid); free(over); free(ord); free(keys);
    *out = pairs;
    return (uint32_t)u;
}

/* ------------------------------------------------------------------ contacts buffer */
typedef struct { Contact *v; size_t n, cap; } CVec;
static void cpush(CVec *cv, Contact c) {
    if (cv->n == cv->cap) { cv->cap = cv->cap ? cv->cap * 2 : 256; cv->v = xrealloc(cv->v, cv->cap * sizeof(Contact)); }
    cv->v[cv->n++] = c;
}

/* terrain.rs:300-359 query_aabb + step.rs:121-133 per collider */
static uint32_t terrain_contacts(OrWorld *w, uint32_t ci, float sweep_ci, CVec *out) {
    if (!w->has_hf) return 0;
    int nx = (int)w->hnx, nz = (int)w->hnz;
    if (nx < 2 || nz < 2) return 0;
    const Aabb *bb = &w->caabb[ci];
    float lminx = bb->min.x - w->hox, lmaxx = bb->max.x - w->hox;
    float lminz = bb->min.z - w->hoz, lmaxz = bb->max.z - w->hoz;
    int gx0 = (int)floorf(lminx / w->hsx) - 1, gx1 = (int)ceilf(lmaxx / w->hsx) + 1;
    int gz0 = (int)floorf(lminz / w->hsz) - 1, gz1 = (int)ceilf(lmaxz / w->hsz) + 1;
#define ICLAMP(v, lo, hi) ((v) < (lo) ? (lo) : ((v) > (hi) ? (hi) : (v)))
    gx0 = ICLAMP(gx0, 0, nx - 2); gx1 = ICLAMP(gx1, 0, nx - 2);
    gz0 = ICLAMP(gz0, 0, nz - 2); gz1 = ICLAMP(gz1, 0, nz - 2);
    if (bb->min.y > w->hmax + 1.0f || bb->max.y < w->hmin - 1.0f) return 0;
    uint32_t tcount = 0;
    uint32_t b = w->cbody[ci];
    for (int gz = gz0; gz <= gz1; gz++) {
        for (int gx = gx0; gx <= gx1; gx++) {
            float h00 = w->hts[gx + gz * nx], h10 = w->hts[gx + 1 + gz * nx];
            float h01 = w->hts[gx + (gz + 1) * nx], h11 = w->hts[gx + 1 + (gz + 1) * nx];
            float x0 = (float)gx * w->hsx + w->hox, x1 = x0 + w->hsx;
            float z0 = (float)gz * w->hsz + w->hoz, z1 = z0 + w->hsz;
            V3 p00 = v3(x0, h00, z0), p10 = v3(x1, h10, z0), p01 = v3(x0, h01, z1), p11 = v3(x1, h11, z1);
            Tri tris[2] = {{p00, p01, p10}, {p10, p01, p11}};
            for (int t = 0; t < 2; t++) {
                tcount++;
                Manifold m;
                if (!collide_shape_tri(w->cshape[ci], w->pos[b], w->rot[b], tris[t], sweep_ci, &m)) continue;
                for (int k = 0; k < m.len; k++) {
                    Contact c;
                    c.a = ci; c.b = TERRAIN_CI; c.point = m.p[k].pos; c.normal = m.p[k].n;
                    c.depth = m.p[k].pen - sweep_ci + MARGIN;
                    c.impulse = 0; c.env_mat = w->hmat; c.feat = m.p[k].feat;
                    c.cell = ((uint32_t)gz * (uint32_t)(nx - 1) + (uint32_t)gx) * 2u + (uint32_t)t;
                    cpush(out, c);
                }
            }
        }
    }
    return tcount;
}

/* step.rs:499-542 */
typedef struct { uint64_t key; uint32_t idx; } KI;
static int cmp_ki(const void *x, const void *y) {
    const KI *a = x, *b = y;
    if (a->key != b->key) return a->key < b->key ? -1 : 1;
    return a->idx < b->idx ? -1 : (a->idx > b->idx);
}
typedef struct { int64_t k; uint32_t idx; } SK;
static int cmp_sk_desc(const void *x, const void *y) { /* Reverse(sort_key6), stable by idx */
    const SK *a = x, *b = y;
    if (a->k != b->k) return a->k > b->k ? -1 : 1;
    return a->idx < b->idx ? -1 : (a->idx > b->idx);
}
static int cmp_sk_asc(const void *x, const void *y) {
    const SK *a = x, *b = y;
    if (a->k != b->k) return a->k < b->k ? -1 : 1;
    return a->idx < b->idx ? -1 : (a->idx > b->idx);
}
static void cull_and_orient(OrWorld *w, CVec *raw, CVec *out) {
    size_t n = raw->n;
    KI *ki = malloc(sizeof(KI) * (n + 1));
    for (size_t i = 0; i < n; i++) {
        uint32_t a = raw->v[i].a, b = raw->v[i].b;
        uint32_t x = a <= b ? a : b, y = a <= b ? b : a;
        ki[i].key = ((uint64_t)x << 32) | y; ki[i].idx = (uint32_t)i;
    }
    qsort(ki, n, sizeof(KI), cmp_ki);
    size_t s = 0;
    while (s < n) {
        size_t e = s;
        while (e < n && ki[e].key == ki[s].key) e++;
        size_t len = e - s;
        if (len <= 4) {
            for (size_t k = s; k < e; k++) cpush(out, raw->v[ki[k].idx]);
        } else if (w->cshape[raw->v[ki[s].idx].a].type == OR_BOX) {
            Cand *cands = malloc(sizeof(Cand) * len);
            size_t deep = 0;
            for (size_t k = 0; k < len; k++) { cands[k].depth = raw->v[ki[s + k].idx].depth; cands[k].p = raw->v[ki[s + k].idx].point; }
            for (size_t k = 1; k < len; k++)
                if (sort_key6(raw->v[ki[s + k].idx].depth) > sort_key6(raw->v[ki[s + deep].idx].depth)) deep = k;
            int idx[4];
            int m = reduce_to_4(cands, (int)len, raw->v[ki[s + deep].idx].normal, idx);
            for (int k = 0; k < m; k++) cpush(out, raw->v[ki[s + idx[k]].idx]);
            free(cands);
        } else {
            SK *sk = malloc(sizeof(SK) * len);
            for (size_t k = 0; k < len; k++) { sk[k].k = sort_key6(raw->v[ki[s + k].idx].depth); sk[k].idx = (uint32_t)k; }
            qsort(sk, len, sizeof(SK), cmp_sk_desc);
            for (size_t k = 0; k < 4; k++) cpush(out, raw->v[ki[s + sk[k].idx].idx]);
            free(sk);
        }
        s = e;
    }
    free(ki);
    for (size_t i = 0; i < out->n; i++) {
        Contact *c = &out->v[i];
        if (c->b == TERRAIN_CI) continue;
        V3 pa = w->pos[w->cbody[c->a]], pb = w->pos[w->cbody[c->b]];
        if (vdot(c->normal, vsub(pb, pa)) < 0.0f) c->normal = vneg(c->normal);
    }
}

/* ------------------------------------------------------------------ solver */
typedef struct {
    uint32_t ai, bi; int terrain;
    float inv_a, inv_b; M3 ii_a, ii_b; V3 ra, rb;
    float m_n, m_t1, m_t2; V3 n, t1, t2;
    float mu_s, mu_k, vs, depth_bias, rest_bias, approach;
    /* W6 time of impact (solver.rs:569-588): the normal velocities of each side at the step's
     * start and after this step's forces, the start gap, the pair's restitution */
    int toi; float toi_s0, toi_e, toi_va0, toi_va1, toi_vb0, toi_vb1;
    /* compliant pair (solver.rs:64-77, 545-567): no rigid normal row, shock or restitution; the
     * implicit soft row dlam = soft_m ((soft_bias - v_n) - soft_gamma lam), lam >= 0, or none (a
     * row whose effective mass or spring is zero keeps lam at 0) */
    int soft, soft_implicit; float soft_gamma, soft_bias, soft_m;
    int patch;  /* friction solved by the manifold's patch, not by this point */
} CC;
/* One manifold's friction patch (solver.rs:219-229, 592-619): a constraint at the points' centroid,
 * the points (contact indices, ascending), the mean radius of the patch and the torsional mass. */
typedef struct { CC at; uint32_t p0, np; float radius, m_tw; } Patch;

/* world/src/world/geom.rs:30-38 */
static void orthonormal_basis(V3 n, V3 *t1, V3 *t2) {
    float ax = fabsf(n.x), ay = fabsf(n.y), az = fabsf(n.z);
    V3 base = (ax <= ay && ax <= az) ? v3(1, 0, 0) : (ay <= az ? v3(0, 1, 0) : v3(0, 0, 1));
    *t1 = vnormalize_or_zero(vcross(base, n));
    *t2 = vcross(n, *t1);
}
/* solver.rs:443-451 */
static V3 rel_vel(const OrWorld *w, const CC *c) {
    V3 va = vadd(w->lin[c->ai], vcross(w->ang[c->ai], c->ra));
    V3 vb = c->terrain ? v3s(0) : vadd(w->lin[c->bi], vcross(w->ang[c->bi], c->rb));
    return vsub(vb, va);
}
/* solver.rs:429-431 */
static float mu_dyn(const CC *c, float slip) {
    if (c->mu_s == c->mu_k && slip == slip) return c->mu_k + 0.0f; /* (+0) * exp(..) = +0 exactly */
    float x = slip / c->vs;
    return c->mu_k + (c->mu_s - c->mu_k) * or_expf_d(-(x * x));
}
/* solver.rs:455-464 */
static void apply_cimp(OrWorld *w, const CC *c, V3 j) {
    if (c->inv_a > 0.0f) {
        V3 nj = vneg(j);
        apply_impulse(w, c->ai, nj);
        apply_ang_impulse(w, c->ai, vcross(c->ra, nj));
    }
    if (c->inv_b > 0.0f) {
        apply_impulse(w, c->bi, j);
        apply_ang_impulse(w, c->bi, vcross(c->rb, j));
    }
}
static float eff_mass(const CC *c, V3 d) {
    V3 ca = vcross(c->ra, d), cb = vcross(c->rb, d);
    float k = ((c->inv_a + c->inv_b) + vdot(ca, m3v(c->ii_a, ca))) + vdot(cb, m3v(c->ii_b, cb));
    return k > 1.0e-9f ? 1.0f / k : 0.0f;
}
/* solver.rs:233-260 patch_radius: the mean distance from the origin over the convex polygon of the
 * plane points (sorted by angle; each fan triangle integrated on a 6 x 6 subdivision by its
 * centroids); collinear points: a quarter of their longest distance */
typedef struct { float x, y; } P2;
static float patch_radius(P2 *p, uint32_t n) {
    if (n < 2) return 0.0f;
    int64_t *key = malloc(sizeof(int64_t) * n);
    for (uint32_t i = 0; i < n; i++) key[i] = sort_key6(or_atan2f(p[i].y, p[i].x));
    for (uint32_t x = 1; x < n; x++) {  /* stable */
        P2 v = p[x]; int64_t kv = key[x]; uint32_t y = x;
        while (y > 0 && key[y - 1] > kv) { p[y] = p[y - 1]; key[y] = key[y - 1]; y--; }
        p[y] = v; key[y] = kv;
    }
    free(key);
    float area = 0.0f, moment = 0.0f;
    const int N = 6;
    for (uint32_t i = 0; i < n; i++) {
        P2 a = p[i], b = p[(i + 1) % n];
        float tri = 0.5f * fabsf(a.x * b.y - a.y * b.x);
        if (tri <= 0.0f) continue;
        float sub = tri / (float)(N * N);
        for (int r = 0; r < N; r++)
            for (int c = 0; c < N - r; c++)
                for (int up = 1; up >= 0; up--) {
                    if (!up && c + r + 1 >= N) continue;
                    float u = up ? (float)r + 1.0f / 3.0f : (float)r + 2.0f / 3.0f;
                    float v = up ? (float)c + 1.0f / 3.0f : (float)c + 2.0f / 3.0f;
                    u = u / (float)N; v = v / (float)N;
                    float x0 = a.x * u + b.x * v, x1 = a.y * u + b.y * v;
                    moment += sqrtf(x0 * x0 + x1 * x1) * sub;
                }
        area += tri;
    }
    if (area > 1.0e-12f) return moment / area;
    float len = 0.0f;
    for (uint32_t i = 0; i < n; i++)
        for (uint32_t j = 0; j < n; j++) {
            float dx = p[i].x - p[j].x, dy = p[i].y - p[j].y;
            len = fmaxf(len, sqrtf(dx * dx + dy * dy));
        }
    return 0.25f * len;
}
/* solver.rs:596-617: the patch of the manifold whose points (contact indices, ascending) are idx:
 * a constraint at their centroid along their mean normal, with the first point's bodies, masses and
 * pair properties; marks the points.  0 when the normals cancel. */
static int build_patch(const OrWorld *w, const Contact *cts, CC *cs, const uint32_t *idx, uint32_t n, Patch *out) {
    const CC first = cs[idx[0]];
    V3 sum = v3s(0), nsum = v3s(0);
    for (uint32_t k = 0; k < n; k++) sum = vadd(sum, cts[idx[k]].point);
    V3 centroid = v3(sum.x / (float)n, sum.y / (float)n, sum.z / (float)n);
    for (uint32_t k = 0; k < n; k++) nsum = vadd(nsum, cts[idx[k]].normal);
    V3 normal = vnormalize_or_zero(nsum);
    if (normal.x == 0.0f && normal.y == 0.0f && normal.z == 0.0f) return 0;
    V3 t1, t2;
    orthonormal_basis(normal, &t1, &t2);
    P2 *plane = malloc(sizeof(P2) * n);
    for (uint32_t k = 0; k < n; k++) {
        V3 d = vsub(cts[idx[k]].point, centroid);
        plane[k].x = vdot(d, t1); plane[k].y = vdot(d, t2);
    }
    CC at = first;
    at.ra = vsub(centroid, w->pos[first.ai]);
    at.rb = first.terrain ? v3s(0) : vsub(centroid, w->pos[first.bi]);
    at.n = normal; at.t1 = t1; at.t2 = t2;
    at.m_n = eff_mass(&at, normal); at.m_t1 = eff_mass(&at, t1); at.m_t2 = eff_mass(&at, t2);
    float kt = vdot(normal, m3v(first.ii_a, normal)) + vdot(normal, m3v(first.ii_b, normal));
    for (uint32_t k = 0; k < n; k++) cs[idx[k]].patch = 1;
    out->at = at; out->np = n;
    out->radius = patch_radius(plane, n);
    free(plane);
    out->m_tw = kt > 1.0e-9f ? 1.0f / kt : 0.0f;
    return 1;
}
/* slip_speed (solver.rs:435-439) */
static float slip_speed(const OrWorld *w, const CC *c) {
    V3 v = rel_vel(w, c);
    float vt1 = vdot(v, c->t1), vt2 = vdot(v, c->t2);
    return sqrtf(vt1 * vt1 + vt2 * vt2);
}
/* f32::clamp */
static float clampf_rs(float x, float lo, float hi) { if (x < lo) x = lo; if (x > hi) x = hi; return x; }
/* sum of the patch's points' normal impulses (Iterator::sum, from -0) */
static float patch_jn(const float *jn, const uint32_t *pts, uint32_t n) {
    float s = -0.0f;
    for (uint32_t k = 0; k < n; k++) s += jn[pts[k]];
    return s;
}
/* torsion impulse about the normal on both sides (solver.rs:778-781, 908-911) */
static void patch_twist(OrWorld *w, const CC *c, float tw) {
    if (c->inv_a > 0.0f) apply_ang_impulse(w, c->ai, vmul(vneg(c->n), tw));
    if (!c->terrain && c->inv_b > 0.0f) apply_ang_impulse(w, c->bi, vmul(c->n, tw));
}
static int key_cmp(const uint32_t *a, const uint32_t *b) {
    for (int i = 0; i < 4; i++) if (a[i] != b[i]) return a[i] < b[i] ? -1 : 1;
    return 0;
}
/* One pass of a friction patch (solver.rs:887-913): sliding at the centroid, then torsion about the
 * normal, both limited by the points' current normal impulse; acc = (t1, t2, torsion).  Returns the
 * residual-mode residual: the largest of |d jt| / m_t and |d tw| / m_tw. */
static float patch_row(OrWorld *w, const Patch *P, const uint32_t *pp, const float *jn, float *acc) {
    const CC *c = &P->at;
    float jsum = patch_jn(jn, pp, P->np);
    V3 v = rel_vel(w, c);
    float vt1 = vdot(v, c->t1), vt2 = vdot(v, c->t2);
    float mu = mu_dyn(c, sqrtf(vt1 * vt1 + vt2 * vt2));
    V3 ot = v3(acc[0], acc[1], 0.0f);
    V3 nt = vclamp_len_max(vsub(ot, v3(vt1 * c->m_t1, vt2 * c->m_t2, 0.0f)), mu * jsum);
    acc[0] = nt.x; acc[1] = nt.y;
    V3 d = vsub(nt, ot);
    apply_cimp(w, c, vadd(vmul(c->t1, d.x), vmul(c->t2, d.y)));
    V3 wa = c->inv_a > 0.0f ? w->ang[c->ai] : v3s(0);
    V3 wb = (c->terrain || c->inv_b <= 0.0f) ? v3s(0) : w->ang[c->bi];
    float w_rel = vdot(vsub(wb, wa), c->n);
    float old_tw = acc[2];
    float lim = (mu * jsum) * P->radius;
    float new_tw = clampf_rs(old_tw - w_rel * P->m_tw, -lim, lim);
    acc[2] = new_tw;
    float dtw = new_tw - old_tw;
    if (dtw != 0.0f) patch_twist(w, c, dtw);
    float r1 = c->m_t1 > 0.0f ? fabsf(d.x) / c->m_t1 : 0.0f;
    float r2 = c->m_t2 > 0.0f ? fabsf(d.y) / c->m_t2 : 0.0f;
    float r3 = P->m_tw > 0.0f ? fabsf(dtw) / P->m_tw : 0.0f;
    return fmaxf(r1, fmaxf(r2, r3));
}
static int wpatch_find(const OrWorld *w, const uint32_t key[4]) {
    uint32_t lo = 0, hi = w->nwpatch;
    while (lo < hi) { uint32_t mid = (lo + hi) / 2; if (key_cmp(w->wpatch[mid].key, key) < 0) lo = mid + 1; else hi = mid; }
    return (lo < w->nwpatch && key_cmp(w->wpatch[lo].key, key) == 0) ? (int)lo : -1;
}
static int is_idle(const OrWorld *w, uint32_t b);
/* warm_patches store (solver.rs:927-944): this step's rigid patches by their first point's key,
 * plus the entries of pairs neither of whose bodies is awake that were not solved */
static void store_patches(OrWorld *w, const Patch *pt, uint32_t npt, const uint32_t *ppts, const uint32_t (*keys)[4], const float *pacc) {
    WarmP *nw = malloc(sizeof(WarmP) * (npt + w->nwpatch + 1));
    uint32_t n = 0;
    for (uint32_t k = 0; k < npt; k++) {
        if (pt[k].at.soft) continue;
        WarmP e;
        memcpy(e.key, keys[ppts[pt[k].p0]], 16);
        e.jt = vadd(vmul(pt[k].at.t1, pacc[3 * k]), vmul(pt[k].at.t2, pacc[3 * k + 1]));
        e.tw = pacc[3 * k + 2];
        uint32_t y = n;  /* key order (a map insert) */
        while (y > 0 && key_cmp(nw[y - 1].key, e.key) > 0) { nw[y] = nw[y - 1]; y--; }
        if (y > 0 && key_cmp(nw[y - 1].key, e.key) == 0) { nw[y - 1] = e; memmove(nw + y, nw + y + 1, sizeof(WarmP) * (n - y)); continue; }
        nw[y] = e; n++;
    }
    WarmP *merged = malloc(sizeof(WarmP) * (n + w->nwpatch + 1));
    uint32_t m = 0, p = 0;
    for (uint32_t q = 0; q < w->nwpatch; q++) {
        const WarmP *o = &w->wpatch[q];
        if (!(is_idle(w, o->key[0]) && is_idle(w, o->key[1]))) continue;
        while (p < n && key_cmp(nw[p].key, o->key) < 0) merged[m++] = nw[p++];
        if (p < n && key_cmp(nw[p].key, o->key) == 0) continue;
        merged[m++] = *o;
    }
    while (p < n) merged[m++] = nw[p++];
    free(nw); free(w->wpatch);
    w->wpatch = merged; w->nwpatch = m;
}
/* solver.rs:730-745: among the unused cached entries of the contact's key (the cache is sorted by
 * key), the one nearest to its anchor with |dla|^2 <= 0.01^2; ties go to the lowest index.  The
 * chosen entry is marked used.  Returns its index, or -1. */
static int warm_pick(const Warm *warm, uint32_t nwarm, const uint32_t key[4], V3 anchor, uint8_t *taken) {
    const float tol2 = 0.01f * 0.01f;
    uint32_t lo = 0, hi = nwarm;
    while (lo < hi) { uint32_t mid = (lo + hi) / 2; if (key_cmp(warm[mid].key, key) < 0) lo = mid + 1; else hi = mid; }
    if (lo >= nwarm || key_cmp(warm[lo].key, key) != 0) return -1;
    int best = -1; float bd = 0;
    for (uint32_t e = lo; e < nwarm && key_cmp(warm[e].key, key) == 0; e++) {
        if (taken[e]) continue;
        float d = vlen2(vsub(warm[e].la, anchor));
        if (d <= tol2 && (best < 0 || d < bd)) { best = (int)e; bd = d; }
    }
    if (best >= 0) taken[best] = 1;
    return best;
}
/* Unit entry of warm_pick: contacts (keys, anchors) matched in order against a key-sorted cache. */
void or_warm_pick_sequence(const uint32_t *ckeys4, const float *cla3, uint32_t nwarm, const uint32_t *qkeys4, const float *anchors3,
                           uint32_t nq, int32_t *out) {
    Warm *warm = calloc(nwarm + 1, sizeof(Warm));
    uint8_t *taken = calloc(nwarm + 1, 1);
    for (uint32_t e = 0; e < nwarm; e++) {
        memcpy(warm[e].key, ckeys4 + 4 * e, 16);
        warm[e].la = v3(cla3[3 * e], cla3[3 * e + 1], cla3[3 * e + 2]);
    }
    for (uint32_t i = 0; i < nq; i++) out[i] = warm_pick(warm, nwarm, qkeys4 + 4 * i, v3(anchors3[3 * i], anchors3[3 * i + 1], anchors3[3 * i + 2]), taken);
    free(warm); free(taken);
}

static int is_idle(const OrWorld *w, uint32_t b) {
    if (b == 0xFFFFFFFFu) return 1;
    if (b >= w->cap || !w->alive[b]) return 0;
    return !w->dyn[b] || !w->awake[b];
}

static uint32_t uf_find(uint32_t *p, uint32_t x) {
    while (p[x] != x) { p[x] = p[p[x]]; x = p[x]; }
    return x;
}

/* solver.rs:477-1008, 1247-1271 */
/* ---- ball joints at velocity level: JointRow (solver.rs:272-289), joint_row_vel / joint_row_apply
 * (291-307), d6_rows for three locked translational axes (336-424: no limits, drives, damping,
 * friction or armature on a ball joint, so each axis gives one equality row, target 0, unbounded) */
/* JointRow (solver.rs:262-289): bodies, angular or linear, direction, lever arms, inverse masses,
 * 1 / (K + gamma), target velocity, accumulated-impulse bounds, softness gamma, tag */
typedef struct { uint32_t a, b; int angular; V3 u, ra, rb; float wa, wb, m, target, lo, hi, gamma; uint8_t tag; } JRow;
static float jrow_vel(const OrWorld *w, const JRow *r) {
    if (r->angular) return vdot(vsub(w->ang[r->b], w->ang[r->a]), r->u);
    V3 va = vadd(w->lin[r->a], vcross(w->ang[r->a], r->ra));
    V3 vb = vadd(w->lin[r->b], vcross(w->ang[r->b], r->rb));
    return vdot(vsub(vb, va), r->u);
}
static void jrow_apply(OrWorld *w, const JRow *r, float lam) {
    V3 p = vmul(r->u, lam);
    if (r->angular) {
        if (r->wa > 0.0f) apply_ang_impulse(w, r->a, vneg(p));
        if (r->wb > 0.0f) apply_ang_impulse(w, r->b, p);
        return;
    }
    if (r->wa > 0.0f) { apply_impulse(w, r->a, vneg(p)); apply_ang_impulse(w, r->a, vcross(r->ra, vneg(p))); }
    if (r->wb > 0.0f) { apply_impulse(w, r->b, p); apply_ang_impulse(w, r->b, vcross(r->rb, p)); }
}
static JRow jrow_make(uint32_t a, uint32_t b, int angular, V3 u, V3 ra, V3 rb, float wa, float wb, float m, uint8_t tag) {
    JRow r = {a, b, angular, u, ra, rb, wa, wb, m, 0.0f, -INFINITY, INFINITY, 0.0f, tag};
    return r;
}
static uint32_t ball_rows(const OrWorld *w, const Joint *j, JRow *out) {
    uint32_t a = j->a, b = j->b;
    V3 pa = w->pos[a], pb = w->pos[b];
    Q4 qa = w->rot[a], qb = w->rot[b];
    Q4 rot_a = qmul(qa, j->fra);
    V3 xa = vadd(pa, qrot(qa, j->fap)), xb = vadd(pb, qrot(qb, j->fbp));
    float wa = w->inv_mass[a], wb = w->inv_mass[b];
    if (wa == 0.0f && wb == 0.0f) return 0;
    M3 ia = inv_inertia_world(w, a), ib = inv_inertia_world(w, b);
    V3 ra = vsub(xa, pa), rb = vsub(xb, pb);
    uint32_t n = 0;
    for (int i = 0; i < 3; i++) {
        V3 u = qrot(rot_a, v3(i == 0 ? 1.0f : 0.0f, i == 1 ? 1.0f : 0.0f, i == 2 ? 1.0f : 0.0f));
        V3 ca = vcross(ra, u), cb = vcross(rb, u);
        float k = ((wa + wb) + vdot(ca, m3v(ia, ca))) + vdot(cb, m3v(ib, cb));
        if (k <= 1.0e-9f) continue;
        out[n++] = jrow_make(a, b, 0, u, ra, rb, wa, wb, 1.0f / k, (uint8_t)(i << 3));
    }
    return n;
}

/* bodies.rs:50-53 position_damped: a compliant damped axis with no limit or velocity drive keeps
 * its damping at position level */
static int d6_position_damped(const D6Ax *x) {
    return x->enabled && x->compliance > 0.0f && x->damping > 0.0f && !x->has_limit && (!x->has_drive || !x->vel_mode);
}
/* bodies.rs:895-906 d6_hinge_axis: the stored axis, else the one rotational axis that is free,
 * limited or driven, else the one compliant axis at least 100x softer than the others */
static int d6_hinge_axis(const Joint *j) {
    if (j->hinge >= 0) return j->hinge < 3 ? j->hinge : -1;
    int freeax = -1, n = 0;
    for (int i = 0; i < 3; i++) {
        const D6Ax *r = &j->r[i];
        if (!(r->enabled && !r->has_limit && !r->has_drive)) { freeax = i; n++; }
    }
    if (n == 1) return freeax;
    float c[3];
    for (int i = 0; i < 3; i++) c[i] = j->r[i].enabled ? fmaxf(dfin(j->r[i].compliance), 0.0f) : 0.0f;
    for (int i = 0; i < 3; i++) {
        if (!(j->r[i].enabled && c[i] > 0.0f)) continue;
        int all = 1;
        for (int k = 0; k < 3; k++) if (k != i && !(c[i] >= 100.0f * c[k])) all = 0;
        if (all) return i;
    }
    return -1;
}
static V3 unit3(int i) { return v3(i == 0 ? 1.0f : 0.0f, i == 1 ? 1.0f : 0.0f, i == 2 ? 1.0f : 0.0f); }
static float v3at(V3 v, int i) { return i == 0 ? v.x : (i == 1 ? v.y : v.z); }
/* bodies.rs:913-922 d6_row_angle: the hinge's own axis measures its twist (2 atan2, radians);
 * another rotational axis the small-angle measure (rot_a (2 q.xyz)) . (rot_a e_i).  The atan2 is
 * the once-rounded double evaluation (or_atan2f). */
static float d6_row_angle(const Joint *j, Q4 rot_a, Q4 rot_b, int i) {
    Q4 q = qmul(qconj(rot_a), rot_b);
    if (q.w < 0.0f) q = q4(-q.x, -q.y, -q.z, -q.w);
    V3 v = v3(q.x, q.y, q.z);
    if (d6_hinge_axis(j) == i) return 2.0f * or_atan2f(v3at(v, i), q.w);
    return vdot(qrot(rot_a, vmul(v, 2.0f)), qrot(rot_a, unit3(i)));
}
/* solver.rs:336-424 d6_rows (no armature): per axis, a damper (v + gamma lam = 0) and a Coulomb
 * friction row, then a locked axis's equality, a velocity drive's motor row, or a limit's two
 * speculative rows; position drives and compliant axes stay at position level */
static uint32_t d6_rows(const OrWorld *w, const Joint *j, float dt, JRow *out) {
    uint32_t a = j->a, b = j->b;
    V3 pa = w->pos[a], pb = w->pos[b];
    Q4 qa = w->rot[a], qb = w->rot[b];
    Q4 rot_a = qmul(qa, j->fra), rot_b = qmul(qb, j->frb);
    V3 xa = vadd(pa, qrot(qa, j->fap)), xb = vadd(pb, qrot(qb, j->fbp));
    V3 axes[3] = {qrot(rot_a, unit3(0)), qrot(rot_a, unit3(1)), qrot(rot_a, unit3(2))};
    float wa = w->inv_mass[a], wb = w->inv_mass[b];
    if (wa == 0.0f && wb == 0.0f) return 0;
    M3 ia = inv_inertia_world(w, a), ib = inv_inertia_world(w, b);
    V3 ra = vsub(xa, pa), rb = vsub(xb, pb), d = vsub(xb, xa);
    uint32_t n = 0;
    for (int angular = 0; angular < 2; angular++)
        for (int i = 0; i < 3; i++) {
            const D6Ax *row = angular ? &j->r[i] : &j->t[i];
            V3 u = axes[i];
            float kk;
            if (angular) kk = vdot(u, m3v(ia, u)) + vdot(u, m3v(ib, u));
            else { V3 ca = vcross(ra, u), cb = vcross(rb, u); kk = ((wa + wb) + vdot(ca, m3v(ia, ca))) + vdot(cb, m3v(ib, cb)); }
            uint8_t tag = (uint8_t)((angular << 5) | (i << 3));
            float damp = d6_position_damped(row) ? 0.0f : row->damping;
            if ((damp > 0.0f || row->friction > 0.0f) && kk > 1.0e-9f) {
                JRow base = jrow_make(a, b, angular, u, ra, rb, wa, wb, 1.0f / kk, tag);
                if (damp > 0.0f) {
                    JRow r = base;
                    r.gamma = 1.0f / (damp * dt);
                    r.m = 1.0f / (kk + r.gamma);
                    r.tag = tag | 4;
                    out[n++] = r;
                }
                if (row->friction > 0.0f) {
                    JRow r = base;
                    float cap = row->friction * dt;
                    r.lo = -cap; r.hi = cap; r.tag = tag | 5;
                    out[n++] = r;
                }
            }
            if (!row->enabled || row->compliance > 0.0f) continue;
            if (row->has_drive && !row->vel_mode) continue;
            if (kk <= 1.0e-9f) continue;
            JRow base = jrow_make(a, b, angular, u, ra, rb, wa, wb, 1.0f / kk, tag);
            if (row->has_drive) {
                JRow r = base;
                float cap = fabsf(row->d_max);
                r.target = row->d_target; r.lo = -cap; r.hi = cap; r.tag = tag | 1;
                out[n++] = r;
            }
            if (!row->has_limit) {
                if (!row->has_drive) out[n++] = base;
            } else {
                float x = angular ? d6_row_angle(j, rot_a, rot_b, i) : vdot(d, u);
                JRow lo = base, hi = base;
                lo.target = fminf((row->lim_min - x) / dt, 0.0f); lo.lo = 0.0f; lo.tag = tag | 2;
                hi.target = fmaxf((row->lim_max - x) / dt, 0.0f); hi.hi = 0.0f; hi.lo = -INFINITY; hi.tag = tag | 3;
                out[n++] = lo;
                out[n++] = hi;
            }
        }
    return n;
}

/* solver.rs:795-820: the legacy row of a rigid distance joint without anchors or damping, run in
 * each velocity iteration with the joint batches: a Baumgarte velocity row along the centre line,
 * not warm-started and not accumulated.  Returns its velocity change |j| (wa + wb), the row's
 * residual in residual mode. */
static float dist_vel_row(OrWorld *w, const Joint *j, float dt) {
    V3 pa = w->pos[j->a], pb = w->pos[j->b];
    V3 dir = vsub(pb, pa);
    float len = vlen(dir);
    if (len < 1.0e-6f) return 0.0f;
    V3 n = vdivs(dir, len);
    float ia = w->inv_mass[j->a], ib = w->inv_mass[j->b];
    float isum = ia + ib;
    if (isum < 1.0e-9f) return 0.0f;
    float vrel = vdot(vsub(w->lin[j->b], w->lin[j->a]), n);
    float bias = (len - j->rest) * (0.1f / dt);
    float jm = -(vrel + bias) / isum;
    apply_impulse(w, j->a, vmul(vneg(n), jm));
    apply_impulse(w, j->b, vmul(n, jm));
    return fabsf(jm) * isum;
}

/* ---- W6 bounce passes (solver.rs:1009-1245) ---- */
static int loaded_at(const float *jn, const float *jt, uint32_t i) { return jn[i] > 0.0f || jt[2 * i] != 0.0f || jt[2 * i + 1] != 0.0f; }
/* grow (solver.rs:1018-1030): involved spreads through rigid loaded contacts to a fixed point,
 * never onto static, sleeping or jointed bodies (the fixed point is the closure, so the scan order
 * does not matter) */
static void w6_grow(const CC *cs, uint32_t nct, const float *jn, const float *jt, const uint8_t *jointed, uint8_t *inv) {
    for (;;) {
        int changed = 0;
        for (uint32_t i = 0; i < nct; i++) {
            const CC *c = &cs[i];
            if (c->soft || c->terrain || !loaded_at(jn, jt, i)) continue;
            if (inv[c->ai] && !inv[c->bi] && c->inv_b > 0.0f && !jointed[c->bi]) { inv[c->bi] = 1; changed = 1; }
            if (inv[c->bi] && !inv[c->ai] && c->inv_a > 0.0f && !jointed[c->ai]) { inv[c->ai] = 1; changed = 1; }
        }
        if (!changed) break;
    }
}
/* free_only (solver.rs:1034-1041): jointed sides immovable, the normal effective mass again;
 * 0 if neither side can move */
static int w6_free_only(const CC *c, const uint8_t *jointed, CC *out) {
    CC k = *c;
    if (jointed[k.ai]) { k.inv_a = 0.0f; k.ii_a = m3zero(); }
    if (k.terrain || jointed[k.bi]) { k.inv_b = 0.0f; k.ii_b = m3zero(); }
    V3 ca = vcross(k.ra, k.n), cb = vcross(k.rb, k.n);
    float kk = ((k.inv_a + k.inv_b) + vdot(ca, m3v(k.ii_a, ca))) + vdot(cb, m3v(k.ii_b, cb));
    if (!(kk > 1.0e-9f)) return 0;
    k.m_n = 1.0f / kk;
    *out = k;
    return 1;
}
static void w6_sides(const CC *c, BSide *sd) {
    sd[0] = or_bside(c->ai, -1.0, c->ra, c->inv_a, c->ii_a, c->n);
    sd[1] = or_bside(c->terrain ? BN_NONE : c->bi, 1.0, c->rb, c->inv_b, c->ii_b, c->n);
}
typedef struct { CC c; float target; int32_t cidx; } W6Row;
/* rows of each component, ascending, components in first-row order */
static uint32_t w6_comps(const BSide (*sides)[2], uint32_t n, uint32_t **order, uint32_t **start) {
    uint32_t *cof = malloc(4 * (n + 1));
    uint32_t nc = bn_components(sides, n, cof);
    *start = calloc(nc + 2, 4);
    *order = malloc(4 * (n + 1));
    for (uint32_t i = 0; i < n; i++) (*start)[cof[i] + 1]++;
    for (uint32_t c = 0; c < nc; c++) (*start)[c + 1] += (*start)[c];
    uint32_t *fill = calloc(nc + 1, 4);
    for (uint32_t i = 0; i < n; i++) (*order)[(*start)[cof[i]] + fill[cof[i]]++] = i;
    free(fill); free(cof);
    return nc;
}
static V3 pseudo_rel(const OrWorld *w, const CC *c) {
    V3 la = w->pl[c->ai], wa = w->pa[c->ai];
    V3 lb = c->terrain ? v3s(0) : w->pl[c->bi], wb = c->terrain ? v3s(0) : w->pa[c->bi];
    return v3(vdot(vsub(vadd(lb, vcross(wb, c->rb)), vadd(la, vcross(wa, c->ra))), c->n), 0.0f, 0.0f);
}
static void w6_passes(OrWorld *w, const CC *cs, uint32_t nct, float *jn, const float *jt, const uint8_t *jointed, float dt) {
    uint32_t nb = w->n;
    uint8_t *toi_set = calloc(nct + 1, 1);
    int any = 0;
    for (uint32_t i = 0; i < nct; i++) {
        toi_set[i] = cs[i].toi && cs[i].rest_bias > 0.0f && jn[i] > 0.0f;
        any |= toi_set[i];
    }
    W6Row *prow_src = malloc(sizeof(W6Row) * (nct + 1));  /* (constraint, pseudo u) */
    uint32_t npseudo = 0;
    if (any) {
        /* loaded contacts outside the bounce, per body: who each one is with (solver.rs:1051-1057);
         * only "is there a partner other than p" is asked, so a body keeps its first partner and
         * whether a second, different one exists */
        uint32_t *p0 = malloc(4 * (nb + 1));
        uint8_t *pc = calloc(nb + 1, 1), *pm = calloc(nb + 1, 1);
#define W6_REST(id, p) do { uint32_t _i = (id), _p = (p); if (!pc[_i]) { pc[_i] = 1; p0[_i] = _p; } else if (p0[_i] != _p) pm[_i] = 1; } while (0)
        for (uint32_t i = 0; i < nct; i++) {
            if (toi_set[i] || !loaded_at(jn, jt, i)) continue;
            W6_REST(cs[i].ai, cs[i].terrain ? TERRAIN_CI : cs[i].bi);
            if (!cs[i].terrain) W6_REST(cs[i].bi, cs[i].ai);
        }
#undef W6_REST
        W6Row *rows = malloc(sizeof(W6Row) * (nct + 1));
        uint32_t nr = 0;
        for (uint32_t i = 0; i < nct; i++) {
            if (!toi_set[i]) continue;
            const CC *c = &cs[i];
            /* a side resting on something other than its partner here is struck from its start
             * velocity (solver.rs:1063-1068) */
            uint32_t oa = c->terrain ? TERRAIN_CI : c->bi;
            int sup_a = pc[c->ai] && (pm[c->ai] || p0[c->ai] != oa);
            float a0 = c->toi_va0, a1 = sup_a ? c->toi_va0 : c->toi_va1;
            float b0 = 0.0f, b1 = 0.0f;
            if (!c->terrain) {
                int sup_b = pc[c->bi] && (pm[c->bi] || p0[c->bi] != c->ai);
                b0 = c->toi_vb0; b1 = sup_b ? c->toi_vb0 : c->toi_vb1;
            }
            float v0 = b0 - a0;
            float acc = ((b1 - a1) - v0) / dt;
            float u; int hp;
            float v_end = bn_toi_bounce(c->toi_s0, v0, acc, c->toi_e, dt, &u, &hp);
            rows[nr].c = *c; rows[nr].target = v_end; rows[nr].cidx = (int32_t)i; nr++;
            if (hp) { prow_src[npseudo].c = *c; prow_src[npseudo].target = u; prow_src[npseudo].cidx = (int32_t)i; npseudo++; }
        }
        free(p0); free(pc); free(pm);
        uint32_t ntoi = nr;
        uint8_t *inv = calloc(nb + 1, 1);
        for (uint32_t r = 0; r < ntoi; r++) { inv[rows[r].c.ai] = 1; if (!rows[r].c.terrain) inv[rows[r].c.bi] = 1; }
        w6_grow(cs, nct, jn, jt, jointed, inv);
        /* the bounced bodies' other loaded contacts: push-only support rows with fresh
         * accumulators and target min(depth_bias, 0) (solver.rs:1081-1092) */
        for (uint32_t i = 0; i < nct; i++) {
            if (toi_set[i] || cs[i].soft || jn[i] <= 0.0f) continue;
            const CC *c = &cs[i];
            if (!inv[c->ai] && (c->terrain || !inv[c->bi])) continue;
            CC k;
            if (!w6_free_only(c, jointed, &k)) continue;
            rows[nr].c = k; rows[nr].target = fminf(c->depth_bias, 0.0f); rows[nr].cidx = -1; nr++;
        }
        free(inv);
        float *sup = calloc(nr + 1, 4);
        BSide (*sides)[2] = malloc(sizeof(BSide) * 2 * (nr + 1));
        for (uint32_t r = 0; r < nr; r++) w6_sides(&rows[r].c, sides[r]);
        uint32_t *ord, *cst;
        uint32_t ncomp = w6_comps((const BSide(*)[2])sides, nr, &ord, &cst);
        w->stats.toi_rows = ntoi;
        w->stats.support_rows = nr - ntoi;
        w->stats.bounce_components = ncomp;
        for (uint32_t cc = 0; cc < ncomp; cc++) {
            const uint32_t *comp = ord + cst[cc];
            uint32_t n = cst[cc + 1] - cst[cc];
            int sweep = n > BN_DIRECT_MAX;
            if (!sweep) {
                BSide (*cs2)[2] = malloc(sizeof(BSide) * 2 * (n + 1));
                double *b = malloc(sizeof(double) * (n + 1)), *acc0 = malloc(sizeof(double) * (n + 1)), *delta = malloc(sizeof(double) * (n + 1));
                for (uint32_t q = 0; q < n; q++) {
                    const W6Row *r = &rows[comp[q]];
                    cs2[q][0] = sides[comp[q]][0]; cs2[q][1] = sides[comp[q]][1];
                    b[q] = (double)(r->target - vdot(rel_vel(w, &r->c), r->c.n));
                    acc0[q] = (double)(r->cidx >= 0 ? jn[r->cidx] : sup[comp[q]]);
                }
                BCsr K = bn_contact_csr((const BSide(*)[2])cs2, n);
                uint32_t solves = 0;
                int conv = bn_bounce_qp(&K, b, acc0, BN_SOLVES, delta, &solves);
                if (solves > w->stats.bounce_solves) w->stats.bounce_solves = solves;
                sweep = !conv;
                for (uint32_t q = 0; q < n; q++) {
                    const W6Row *r = &rows[comp[q]];
                    float *slot = r->cidx >= 0 ? &jn[r->cidx] : &sup[comp[q]];
                    float old = *slot;
                    float nv = fmaxf((float)((double)old + delta[q]), 0.0f);
                    *slot = nv;
                    apply_cimp(w, &r->c, vmul(r->c.n, nv - old));
                }
                bn_csr_free(&K);
                free(cs2); free(b); free(acc0); free(delta);
            }
            if (!sweep) continue;
            w->stats.bounce_fallbacks++;
            /* above the direct limit, or not converged: Gauss-Seidel from where it got to,
             * alternating direction (solver.rs:1135-1154) */
            for (uint32_t it = 0; it < BN_SWEEPS; it++) {
                float change = 0.0f;
                for (uint32_t m = 0; m < n; m++) {
                    uint32_t ri = comp[it % 2 == 0 ? m : n - 1 - m];
                    const W6Row *r = &rows[ri];
                    float v = vdot(rel_vel(w, &r->c), r->c.n);
                    float *slot = r->cidx >= 0 ? &jn[r->cidx] : &sup[ri];
                    float old = *slot;
                    float nv = fmaxf(old + (r->target - v) * r->c.m_n, 0.0f);
                    *slot = nv;
                    apply_cimp(w, &r->c, vmul(r->c.n, nv - old));
                    change = fmaxf(change, fabsf(nv - old) / r->c.m_n / (1.0f + fabsf(r->target)));
                }
                if (change < BN_TOL) break;
            }
        }
        free(ord); free(cst); free(sides); free(sup); free(rows);
    }
    /* the bounce's position: pseudo-velocities along the bounced contacts' normals, the involved
     * bodies' other loaded contacts holding their gaps both ways (solver.rs:1163-1243) */
    if (npseudo > 0) {
        uint8_t *in_set = calloc(nct + 1, 1), *inv = calloc(nb + 1, 1);
        for (uint32_t r = 0; r < npseudo; r++) {
            const CC *c = &prow_src[r].c;
            in_set[prow_src[r].cidx] = 1;
            inv[c->ai] = 1;
            if (!c->terrain) inv[c->bi] = 1;
        }
        w6_grow(cs, nct, jn, jt, jointed, inv);
        w->stats.pseudo_rows = npseudo;
        W6Row *rows = malloc(sizeof(W6Row) * (nct + 1));
        uint32_t nr = 0;
        for (uint32_t r = 0; r < npseudo; r++) {
            CC k;
            if (w6_free_only(&cs[prow_src[r].cidx], jointed, &k)) { rows[nr].c = k; rows[nr].target = prow_src[r].target; rows[nr].cidx = -1; nr++; }
        }
        for (uint32_t i = 0; i < nct; i++) {
            if (in_set[i] || cs[i].soft || !loaded_at(jn, jt, i)) continue;
            const CC *c = &cs[i];
            if (!inv[c->ai] && (c->terrain || !inv[c->bi])) continue;
            CC k;
            if (w6_free_only(c, jointed, &k)) { rows[nr].c = k; rows[nr].target = 0.0f; rows[nr].cidx = -1; nr++; }
        }
        for (uint32_t b = 0; b < nb; b++) { w->pl[b] = v3s(0); w->pa[b] = v3s(0); }
        BSide (*sides)[2] = malloc(sizeof(BSide) * 2 * (nr + 1));
        for (uint32_t r = 0; r < nr; r++) w6_sides(&rows[r].c, sides[r]);
        D3 *accl = calloc(nb + 1, sizeof(D3)), *acca = calloc(nb + 1, sizeof(D3));
        uint8_t *direct = calloc(nb + 1, 1);
        uint32_t *ord, *cst;
        uint32_t ncomp = w6_comps((const BSide(*)[2])sides, nr, &ord, &cst);
        for (uint32_t cc = 0; cc < ncomp; cc++) {
            const uint32_t *comp = ord + cst[cc];
            uint32_t n = cst[cc + 1] - cst[cc];
            if (n <= BN_DIRECT_MAX) {
                BSide (*cs2)[2] = malloc(sizeof(BSide) * 2 * (n + 1));
                double *u = malloc(sizeof(double) * (n + 1)), *x = malloc(sizeof(double) * (n + 1));
                for (uint32_t q = 0; q < n; q++) {
                    cs2[q][0] = sides[comp[q]][0]; cs2[q][1] = sides[comp[q]][1];
                    u[q] = (double)rows[comp[q]].target;
                }
                BCsr K = bn_contact_csr((const BSide(*)[2])cs2, n);
                bn_pseudo_solve(&K, u, x);
                for (uint32_t q = 0; q < n; q++)
                    for (int t = 0; t < 2; t++) {
                        const BSide *sd = &cs2[q][t];
                        if (sd->body == BN_NONE) continue;
                        accl[sd->body] = d3add(accl[sd->body], d3mul(sd->ml, x[q]));
                        acca[sd->body] = d3add(acca[sd->body], d3mul(sd->ma, x[q]));
                        direct[sd->body] = 1;
                    }
                bn_csr_free(&K);
                free(cs2); free(u); free(x);
                continue;
            }
            w->stats.bounce_fallbacks++;
            for (uint32_t it = 0; it < BN_SWEEPS; it++) {
                float change = 0.0f;
                for (uint32_t m = 0; m < n; m++) {
                    const W6Row *r = &rows[comp[it % 2 == 0 ? m : n - 1 - m]];
                    const CC *c = &r->c;
                    float d = r->target - pseudo_rel(w, c).x;
                    change = fmaxf(change, fabsf(d) / (1.0f + fabsf(r->target)));
                    V3 j = vmul(c->n, d * c->m_n);
                    if (c->inv_a > 0.0f) {
                        w->pl[c->ai] = vsub(w->pl[c->ai], vmul(j, c->inv_a));
                        w->pa[c->ai] = vsub(w->pa[c->ai], m3v(c->ii_a, vcross(c->ra, j)));
                    }
                    if (c->inv_b > 0.0f) {
                        w->pl[c->bi] = vadd(w->pl[c->bi], vmul(j, c->inv_b));
                        w->pa[c->bi] = vadd(w->pa[c->bi], m3v(c->ii_b, vcross(c->rb, j)));
                    }
                }
                if (change < BN_TOL) break;
            }
        }
        for (uint32_t b = 0; b < nb; b++)
            if (direct[b]) {
                w->pl[b] = v3((float)accl[b].x, (float)accl[b].y, (float)accl[b].z);
                w->pa[b] = v3((float)acca[b].x, (float)acca[b].y, (float)acca[b].z);
            }
        w->has_pseudo = 1;
        free(ord); free(cst); free(sides); free(accl); free(acca); free(direct); free(rows); free(in_set); free(inv);
    }
    free(prow_src); free(toi_set);
}

static void solve(OrWorld *w, Contact *cts, uint32_t nct, const V3 *pre_lin, const V3 *pre_ang, float dt) {
    const float beta = 0.05f, slop = 0.005f;
    uint32_t nb = w->n;
    CC *cs = malloc(sizeof(CC) * (nct + 1));
    /* bodies in any joint (solver.rs:483-491): W6 leaves their contacts as they were */
    uint8_t *jointed = calloc(nb + 1, 1);
    for (uint32_t q = 0; q < w->nj; q++) { jointed[w->joints[q].a] = 1; jointed[w->joints[q].b] = 1; }
    for (uint32_t i = 0; i < nct; i++) {
        const Contact *c = &cts[i];
        CC *k = &cs[i];
        k->ai = w->cbody[c->a];
        k->terrain = c->b == TERRAIN_CI;
        k->bi = k->terrain ? 0 : w->cbody[c->b];
        int a_aw = w->awake[k->ai];
        int b_aw = k->terrain ? 0 : w->awake[k->bi];
        k->inv_a = a_aw ? w->inv_mass[k->ai] : 0.0f;
        k->inv_b = b_aw ? w->inv_mass[k->bi] : 0.0f;
        k->ii_a = k->inv_a > 0.0f ? inv_inertia_world(w, k->ai) : m3zero();
        k->ii_b = k->inv_b > 0.0f ? inv_inertia_world(w, k->bi) : m3zero();
        PairP pp = contact_props(w, c);
        k->n = c->normal;
        orthonormal_basis(k->n, &k->t1, &k->t2);
        k->ra = vsub(c->point, w->pos[k->ai]);
        k->rb = k->terrain ? v3s(0) : vsub(c->point, w->pos[k->bi]);
        k->m_n = eff_mass(k, k->n); k->m_t1 = eff_mass(k, k->t1); k->m_t2 = eff_mass(k, k->t2);
        k->mu_s = pp.mu_s; k->mu_k = pp.mu_k; k->vs = fmaxf(pp.vs, 1.0e-4f);
        float pen = c->depth - MARGIN;
        k->depth_bias = pen > 0.0f ? fminf(fmaxf(pen - slop, 0.0f) * (beta / dt), 2.0f) : pen / dt;
        V3 pva = vadd(pre_lin[k->ai], vcross(pre_ang[k->ai], k->ra));
        V3 pvb = k->terrain ? v3s(0) : vadd(pre_lin[k->bi], vcross(pre_ang[k->bi], k->rb));
        float v_pre = vdot(vsub(pvb, pva), k->n);
        float vrn = 0.5f * (v_pre + vdot(rel_vel(w, k), k->n));
        k->approach = vrn;
        k->rest_bias = 0.0f;
        k->toi = 0; k->toi_s0 = k->toi_e = 0.0f; k->toi_va0 = k->toi_va1 = k->toi_vb0 = k->toi_vb1 = 0.0f;
        k->soft = pp.stiff > 0.0f; k->soft_implicit = 0; k->soft_gamma = k->soft_bias = k->soft_m = 0.0f; k->patch = 0;
        if (k->soft) {
            /* implicit Hunt-Crossley (solver.rs:545-566): the spring acts at the end-of-step
             * penetration, the damping linearised at the predicted mid-step penetration (or the
             * step-start one); no Baumgarte bias */
            k->depth_bias = 0.0f;
            float kk = pp.stiff;
            float vn0 = vdot(rel_vel(w, k), k->n);
            float pen_d = w->damp_at_start ? pen : pen - (0.5f * dt) * vn0;
            float cdamp = (kk * fmaxf(pp.diss, 0.0f)) * fmaxf(pen_d, 0.0f);
            float denom = cdamp + dt * kk;
            if (k->m_n > 0.0f && denom > 0.0f) {
                k->soft_implicit = 1;
                k->soft_gamma = 1.0f / (dt * denom);
                k->soft_bias = (((dt * kk) / denom) * pen) / dt;
                k->soft_m = 1.0f / (1.0f / k->m_n + k->soft_gamma);
            }
            continue;
        }
        if (vrn < -1.0f) {
            k->rest_bias = -pp.rest * vrn;
            /* W6: free bodies bounce at their time of impact (solver.rs:572-587); a side's start
             * velocity is last step's end linear velocity with the pre-step angular one */
            if (!jointed[k->ai] && (k->terrain || !jointed[k->bi])) {
                k->toi = 1;
                k->toi_s0 = -pen;
                k->toi_e = pp.rest;
                k->toi_va0 = vdot(vadd(w->last_lin[k->ai], vcross(pre_ang[k->ai], k->ra)), k->n);
                k->toi_va1 = vdot(vadd(w->lin[k->ai], vcross(w->ang[k->ai], k->ra)), k->n);
                if (!k->terrain) {
                    k->toi_vb0 = vdot(vadd(w->last_lin[k->bi], vcross(pre_ang[k->bi], k->rb)), k->n);
                    k->toi_vb1 = vdot(vadd(w->lin[k->bi], vcross(w->ang[k->bi], k->rb)), k->n);
                }
            }
        }
    }
    /* friction patches (solver.rs:592-619): manifolds of two or more points on patch-friction pairs,
     * in (collider a, collider b) order */
    uint32_t *ppts = malloc(4 * (nct + 1)), npp = 0;
    for (uint32_t i = 0; i < nct; i++) if (contact_props(w, &cts[i]).patch) ppts[npp++] = i;
    for (uint32_t x = 1; x < npp; x++) {  /* stable insertion sort by the collider pair */
        uint32_t v = ppts[x], y = x;
        while (y > 0 && (cts[ppts[y - 1]].a > cts[v].a || (cts[ppts[y - 1]].a == cts[v].a && cts[ppts[y - 1]].b > cts[v].b))) { ppts[y] = ppts[y - 1]; y--; }
        ppts[y] = v;
    }
    Patch *pt = malloc(sizeof(Patch) * (npp / 2 + 1));
    uint32_t npt = 0;
    for (uint32_t g = 0; g < npp;) {
        uint32_t e = g;
        while (e < npp && cts[ppts[e]].a == cts[ppts[g]].a && cts[ppts[e]].b == cts[ppts[g]].b) e++;
        if (e - g >= 2 && build_patch(w, cts, cs, ppts + g, e - g, &pt[npt])) { pt[npt].p0 = g; npt++; }
        g = e;
    }
    /* level schedule, for the stats (solver.rs:621-642) */
    uint32_t *blev = calloc(nb + 1, 4);
    uint32_t levels = 0;
    for (uint32_t i = 0; i < nct; i++) {
        int wa = cs[i].inv_a > 0.0f, wb = !cs[i].terrain && cs[i].inv_b > 0.0f;
        if (!wa && !wb) continue;
        uint32_t la = wa ? blev[cs[i].ai] : 0, lb = wb ? blev[cs[i].bi] : 0;
        uint32_t t = la > lb ? la : lb;
        if (t + 1 > levels) levels = t + 1;
        if (wa) blev[cs[i].ai] = t + 1;
        if (wb) blev[cs[i].bi] = t + 1;
    }
    free(blev);
    w->stats.levels = levels;

    /* joint rows of the active joints (a dynamic body on either side, islands.rs:96-101;
     * solver.rs:647-665), then their warm start from last step's rows when the tags agree, serial in
     * joint order (solver.rs:700-715) */
    JRow *jr = malloc(sizeof(JRow) * (OR_JROWS_MAX * w->nj + 1));
    uint32_t *jr0 = malloc(4 * (w->nj + 1)), *jr1 = malloc(4 * (w->nj + 1)), njr = 0;
    for (uint32_t q = 0; q < w->nj; q++) {
        const Joint *j = &w->joints[q];
        jr0[q] = njr;
        if (j->kind == J_BALL && (w->inv_mass[j->a] > 0.0f || w->inv_mass[j->b] > 0.0f)) njr += ball_rows(w, j, jr + njr);
        if (j->kind == J_D6 && (w->inv_mass[j->a] > 0.0f || w->inv_mass[j->b] > 0.0f)) njr += d6_rows(w, j, dt, jr + njr);
        jr1[q] = njr;
    }
    float *racc = calloc(njr + 1, 4);
    for (uint32_t q = 0; q < w->nj; q++) {
        uint32_t s0 = jr0[q], e0 = jr1[q];
        if (e0 == s0 || !w->wj[q].has || w->wj[q].n != e0 - s0) continue;
        int same = 1;
        for (uint32_t k = s0; k < e0; k++) if (w->wj[q].tag[k - s0] != jr[k].tag) same = 0;
        if (!same) continue;
        for (uint32_t k = s0; k < e0; k++) {
            float lam = clampf_rs(w->wj[q].lam[k - s0], jr[k].lo, jr[k].hi);
            racc[k] = lam;
            if (lam != 0.0f) jrow_apply(w, &jr[k], lam);
        }
    }

    float *jn = calloc(nct + 1, 4), *jt = calloc(2 * nct + 2, 4);
    /* warm start: keys and anchors (solver.rs:717-762) */
    uint32_t (*keys)[4] = malloc(sizeof(uint32_t) * 4 * (nct + 1));
    V3 *anch = malloc(sizeof(V3) * (nct + 1));
    for (uint32_t i = 0; i < nct; i++) {
        keys[i][0] = cs[i].ai; keys[i][1] = cs[i].terrain ? 0xFFFFFFFFu : cs[i].bi; keys[i][2] = 0; keys[i][3] = 0;
        anch[i] = qrot(qconj(w->rot[cs[i].ai]), vsub(cts[i].point, w->pos[cs[i].ai]));
    }
    uint8_t *taken = calloc(w->nwarm + 1, 1);
    uint32_t matched = 0;
    for (uint32_t i = 0; i < nct; i++) {
        if (cs[i].soft) continue;  /* compliant rows start cold */
        int best = warm_pick(w->warm, w->nwarm, keys[i], anch[i], taken);
        if (best < 0) continue;
        matched++;
        const Warm *e = &w->warm[best];
        const CC *k = &cs[i];
        jn[i] = e->jn;
        V3 imp = vmul(k->n, e->jn);
        if (!k->patch) {
            V3 v = rel_vel(w, k);
            float vt1 = vdot(v, k->t1), vt2 = vdot(v, k->t2);
            float mu = mu_dyn(k, sqrtf(vt1 * vt1 + vt2 * vt2));
            V3 t = vclamp_len_max(v3(vdot(e->jt, k->t1), vdot(e->jt, k->t2), 0.0f), mu * e->jn);
            jt[2 * i] = t.x; jt[2 * i + 1] = t.y;
            imp = vadd(imp, vadd(vmul(k->t1, t.x), vmul(k->t2, t.y)));
        }
        apply_cimp(w, k, imp);
    }
    free(taken);
    /* patches (solver.rs:763-782): sliding and torsion from the carried patch impulses, limited by
     * the points' warm normal load; compliant pairs start cold */
    float *pacc = calloc(3 * npt + 3, 4);
    for (uint32_t k = 0; k < npt; k++) {
        const Patch *P = &pt[k];
        if (P->at.soft) continue;
        const uint32_t *pp = ppts + P->p0;
        int e = wpatch_find(w, keys[pp[0]]);
        if (e < 0) continue;
        const WarmP *wp = &w->wpatch[e];
        const CC *c = &P->at;
        float jsum = patch_jn(jn, pp, P->np);
        float mu = mu_dyn(c, slip_speed(w, c));
        V3 t = vclamp_len_max(v3(vdot(wp->jt, c->t1), vdot(wp->jt, c->t2), 0.0f), mu * jsum);
        float lim = (mu * jsum) * P->radius;
        float tw = clampf_rs(wp->tw, -lim, lim);
        pacc[3 * k] = t.x; pacc[3 * k + 1] = t.y; pacc[3 * k + 2] = tw;
        apply_cimp(w, c, vadd(vmul(c->t1, t.x), vmul(c->t2, t.y)));
        if (tw != 0.0f) patch_twist(w, c, tw);
    }
    w->stats.warm_matched = matched;

    /* islands of the written bodies (residual mode stops each island on its own residual).  A joint
     * writes a sleeping body too; a contact that reads such a body joins its island. */
    uint32_t *par = malloc(sizeof(uint32_t) * (nb + 1));
    uint8_t *jw = calloc(nb + 1, 1);
    for (uint32_t k = 0; k < njr; k++) { if (jr[k].wa > 0.0f) jw[jr[k].a] = 1; if (jr[k].wb > 0.0f) jw[jr[k].b] = 1; }
    /* a rigid distance joint's legacy row (solver.rs:795-820) writes each side with mass; a compliant
     * one acts at position level only */
    uint8_t *dact = calloc(w->nj + 1, 1);
    uint32_t *disl = calloc(w->nj + 1, 4);
    for (uint32_t q = 0; q < w->nj; q++) {
        const Joint *j = &w->joints[q];
        if (j->kind != J_DIST || j->compliance > 0.0f) continue;
        if (!(w->inv_mass[j->a] > 0.0f || w->inv_mass[j->b] > 0.0f)) continue;
        dact[q] = 1;
        if (w->inv_mass[j->a] > 0.0f) jw[j->a] = 1;
        if (w->inv_mass[j->b] > 0.0f) jw[j->b] = 1;
    }
    for (uint32_t b = 0; b <= nb; b++) par[b] = b;
    for (uint32_t i = 0; i < nct; i++) {
        int wa = cs[i].inv_a > 0.0f, wb = !cs[i].terrain && cs[i].inv_b > 0.0f;
        if (wa || wb) { wa = wa || jw[cs[i].ai]; wb = wb || (!cs[i].terrain && jw[cs[i].bi]); }
        if (wa && wb) {
            uint32_t x = uf_find(par, cs[i].ai), y = uf_find(par, cs[i].bi);
            if (x != y) par[x > y ? x : y] = x < y ? x : y;
        }
    }
    for (uint32_t k = 0; k < njr; k++) {
        if (jr[k].wa > 0.0f && jr[k].wb > 0.0f) {
            uint32_t x = uf_find(par, jr[k].a), y = uf_find(par, jr[k].b);
            if (x != y) par[x > y ? x : y] = x < y ? x : y;
        }
    }
    for (uint32_t q = 0; q < w->nj; q++) {
        const Joint *j = &w->joints[q];
        if (dact[q] && w->inv_mass[j->a] > 0.0f && w->inv_mass[j->b] > 0.0f) {
            uint32_t x = uf_find(par, j->a), y = uf_find(par, j->b);
            if (x != y) par[x > y ? x : y] = x < y ? x : y;
        }
    }
    uint32_t *isl = malloc(sizeof(uint32_t) * (nct + 1));
    uint32_t *isl_n = calloc(nb + 1, 4), *isl_it = calloc(nb + 1, 4);
    uint8_t *isl_done = calloc(nb + 1, 1);
    float *isl_r = calloc(nb + 1, 4);
    for (uint32_t i = 0; i < nct; i++) {
        int wa = cs[i].inv_a > 0.0f, wb = !cs[i].terrain && cs[i].inv_b > 0.0f;
        isl[i] = wa ? uf_find(par, cs[i].ai) : (wb ? uf_find(par, cs[i].bi) : 0xFFFFFFFFu);
        if (isl[i] != 0xFFFFFFFFu) isl_n[isl[i]]++;
    }
    /* rows per island (contacts and joint rows): an island with rows takes part in the sweeps */
    uint32_t *isl_rows = calloc(nb + 1, 4), *jisl = malloc(4 * (njr + 1));
    for (uint32_t i = 0; i < nct; i++) if (isl[i] != 0xFFFFFFFFu) isl_rows[isl[i]]++;
    for (uint32_t k = 0; k < njr; k++) {
        jisl[k] = uf_find(par, jr[k].wa > 0.0f ? jr[k].a : jr[k].b);
        isl_rows[jisl[k]]++;
    }
    for (uint32_t q = 0; q < w->nj; q++) {
        if (!dact[q]) continue;
        const Joint *j = &w->joints[q];
        disl[q] = uf_find(par, w->inv_mass[j->a] > 0.0f ? j->a : j->b);
        isl_rows[disl[q]]++;
    }
    /* the sweep statistics count a joint with rows as one entry of its island, like a contact, and
     * so a friction patch */
    for (uint32_t q = 0; q < w->nj; q++) {
        if (jr1[q] > jr0[q]) isl_n[jisl[jr0[q]]]++;
        else if (dact[q]) isl_n[disl[q]]++;
    }
    for (uint32_t k = 0; k < npt; k++)
        if (isl[ppts[pt[k].p0]] != 0xFFFFFFFFu) isl_n[isl[ppts[pt[k].p0]]]++;

    /* PGS iterations (solver.rs:788-880) as the reference's level batches (solver.rs:621-687):
     * joints by joint level, then contacts by contact level, each batch in index order.  Rows of
     * one batch write pairwise distinct bodies, so the batched order gives the serial order's bits,
     * and a batch of at least PAR_MIN rows runs in parallel as run_batch does (solver.rs:21-23). */
    uint32_t *jlev = calloc(nb + 1, 4), *clev = calloc(nb + 1, 4);
    uint32_t *jord = malloc(4 * (w->nj + 1)), *jlo = calloc(w->nj + 2, 4), njl = 0;
    uint32_t *cord = malloc(4 * (nct + 1)), *clo = calloc(nct + 2, 4), ncl = 0;
    {
        uint32_t *lv = malloc(4 * (w->nj + 1));
        for (uint32_t q = 0; q < w->nj; q++) {
            lv[q] = 0xFFFFFFFFu;
            if (!dact[q] && jr1[q] == jr0[q]) continue;
            uint32_t a = w->joints[q].a, b = w->joints[q].b;
            uint32_t t = jlev[a] > jlev[b] ? jlev[a] : jlev[b];
            lv[q] = t; jlev[a] = jlev[b] = t + 1;
            if (t + 1 > njl) njl = t + 1;
        }
        for (uint32_t q = 0; q < w->nj; q++) if (lv[q] != 0xFFFFFFFFu) jlo[lv[q] + 1]++;
        for (uint32_t l = 0; l < njl; l++) jlo[l + 1] += jlo[l];
        uint32_t *fill = calloc(njl + 1, 4);
        for (uint32_t q = 0; q < w->nj; q++) if (lv[q] != 0xFFFFFFFFu) jord[jlo[lv[q]] + fill[lv[q]]++] = q;
        free(fill); free(lv);
    }
    {
        uint32_t *lv = malloc(4 * (nct + 1));
        for (uint32_t i = 0; i < nct; i++) {
            int wa = cs[i].inv_a > 0.0f, wb = !cs[i].terrain && cs[i].inv_b > 0.0f;
            lv[i] = 0xFFFFFFFFu;
            if (isl[i] == 0xFFFFFFFFu) continue;
            uint32_t la = wa ? clev[cs[i].ai] : 0, lb = wb ? clev[cs[i].bi] : 0;
            uint32_t t = la > lb ? la : lb;
            lv[i] = t;
            if (wa) clev[cs[i].ai] = t + 1;
            if (wb) clev[cs[i].bi] = t + 1;
            if (t + 1 > ncl) ncl = t + 1;
        }
        for (uint32_t i = 0; i < nct; i++) if (lv[i] != 0xFFFFFFFFu) clo[lv[i] + 1]++;
        for (uint32_t l = 0; l < ncl; l++) clo[l + 1] += clo[l];
        uint32_t *fill = calloc(ncl + 1, 4);
        for (uint32_t i = 0; i < nct; i++) if (lv[i] != 0xFFFFFFFFu) cord[clo[lv[i]] + fill[lv[i]]++] = i;
        free(fill); free(lv);
    }
    float *jres = calloc(w->nj + 1, 4), *cres = calloc(nct + 1, 4);
    uint8_t *jran = calloc(w->nj + 1, 1), *cran = calloc(nct + 1, 1);
    uint32_t iters = w->resid_on ? w->resid_max : w->iv;
    uint32_t used = 0;
    float resid = 0;
    for (uint32_t it = 0; it < iters; it++) {
        int active = 0;
        for (uint32_t b = 0; b <= nb; b++) isl_r[b] = 0.0f;
        /* A. joints (solver.rs:789-835) */
        for (uint32_t l = 0; l < njl; l++) {
            uint32_t lo = jlo[l], hi = jlo[l + 1];
#pragma omp parallel for schedule(static) if (hi - lo >= PAR_MIN) num_threads(g_threads)
            for (uint32_t o = lo; o < hi; o++) {
                uint32_t q = jord[o];
                jran[q] = 0; jres[q] = 0.0f;
                if (dact[q]) {
                    if (isl_done[disl[q]]) continue;
                    jran[q] = 1;
                    jres[q] = dist_vel_row(w, &w->joints[q], dt);
                    continue;
                }
                for (uint32_t k = jr0[q]; k < jr1[q]; k++) {
                    const JRow *r = &jr[k];
                    if (isl_done[jisl[k]]) continue;
                    jran[q] = 1;
                    float v = jrow_vel(w, r);
                    float old = racc[k];
                    /* rigid rows old + (target - v) m; soft rows (dampers) add gamma lam */
                    float nwv = r->gamma == 0.0f ? old + (r->target - v) * r->m : old + ((r->target - v) - r->gamma * old) * r->m;
                    nwv = clampf_rs(nwv, r->lo, r->hi);
                    racc[k] = nwv;
                    if (nwv != old) jrow_apply(w, r, nwv - old);
                    float rr = fabsf(nwv - old) / r->m;
                    if (rr > jres[q]) jres[q] = rr;
                }
            }
        }
        for (uint32_t q = 0; q < w->nj; q++) {
            if (!jran[q]) continue;
            active = 1;
            uint32_t b = dact[q] ? disl[q] : jisl[jr0[q]];
            if (w->resid_on && jres[q] > isl_r[b]) isl_r[b] = jres[q];
        }
        /* B. contacts (solver.rs:838-880) */
        for (uint32_t l = 0; l < ncl; l++) {
            uint32_t lo = clo[l], hi = clo[l + 1];
#pragma omp parallel for schedule(static) if (hi - lo >= PAR_MIN) num_threads(g_threads)
            for (uint32_t o = lo; o < hi; o++) {
                uint32_t i = cord[o];
                const CC *k = &cs[i];
                cran[i] = 0;
                if (isl_done[isl[i]]) continue;
                cran[i] = 1;
                float old = jn[i], nw = old;
                if (k->soft_implicit) {  /* the implicit compliant row (solver.rs:848-855) */
                    float vrn = vdot(rel_vel(w, k), k->n);
                    nw = fmaxf(old + k->soft_m * ((k->soft_bias - vrn) - k->soft_gamma * old), 0.0f);
                    jn[i] = nw;
                    if (nw != old) apply_cimp(w, k, vmul(k->n, nw - old));
                } else if (!k->soft) {
                    float vrn = vdot(rel_vel(w, k), k->n);
                    float jraw = (k->depth_bias - vrn) * k->m_n;
                    nw = fmaxf(old + jraw, 0.0f);
                    jn[i] = nw;
                    apply_cimp(w, k, vmul(k->n, nw - old));
                }
                const float rm = k->soft_implicit ? k->soft_m : k->m_n, rt = k->soft_implicit ? k->soft_bias : k->depth_bias;
                float r0 = rm > 0.0f ? fabsf(nw - old) / rm / (1.0f + fabsf(rt)) : 0.0f;
                if (k->patch) { cres[i] = r0; continue; }  /* friction by the patch (solver.rs:865) */
                V3 vf = rel_vel(w, k);
                float vt1 = vdot(vf, k->t1), vt2 = vdot(vf, k->t2);
                float slip = sqrtf(vt1 * vt1 + vt2 * vt2);
                float mu = mu_dyn(k, slip);
                V3 ot = v3(jt[2 * i], jt[2 * i + 1], 0.0f);
                V3 nt = vclamp_len_max(vsub(ot, v3(vt1 * k->m_t1, vt2 * k->m_t2, 0.0f)), mu * nw);
                jt[2 * i] = nt.x; jt[2 * i + 1] = nt.y;
                V3 d = vsub(nt, ot);
                apply_cimp(w, k, vadd(vmul(k->t1, d.x), vmul(k->t2, d.y)));
                float r1 = k->m_t1 > 0.0f ? fabsf(d.x) / k->m_t1 : 0.0f;
                float r2 = k->m_t2 > 0.0f ? fabsf(d.y) / k->m_t2 : 0.0f;
                cres[i] = fmaxf(r0, fmaxf(r1, r2));
            }
        }
        for (uint32_t o = 0; o < clo[ncl]; o++) {
            uint32_t i = cord[o];
            if (!cran[i]) continue;
            active = 1;
            if (w->resid_on && cres[i] > isl_r[isl[i]]) isl_r[isl[i]] = cres[i];
        }
        /* C. friction patches (solver.rs:882-915), in patch order; a patch with no writable body
         * runs after the loop */
        for (uint32_t k = 0; k < npt; k++) {
            uint32_t pis = isl[ppts[pt[k].p0]];
            if (pis == 0xFFFFFFFFu || isl_done[pis]) continue;
            float r = patch_row(w, &pt[k], ppts + pt[k].p0, jn, pacc + 3 * k);
            active = 1;
            if (w->resid_on && r > isl_r[pis]) isl_r[pis] = r;
        }
        if (!active) break;
        used++;
        for (uint32_t b = 0; b <= nb; b++) {
            if (!isl_rows[b] || isl_done[b]) continue;
            isl_it[b]++;
            if (w->resid_on && isl_r[b] <= w->resid_tol) isl_done[b] = 1;
        }
    }
    uint32_t isum = 0;
    resid = 0.0f;
    for (uint32_t b = 0; b <= nb; b++) {
        isum += isl_it[b] * isl_n[b];
        if (isl_rows[b] && isl_r[b] > resid) resid = isl_r[b];
    }
    for (uint32_t it = 0; it < iters; it++)
        for (uint32_t k = 0; k < npt; k++)
            if (isl[ppts[pt[k].p0]] == 0xFFFFFFFFu) (void)patch_row(w, &pt[k], ppts + pt[k].p0, jn, pacc + 3 * k);
    w->stats.iters_used = used;
    w->stats.iters_sum = isum;
    w->stats.residual = resid;
    free(par); free(isl); free(isl_n); free(isl_it); free(isl_done); free(isl_r); free(isl_rows); free(jisl); free(jw);
    free(dact); free(disl);
    free(jlev); free(clev); free(jord); free(jlo); free(cord); free(clo); free(jres); free(cres); free(jran); free(cran);
    /* joints solved this step replace their warm rows (solver.rs:945-952) */
    for (uint32_t q = 0; q < w->nj; q++) {
        uint32_t s0 = jr0[q], e0 = jr1[q];
        if (e0 == s0) continue;
        WarmJ *e = &w->wj[q];
        e->has = 1; e->n = (uint8_t)(e0 - s0);
        for (uint32_t k = s0; k < e0; k++) { e->tag[k - s0] = jr[k].tag; e->lam[k - s0] = racc[k]; }
    }
    free(jr); free(jr0); free(jr1); free(racc);

    /* warm-cache store (solver.rs:917-944) */
    Warm *nw = malloc(sizeof(Warm) * (nct + w->nwarm + 1));
    uint32_t nn = 0;
    for (uint32_t i = 0; i < nct; i++) {
        if (cs[i].soft) continue;
        V3 j = vadd(vmul(cs[i].t1, jt[2 * i]), vmul(cs[i].t2, jt[2 * i + 1]));
        if (jn[i] == 0.0f && viszero(j)) continue;
        Warm e; memcpy(e.key, keys[i], 16); e.la = anch[i]; e.jn = jn[i]; e.jt = j;
        nw[nn++] = e;
    }
    /* contacts arrive in key order (bucket order); merge the idle pairs' old entries */
    Warm *merged = malloc(sizeof(Warm) * (nn + w->nwarm + 1));
    uint32_t mcount = 0, p = 0, q = 0;
    while (p < nn || q < w->nwarm) {
        if (q < w->nwarm) {
            const uint32_t *ok = w->warm[q].key;
            int keep = is_idle(w, ok[0]) && is_idle(w, ok[1]);
            if (keep) {
                /* skip if the new map has this key */
                uint32_t lo = 0, hi = nn;
                while (lo < hi) { uint32_t mid = (lo + hi) / 2; if (key_cmp(nw[mid].key, ok) < 0) lo = mid + 1; else hi = mid; }
                if (lo < nn && key_cmp(nw[lo].key, ok) == 0) keep = 0;
            }
            if (!keep) { q++; continue; }
            if (p < nn && key_cmp(nw[p].key, ok) < 0) { merged[mcount++] = nw[p++]; continue; }
            merged[mcount++] = w->warm[q++];
            continue;
        }
        merged[mcount++] = nw[p++];
    }
    free(nw);
    free(w->warm);
    w->warm = merged; w->nwarm = mcount;
    store_patches(w, pt, npt, ppts, keys, pacc);

    /* shock propagation (solver.rs:964-993) */
    V3 up = vlen2(w->gravity) > 0.0f ? vneg(vnormalize(w->gravity)) : v3(0, 1, 0);
    SK *ord = malloc(sizeof(SK) * (nct + 1));
    uint32_t no = 0;
    for (uint32_t i = 0; i < nct; i++) {
        if (cs[i].soft || !(cs[i].inv_a > 0.0f || cs[i].inv_b > 0.0f)) continue;
        float ha = vdot(w->pos[cs[i].ai], up);
        float h = cs[i].terrain ? -INFINITY : fminf(ha, vdot(w->pos[cs[i].bi], up));
        ord[no].k = sort_key6(h); ord[no].idx = i; no++;
    }
    qsort(ord, no, sizeof(SK), cmp_sk_asc);
    for (uint32_t o = 0; o < no; o++) {
        uint32_t i = ord[o].idx;
        CC k = cs[i];
        if (k.inv_a > 0.0f && k.inv_b > 0.0f) {
            if (vdot(w->pos[k.ai], up) < vdot(w->pos[k.bi], up)) { k.inv_a = 0.0f; k.ii_a = m3zero(); }
            else { k.inv_b = 0.0f; k.ii_b = m3zero(); }
            k.m_n = eff_mass(&k, k.n);
        }
        float v = vdot(rel_vel(w, &k), k.n);
        float old = jn[i];
        float nwv = fmaxf(old + (k.depth_bias - v) * k.m_n, 0.0f);
        jn[i] = nwv;
        apply_cimp(w, &k, vmul(k.n, nwv - old));
    }
    free(ord);
    /* restitution relax (solver.rs:995-1008): contacts that carried load and bounce, other than
     * the time-of-impact ones, one sweep in index order */
    for (uint32_t i = 0; i < nct; i++) {
        const CC *k = &cs[i];
        if (k->toi || k->rest_bias <= 0.0f || jn[i] <= 0.0f) continue;
        w->stats.relax_rows++;
        float v = vdot(rel_vel(w, k), k->n);
        float old = jn[i];
        float nwv = fmaxf(old + (k->rest_bias - v) * k->m_n, 0.0f);
        jn[i] = nwv;
        apply_cimp(w, k, vmul(k->n, nwv - old));
    }
    w6_passes(w, cs, nct, jn, jt, jointed, dt);
    free(jointed);
    /* loaded flags + impulses (solver.rs:1247-1270) */
    memset(w->loaded, 0, w->cap + 1);
    w->loaded_valid = 1;
    for (uint32_t i = 0; i < nct; i++) {
        if (jn[i] > 0.0f || jt[2 * i] != 0.0f || jt[2 * i + 1] != 0.0f) {
            w->loaded[cs[i].ai] = 1;
            if (!cs[i].terrain) w->loaded[cs[i].bi] = 1;
        }
        cts[i].impulse = jn[i];
    }
    for (uint32_t k = 0; k < npt; k++) {
        if (pacc[3 * k] == 0.0f && pacc[3 * k + 1] == 0.0f && pacc[3 * k + 2] == 0.0f) continue;
        for (uint32_t m = 0; m < pt[k].np; m++) {
            const CC *c = &cs[ppts[pt[k].p0 + m]];
            w->loaded[c->ai] = 1;
            if (!c->terrain) w->loaded[c->bi] = 1;
        }
    }
    w->stats.patches = npt;
    free(pacc); free(pt); free(ppts);
    free(jn); free(jt); free(keys); free(anch); free(cs);
}

/* islands.rs:27-101: only emptiness matters here (route_to_island keeps a joint or contact with a
 * dynamic body on either side) */
static int islands_nonempty(const OrWorld *w, const Contact *cts, uint32_t n) {
    for (uint32_t q = 0; q < w->nj; q++)
        if (w->inv_mass[w->joints[q].a] > 0.0f || w->inv_mass[w->joints[q].b] > 0.0f) return 1;
    for (uint32_t i = 0; i < n; i++) {
        uint32_t a = w->cbody[cts[i].a];
        int da = w->inv_mass[a] > 0.0f;
        int db = cts[i].b != TERRAIN_CI && w->inv_mass[w->cbody[cts[i].b]] > 0.0f;
        if (da || db) return 1;
    }
    return 0;
}

/* ------------------------------------------------------------------ sin/cos (glam from_axis_angle) */
void or_sincosf(float x, float *s, float *c) { or_sincos_libm(x, s, c); }
/* Quat::from_axis_angle: (axis * sin(a/2), cos(a/2)) */
static Q4 from_axis_angle(V3 axis, float angle) {
    float s, c;
    or_sincosf(angle * 0.5f, &s, &c);
    V3 v = vmul(axis, s);
    return q4(v.x, v.y, v.z, c);
}
/* step.rs:230-233 turn */
static Q4 turn(Q4 q, V3 wv, float h) {
    if (vlen2(wv) <= 1.0e-18f) return q;
    return qnormalize(qmul(from_axis_angle(vnormalize(wv), vlen(wv) * h), q));
}

/* ---- ball joints at position level: solve_d6's translational rows (articulation/src/bodies.rs:616-739)
 * with SHAKE directions from the step's start poses (world.rs:1072-1076, JointProjection::Shake by
 * default; bodies.rs:422-499 solve_step_with, one iteration per call).  A ball joint's rows are
 * rigid (compliance 0): alpha = 0 and the multiplier's history drops out. */
static V3 normalize_fin(V3 v) {
    v = dfin_v(v);
    float l2 = vlen2(v);
    return l2 > 1.0e-12f ? vmul(v, 1.0f / sqrtf(l2)) : v3s(0);
}
/* bodies.rs:562-566 turned: rotation between two orientations, radians */
static float turned(Q4 a, Q4 b) { return 2.0f * acosf(fminf(fabsf(qmul(a, qconj(b)).w), 1.0f)); }
/* bodies.rs:569-573 inv_inertia_at */
static M3 inv_inertia_at(const OrWorld *w, uint32_t i, Q4 q) { M3 r = m3quat(q); return m3mul(m3mul(r, inv_inertia_local_of(w, i)), m3t(r)); }
/* dynamics.rs:207-210, 71-89 */
static void pos_delta(OrWorld *w, uint32_t i, V3 dp) {
    if (w->inv_mass[i] > 0.0f && vfinite(dp)) w->pos[i] = dfin_v(vadd(w->pos[i], dp));
}
static void orient_delta(OrWorld *w, uint32_t i, V3 dth) {
    if (w->inv_mass[i] <= 0.0f || !vfinite(dth)) return;
    V3 d = dfin_v(dth);
    if (vlen2(d) <= 1.0e-24f) return;
    Q4 dq = qnormalize(q4(d.x * 0.5f, d.y * 0.5f, d.z * 0.5f, 1.0f));
    w->rot[i] = fin_q(qnormalize(qmul(dq, w->rot[i])));
}
static void ball_pos(OrWorld *w, const Joint *j, const V3 *spos, const Q4 *srot, float dt) {
    (void)spos;
    uint32_t a = j->a, b = j->b;
    float wa = dfin(w->inv_mass[a]), wb = dfin(w->inv_mass[b]);
    Q4 sa = srot[a], sb = srot[b];
    /* SHAKE_MAX_TURN = 0.5 rad: beyond it this call uses the current iterate's directions */
    int shake = turned(sa, w->rot[a]) < 0.5f && turned(sb, w->rot[b]) < 0.5f;
    V3 ax0[3] = {v3s(0), v3s(0), v3s(0)}, la0 = v3s(0), lb0 = v3s(0);
    M3 ia0 = m3zero(), ib0 = m3zero();
    if (shake) {
        Q4 ra0 = qmul(sa, j->fra);
        for (int i = 0; i < 3; i++) ax0[i] = normalize_fin(qrot(ra0, v3(i == 0 ? 1.0f : 0.0f, i == 1 ? 1.0f : 0.0f, i == 2 ? 1.0f : 0.0f)));
        la0 = dfin_v(qrot(sa, j->fap)); lb0 = dfin_v(qrot(sb, j->fbp));
        ia0 = inv_inertia_at(w, a, sa); ib0 = inv_inertia_at(w, b, sb);
    }
    for (int i = 0; i < 3; i++) {
        V3 pa = w->pos[a], pb = w->pos[b];
        Q4 qa = w->rot[a], qb = w->rot[b];
        V3 u = normalize_fin(qrot(qmul(qa, j->fra), v3(i == 0 ? 1.0f : 0.0f, i == 1 ? 1.0f : 0.0f, i == 2 ? 1.0f : 0.0f)));
        if (viszero(u)) continue;
        V3 xa = dfin_v(vadd(pa, qrot(qa, j->fap)));
        V3 xb = dfin_v(vadd(pb, qrot(qb, j->fbp)));
        V3 d = dfin_v(vsub(xb, xa));
        V3 un, lva, lvb; M3 iaw, ibw;
        if (shake) { un = ax0[i]; lva = la0; lvb = lb0; iaw = ia0; ibw = ib0; }
        else { un = u; lva = vsub(xa, pa); lvb = vsub(xb, pb); iaw = inv_inertia_world(w, a); ibw = inv_inertia_world(w, b); }
        float c = dfin(dfin(vdot(d, u)) - 0.0f);
        V3 rau = dfin_v(vcross(lva, un)), rbu = dfin_v(vcross(lvb, un));
        float wra = dfin(vdot(rau, m3v(iaw, rau))), wrb = dfin(vdot(rbu, m3v(ibw, rbu)));
        float wsum = dfin(((wa + wb) + wra) + wrb);
        float denom = dfin(wsum + 0.0f);
        if (denom <= 0.0f) continue;
        c = dfin(c + dfin(0.0f * dt));
        float dl = dfin(-(c + 0.0f) / denom);
        pos_delta(w, a, dfin_v(vmul(vneg(un), dfin(dl * wa))));
        pos_delta(w, b, dfin_v(vmul(un, dfin(dl * wb))));
        orient_delta(w, a, dfin_v(vmul(vneg(m3v(iaw, rau)), dl)));
        orient_delta(w, b, dfin_v(vmul(m3v(ibw, rbu), dl)));
    }
}

/* bodies.rs:1010-1030 row_error: locked at 0, a limit only while violated, a position drive's target
 * (clamped into the limit), a velocity drive free at position level.  Returns 0 for no row. */
static int d6_row_error(const D6Ax *x, float coord, float *c) {
    int has = 1;
    *c = dfin(coord - 0.0f);
    if (x->has_limit) {
        float lo = dfin(x->lim_min), hi = dfin(x->lim_max);
        if (coord < lo) *c = dfin(coord - lo);
        else if (coord > hi) *c = dfin(coord - hi);
        else has = 0;
    }
    if (x->has_drive) {
        if (!x->vel_mode) {
            float t = dfin(x->d_target);
            if (x->has_limit) t = clampf_rs(t, dfin(x->lim_min), dfin(x->lim_max));
            *c = dfin(coord - t);
            has = 1;
        } else if (!x->has_limit) has = 0;
    }
    return has;
}
/* bodies.rs:433-438 alpha~ = compliance / dt^2 (0 for rigid rows) */
static float d6_alpha(float c, float dt2) { c = fmaxf(dfin(c), 0.0f); return c <= 0.0f ? 0.0f : dfin(c / dt2); }
/* the angle a rotational row measures at frame rotations ra, rb (bodies.rs:786-795) */
static float d6_pos_angle(int hinge, int i, Q4 ra, Q4 rb, V3 u) {
    Q4 q = qmul(qconj(ra), rb);
    if (q.w < 0.0f) q = q4(-q.x, -q.y, -q.z, -q.w);
    V3 v = v3(q.x, q.y, q.z);
    if (hinge == i) return dfin(2.0f * or_atan2f(v3at(v, i), q.w));
    if (hinge >= 0) return dfin(vdot(dfin_v(vcross(normalize_fin(qrot(ra, unit3(hinge))), normalize_fin(qrot(rb, unit3(hinge))))), u));
    return dfin(vdot(dfin_v(qrot(ra, dfin_v(vmul(v, 2.0f)))), u));
}
/* bodies.rs:616-873 solve_d6 (no rotor rows): each translational row, then each rotational row,
 * measured afresh from the current poses and corrected as an XPBD row along the step-start
 * directions (SHAKE) unless a body turned 0.5 rad or more; a damped compliant row's change since
 * the step's start joins as XPBD's damping term. */
static void d6_pos(OrWorld *w, Joint *j, const V3 *spos, const Q4 *srot, float dt) {
    uint32_t a = j->a, b = j->b;
    float wa = dfin(w->inv_mass[a]), wb = dfin(w->inv_mass[b]);
    float dt2 = fmaxf(dt * dt, 1.0e-12f);
    V3 sap = spos[a], sbp = spos[b];
    Q4 sa = srot[a], sb = srot[b];
    int shake = turned(sa, w->rot[a]) < 0.5f && turned(sb, w->rot[b]) < 0.5f;
    V3 ax0[3] = {v3s(0), v3s(0), v3s(0)}, la0 = v3s(0), lb0 = v3s(0);
    M3 ia0 = m3zero(), ib0 = m3zero();
    if (shake) {
        Q4 ra0 = qmul(sa, j->fra);
        for (int i = 0; i < 3; i++) ax0[i] = normalize_fin(qrot(ra0, unit3(i)));
        la0 = dfin_v(qrot(sa, j->fap)); lb0 = dfin_v(qrot(sb, j->fbp));
        ia0 = inv_inertia_at(w, a, sa); ib0 = inv_inertia_at(w, b, sb);
    }
    for (int i = 0; i < 3; i++) {
        D6Ax *row = &j->t[i];
        if (!row->enabled) continue;
        V3 pa = w->pos[a], pb = w->pos[b];
        Q4 qa = w->rot[a], qb = w->rot[b];
        V3 u = normalize_fin(qrot(qmul(qa, j->fra), unit3(i)));
        if (viszero(u)) continue;
        V3 xa = dfin_v(vadd(pa, qrot(qa, j->fap)));
        V3 xb = dfin_v(vadd(pb, qrot(qb, j->fbp)));
        V3 d = dfin_v(vsub(xb, xa));
        V3 un, lva, lvb; M3 iaw, ibw;
        if (shake) { un = ax0[i]; lva = la0; lvb = lb0; iaw = ia0; ibw = ib0; }
        else { un = u; lva = vsub(xa, pa); lvb = vsub(xb, pb); iaw = inv_inertia_world(w, a); ibw = inv_inertia_world(w, b); }
        float coord = dfin(vdot(d, u)), c;
        if (!d6_row_error(row, coord, &c)) continue;
        V3 rau = dfin_v(vcross(lva, un)), rbu = dfin_v(vcross(lvb, un));
        float wra = dfin(vdot(rau, m3v(iaw, rau))), wrb = dfin(vdot(rbu, m3v(ibw, rbu)));
        float wsum = dfin(((wa + wb) + wra) + wrb);
        float alpha = d6_alpha(row->compliance, dt2);
        float denom = dfin(wsum + alpha);
        if (denom <= 0.0f) continue;
        c = dfin(c + dfin(0.0f * dt));
        float lold = dfin(row->lam), dl;
        if (d6_position_damped(row)) {
            V3 u0 = normalize_fin(qrot(qmul(sa, j->fra), unit3(i)));
            V3 d0 = dfin_v(vsub(vadd(sbp, qrot(sb, j->fbp)), vadd(sap, qrot(sa, j->fap))));
            float gamma = dfin(row->compliance * row->damping / dt);
            float dc = dfin(coord - vdot(d0, u0));
            dl = dfin(-((c + dfin(alpha * lold)) + dfin(gamma * dc)) / dfin((1.0f + gamma) * wsum + alpha));
        } else dl = dfin(-(c + dfin(alpha * lold)) / denom);
        row->lam = dfin(lold + dl);
        pos_delta(w, a, dfin_v(vmul(vneg(un), dfin(dl * wa))));
        pos_delta(w, b, dfin_v(vmul(un, dfin(dl * wb))));
        orient_delta(w, a, dfin_v(vmul(vneg(m3v(iaw, rau)), dl)));
        orient_delta(w, b, dfin_v(vmul(m3v(ibw, rbu), dl)));
    }
    int hinge = d6_hinge_axis(j);
    for (int i = 0; i < 3; i++) {
        D6Ax *row = &j->r[i];
        if (!row->enabled) continue;
        Q4 ra = qmul(w->rot[a], j->fra), rb = qmul(w->rot[b], j->frb);
        V3 u = normalize_fin(qrot(ra, unit3(i)));
        if (viszero(u)) continue;
        float angle = d6_pos_angle(hinge, i, ra, rb, u);
        V3 un; M3 iaw, ibw;
        if (shake) { un = ax0[i]; iaw = ia0; ibw = ib0; }
        else { un = u; iaw = inv_inertia_world(w, a); ibw = inv_inertia_world(w, b); }
        float c;
        if (!d6_row_error(row, angle, &c)) continue;
        float wsum = dfin(vdot(un, m3v(iaw, un)) + vdot(un, m3v(ibw, un)));
        float alpha = d6_alpha(row->compliance, dt2);
        float denom = dfin(wsum + alpha);
        if (denom <= 0.0f) continue;
        c = dfin(c + dfin(0.0f * dt));
        float lold = dfin(row->lam), dl;
        if (d6_position_damped(row)) {
            Q4 ra0 = qmul(sa, j->fra), rb0 = qmul(sb, j->frb);
            V3 u0 = normalize_fin(qrot(ra0, unit3(i)));
            float angle0 = d6_pos_angle(hinge, i, ra0, rb0, u0);
            float gamma = dfin(row->compliance * row->damping / dt);
            float dc = dfin(angle - angle0);
            dl = dfin(-((c + dfin(alpha * lold)) + dfin(gamma * dc)) / dfin((1.0f + gamma) * wsum + alpha));
        } else dl = dfin(-(c + dfin(alpha * lold)) / denom);
        row->lam = dfin(lold + dl);
        orient_delta(w, a, dfin_v(vmul(vneg(m3v(iaw, un)), dl)));
        orient_delta(w, b, dfin_v(vmul(m3v(ibw, un), dl)));
    }
}

/* bodies.rs:501-548 solve_distance_row: the XPBD row of a distance joint without anchors or
 * damping, corrected along the start-of-step centre line (SHAKE) unless the rod turned more than
 * SHAKE_MAX_TURN = 0.5 rad this step (cos 0.5 = 0x3f60a940 as f32), the multiplier accumulated
 * over the step's iterations. */
static void dist_pos(OrWorld *w, Joint *j, const V3 *spos, float dt) {
    const float cos_max_turn = 0.87758255f;
    uint32_t a = j->a, b = j->b;
    float wa = w->inv_mass[a], wb = w->inv_mass[b];
    float wsum = dfin(wa + wb);
    if (wsum <= 0.0f) return;
    V3 d = vsub(w->pos[b], w->pos[a]);
    float len2 = vlen2(d);
    if (!isfinite(len2) || len2 <= 1.0e-20f) return;
    float len = sqrtf(len2);
    if (len <= 1.0e-6f) return;
    V3 ncur = normalize_fin(vdivs(d, len));
    if (viszero(ncur)) return;
    V3 n = ncur;
    V3 n0 = normalize_fin(vsub(spos[b], spos[a]));
    if (!viszero(n0) && vdot(n0, ncur) > cos_max_turn) n = n0;
    float c = dfin(len - dfin(j->rest));
    float dt2 = fmaxf(dt * dt, 1.0e-12f);
    float alpha = j->compliance <= 0.0f ? 0.0f : dfin(fmaxf(dfin(j->compliance), 0.0f) / dt2);
    float denom = dfin(wsum + alpha);
    if (denom <= 0.0f) return;
    float lold = dfin(j->lam);
    float lam = dfin(-(c + dfin(alpha * lold)) / (denom * vdot(n, ncur)));
    if (fabsf(lam) <= 1.0e-9f) return;
    j->lam = dfin(lold + lam);
    pos_delta(w, a, dfin_v(vmul(vneg(n), dfin(lam * wa))));
    pos_delta(w, b, dfin_v(vmul(n, dfin(lam * wb))));
}

typedef struct { uint32_t a, b; int has_b; V3 la, lb, pbw, n; float pen0; } PosC;

/* step.rs:447-497 */
static void project(OrWorld *w, const PosC *pc, uint32_t npc, V3 *slin, V3 *sang, int shock) {
    V3 up = vlen2(w->gravity) > 0.0f ? vneg(vnormalize(w->gravity)) : v3(0, 1, 0);
    SK *ord = malloc(sizeof(SK) * (npc + 1));
    for (uint32_t i = 0; i < npc; i++) {
        ord[i].idx = i;
        if (shock) {
            float h = pc[i].has_b ? fminf(vdot(w->pos[pc[i].a], up), vdot(w->pos[pc[i].b], up)) : -INFINITY;
            ord[i].k = sort_key6(h);
        } else ord[i].k = 0;
    }
    if (shock) qsort(ord, npc, sizeof(SK), cmp_sk_asc);
    for (uint32_t o = 0; o < npc; o++) {
        const PosC *c = &pc[ord[o].idx];
        int am = w->inv_mass[c->a] > 0.0f && w->awake[c->a];
        int bm = c->has_b && w->inv_mass[c->b] > 0.0f && w->awake[c->b];
        if (shock && am && bm) {
            float ha = vdot(w->pos[c->a], up), hb = vdot(w->pos[c->b], up);
            if (ha < hb) am = 0; else bm = 0;
        }
        V3 pa = w->pos[c->a]; Q4 qa = w->rot[c->a];
        V3 wa = vadd(pa, qrot(qa, c->la));
        V3 wb, pb = v3s(0); Q4 qb = qident();
        if (c->has_b) { pb = w->pos[c->b]; qb = w->rot[c->b]; wb = vadd(pb, qrot(qb, c->lb)); }
        else wb = c->pbw;
        float pen = c->pen0 - vdot(vsub(wb, wa), c->n);
        float err = pen - POS_SLOP;
        if (err <= 0.0f) continue;
        V3 ra = vsub(wa, pa);
        V3 rb = c->has_b ? vsub(wb, pb) : v3s(0);
        float ima = 0, imb = 0; M3 iia = m3zero(), iib = m3zero();
        if (am) { ima = w->inv_mass[c->a]; iia = inv_inertia_world(w, c->a); }
        if (c->has_b && bm) { imb = w->inv_mass[c->b]; iib = inv_inertia_world(w, c->b); }
        V3 ca = vcross(ra, c->n), cb = vcross(rb, c->n);
        float wsum = ((ima + imb) + vdot(ca, m3v(iia, ca))) + vdot(cb, m3v(iib, cb));
        if (wsum <= 1.0e-9f) continue;
        float lambda = err / wsum;
        for (int side = 0; side < 2; side++) {
            uint32_t id; V3 p; Q4 q; float im; M3 ii; V3 r; V3 j;
            if (side == 0) { if (!am) continue; id = c->a; p = pa; q = qa; im = ima; ii = iia; r = ra; j = vmul(vneg(c->n), lambda); }
            else { if (!(c->has_b && bm)) continue; id = c->b; p = pb; q = qb; im = imb; ii = iib; r = rb; j = vmul(c->n, lambda); }
            V3 dth = m3v(ii, vcross(r, j));
            slin[id] = vadd(slin[id], vmul(j, im));
            sang[id] = vadd(sang[id], dth);
            Q4 dq = qscale(qmul(q4(dth.x, dth.y, dth.z, 0.0f), q), 0.5f);
            Q4 nr = qnormalize(q4(q.x + dq.x, q.y + dq.y, q.z + dq.z, q.w + dq.w));
            b_set_pose(w, id, vadd(p, vmul(j, im)), nr);
        }
    }
    free(ord);
}

/* ------------------------------------------------------------------ the step */
/* ---- kinematic players: World::update_players (step.rs:544-639), run first in World::step
 * (step.rs:51).  Each player in order: gravity or the sticky floor, the input direction at its
 * speed, then up to four resolution passes of the capsule (radius, height / 2) at the proposed
 * position: the first terrain triangle within 0.05 m, else the first collider (other than its own
 * body) whose stored box meets the capsule's box and whose shape is within 0.05 m; push out along
 * the contact normal by the penetration + 1 mm, drop the part of the move into the normal, kick a
 * dynamic body hit (wake, impulse -10 n); a normal with y > 0.6 grounds the player.  The stored
 * boxes are those before any player moved (the step refreshes them only at its end). */
static int q_hf_range(const OrWorld *w, Aabb bb, int *gx0, int *gx1, int *gz0, int *gz1);
static void q_hf_tris(const OrWorld *w, int gx, int gz, Tri out[2]);
static void update_players(OrWorld *w, float dt) {
    if (!w->np) return;
    for (uint32_t c = 0; c < w->nc; c++) { uint32_t b = w->cbody[c]; w->caabb[c] = aabb_of(w->cshape[c], w->pos[b], w->rot[b]); }
    for (uint32_t pi = 0; pi < w->np; pi++) {
        OrPlayer *pl = &w->players[pi];
        V3 pos = w->pos[pl->body];
        if (!pl->grounded) pl->vvel -= 9.81f * dt;
        else if (pl->vvel <= 0.0f) pl->vvel = -0.5f;
        float m2 = vlen2(pl->input);
        V3 dir = m2 > 1.0e-6f ? vdivs(pl->input, sqrtf(m2)) : v3s(0);
        V3 desired = vadd(vmul(vmul(dir, pl->speed), dt), v3(0.0f, pl->vvel * dt, 0.0f));
        pl->grounded = 0;
        ShapeD cap; cap.type = OR_CAPSULE; cap.d[0] = pl->radius; cap.d[1] = pl->height * 0.5f; cap.d[2] = 0.0f;
        Q4 qi = qident();
        for (int it = 0; it < 4; it++) {
            if (vlen2(desired) < 1.0e-6f) break;
            V3 xp = vadd(pos, desired);
            Aabb bb = aabb_of(cap, xp, qi);
            int hit = 0; V3 n = v3s(0); float pen = 0.0f; int32_t hb = -1;
            int gx0, gx1, gz0, gz1;
            if (q_hf_range(w, bb, &gx0, &gx1, &gz0, &gz1))
                for (int gz = gz0; gz <= gz1 && !hit; gz++)
                    for (int gx = gx0; gx <= gx1 && !hit; gx++) {
                        Tri tr[2];
                        q_hf_tris(w, gx, gz, tr);
                        for (int k = 0; k < 2 && !hit; k++) {
                            Manifold m;
                            if (collide_shape_tri(cap, xp, qi, tr[k], 0.05f, &m)) { hit = 1; n = m.p[0].n; pen = m.p[0].pen; }
                        }
                    }
            if (!hit)
                for (uint32_t c = 0; c < w->nc && !hit; c++) {
                    uint32_t b = w->cbody[c];
                    const Aabb *cb = &w->caabb[c];  /* Aabb::intersects (aabb.rs:75-79) */
                    if (b == pl->body || !(bb.min.x <= cb->max.x && bb.max.x >= cb->min.x && bb.min.y <= cb->max.y &&
                                           bb.max.y >= cb->min.y && bb.min.z <= cb->max.z && bb.max.z >= cb->min.z)) continue;
                    Manifold m;
                    if (collide_shapes(cap, xp, qi, w->cshape[c], w->pos[b], w->rot[b], 0.05f, &m)) { hit = 1; n = m.p[0].n; pen = m.p[0].pen; hb = (int32_t)b; }
                }
            if (!hit) break;
            if (n.y > 0.6f) { pl->grounded = 1; pl->vvel = 0.0f; }
            desired = vadd(desired, vmul(n, pen + 0.001f));
            float dd = vdot(desired, n);
            if (dd < 0.0f) desired = vsub(desired, vmul(n, dd));
            if (hb >= 0 && w->dyn[hb]) {
                w->awake[hb] = 1; w->sticks[hb] = 0;
                apply_impulse(w, (uint32_t)hb, vmul(vneg(n), 10.0f));
            }
        }
        pos = vadd(pos, desired);
        b_set_pose(w, pl->body, pos, qident());
    }
}

/* world.rs:466-475 add_player with the FFI's dimensions (lib.rs:1580-1626): a static body with the
 * controller's default capsule collider (r 0.4, half height 0.9, material Default); the controller
 * then takes the given radius, height and speed.  Returns the player index, or -1. */
int32_t or_add_player(OrWorld *w, const float *pose, float radius, float height, float speed) {
    float dims[3] = {0.4f, 1.8f * 0.5f, 0.0f}, vel[6] = {0, 0, 0, 0, 0, 0};
    int32_t b = add_body(w, OR_CAPSULE, dims, pose, vel, 0, 0.0f, 1);
    if (b < 0) return -1;
    if (w->np == w->capp) { w->capp = w->capp ? 2 * w->capp : 4; w->players = xrealloc(w->players, sizeof(OrPlayer) * w->capp); }
    OrPlayer *pl = &w->players[w->np];
    memset(pl, 0, sizeof(*pl));
    pl->body = (uint32_t)b; pl->radius = radius; pl->height = height; pl->speed = speed;
    return (int32_t)w->np++;
}
void or_player_set_input(OrWorld *w, uint32_t idx, const float *dir3) { if (idx < w->np) w->players[idx].input = v3(dir3[0], dir3[1], dir3[2]); }
/* pose {p, q}, grounded; returns the player's body index */
uint32_t or_player_get(const OrWorld *w, uint32_t idx, float *pose7, int *grounded, float *vvel) {
    const OrPlayer *pl = &w->players[idx];
    V3 p = w->pos[pl->body]; Q4 q = w->rot[pl->body];
    pose7[0] = p.x; pose7[1] = p.y; pose7[2] = p.z; pose7[3] = q.x; pose7[4] = q.y; pose7[5] = q.z; pose7[6] = q.w;
    *grounded = pl->grounded; *vvel = pl->vvel;
    return pl->body;
}

void or_step(OrWorld *w, float dt) {
    uint32_t count = w->n;
    w->tick += 1;
    memset(&w->stats, 0, sizeof(w->stats));
    w->loaded_valid = 0;
    w->has_pseudo = 0;
    update_players(w, dt);
    /* S1 (step.rs:53-68) */
    V3 *pre_lin = malloc(sizeof(V3) * (count + 1)), *pre_ang = malloc(sizeof(V3) * (count + 1));
    for (uint32_t i = 0; i < count; i++) { pre_lin[i] = w->lin[i]; pre_ang[i] = w->ang[i]; }
    V3 g = v3(gq6(w->gravity.x), gq6(w->gravity.y), gq6(w->gravity.z));
    for (uint32_t i = 0; i < count; i++) {
        if (!w->dyn[i] || !w->awake[i]) continue;
        V3 a = vfinite(w->pos[i]) ? g : v3s(0);
        b_set_vel(w, i, vadd(w->lin[i], vmul(a, dt)), w->ang[i]);
    }
    /* S2 AABBs, sweep, swept (step.rs:71-82) */
    uint32_t nc = w->nc;
    for (uint32_t c = 0; c < nc; c++) { uint32_t b = w->cbody[c]; w->caabb[c] = aabb_of(w->cshape[c], w->pos[b], w->rot[b]); }
    float *sweep = malloc(sizeof(float) * (nc + 1));
    Aabb *sw = malloc(sizeof(Aabb) * (nc + 1));
    for (uint32_t c = 0; c < nc; c++) {
        uint32_t b = w->cbody[c];
        float reach = 0.5f * vlen(vsub(w->caabb[c].max, w->caabb[c].min));
        sweep[c] = MARGIN + (vlen(w->lin[b]) + vlen(w->ang[b]) * reach) * dt;
        sw[c].min = vsub(w->caabb[c].min, v3s(sweep[c]));
        sw[c].max = vadd(w->caabb[c].max, v3s(sweep[c]));
    }
    free(w->pairs);
    w->npairs = broadphase(sw, nc, &w->pairs);
    w->stats.pairs = w->npairs;
    /* narrowphase, pairs (step.rs:95-119); bodies joined directly by a joint do not collide
     * (step.rs:85-93,102) */
    uint64_t *jkeys = malloc(8 * (w->nj + 1));
    for (uint32_t q = 0; q < w->nj; q++) {
        uint32_t a = w->joints[q].a, b = w->joints[q].b;
        jkeys[q] = ((uint64_t)(a < b ? a : b) << 32) | (a < b ? b : a);
    }
    qsort(jkeys, w->nj, 8, cmp_u64);
    CVec raw = {0};
    /* pairs in parallel above PAR_MIN (step.rs:113-117), each into its own manifold, then appended
     * in pair order: the output is the serial one */
    uint32_t np = w->npairs;
    Manifold *pm = malloc(sizeof(Manifold) * (np + 1));
    float *preach = malloc(sizeof(float) * (np + 1));
#pragma omp parallel for schedule(dynamic, 64) if (np >= PAR_MIN) num_threads(g_threads)
    for (uint32_t p = 0; p < np; p++) {
        pm[p].len = 0;
        uint32_t i = w->pairs[2 * p], j = w->pairs[2 * p + 1];
        uint32_t ba = w->cbody[i], bb = w->cbody[j];
        if (ba == bb) continue;
        if (w->nj) {
            uint64_t key = ((uint64_t)(ba < bb ? ba : bb) << 32) | (ba < bb ? bb : ba);
            if (bsearch(&key, jkeys, w->nj, 8, cmp_u64)) continue;
        }
        preach[p] = sweep[i] + sweep[j] - MARGIN;
        if (!collide_shapes(w->cshape[i], w->pos[ba], w->rot[ba], w->cshape[j], w->pos[bb], w->rot[bb], preach[p], &pm[p])) pm[p].len = 0;
    }
    for (uint32_t p = 0; p < np; p++) {
        const Manifold *m = &pm[p];
        for (int k = 0; k < m->len; k++) {
            Contact c; c.a = w->pairs[2 * p]; c.b = w->pairs[2 * p + 1]; c.point = m->p[k].pos; c.normal = m->p[k].n;
            c.depth = m->p[k].pen - preach[p] + MARGIN; c.impulse = 0; c.env_mat = 0; c.feat = m->p[k].feat; c.cell = 0xFFFFFFFFu;
            cpush(&raw, c);
        }
    }
    free(pm); free(preach);
    free(jkeys);
    /* terrain (step.rs:121-142): colliders in parallel above PAR_MIN, each into its own buffer,
     * appended in collider order */
    uint32_t tc = 0;
    CVec *tcv = calloc(nc + 1, sizeof(CVec));
    uint32_t *tcn = calloc(nc + 1, sizeof(uint32_t));
#pragma omp parallel for schedule(dynamic, 16) if (nc >= PAR_MIN) num_threads(g_threads)
    for (uint32_t c = 0; c < nc; c++) tcn[c] = terrain_contacts(w, c, sweep[c], &tcv[c]);
    for (uint32_t c = 0; c < nc; c++) {
        tc += tcn[c];
        for (size_t k = 0; k < tcv[c].n; k++) cpush(&raw, tcv[c].v[k]);
        free(tcv[c].v);
    }
    free(tcv); free(tcn);
    w->stats.terrain_candidates = tc;
    /* S3 (step.rs:144-150) */
    CVec cul = {0};
    cull_and_orient(w, &raw, &cul);
    free(raw.v);
    if (islands_nonempty(w, cul.v, (uint32_t)cul.n)) {
        w->stats.islands = 1;
        solve(w, cul.v, (uint32_t)cul.n, pre_lin, pre_ang, dt);
    }
    /* S4 (step.rs:154-280) */
    V3 *opos = calloc(count + 1, sizeof(V3)); Q4 *orot = calloc(count + 1, sizeof(Q4));
    V3 *ppos = calloc(count + 1, sizeof(V3)); Q4 *prot = calloc(count + 1, sizeof(Q4));
    for (uint32_t i = 0; i < count; i++) { orot[i] = qident(); prot[i] = qident(); }
    uint8_t *constrained = calloc(count + 1, 1), *ball = calloc(count + 1, 1);
    for (uint32_t k = 0; k < cul.n; k++) {
        constrained[w->cbody[cul.v[k].a]] = 1;
        if (cul.v[k].b != TERRAIN_CI) constrained[w->cbody[cul.v[k].b]] = 1;
    }
    uint8_t *jointed = calloc(count + 1, 1);
    for (uint32_t q = 0; q < w->nj; q++) {
        uint32_t a = w->joints[q].a, b = w->joints[q].b;
        constrained[a] = constrained[b] = 1;
        jointed[a] = jointed[b] = 1;
    }
    for (uint32_t i = 0; i < count; i++) {
        int ld = w->loaded_valid ? w->loaded[i] : 0;
        ball[i] = !constrained[i] || (!jointed[i] && !ld);
    }
    for (uint32_t k = 0; k < cul.n; k++) {
        const Contact *c = &cul.v[k];
        uint32_t a = w->cbody[c->a];
        int hb = c->b != TERRAIN_CI;
        uint32_t b = hb ? w->cbody[c->b] : 0;
        if (!ball[a] && (!hb || !ball[b])) continue;
        V3 la = ball[a] ? vmul(vadd(w->last_lin[a], w->lin[a]), 0.5f) : w->lin[a];
        V3 wa = w->ang[a];
        V3 ra = vsub(c->point, w->pos[a]);
        V3 lb = v3s(0), wb = v3s(0), rb = v3s(0);
        if (hb) { lb = ball[b] ? vmul(vadd(w->last_lin[b], w->lin[b]), 0.5f) : w->lin[b]; wb = w->ang[b]; rb = vsub(c->point, w->pos[b]); }
        float approach = vdot(vsub(vadd(lb, vcross(wb, rb)), vadd(la, vcross(wa, ra))), c->normal);
        V3 val = w->lin[a], vaa = w->ang[a], vbl = v3s(0), vba = v3s(0);
        if (hb) { vbl = w->lin[b]; vba = w->ang[b]; }
        float vend = vdot(vsub(vadd(vbl, vcross(vba, rb)), vadd(val, vcross(vaa, ra))), c->normal);
        float s0 = -(c->depth - MARGIN);
        if (s0 + (2.0f * vend) * dt < 0.0f || s0 + approach * dt < 0.0f) { ball[a] = 0; if (hb) ball[b] = 0; }
    }
    for (uint32_t i = 0; i < count; i++) {
        if (!w->dyn[i] || !w->awake[i]) continue;
        V3 p = w->pos[i]; Q4 q = w->rot[i];
        opos[i] = p; orot[i] = q;
        V3 vl = w->lin[i], va = w->ang[i];
        p = vadd(p, vmul(vl, dt));
        if (ball[i]) p = vadd(p, vmul(vsub(w->last_lin[i], vl), 0.5f * dt));
        if (vlen2(va) > 0.0f) {
            M3 l = inv_inertia_world(w, i);
            float det = m3det(l);
            int has_m = isfinite(det) && fabsf(det) > 1.0e-20f;
            V3 mom = has_m ? m3v(m3inv(l), va) : v3s(0);
            Q4 start = q;
            q = turn(start, va, dt);
            if (has_m) {
                M3 il = inv_inertia_local_of(w, i);
                float mean = ((il.c0.x + il.c1.y) + il.c2.z) / 3.0f;
                M3 dev = m3sub(il, m3diag(v3s(mean)));
                int principal = vlen(vcross(mom, va)) < (1.0e-5f * vlen(mom)) * vlen(va);
                float gain = principal ? 0.0f : (dt * fmaxf(fmaxf(vlen(dev.c0), vlen(dev.c1)), vlen(dev.c2))) * vlen(mom);
                float gq = ceilf(gain / 0.5f);
                uint32_t n;
                if (!(gq >= 1.0f)) n = 1; /* NaN, 0 and negatives saturate to 0 in `as usize`, then clamp to 1 */
                else if (gq > 256.0f) n = 256;
                else n = (uint32_t)gq;
                float h = dt / (float)n;
                Q4 qq = start; V3 wv = va;
                for (uint32_t s = 0; s < n; s++) {
                    Q4 q0 = qq;
                    qq = turn(q0, wv, h);
                    M3 r = m3quat(qq);
                    V3 wend = m3v(m3mul(m3mul(r, il), m3t(r)), mom);
                    for (int t = 0; t < 2; t++) {
                        qq = turn(q0, vmul(vadd(wv, wend), 0.5f), h);
                        r = m3quat(qq);
                        wend = m3v(m3mul(m3mul(r, il), m3t(r)), mom);
                    }
                    wv = wend;
                }
                q = qq;
                va = wv;
                b_set_vel(w, i, vl, va);
            }
        }
        /* a bounced free body's pseudo-velocity moves it and never enters its velocity
         * (step.rs:271-277) */
        if (w->has_pseudo) {
            p = vadd(p, vmul(w->pl[i], dt));
            if (vlen2(w->pa[i]) > 0.0f) q = qnormalize(qmul(from_axis_angle(vnormalize(w->pa[i]), vlen(w->pa[i]) * dt), q));
        }
        b_set_pose(w, i, p, q);
        ppos[i] = w->pos[i]; prot[i] = w->rot[i];
    }
    /* S5 (step.rs:284-310, 423-443) */
    PosC *pc = malloc(sizeof(PosC) * (cul.n + 1));
    uint32_t npc = 0;
    for (uint32_t k = 0; k < cul.n; k++) {
        const Contact *c = &cul.v[k];
        uint32_t a = w->cbody[c->a];
        int hb = c->b != TERRAIN_CI;
        uint32_t b = hb ? w->cbody[c->b] : 0;
        int ma = w->inv_mass[a] > 0.0f && w->awake[a];
        int mb = hb && w->inv_mass[b] > 0.0f && w->awake[b];
        if (!ma && !mb) continue;
        if (contact_props(w, c).stiff > 0.0f) continue;  /* compliant pairs penetrate by design (step.rs:433) */
        PosC x; x.a = a; x.b = b; x.has_b = hb; x.n = c->normal; x.pen0 = c->depth - MARGIN;
        V3 pap = (w->dyn[a] && w->awake[a]) ? opos[a] : w->pos[a];
        Q4 paq = (w->dyn[a] && w->awake[a]) ? orot[a] : w->rot[a];
        x.la = qrot(qconj(paq), vsub(c->point, pap));
        if (hb) {
            V3 pbp = (w->dyn[b] && w->awake[b]) ? opos[b] : w->pos[b];
            Q4 pbq = (w->dyn[b] && w->awake[b]) ? orot[b] : w->rot[b];
            x.lb = qrot(qconj(pbq), vsub(c->point, pbp));
        } else x.lb = v3s(0);
        x.pbw = c->point;
        pc[npc++] = x;
    }
    V3 *slin = calloc(count + 1, sizeof(V3)), *sang = calloc(count + 1, sizeof(V3));
    if (w->nj > 0 || npc > 0) {
        /* step.rs:293-308: start poses (the velocity solve's) for SHAKE; joints between two static
         * or sleeping bodies are not solved; each joint pass is followed by a contact projection */
        V3 *spos = malloc(sizeof(V3) * (count + 1)); Q4 *srot = malloc(sizeof(Q4) * (count + 1));
        uint8_t *frozen = malloc(count + 1);
        for (uint32_t i = 0; i < count; i++) {
            int live = w->dyn[i] && w->awake[i];
            spos[i] = live ? opos[i] : w->pos[i];
            srot[i] = live ? orot[i] : w->rot[i];
            frozen[i] = !live;
        }
        /* Joints::begin_step (bodies.rs:308-318): the step's multipliers start at zero */
        for (uint32_t q = 0; q < w->nj; q++) {
            w->joints[q].lam = 0.0f;
            for (int i = 0; i < 3; i++) { w->joints[q].t[i].lam = 0.0f; w->joints[q].r[i].lam = 0.0f; }
        }
        for (uint32_t it = 0; it < w->ip; it++) {
            for (uint32_t q = 0; q < w->nj; q++) {
                Joint *j = &w->joints[q];
                if (frozen[j->a] && frozen[j->b]) continue;
                if (j->kind == J_DIST) dist_pos(w, j, spos, dt);
                else if (j->kind == J_D6) d6_pos(w, j, spos, srot, dt);
                else ball_pos(w, j, spos, srot, dt);
            }
            project(w, pc, npc, slin, sang, 0);
        }
        project(w, pc, npc, slin, sang, 1);
        free(spos); free(srot); free(frozen);
    }
    free(pc);
    /* S6 (step.rs:316-379): the grounding scan is restated as a per-body flag (same result) */
    uint8_t *ground = calloc(count + 1, 1);
    for (uint32_t k = 0; k < cul.n; k++) {
        const Contact *c = &cul.v[k];
        if (c->b == TERRAIN_CI && c->depth - MARGIN > -POS_SLOP && contact_props(w, c).stiff <= 0.0f) ground[w->cbody[c->a]] = 1;
    }
    /* step.rs:353-365: a body of a D6 joint with a rotational velocity drive keeps the velocity
     * solve's angular velocity (no derivation from its rotation change) */
    uint8_t *driven = calloc(count + 1, 1);
    for (uint32_t q = 0; q < w->nj; q++) {
        const Joint *j = &w->joints[q];
        if (j->kind != J_D6) continue;
        int vd = 0;
        for (int k = 0; k < 3; k++) if (j->r[k].has_drive && j->r[k].vel_mode) vd = 1;
        if (vd) { driven[j->a] = 1; driven[j->b] = 1; }
    }
    for (uint32_t i = 0; i < count; i++) {
        if (!w->dyn[i] || !w->awake[i]) continue;
        V3 vl = w->lin[i], va = w->ang[i];
        V3 np = w->pos[i]; Q4 nq = w->rot[i];
        V3 sl = jointed[i] ? v3s(0) : slin[i], sa = jointed[i] ? v3s(0) : sang[i];
        V3 dl = vsub(vsub(np, ppos[i]), sl);
        if (!viszero(dl)) vl = vadd(vl, vdivs(dl, dt));
        if (ground[i] && vl.y < 0.0f) {
            vl.y = 0.0f;
            np.y = opos[i].y;
            b_set_pose(w, i, np, nq);
        }
        if (!driven[i] && !qeq(nq, prot[i])) {
            Q4 dq = qmul(nq, qconj(prot[i]));
            V3 da = vmul(v3(dq.x, dq.y, dq.z), 2.0f);
            if (dq.w < 0.0f) da = vneg(da);
            va = vadd(va, vdivs(vsub(da, sa), dt));
        }
        b_set_vel(w, i, vl, va);
    }
    free(ground); free(driven);
    for (uint32_t i = 0; i < count; i++) w->last_lin[i] = w->lin[i];
    /* S7 (sleep.rs:59-80, step.rs:381-384) */
    for (uint32_t i = 0; i < count; i++) {
        if (!w->sallow[i]) continue;
        if (!w->awake[i] || !w->dyn[i]) continue;
        if (vlen2(w->lin[i]) < 0.01f && vlen2(w->ang[i]) < 0.01f) {
            w->sticks[i] = (uint16_t)(w->sticks[i] + 1);
            if (w->sticks[i] > 30) w->awake[i] = 0;
        } else w->sticks[i] = 0;
    }
    for (uint32_t c = 0; c < nc; c++) { uint32_t b = w->cbody[c]; w->caabb[c] = aabb_of(w->cshape[c], w->pos[b], w->rot[b]); }
    free(w->contacts);
    w->contacts = cul.v; w->ncontacts = (uint32_t)cul.n;
    w->stats.contacts = w->ncontacts;
    free(pre_lin); free(pre_ang); free(sweep); free(sw);
    free(opos); free(orot); free(ppos); free(prot); free(constrained); free(ball); free(slin); free(sang); free(jointed);
}

/* ------------------------------------------------------------------ hash (provenance.rs:6-24, world.rs:882-909) */
/* The hash input goes to a BLAKE3 state, or (for diagnosing a differing hash) into a byte buffer. */
typedef struct { OrBlake3 *b3; uint8_t *buf; size_t n, cap; } HSink;
static void hput(HSink *h, const void *p, size_t n) {
    if (h->b3) { or_blake3_update(h->b3, p, n); return; }
    if (h->buf && h->n + n <= h->cap) memcpy(h->buf + h->n, p, n);
    h->n += n;
}
static void hu32(HSink *h, uint32_t v) { uint8_t b[4] = {(uint8_t)v, (uint8_t)(v >> 8), (uint8_t)(v >> 16), (uint8_t)(v >> 24)}; hput(h, b, 4); }
static void hu64(HSink *h, uint64_t v) { hu32(h, (uint32_t)v); hu32(h, (uint32_t)(v >> 32)); }
static void hf(HSink *h, float x) { hu32(h, canon_bits(x)); }
static void hv3(HSink *h, V3 v) { hf(h, v.x); hf(h, v.y); hf(h, v.z); }

static void step_hash_feed(const OrWorld *w, HSink *hs) {
#define h (*hs)
    hu64(&h, w->epoch);
    hu64(&h, w->rng);
    uint8_t empty[32];
    or_blake3(NULL, 0, empty); /* ScheduleRecorder::digest of no stages */
    hput(&h, empty, 32);
    for (uint32_t i = 0; i < w->n; i++) {
        hu32(&h, i);
        hv3(&h, w->pos[i]);
        hf(&h, w->rot[i].x); hf(&h, w->rot[i].y); hf(&h, w->rot[i].z); hf(&h, w->rot[i].w);
        hv3(&h, w->lin[i]); hv3(&h, w->ang[i]);
    }
    /* warm: u64 len (keys), then per key: 4 u32, u32 count, entries */
    uint64_t nkeys = 0;
    for (uint32_t i = 0; i < w->nwarm; i++) if (i == 0 || key_cmp(w->warm[i].key, w->warm[i - 1].key) != 0) nkeys++;
    hu64(&h, nkeys);
    for (uint32_t i = 0; i < w->nwarm;) {
        uint32_t e = i;
        while (e < w->nwarm && key_cmp(w->warm[e].key, w->warm[i].key) == 0) e++;
        for (int k = 0; k < 4; k++) hu32(&h, w->warm[i].key[k]);
        hu32(&h, e - i);
        for (uint32_t j = i; j < e; j++) { hv3(&h, w->warm[j].la); hf(&h, w->warm[j].jn); hv3(&h, w->warm[j].jt); }
        i = e;
    }
    hu64(&h, w->nwpatch); /* warm_patches, in key order */
    for (uint32_t i = 0; i < w->nwpatch; i++) {
        for (int k = 0; k < 4; k++) hu32(&h, w->wpatch[i].key[k]);
        hv3(&h, w->wpatch[i].jt); hf(&h, w->wpatch[i].tw);
    }
    uint64_t nwj = 0;
    for (uint32_t q = 0; q < w->nj; q++) nwj += w->wj[q].has;
    hu64(&h, nwj); /* warm_joints, in joint order */
    for (uint32_t q = 0; q < w->nj; q++) {
        if (!w->wj[q].has) continue;
        hu32(&h, q);
        hu32(&h, w->wj[q].n);
        for (uint32_t k = 0; k < w->wj[q].n; k++) { hput(&h, &w->wj[q].tag[k], 1); hf(&h, w->wj[q].lam[k]); }
    }
    hu64(&h, 0); /* rotor */
    hu64(&h, w->n);
    for (uint32_t i = 0; i < w->n; i++) hv3(&h, w->last_lin[i]);
    uint32_t sl = w->cap > w->n ? w->cap : w->n;
    hu64(&h, sl);
    hput(&h, w->awake, sl);
    hu64(&h, sl);
    for (uint32_t i = 0; i < sl; i++) { uint8_t b[2] = {(uint8_t)w->sticks[i], (uint8_t)(w->sticks[i] >> 8)}; hput(&h, b, 2); }
#undef h
}
void or_step_hash(const OrWorld *w, uint8_t out[32]) {
    OrBlake3 b3;
    or_blake3_init(&b3);
    HSink hs = {&b3, NULL, 0, 0};
    step_hash_feed(w, &hs);
    or_blake3_final(&b3, out);
}
/* The step hash's input bytes (up to cap of them into out); returns their count. */
size_t or_step_hash_preimage(const OrWorld *w, uint8_t *out, size_t cap) {
    HSink hs = {NULL, out, 0, cap};
    step_hash_feed(w, &hs);
    return hs.n;
}

uint32_t or_export_carried(const OrWorld *w, float *last3, uint8_t *awake, uint16_t *sticks, uint32_t *wkeys4, float *wla3, float *wjn,
                           float *wjt3, uint32_t cap, uint32_t *jq, uint8_t *jn_rows, uint8_t *jtags3, float *jlam3, uint32_t *njw) {
    for (uint32_t i = 0; last3 && i < w->n; i++) { last3[3 * i] = w->last_lin[i].x; last3[3 * i + 1] = w->last_lin[i].y; last3[3 * i + 2] = w->last_lin[i].z; }
    uint32_t sl = w->cap > w->n ? w->cap : w->n;
    for (uint32_t i = 0; awake && i < sl; i++) { awake[i] = w->awake[i]; sticks[i] = w->sticks[i]; }
    for (uint32_t e = 0; wkeys4 && e < w->nwarm && e < cap; e++) {
        memcpy(wkeys4 + 4 * e, w->warm[e].key, 16);
        wla3[3 * e] = w->warm[e].la.x; wla3[3 * e + 1] = w->warm[e].la.y; wla3[3 * e + 2] = w->warm[e].la.z;
        wjn[e] = w->warm[e].jn;
        wjt3[3 * e] = w->warm[e].jt.x; wjt3[3 * e + 1] = w->warm[e].jt.y; wjt3[3 * e + 2] = w->warm[e].jt.z;
    }
    uint32_t k = 0;
    for (uint32_t q = 0; q < w->nj; q++) {
        if (!w->wj[q].has) continue;
        if (jq) {
            jq[k] = q; jn_rows[k] = w->wj[q].n;
            for (uint32_t r = 0; r < w->wj[q].n; r++) { jtags3[3 * k + r] = w->wj[q].tag[r]; jlam3[3 * k + r] = w->wj[q].lam[r]; }
        }
        k++;
    }
    if (njw) *njw = k;
    return w->nwarm;
}
uint32_t or_capacity(const OrWorld *w) { return w->cap; }
uint32_t or_num_bodies(const OrWorld *w) { return w->n; }
uint64_t or_tick(const OrWorld *w) { return w->tick; }
void or_get_state(const OrWorld *w, float *pos, float *rot, float *lin, float *ang) {
    for (uint32_t i = 0; i < w->n; i++) {
        if (pos) { pos[3 * i] = w->pos[i].x; pos[3 * i + 1] = w->pos[i].y; pos[3 * i + 2] = w->pos[i].z; }
        if (rot) { rot[4 * i] = w->rot[i].x; rot[4 * i + 1] = w->rot[i].y; rot[4 * i + 2] = w->rot[i].z; rot[4 * i + 3] = w->rot[i].w; }
        if (lin) { lin[3 * i] = w->lin[i].x; lin[3 * i + 1] = w->lin[i].y; lin[3 * i + 2] = w->lin[i].z; }
        if (ang) { ang[3 * i] = w->ang[i].x; ang[3 * i + 1] = w->ang[i].y; ang[3 * i + 2] = w->ang[i].z; }
    }
}
void or_get_stats(const OrWorld *w, OrStats *s) { *s = w->stats; }
uint32_t or_get_pairs(const OrWorld *w, uint32_t *ij, uint32_t cap) {
    uint32_t n = w->npairs < cap ? w->npairs : cap;
    if (ij) memcpy(ij, w->pairs, sizeof(uint32_t) * 2 * n);
    return w->npairs;
}
uint32_t or_get_contacts(const OrWorld *w, OrContact *out, uint32_t cap) {
    uint32_t n = w->ncontacts < cap ? w->ncontacts : cap;
    for (uint32_t i = 0; out && i < n; i++) {
        const Contact *c = &w->contacts[i];
        out[i].a_col = c->a; out[i].b_col = c->b;
        out[i].point[0] = c->point.x; out[i].point[1] = c->point.y; out[i].point[2] = c->point.z;
        out[i].normal[0] = c->normal.x; out[i].normal[1] = c->normal.y; out[i].normal[2] = c->normal.z;
        out[i].depth = c->depth; out[i].impulse = c->impulse; out[i].feat = c->feat; out[i].cell = c->cell;
    }
    return w->ncontacts;
}
uint32_t or_get_awake(const OrWorld *w, uint8_t *out, uint32_t cap) {
    uint32_t n = w->n < cap ? w->n : cap;
    if (out) memcpy(out, w->awake, n);
    return w->n;
}

/* ------------------------------------------------------------------ snapshots (ffi/src/lib.rs:1696-1800) */
typedef struct {
    uint64_t tick; uint32_t n, cap;
    V3 *pos; Q4 *rot; V3 *lin, *ang;
    Warm *warm; uint32_t nwarm;
    WarmP *wpatch; uint32_t nwpatch;
    V3 *last_lin; uint8_t *awake; uint16_t *sticks;
    WarmJ *wj; uint32_t nj;
    OrPlayer *players; uint32_t np;
} Snap;
void *or_snapshot(const OrWorld *w) {
    Snap *s = calloc(1, sizeof(Snap));
    s->tick = w->tick; s->n = w->n; s->cap = w->cap;
    s->pos = malloc(sizeof(V3) * (w->n + 1)); s->rot = malloc(sizeof(Q4) * (w->n + 1));
    s->lin = malloc(sizeof(V3) * (w->n + 1)); s->ang = malloc(sizeof(V3) * (w->n + 1));
    for (uint32_t i = 0; i < w->n; i++) {
        if (w->alive[i]) { s->pos[i] = w->pos[i]; s->rot[i] = w->rot[i]; s->lin[i] = w->lin[i]; s->ang[i] = w->ang[i]; }
        else { s->pos[i] = v3s(0); s->rot[i] = qident(); s->lin[i] = v3s(0); s->ang[i] = v3s(0); }
    }
    s->warm = malloc(sizeof(Warm) * (w->nwarm + 1)); memcpy(s->warm, w->warm, sizeof(Warm) * w->nwarm); s->nwarm = w->nwarm;
    s->wpatch = malloc(sizeof(WarmP) * (w->nwpatch + 1)); memcpy(s->wpatch, w->wpatch, sizeof(WarmP) * w->nwpatch); s->nwpatch = w->nwpatch;
    s->last_lin = malloc(sizeof(V3) * (w->n + 1)); memcpy(s->last_lin, w->last_lin, sizeof(V3) * w->n);
    s->awake = malloc(w->cap + 1); memcpy(s->awake, w->awake, w->cap);
    s->sticks = malloc(2 * (w->cap + 1)); memcpy(s->sticks, w->sticks, 2 * w->cap);
    s->nj = w->nj; s->wj = malloc(sizeof(WarmJ) * (w->nj + 1)); memcpy(s->wj, w->wj, sizeof(WarmJ) * w->nj);
    s->np = w->np; s->players = malloc(sizeof(OrPlayer) * (w->np + 1)); memcpy(s->players, w->players, sizeof(OrPlayer) * w->np);
    return s;
}
void or_restore(OrWorld *w, const void *sp) {
    const Snap *s = sp;
    w->tick = s->tick;
    /* lib.rs:1760-1761: the players as snapshot (players added since are dropped) */
    w->np = s->np < w->np ? s->np : w->np;
    memcpy(w->players, s->players, sizeof(OrPlayer) * w->np);
    for (uint32_t i = 0; i < s->n && i < w->n; i++) {
        if (!w->alive[i]) continue;
        float pose[7] = {s->pos[i].x, s->pos[i].y, s->pos[i].z, s->rot[i].x, s->rot[i].y, s->rot[i].z, s->rot[i].w};
        float vel[6] = {s->lin[i].x, s->lin[i].y, s->lin[i].z, s->ang[i].x, s->ang[i].y, s->ang[i].z};
        or_set_body_pose(w, i, pose);
        or_set_body_vel(w, i, vel);
    }
    free(w->warm);
    w->warm = malloc(sizeof(Warm) * (s->nwarm + 1)); memcpy(w->warm, s->warm, sizeof(Warm) * s->nwarm); w->nwarm = s->nwarm;
    for (uint32_t i = 0; i < s->n && i < w->n; i++) w->last_lin[i] = s->last_lin[i];
    for (uint32_t i = 0; i < s->cap && i < w->cap; i++) { w->awake[i] = s->awake[i]; w->sticks[i] = s->sticks[i]; }
    uint32_t k = 0;
    for (uint32_t i = 0; i < w->nwarm; i++) {
        uint32_t a = w->warm[i].key[0], b = w->warm[i].key[1];
        int la = a == 0xFFFFFFFFu || (a < w->cap && w->alive[a]);
        int lb = b == 0xFFFFFFFFu || (b < w->cap && w->alive[b]);
        if (la && lb) w->warm[k++] = w->warm[i];
    }
    w->nwarm = k;
    free(w->wpatch);
    w->wpatch = malloc(sizeof(WarmP) * (s->nwpatch + 1));
    k = 0;
    for (uint32_t i = 0; i < s->nwpatch; i++) {
        uint32_t a = s->wpatch[i].key[0], b = s->wpatch[i].key[1];
        int la = a == 0xFFFFFFFFu || (a < w->cap && w->alive[a]);
        int lb = b == 0xFFFFFFFFu || (b < w->cap && w->alive[b]);
        if (la && lb) w->wpatch[k++] = s->wpatch[i];
    }
    w->nwpatch = k;
    /* warm joint rows as snapshot (joints added since start cold); a dead body's joints drop theirs */
    for (uint32_t q = 0; q < w->nj; q++) {
        if (q < s->nj) w->wj[q] = s->wj[q]; else memset(&w->wj[q], 0, sizeof(WarmJ));
        if (!w->alive[w->joints[q].a] || !w->alive[w->joints[q].b]) w->wj[q].has = 0;
    }
}
void or_snapshot_free(void *sp) {
    Snap *s = sp;
    if (!s) return;
    free(s->pos); free(s->rot); free(s->lin); free(s->ang); free(s->warm); free(s->wpatch); free(s->last_lin); free(s->awake); free(s->sticks); free(s->wj);
    free(s->players);
    free(s);
}

/* ------------------------------------------------------------------ unit entry points */
/* patch_radius of n plane points (x, y pairs) */
float or_patch_radius(const float *xy, uint32_t n) {
    P2 *p = malloc(sizeof(P2) * (n + 1));
    for (uint32_t i = 0; i < n; i++) { p[i].x = xy[2 * i]; p[i].y = xy[2 * i + 1]; }
    float r = patch_radius(p, n);
    free(p);
    return r;
}
float or_atan2f_unit(float y, float x) { return or_atan2f(y, x); }
static void write_m(const Manifold *m, float *out) {
    for (int i = 0; i < m->len; i++) {
        out[7 * i] = m->p[i].pos.x; out[7 * i + 1] = m->p[i].pos.y; out[7 * i + 2] = m->p[i].pos.z;
        out[7 * i + 3] = m->p[i].n.x; out[7 * i + 4] = m->p[i].n.y; out[7 * i + 5] = m->p[i].n.z;
        out[7 * i + 6] = m->p[i].pen;
    }
}
int or_collide(int sa, const float *da, const float *pa, int sb, const float *db, const float *pb, float cd, float *out) {
    ShapeD A = {sa, {da[0], da[1], da[2]}}, B = {sb, {db[0], db[1], db[2]}};
    Manifold m;
    if (!collide_shapes(A, v3(pa[0], pa[1], pa[2]), q4(pa[3], pa[4], pa[5], pa[6]), B, v3(pb[0], pb[1], pb[2]),
                        q4(pb[3], pb[4], pb[5], pb[6]), cd, &m)) return 0;
    write_m(&m, out);
    return m.len;
}
int or_collide_box_triangle(const float *he, const float *pose, const float *t9, float cd, float *out) {
    Tri t = {v3(t9[0], t9[1], t9[2]), v3(t9[3], t9[4], t9[5]), v3(t9[6], t9[7], t9[8])};
    ShapeD s = {OR_BOX, {he[0], he[1], he[2]}};
    Manifold m;
    if (!collide_shape_tri(s, v3(pose[0], pose[1], pose[2]), q4(pose[3], pose[4], pose[5], pose[6]), t, cd, &m)) return 0;
    write_m(&m, out);
    return m.len;
}
int or_collide_shape_triangle(int shape, const float *dims, const float *pose, const float *t9, float cd, float *out) {
    ShapeD s; s.type = shape; s.d[0] = dims[0]; s.d[1] = dims[1]; s.d[2] = dims[2];
    Tri t = {v3(t9[0], t9[1], t9[2]), v3(t9[3], t9[4], t9[5]), v3(t9[6], t9[7], t9[8])};
    Manifold m;
    if (!collide_shape_tri(s, v3(pose[0], pose[1], pose[2]), q4(pose[3], pose[4], pose[5], pose[6]), t, cd, &m)) return 0;
    write_m(&m, out);
    return m.len;
}
int or_collide_sphere_triangle(float r, const float *c, const float *t9, float cd, float *out) {
    Tri t = {v3(t9[0], t9[1], t9[2]), v3(t9[3], t9[4], t9[5]), v3(t9[6], t9[7], t9[8])};
    ShapeD s = {OR_SPHERE, {r, 0, 0}};
    Manifold m;
    if (!collide_shape_tri(s, v3(c[0], c[1], c[2]), qident(), t, cd, &m)) return 0;
    write_m(&m, out);
    return m.len;
}

/* Broadphase over caller-given (swept) AABBs {minx,miny,minz,maxx,maxy,maxz} per collider;
 * returns the pair count, writes up to cap (i, j) pairs. */
uint32_t or_broadphase(const float *bb, uint32_t n, uint32_t *out, uint32_t cap) {
    Aabb *a = malloc(sizeof(Aabb) * (n + 1));
    for (uint32_t i = 0; i < n; i++) { a[i].min = v3(bb[6 * i], bb[6 * i + 1], bb[6 * i + 2]); a[i].max = v3(bb[6 * i + 3], bb[6 * i + 4], bb[6 * i + 5]); }
    uint32_t *pairs = NULL;
    uint32_t np = broadphase(a, n, &pairs);
    for (uint32_t k = 0; k < np && k < cap; k++) { out[2 * k] = pairs[2 * k]; out[2 * k + 1] = pairs[2 * k + 1]; }
    free(pairs); free(a);
    return np;
}

/* Loads a full state (e.g. one the device library reached): bodies, sleep state, last linear
 * velocities, tick and the warm cache (keys (a<<32|b), anchors, jn, jt; sorted by key). */
void or_load_state(OrWorld *w, const float *pos, const float *rot, const float *lin, const float *ang, const float *last_lin,
                   const uint32_t *awake, const uint32_t *sticks, uint64_t tick, const uint64_t *wkeys, const float *wla,
                   const float *wjn, const float *wjt, uint32_t nwarm) {
    for (uint32_t i = 0; i < w->n; i++) {
        w->pos[i] = v3(pos[3 * i], pos[3 * i + 1], pos[3 * i + 2]);
        w->rot[i] = q4(rot[4 * i], rot[4 * i + 1], rot[4 * i + 2], rot[4 * i + 3]);
        w->lin[i] = v3(lin[3 * i], lin[3 * i + 1], lin[3 * i + 2]);
        w->ang[i] = v3(ang[3 * i], ang[3 * i + 1], ang[3 * i + 2]);
        w->last_lin[i] = v3(last_lin[3 * i], last_lin[3 * i + 1], last_lin[3 * i + 2]);
        w->awake[i] = (uint8_t)awake[i];
        w->sticks[i] = (uint16_t)sticks[i];
    }
    w->tick = tick;
    free(w->warm);
    w->warm = malloc(sizeof(Warm) * (nwarm + 1));
    for (uint32_t k = 0; k < nwarm; k++) {
        Warm *e = &w->warm[k];
        e->key[0] = (uint32_t)(wkeys[k] >> 32); e->key[1] = (uint32_t)wkeys[k]; e->key[2] = 0; e->key[3] = 0;
        e->la = v3(wla[3 * k], wla[3 * k + 1], wla[3 * k + 2]);
        e->jn = wjn[k];
        e->jt = v3(wjt[3 * k], wjt[3 * k + 1], wjt[3 * k + 2]);
    }
    w->nwarm = nwarm;
}

/* ------------------------------------------------------------------ queries
 * world/src/world/queries.rs:50-261 (World::raycast against collider AABBs and terrain triangles;
 * World::sweep_capsule: overlap at the start against true shapes, then tip and base spheres swept
 * against AABBs), with the geometry they call: ray_aabb / normal_from_aabb_point
 * (world/src/world/geom.rs:40-93), Triangle::intersect_ray / barycentric / normal
 * (geom/src/triangle.rs:25-232, plane.rs:15-86), the query crate's sphere sweeps
 * (query/src/lib.rs: ray_aabb_slab_unit, sweep_sphere_aabb, sweep_two_spheres_aabbs, pick_better),
 * HeightField::query_aabb (terrain.rs:300-357).  Body indices; TERRAIN_CI for the terrain. */
static V3 q_normal_from_aabb_point(Aabb bb, V3 p) {
    const float eps = 1.0e-4f;
    if (fabsf(p.x - bb.min.x) < eps) return v3(-1, 0, 0);
    if (fabsf(p.x - bb.max.x) < eps) return v3(1, 0, 0);
    if (fabsf(p.y - bb.min.y) < eps) return v3(0, -1, 0);
    if (fabsf(p.y - bb.max.y) < eps) return v3(0, 1, 0);
    if (fabsf(p.z - bb.min.z) < eps) return v3(0, 0, -1);
    if (fabsf(p.z - bb.max.z) < eps) return v3(0, 0, 1);
    V3 c = vmul(vadd(bb.min, bb.max), 0.5f);
    V3 n = vsub(p, c);
    float l2 = vlen2(n);
    if (l2 > 1.0e-20f) return vdivs(n, sqrtf(l2));
    return v3(0, 1, 0);
}
static int q_ray_aabb(V3 o, V3 d, Aabb bb, float max_dist, float *t_out, V3 *p_out, V3 *n_out) {
    const float eps = 1.0e-8f;
    float tmin = 0.0f, tmax = max_dist;
    float oo[3] = {o.x, o.y, o.z}, dd[3] = {d.x, d.y, d.z};
    float mn[3] = {bb.min.x, bb.min.y, bb.min.z}, mx[3] = {bb.max.x, bb.max.y, bb.max.z};
    for (int a = 0; a < 3; a++) {
        if (fabsf(dd[a]) < eps) {
            if (oo[a] < mn[a] || oo[a] > mx[a]) return 0;
        } else {
            float inv = 1.0f / dd[a];
            float t1 = (mn[a] - oo[a]) * inv, t2 = (mx[a] - oo[a]) * inv;
            if (t1 > t2) { float x = t1; t1 = t2; t2 = x; }
            if (t1 > tmin) tmin = t1;
            if (t2 < tmax) tmax = t2;
            if (tmin > tmax) return 0;
        }
    }
    if (tmin < 0.0f || tmin > max_dist) return 0;
    *t_out = tmin;
    *p_out = vadd(o, vmul(d, tmin));
    *n_out = q_normal_from_aabb_point(bb, *p_out);
    return 1;
}
static int q_finite3(V3 v) { return vfinite(v); }
/* Triangle::intersect_ray(ray, 0, t_max): plane through the triangle, then barycentric inside test */
static int q_tri_ray(Tri t, V3 o, V3 d, float t_max, float *t_out) {
    if (0.0f > t_max) return 0;
    V3 nu = vcross(vsub(t.b, t.a), vsub(t.c, t.a));
    float a2 = vdot(nu, nu);
    if (!isfinite(a2) || a2 <= 1.0e-16f) return 0;                       /* is_degenerate */
    if (!(q_finite3(t.a) && q_finite3(t.b) && q_finite3(t.c))) return 0;  /* Plane::from_points */
    float len2 = vlen2(nu);
    if (!isfinite(len2) || len2 <= 1.0e-6f * 1.0e-6f) return 0;
    float inv_len = 1.0f / sqrtf(len2);
    V3 pn = vmul(nu, inv_len);
    float pd = -vdot(pn, t.a);
    if (!(q_finite3(pn) && isfinite(pd))) return 0;
    if (!(q_finite3(o) && q_finite3(d))) return 0;                       /* Plane::intersect_ray */
    float denom = vdot(pn, d);
    if (!isfinite(denom) || fabsf(denom) < 1.0e-6f) return 0;
    float numer = -(vdot(pn, o) + pd);
    if (!isfinite(numer)) return 0;
    float tt = numer / denom;
    if (!isfinite(tt)) return 0;
    if (!(tt >= 0.0f && tt <= t_max)) return 0;
    V3 p = vadd(o, vmul(d, tt));                                         /* Ray::point_at */
    /* barycentric */
    if (!(q_finite3(t.a) && q_finite3(t.b) && q_finite3(t.c) && q_finite3(p))) return 0;
    V3 v0 = vsub(t.b, t.a), v1 = vsub(t.c, t.a), v2 = vsub(p, t.a);
    float d00 = vdot(v0, v0), d01 = vdot(v0, v1), d11 = vdot(v1, v1), d20 = vdot(v2, v0), d21 = vdot(v2, v1);
    float den = d00 * d11 - d01 * d01;
    float scale = fmaxf(fabsf(d00 * d11), 1.0f);
    if (!isfinite(den) || fabsf(den) <= 1.0e-12f * scale) return 0;
    float id = 1.0f / den;
    float bv = (d11 * d20 - d01 * d21) * id, bw = (d00 * d21 - d01 * d20) * id;
    float bu = 1.0f - bv - bw;
    if (!(isfinite(bu) && isfinite(bv) && isfinite(bw))) return 0;
    if (!(bu >= -1.0e-5f && bv >= -1.0e-5f && bw >= -1.0e-5f)) return 0;  /* contains_barycentric */
    *t_out = tt;
    return 1;
}
static V3 q_tri_normal(Tri t) {
    V3 n = vcross(vsub(t.b, t.a), vsub(t.c, t.a));
    if (!vfinite(n)) return v3s(0);
    return vnormalize_or_zero(n);
}
/* HeightField::query_aabb's cell range (terrain.rs:300-319); 0 when nothing is emitted */
static int q_hf_range(const OrWorld *w, Aabb bb, int *gx0, int *gx1, int *gz0, int *gz1) {
    if (!w->has_hf) return 0;
    int nx = (int)w->hnx, nz = (int)w->hnz;
    if (nx < 2 || nz < 2) return 0;
    float lminx = bb.min.x - w->hox, lmaxx = bb.max.x - w->hox;
    float lminz = bb.min.z - w->hoz, lmaxz = bb.max.z - w->hoz;
    *gx0 = ICLAMP((int)floorf(lminx / w->hsx) - 1, 0, nx - 2); *gx1 = ICLAMP((int)ceilf(lmaxx / w->hsx) + 1, 0, nx - 2);
    *gz0 = ICLAMP((int)floorf(lminz / w->hsz) - 1, 0, nz - 2); *gz1 = ICLAMP((int)ceilf(lmaxz / w->hsz) + 1, 0, nz - 2);
    if (bb.min.y > w->hmax + 1.0f || bb.max.y < w->hmin - 1.0f) return 0;
    return 1;
}
static void q_hf_tris(const OrWorld *w, int gx, int gz, Tri out[2]) {
    int nx = (int)w->hnx;
    float h00 = w->hts[gx + gz * nx], h10 = w->hts[gx + 1 + gz * nx];
    float h01 = w->hts[gx + (gz + 1) * nx], h11 = w->hts[gx + 1 + (gz + 1) * nx];
    float x0 = (float)gx * w->hsx + w->hox, x1 = x0 + w->hsx;
    float z0 = (float)gz * w->hsz + w->hoz, z1 = z0 + w->hsz;
    V3 p00 = v3(x0, h00, z0), p10 = v3(x1, h10, z0), p01 = v3(x0, h01, z1), p11 = v3(x1, h11, z1);
    Tri a = {p00, p01, p10}, b = {p10, p01, p11};
    out[0] = a; out[1] = b;
}
static void q_put(float *o3, V3 v) { o3[0] = v.x; o3[1] = v.y; o3[2] = v.z; }
static Aabb q_aabb_of_body(const OrWorld *w, uint32_t c) { uint32_t b = w->cbody[c]; return aabb_of(w->cshape[c], w->pos[b], w->rot[b]); }

/* World::raycast (queries.rs:50-116); dir is normalised by the caller as the FFI does */
int or_raycast(const OrWorld *w, const float *origin3, const float *dir3, float max_dist, int32_t ignore, OrQueryHit *out) {
    memset(out, 0, sizeof(*out));
    V3 o = v3(origin3[0], origin3[1], origin3[2]), dr = v3(dir3[0], dir3[1], dir3[2]);
    float max_d = fmaxf(max_dist, 0.0f);
    if (max_d <= 0.0f) return 0;
    float len = vlen(dr);
    if (!isfinite(len) || len < 1.0e-12f) return 0;
    V3 d = vdivs(dr, len);
    int have = 0; float bt = 0; uint32_t bi = 0; V3 bp = v3s(0), bn = v3s(0);
    for (uint32_t c = 0; c < w->nc; c++) {
        if (ignore >= 0 && w->cbody[c] == (uint32_t)ignore) continue;
        float t; V3 p, n;
        if (!q_ray_aabb(o, d, q_aabb_of_body(w, c), max_d, &t, &p, &n)) continue;
        if (!have || (t < bt - 1.0e-9f) || (fabsf(t - bt) <= 1.0e-9f && c < bi)) { have = 1; bt = t; bi = c; bp = p; bn = n; }
    }
    V3 end = vadd(o, vmul(d, max_d));
    Aabb qb = aabb_new(vmin(o, end), vmax(o, end));
    int gx0, gx1, gz0, gz1;
    if (q_hf_range(w, qb, &gx0, &gx1, &gz0, &gz1))
        for (int gz = gz0; gz <= gz1; gz++)
            for (int gx = gx0; gx <= gx1; gx++) {
                Tri tr[2];
                q_hf_tris(w, gx, gz, tr);
                for (int k = 0; k < 2; k++) {
                    float t;
                    if (!q_tri_ray(tr[k], o, d, max_d, &t)) continue;
                    V3 p = isfinite(t) ? vadd(o, vmul(d, t)) : o;
                    if (!have || t < bt - 1.0e-9f) { have = 1; bt = t; bi = TERRAIN_CI; bp = p; bn = q_tri_normal(tr[k]); }
                }
            }
    if (!have) return 0;
    out->hit = 1;
    out->body = bi == TERRAIN_CI ? TERRAIN_CI : w->cbody[bi];
    out->fraction = fclampf(bt / max_d, 0.0f, 1.0f);
    q_put(out->point, bp); q_put(out->normal, bn);
    return 1;
}

/* query/src/lib.rs ray_aabb_slab_unit on the AABB grown by r (sweep_sphere_aabb with dt = 1 and
 * a non-degenerate sweep) */
static int q_sweep_sphere_aabb(V3 p0, V3 dir, float r, Aabb bb, float *t_out, V3 *n_out) {
    V3 mn = vsub(bb.min, v3s(r)), mx = vadd(bb.max, v3s(r));
    V3 inv = v3(p0.x == p0.x && fabsf(dir.x) > 1.0e-9f ? 1.0f / dir.x : 1.0e9f,
                p0.y == p0.y && fabsf(dir.y) > 1.0e-9f ? 1.0f / dir.y : 1.0e9f,
                p0.z == p0.z && fabsf(dir.z) > 1.0e-9f ? 1.0f / dir.z : 1.0e9f);
    V3 t1 = v3((mn.x - p0.x) * inv.x, (mn.y - p0.y) * inv.y, (mn.z - p0.z) * inv.z);
    V3 t2 = v3((mx.x - p0.x) * inv.x, (mx.y - p0.y) * inv.y, (mx.z - p0.z) * inv.z);
    V3 tmin = vmin(t1, t2), tmax = vmax(t1, t2);
    float te = tmin.x;
    V3 n = v3(t1.x > t2.x ? 1.0f : -1.0f, 0, 0);
    if (tmin.y > te) { te = tmin.y; n = v3(0, t1.y > t2.y ? 1.0f : -1.0f, 0); }
    if (tmin.z > te) { te = tmin.z; n = v3(0, 0, t1.z > t2.z ? 1.0f : -1.0f); }
    float tx = fminf(fminf(tmax.x, tmax.y), tmax.z);
    if (te <= tx && tx >= 0.0f && te >= 0.0f && te <= 1.0f) { *t_out = te; *n_out = n; return 1; }
    return 0;
}
/* sweep_two_spheres_aabbs over one AABB: the tip's hit, replaced by the base's when earlier by
 * more than 1e-9 (pick_better) */
static int q_sweep_two(V3 tip0, V3 base0, V3 v, float r, Aabb bb, float *t_out, V3 *n_out, int *kind) {
    float t; V3 n;
    int have = 0;
    if (q_sweep_sphere_aabb(tip0, v, r, bb, &t, &n)) { have = 1; *t_out = t; *n_out = n; *kind = 0; }
    if (q_sweep_sphere_aabb(base0, v, r, bb, &t, &n)) {
        if (!have || t < *t_out - 1.0e-9f) { have = 1; *t_out = t; *n_out = n; *kind = 1; }
    }
    return have;
}
/* World::sweep_capsule (queries.rs:118-261) */
int or_capsule_sweep(const OrWorld *w, const float *from3, const float *to3, float radius, float half_height, int32_t ignore,
                     OrQueryHit *out) {
    memset(out, 0, sizeof(*out));
    V3 from = v3(from3[0], from3[1], from3[2]), to = v3(to3[0], to3[1], to3[2]);
    V3 v = vsub(to, from);
    float dist_sq = vlen2(v);
    ShapeD cap; cap.type = OR_CAPSULE; cap.d[0] = radius; cap.d[1] = half_height; cap.d[2] = 0.0f;
    Q4 qi = qident();
    for (uint32_t c = 0; c < w->nc; c++) {
        uint32_t b = w->cbody[c];
        if (ignore >= 0 && b == (uint32_t)ignore) continue;
        Manifold m;
        if (collide_shapes(cap, from, qi, w->cshape[c], w->pos[b], w->rot[b], 0.0f, &m) && m.len > 0) {
            out->hit = 1; out->body = b; out->fraction = 0.0f; out->started = 1; q_put(out->point, m.p[0].pos); q_put(out->normal, m.normal);
            return 1;
        }
    }
    Aabb sb = aabb_of(cap, from, qi);
    int gx0, gx1, gz0, gz1;
    if (q_hf_range(w, sb, &gx0, &gx1, &gz0, &gz1))
        for (int gz = gz0; gz <= gz1; gz++)
            for (int gx = gx0; gx <= gx1; gx++) {
                Tri tr[2];
                q_hf_tris(w, gx, gz, tr);
                for (int k = 0; k < 2; k++) {
                    Manifold m;
                    if (collide_shape_tri(cap, from, qi, tr[k], 0.0f, &m) && m.len > 0) {
                        out->hit = 1; out->body = TERRAIN_CI; out->fraction = 0.0f; out->started = 1; q_put(out->point, m.p[0].pos);
                        q_put(out->normal, m.normal);
                        return 1;
                    }
                }
            }
    if (dist_sq < 1.0e-12f) return 0;
    V3 tip0 = vadd(from, v3(0.0f, half_height, 0.0f)), base0 = vadd(from, v3(0.0f, -half_height, 0.0f));
    const float eps = 1.0e-9f;
    int have = 0; float bt = 0; uint32_t bi = 0; V3 bp = v3s(0), bn = v3s(0);
    for (uint32_t c = 0; c < w->nc; c++) {
        uint32_t b = w->cbody[c];
        if (ignore >= 0 && b == (uint32_t)ignore) continue;
        float t; V3 n; int kind;
        if (!q_sweep_two(tip0, base0, v, radius, q_aabb_of_body(w, c), &t, &n, &kind)) continue;
        t = fclampf(t, 0.0f, 1.0f);
        V3 p = vsub(vadd(kind == 0 ? tip0 : base0, vmul(v, t)), vmul(n, radius));
        if (!have || (t < bt - eps) || (fabsf(t - bt) <= eps && c < bi)) { have = 1; bt = t; bi = c; bp = p; bn = n; }
    }
    V3 ext = v3s(radius + half_height);
    Aabb swb = aabb_new(vsub(vmin(from, to), ext), vadd(vmax(from, to), ext));
    if (q_hf_range(w, swb, &gx0, &gx1, &gz0, &gz1))
        for (int gz = gz0; gz <= gz1; gz++)
            for (int gx = gx0; gx <= gx1; gx++) {
                Tri tr[2];
                q_hf_tris(w, gx, gz, tr);
                for (int k = 0; k < 2; k++) {
                    /* aabb_of(Triangle, identity): the vertices through the identity pose */
                    V3 wa = vadd(v3s(0), qrot(qi, tr[k].a)), wb = vadd(v3s(0), qrot(qi, tr[k].b)), wc = vadd(v3s(0), qrot(qi, tr[k].c));
                    Aabb eb = aabb_new(vmin(vmin(wa, wb), wc), vmax(vmax(wa, wb), wc));
                    float t; V3 n; int kind;
                    if (!q_sweep_two(tip0, base0, v, radius, eb, &t, &n, &kind)) continue;
                    t = fclampf(t, 0.0f, 1.0f);
                    V3 p = vsub(vadd(kind == 0 ? tip0 : base0, vmul(v, t)), vmul(n, radius));
                    if (!have || t < bt - eps) { have = 1; bt = t; bi = TERRAIN_CI; bp = p; bn = n; }
                }
            }
    if (!have) return 0;
    out->hit = 1;
    out->body = bi == TERRAIN_CI ? TERRAIN_CI : w->cbody[bi];
    out->fraction = fclampf(bt, 0.0f, 1.0f);
    out->started = 0;
    q_put(out->point, bp); q_put(out->normal, bn);
    return 1;
}

/* The W6 numerics alone, for the oracle's own pin tests: a symmetric CSR matrix (both triangles,
 * columns ascending, diagonal present). */
static BCsr bn_wrap(uint32_t n, const uint32_t *ptr, const uint32_t *idx, const double *val) {
    BCsr k;
    k.n = n;
    k.ptr = malloc(4 * (n + 1)); memcpy(k.ptr, ptr, 4 * (n + 1));
    k.idx = malloc(4 * (ptr[n] + 1)); memcpy(k.idx, idx, 4 * ptr[n]);
    k.val = malloc(8 * (ptr[n] + 1)); memcpy(k.val, val, 8 * ptr[n]);
    return k;
}
uint32_t or_bn_min_degree(uint32_t n, const uint32_t *ptr, const uint32_t *idx, uint32_t *order) {
    if (n > BN_DIRECT_MAX) return 0xFFFFFFFFu;
    double *val = calloc(ptr[n] + 1, 8);
    BCsr k = bn_wrap(n, ptr, idx, val);
    BSym s;
    bn_min_degree(&k, &s);
    memcpy(order, s.order, 4 * n);
    uint32_t l = s.cp[n];
    bn_sym_free(&s); bn_csr_free(&k); free(val);
    return l;
}
int or_bn_pseudo_solve(uint32_t n, const uint32_t *ptr, const uint32_t *idx, const double *val, const double *u, double *x) {
    if (n > BN_DIRECT_MAX) return -1;
    BCsr k = bn_wrap(n, ptr, idx, val);
    bn_pseudo_solve(&k, u, x);
    bn_csr_free(&k);
    return 0;
}
int or_bn_bounce_qp(uint32_t n, const uint32_t *ptr, const uint32_t *idx, const double *val, const double *b, const double *acc0,
                    double *x, uint32_t *solves) {
    if (n > BN_DIRECT_MAX) return -1;
    BCsr k = bn_wrap(n, ptr, idx, val);
    int c = bn_bounce_qp(&k, b, acc0, BN_SOLVES, x, solves);
    bn_csr_free(&k);
    return c;
}
float or_toi_bounce(float s0, float v0, float a, float e, float dt, float *pseudo, int *has_pseudo) {
    return bn_toi_bounce(s0, v0, a, e, dt, pseudo, has_pseudo);
}
