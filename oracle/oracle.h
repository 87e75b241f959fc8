/* Oracle (TEST INFRASTRUCTURE ONLY — never shipped, never on the product path).
 *
 * A plain-C restatement of the RiftPhys reference step `World::step`
 * (crates/riftphys-world/src/world/step.rs:40-420) and everything it calls on the contact path,
 * in the reference's evaluation order, single-threaded.  The reference itself is Rust and cannot
 * be compiled here (no Rust toolchain), so this restatement is the checker the GPU path is
 * compared against.  Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
 * load liboracle.so.
 *
 * Parity status of each function is listed in oracle/cpu_ref.c's header.
 */
#ifndef RP_ORACLE_H
#define RP_ORACLE_H
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
    uint32_t cv[8];
    uint64_t chunk_counter;
    uint8_t buf[64];
    uint32_t buf_len;
    uint32_t blocks_compressed;
    uint32_t stack[54][8];
    uint32_t stack_len;
} OrBlake3;

void or_blake3_init(OrBlake3 *h);
void or_blake3_update(OrBlake3 *h, const void *data, size_t len);
void or_blake3_final(const OrBlake3 *h, uint8_t out32[32]);
void or_blake3(const void *data, size_t len, uint8_t out32[32]);

typedef struct OrWorld OrWorld;

/* Shape codes shared with the product's debug API. */
enum { OR_SPHERE = 0, OR_BOX = 1, OR_CAPSULE = 2 };

/* One contact after cull/orient, with the build-defined feature id. */
typedef struct {
    uint32_t a_col, b_col; /* b_col = 0xFFFFFFFF for terrain */
    float point[3];
    float normal[3];
    float depth;
    float impulse;
    uint32_t feat;  /* kind<<24 | which<<16 | k */
    uint32_t cell;  /* terrain: ((gz*(nx-1))+gx)*2+tri ; body pair: 0xFFFFFFFF */
} OrContact;

typedef struct {
    uint32_t pairs, contacts, islands;
    uint32_t terrain_candidates;
    uint32_t levels;       /* contact batches of the velocity solve */
    uint32_t warm_matched;
    uint32_t iters_used;   /* velocity iterations run by the island that ran the most */
    float residual;        /* largest final residual over the islands (m/s), residual mode only */
    uint32_t iters_sum;    /* sum over islands of (iterations run x the island's contacts) */
    /* W6 bounce passes of the last step (BounceStats, world.rs): time-of-impact rows, support rows,
     * pseudo rows, velocity-pass components, components swept instead of solved directly, and the
     * most interior-point solves any component took */
    uint32_t toi_rows, support_rows, pseudo_rows, bounce_components, bounce_fallbacks, bounce_solves;
    uint32_t relax_rows;   /* contacts bounced by the restitution relax pass */
    uint32_t patches;      /* friction patches of the last step */
} OrStats;

OrWorld *or_world_create(uint32_t max_bodies);
void or_world_destroy(OrWorld *w);
void or_set_epoch(OrWorld *w, uint64_t epoch);
void or_set_rng_seed(OrWorld *w, uint64_t seed);
void or_set_gravity(OrWorld *w, float gx, float gy, float gz);
void or_set_iterations(OrWorld *w, uint32_t vel, uint32_t pos);
/* Residual mode (build addition): each island (component of the written bodies) iterates until
 * its sweep residual max(|dJn|/m_n/(1+|bias|), |dJt1|/m_t1, |dJt2|/m_t2) <= tol (the criterion of
 * solver.rs:1149), at most max_iters sweeps.  Islands share no written body, so per-island
 * stopping is independent of their order. */
void or_set_residual_mode(OrWorld *w, int on, float tol, uint32_t max_iters);
void or_set_pair_props(OrWorld *w, uint32_t mat_a, uint32_t mat_b, float mu_s, float mu_k,
                       float restitution, float stribeck_v);
/* materials.rs:75-90: the pair's contact model (stiffness > 0: compliant Hunt-Crossley contact with
 * that dissipation; patch: a manifold's friction solved as one patch); world.rs:992: damping
 * linearised at the step-start penetration instead of the predicted mid-step one */
void or_set_pair_contact_model(OrWorld *w, uint32_t mat_a, uint32_t mat_b, float stiffness, float dissipation, int patch);
void or_set_compliant_damping_at_start(OrWorld *w, int on);
/* unit entry points: a friction patch's mean radius over n plane points (x, y pairs), and the
 * patch point order's atan2 */
float or_patch_radius(const float *xy, uint32_t n);
float or_atan2f_unit(float y, float x);
/* shape: OR_*; dims: sphere {r}, box {hx,hy,hz}, capsule {r,hh}; pose {px,py,pz,qx,qy,qz,qw};
 * vel {lx,ly,lz,ax,ay,az}; density <= 0 means the reference default (1500 kg/m^3). Returns the
 * 0-based body id. */
int32_t or_add_body(OrWorld *w, int shape, const float *dims, const float *pose, const float *vel,
                    int dynamic, float density);
int or_remove_body(OrWorld *w, uint32_t id);
void or_set_heightfield(OrWorld *w, uint32_t nx, uint32_t nz, float sx, float sz, float ox, float oz,
                        const float *heights, uint32_t material);
void or_set_body_pose(OrWorld *w, uint32_t id, const float *pose);
void or_set_body_vel(OrWorld *w, uint32_t id, const float *vel);
void or_set_sleep_allowed(OrWorld *w, uint32_t id, int allowed);
/* The material id (0..15, materials.rs core table) of body id's collider. */
void or_set_body_material(OrWorld *w, uint32_t id, uint32_t mat);
void or_step(OrWorld *w, float dt);
/* Threads of the step's parallel regions (default 1): where the reference uses rayon (narrowphase
 * pairs, terrain colliders, PGS level batches of at least 64 rows), bit-identical for any count. */
void or_set_threads(int n);
void or_step_hash(const OrWorld *w, uint8_t out32[32]);
size_t or_step_hash_preimage(const OrWorld *w, uint8_t *out, size_t cap);
/* world.rs:718 add_ball_joint: bodies a, b (indices), frames {px,py,pz,qx,qy,qz,qw} local to each
 * body.  Returns the joint index, or -1. */
int32_t or_add_ball_joint(OrWorld *w, uint32_t a, uint32_t b, const float *frame_a, const float *frame_b);
/* world.rs:537 add_distance_joint between the centres of mass of a and b: rest length (m),
 * compliance (m/N, 0 = rigid).  Returns the joint index, or -1. */
int32_t or_add_distance_joint(OrWorld *w, uint32_t a, uint32_t b, float rest, float compliance);
/* bodies.rs:321-390 (no armature): a D6 joint with the translational axes of t_mask and the
 * rotational axes of r_mask constrained (bit i: axis i of frame A) and a stored hinge axis (0..2,
 * -1 none); a hinge is t_mask 7, r_mask 7 minus its axis.  Per axis (angular: rotational):
 * a limit (on = 0 removes it), a drive (position or velocity mode), damping and friction, and
 * compliance. */
int32_t or_add_d6_joint(OrWorld *w, uint32_t a, uint32_t b, const float *fa, const float *fb, uint32_t t_mask, uint32_t r_mask, int32_t hinge);
void or_set_d6_limit(OrWorld *w, uint32_t q, int angular, uint32_t axis, int on, float lo, float hi);
void or_set_d6_drive(OrWorld *w, uint32_t q, int angular, uint32_t axis, int on, float target, float kp, float kd, float max_impulse, int vel_mode);
void or_set_d6_damping(OrWorld *w, uint32_t q, int angular, uint32_t axis, float damping, float friction);
void or_set_d6_compliance(OrWorld *w, uint32_t q, int angular, uint32_t axis, float compliance);
/* world.rs:595 add_body without add_collider: the shape only sets the mass properties. */
int32_t or_add_body_nc(OrWorld *w, int shape, const float *dims, const float *pose, const float *vel, int dynamic,
                       float density);
uint32_t or_num_joints(const OrWorld *w);
uint32_t or_get_joint_warm(const OrWorld *w, uint32_t q, uint8_t *tags, float *lams);
int or_shake_ok(float qw);
uint32_t or_num_bodies(const OrWorld *w);
uint64_t or_tick(const OrWorld *w);
/* pos[3n], rot[4n], lin[3n], ang[3n] */
void or_get_state(const OrWorld *w, float *pos, float *rot, float *lin, float *ang);
void or_get_stats(const OrWorld *w, OrStats *s);
/* Debug views of the last step. Each returns the count; copies at most cap entries. */
uint32_t or_get_pairs(const OrWorld *w, uint32_t *ij, uint32_t cap);
uint32_t or_get_contacts(const OrWorld *w, OrContact *out, uint32_t cap);
uint32_t or_get_awake(const OrWorld *w, uint8_t *out, uint32_t cap);
/* Snapshot / restore (rphys_world_snapshot_* semantics). */
void *or_snapshot(const OrWorld *w);
void or_restore(OrWorld *w, const void *snap);
void or_snapshot_free(void *snap);

/* Unit-level entry points used by the oracle's own pin tests. */
int or_collide(int shape_a, const float *dims_a, const float *pose_a, int shape_b, const float *dims_b,
               const float *pose_b, float contact_dist, float *out_points /* 4 x (pos3, n3, pen) */);
int or_collide_box_triangle(const float *he, const float *pose, const float *tri9, float contact_dist,
                            float *out_points);
int or_collide_sphere_triangle(float r, const float *center, const float *tri9, float contact_dist,
                               float *out_points);
void or_sincosf(float x, float *s, float *c);
int or_collide_shape_triangle(int shape, const float *dims, const float *pose, const float *tri9, float contact_dist, float *out_points);
void or_warm_pick_sequence(const uint32_t *cache_keys4, const float *cache_la3, uint32_t nwarm, const uint32_t *keys4,
                           const float *anchors3, uint32_t n, int32_t *out);
/* The carried state the step hash covers (hash_carried, world.rs:884-909): last velocities, sleep
 * slots (capacity entries), warm-cache entries (up to cap) and stored joint rows.  Returns the
 * warm-cache entry count. */
uint32_t or_export_carried(const OrWorld *w, float *last3, uint8_t *awake, uint16_t *sticks, uint32_t *wkeys4, float *wla3, float *wjn,
                           float *wjt3, uint32_t cap, uint32_t *joint_idx, uint8_t *joint_rows, uint8_t *tags3, float *lams3,
                           uint32_t *n_joint_warm);
uint32_t or_capacity(const OrWorld *w);
uint32_t or_broadphase(const float *aabbs6, uint32_t n, uint32_t *out_ij, uint32_t cap);
void or_load_state(OrWorld *w, const float *pos, const float *rot, const float *lin, const float *ang, const float *last_lin,
                   const uint32_t *awake, const uint32_t *sticks, uint64_t tick, const uint64_t *wkeys, const float *wla,
                   const float *wjn, const float *wjt, uint32_t nwarm);

/* Queries (world/src/world/queries.rs): body index of the hit (0xFFFFFFFF terrain), fraction of the
 * ray / sweep, point and normal; started = the capsule overlapped at its start.  ignore = a body
 * index to skip, or -1. */
typedef struct { int hit; uint32_t body; float fraction; float point[3], normal[3]; int started; } OrQueryHit;
int or_raycast(const OrWorld *w, const float *origin3, const float *dir3, float max_dist, int32_t ignore, OrQueryHit *out);
int or_capsule_sweep(const OrWorld *w, const float *from3, const float *to3, float radius, float half_height, int32_t ignore,
                     OrQueryHit *out);

/* Kinematic players (world.rs:466-475, step.rs:544-639, lib.rs:1580-1693): add returns the player
 * index (-1 when the world is full); get writes the pose and grounded flag and returns the body. */
int32_t or_add_player(OrWorld *w, const float *pose, float radius, float height, float speed);
void or_player_set_input(OrWorld *w, uint32_t idx, const float *dir3);
uint32_t or_player_get(const OrWorld *w, uint32_t idx, float *pose7, int *grounded, float *vvel);

/* W6 numerics (bounce_lcp.rs), for the pin tests: K symmetric CSR, n <= 250.  min_degree writes
 * the elimination order and returns the factor's entries below the diagonal; pseudo_solve solves
 * (K + CFM diag K) x = u; bounce_qp solves the push-only LCP and returns 1 if converged. */
uint32_t or_bn_min_degree(uint32_t n, const uint32_t *ptr, const uint32_t *idx, uint32_t *order);
int or_bn_pseudo_solve(uint32_t n, const uint32_t *ptr, const uint32_t *idx, const double *val, const double *u, double *x);
int or_bn_bounce_qp(uint32_t n, const uint32_t *ptr, const uint32_t *idx, const double *val, const double *b, const double *acc0,
                    double *x, uint32_t *solves);
float or_toi_bounce(float s0, float v0, float a, float e, float dt, float *pseudo, int *has_pseudo);

#ifdef __cplusplus
}
#endif
#endif
