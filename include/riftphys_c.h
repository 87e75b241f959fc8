/* riftphys_c.h — C ABI of the MI355X step library (librphys_amd.so).
 *
 * Every declaration in the first part keeps the name, signature, layout and error behaviour of
 * the reference header crates/riftphys-ffi/src/riftphys_c.h (cited per entry, header lines "h:N"
 * and implementation lines "lib.rs:N" of crates/riftphys-ffi/src/lib.rs), so a host written
 * against the reference links against this library unchanged.  The second part holds additive
 * exports (ABI version stays 1; no v1 layout changes).
 */
#ifndef RIFTPHYS_C_H
#define RIFTPHYS_C_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define RPHYS_ABI_VERSION 1u
uint32_t rphys_get_abi_version(void);                 /* h:13, lib.rs:84-94 */
const char *rphys_get_build_id(void);                 /* h:14 */

typedef uint32_t RPhysBool;
#define RPHYS_FALSE 0u
#define RPHYS_TRUE 1u

typedef uint32_t RPhysResult;                         /* h:22-28 */
#define RPHYS_OK 0u
#define RPHYS_ERR_NULL 1u
#define RPHYS_ERR_INVALID_ID 2u
#define RPHYS_ERR_CAPACITY 3u
#define RPHYS_ERR_BAD_ARG 4u
#define RPHYS_ERR_INTERNAL 0x80000000u

#define RPHYS_INVALID_BODY_ID 0u                      /* h:30-35 */
#define RPHYS_TERRAIN_BODY_ID 0xFFFFFFFFu

const char *rphys_last_error_message(void);           /* h:38, lib.rs:56-82 (thread-local) */

typedef struct RPhysWorld RPhysWorld;

typedef struct RPhysVec3 { float x, y, z; } RPhysVec3;             /* h:46-60 */
typedef struct RPhysQuat { float x, y, z, w; } RPhysQuat;
typedef struct RPhysIsometry { RPhysVec3 pos; RPhysQuat rot; } RPhysIsometry;
typedef struct RPhysVelocity { RPhysVec3 lin; RPhysVec3 ang; } RPhysVelocity;
typedef struct RPhysMaterial { float mu_static; float mu_dynamic; float restitution; } RPhysMaterial;

typedef uint32_t RPhysBodyType;                       /* h:66-68 */
#define RPHYS_BODY_STATIC 0u
#define RPHYS_BODY_DYNAMIC 1u

typedef struct RPhysCapsuleBodyDesc {                 /* h:70-79 */
    RPhysIsometry pose; RPhysVelocity vel; float radius; float half_height; float mass;
    RPhysBodyType body_type; RPhysMaterial material; uint32_t user_tag;
} RPhysCapsuleBodyDesc;
typedef struct RPhysBoxBodyDesc {                     /* h:81-89 */
    RPhysIsometry pose; RPhysVelocity vel; float hx, hy, hz; float mass;
    RPhysBodyType body_type; RPhysMaterial material; uint32_t user_tag;
} RPhysBoxBodyDesc;
typedef struct RPhysSphereBodyDesc {                  /* h:91-99 */
    RPhysIsometry pose; RPhysVelocity vel; float radius; float mass;
    RPhysBodyType body_type; RPhysMaterial material; uint32_t user_tag;
} RPhysSphereBodyDesc;

#define RPHYS_STEP_HASH_BYTES 32
typedef struct RPhysHash32 { uint8_t bytes[RPHYS_STEP_HASH_BYTES]; } RPhysHash32;  /* h:103-107 */

typedef struct RPhysStepParams {                      /* h:111-117 */
    uint32_t size_bytes; uint32_t flags; uint32_t substeps; uint32_t solver_iters; float dt;
} RPhysStepParams;

typedef struct RPhysRayQuery { RPhysVec3 origin; RPhysVec3 dir; float max_distance; uint32_t ignore_body; } RPhysRayQuery;
typedef struct RPhysRayHit { RPhysBool hit; uint32_t body_id; float fraction; RPhysVec3 point; RPhysVec3 normal; } RPhysRayHit;
typedef struct RPhysCapsuleSweepQuery { RPhysVec3 from; RPhysVec3 to; float radius; float half_height; uint32_t ignore_body; } RPhysCapsuleSweepQuery;
typedef struct RPhysCapsuleSweepHit { RPhysBool hit; uint32_t body_id; float fraction; RPhysBool started_overlapping; RPhysVec3 point; RPhysVec3 normal; } RPhysCapsuleSweepHit;

typedef struct RPhysHeightfieldDesc {                 /* h:155-166 */
    uint32_t size_bytes; uint32_t flags; uint32_t width; uint32_t height; uint32_t row_stride;
    float cell_size_x; float cell_size_z; float height_scale; float height_offset;
    const int16_t *heights;
} RPhysHeightfieldDesc;

/* h:168-172, lib.rs:1419-1493 */
RPhysResult rphys_world_set_heightfield_i16(RPhysWorld *world, const RPhysHeightfieldDesc *hf, float origin_x, float origin_z, float y_offset);
/* h:174-182, lib.rs:1495-1549 */
RPhysResult rphys_world_set_heightfield_raw32_square(RPhysWorld *world, const uint8_t *bytes, uint32_t bytes_len, float world_size_x,
                                                     float world_size_z, float origin_x, float origin_z, float y_offset);
RPhysResult rphys_world_clear_heightfield(RPhysWorld *world);          /* h:184, lib.rs:1551-1567 */

RPhysWorld *rphys_world_create(uint32_t max_bodies, uint32_t max_colliders);  /* h:189, lib.rs:608-631 */
void rphys_world_destroy(RPhysWorld *world);                         /* h:190, lib.rs:633-639 */
RPhysResult rphys_world_set_epoch(RPhysWorld *world, uint64_t epoch);    /* h:191 */
RPhysResult rphys_world_set_rng_seed(RPhysWorld *world, uint64_t seed);  /* h:192 */

/* h:195-198, lib.rs:677-927 (1-based ids, density of MaterialId::Default) */
RPhysResult rphys_add_capsule_body(RPhysWorld *world, const RPhysCapsuleBodyDesc *desc, uint32_t *out_body_id);
RPhysResult rphys_add_box_body(RPhysWorld *world, const RPhysBoxBodyDesc *desc, uint32_t *out_body_id);
RPhysResult rphys_add_sphere_body(RPhysWorld *world, const RPhysSphereBodyDesc *desc, uint32_t *out_body_id);
RPhysResult rphys_body_remove(RPhysWorld *world, uint32_t body_id);

/* h:201-208, lib.rs:929-1101 */
RPhysResult rphys_body_get_pose(const RPhysWorld *world, uint32_t body_id, RPhysIsometry *out_pose);
RPhysResult rphys_body_set_pose(RPhysWorld *world, uint32_t body_id, const RPhysIsometry *pose);
RPhysResult rphys_body_get_velocity(const RPhysWorld *world, uint32_t body_id, RPhysVelocity *out_vel);
RPhysResult rphys_body_set_velocity(RPhysWorld *world, uint32_t body_id, const RPhysVelocity *vel);
RPhysResult rphys_body_apply_impulse(RPhysWorld *world, uint32_t body_id, RPhysVec3 impulse);

/* h:211-212, lib.rs:1103-1164: every substep is one device tick */
RPhysResult rphys_world_step_ex(RPhysWorld *world, const RPhysStepParams *p);
RPhysResult rphys_world_step(RPhysWorld *world, float dt);

/* h:215-218, lib.rs:1166-1266: the hash is computed on the device */
RPhysResult rphys_world_tick_index(const RPhysWorld *world, uint64_t *out_tick);
RPhysResult rphys_world_num_bodies(const RPhysWorld *world, uint32_t *out_num);
RPhysResult rphys_world_get_step_hash(const RPhysWorld *world, RPhysHash32 *out_hash);
RPhysResult rphys_world_step_and_hash(RPhysWorld *world, float dt, RPhysHash32 *out_hash);

/* h:221-222, lib.rs:1268-1410: World::raycast (collider bounding boxes, terrain triangles) and
 * World::sweep_capsule (queries.rs:50-261), answered on the device from the world's current state */
RPhysResult rphys_world_raycast(const RPhysWorld *world, const RPhysRayQuery *query, RPhysRayHit *out_hit);
RPhysResult rphys_world_capsule_sweep(const RPhysWorld *world, const RPhysCapsuleSweepQuery *query, RPhysCapsuleSweepHit *out_hit);
/* Batched queries (build addition): n rays or capsule sweeps answered in one device launch, each as
 * rphys_world_raycast / rphys_world_capsule_sweep would (same checks; the first invalid query's
 * error is returned with every output a miss). */
RPhysResult rphys_world_raycast_n(const RPhysWorld *world, const RPhysRayQuery *queries, uint32_t n, RPhysRayHit *out_hits);
RPhysResult rphys_world_capsule_sweep_n(const RPhysWorld *world, const RPhysCapsuleSweepQuery *queries, uint32_t n,
                                        RPhysCapsuleSweepHit *out_hits);

RPhysResult rphys_last_error_copy(char *dst, uint32_t cap, uint32_t *out_len);   /* h:225, lib.rs:561-604 */

/* h:227-243, lib.rs:1580-1693: the kinematic player controller (World::add_player,
 * World::update_players; the players move on the device as the first stage of every tick).
 * rphys_add_player validates its arguments as the reference does (out index set to 0xFFFFFFFF
 * first; null -> ERR_NULL; non-finite pose, radius <= 0 or height <= 0 -> ERR_BAD_ARG) and adds
 * the player's body (infinite mass, not dynamic, default capsule collider) and controller; players are indexed
 * 0, 1, ... per world.  set_input stores the move direction (the tick normalises it when its
 * squared length exceeds 1e-6, else uses zero); get_pose returns the controller's pose and grounded flag.  Both answer an
 * unknown index with ERR_BAD_ARG "player_idx out of bounds", the identity pose and RPHYS_FALSE
 * written first. */
typedef struct RPhysPlayerDesc { RPhysIsometry start_pose; float radius; float height; float speed; } RPhysPlayerDesc;
RPhysResult rphys_add_player(RPhysWorld *world, const RPhysPlayerDesc *desc, uint32_t *out_player_idx);
RPhysResult rphys_player_set_input(RPhysWorld *world, uint32_t player_idx, RPhysVec3 move_dir);
RPhysResult rphys_player_get_pose(const RPhysWorld *world, uint32_t player_idx, RPhysIsometry *out_pose, RPhysBool *out_grounded);

/* lib.rs:107-115: the reference uses this type in rphys_world_get_contacts but its header never
 * defines it; 36 bytes, fields in the reference's order. */
typedef struct RPhysContactEvent {
    uint32_t body_id_a; uint32_t body_id_b; RPhysVec3 position; RPhysVec3 normal; float impulse;
} RPhysContactEvent;
/* h:248-253, lib.rs:1801-1878: contacts of the last tick (count only when out_buffer is NULL) */
RPhysResult rphys_world_get_contacts(const RPhysWorld *world, RPhysContactEvent *out_buffer, uint32_t buffer_capacity, uint32_t *out_count);

typedef struct RPhysWorldSnapshot RPhysWorldSnapshot;                    /* h:256-260, lib.rs:1696-1800 */
RPhysWorldSnapshot *rphys_world_snapshot_create(const RPhysWorld *world);
RPhysResult rphys_world_snapshot_restore(RPhysWorld *world, const RPhysWorldSnapshot *snapshot);
void rphys_world_snapshot_free(RPhysWorldSnapshot *snapshot);

/* ============================ additive exports (no counterpart in the reference ABI) ========== */

/* RPhysStepParams.flags bit: take solver_iters (> 0) as the velocity iterations of this call.
 * Without it solver_iters is ignored, as in the reference (lib.rs:1147). */
#define RPHYS_STEP_FLAG_SOLVER_ITERS 1u

typedef uint32_t RPhysShape;
#define RPHYS_SHAPE_SPHERE 0u
#define RPHYS_SHAPE_BOX 1u
#define RPHYS_SHAPE_CAPSULE 2u

/* A body with an explicit density in kg/m^3 (<= 0: the reference default 1500), the only way
 * to express other densities (mass.rs:92-165 takes density; the v1 descs' mass is ignored).
 * dims: sphere {r}, box {hx,hy,hz}, capsule {r, half_height}. */
RPhysResult rphys_add_body_density(RPhysWorld *world, RPhysShape shape, const float dims[3], const RPhysIsometry *pose,
                                   const RPhysVelocity *vel, RPhysBodyType body_type, float density, uint32_t *out_body_id);
/* World::set_pair_props (world.rs:975-979): friction/restitution of a material-id pair.
 * FFI bodies use MaterialId::Default (0); heightfields use Grit (4). */
/* Ball joint between two live bodies (additive: the reference's C ABI has no joint export; it
 * replaces World::add_ball_joint, crates/riftphys-world/src/world/world.rs:718, i.e.
 * crates/riftphys-articulation/src/bodies.rs:326-331 add_ball: the three translational axes of
 * frame A locked, rotation free).  frame_a / frame_b are the joint frames in body A's / B's space.
 * Solved as the reference solves a ball joint: three velocity rows per joint ahead of the contacts
 * in every PGS sweep (solver.rs:336-424, 644-715, 788-835, 945-952) and SHAKE position rows ahead
 * of every contact projection pass (articulation/src/bodies.rs:616-739, step.rs:284-310); bodies
 * joined directly do not collide (step.rs:85-93).  *out_joint_id receives the world's joint index
 * (0-based, creation order, never reused; 0xFFFFFFFF on failure).  Errors: null -> ERR_NULL;
 * unknown/removed/terrain body -> ERR_INVALID_ID; body_a == body_b or a non-finite frame ->
 * ERR_BAD_ARG.  A removed body's joints stay, attached to a static body, as in the reference. */
RPhysResult rphys_world_add_ball_joint(RPhysWorld *world, uint32_t body_a, uint32_t body_b, const RPhysIsometry *frame_a,
                                       const RPhysIsometry *frame_b, uint32_t *out_joint_id);
/* replaces World::add_d6_joint / add_hinge_joint (crates/riftphys-world/src/world/world.rs:719-720;
 * crates/riftphys-articulation/src/bodies.rs:321-338 add_d6 / add_hinge): a D6 joint whose
 * translational axes (t_mask) and rotational axes (r_mask) of frame A are constrained (bit i: axis
 * i), with a stored hinge axis (0..2, or -1: none, derived each step as bodies.rs:895-906 does).
 * Limits, drives, damping, friction and compliance per axis are set by the calls below.  Solved as
 * the reference solves a D6 joint with no armature: velocity rows of dampers, friction, locked axes,
 * velocity motors and speculative limits (solver.rs:336-424), warm-started by row kind, and the
 * position solve of solve_d6 (bodies.rs:616-873: rows of every constrained axis, the hinge's twist
 * in radians, limits, position drives, XPBD compliance).  Islands holding a D6 joint are solved by
 * the workgroup-per-island tier.  Errors as rphys_world_add_ball_joint, and ERR_BAD_ARG for a mask
 * above 7 or a hinge axis outside -1..2. */
RPhysResult rphys_world_add_d6_joint(RPhysWorld *world, uint32_t body_a, uint32_t body_b, const RPhysIsometry *frame_a,
                                     const RPhysIsometry *frame_b, uint32_t t_mask, uint32_t r_mask, int32_t hinge_axis,
                                     uint32_t *out_joint_id);
/* replaces World::add_distance_joint (crates/riftphys-world/src/world/world.rs:537-539;
 * articulation/src/bodies.rs:273-285): the two bodies' centres held `rest` apart (a non-finite
 * rest becomes 0), rigid (compliance 0: the legacy Baumgarte velocity row of solver.rs:795-820 in
 * every velocity sweep, not warm-started, and the XPBD position row of bodies.rs:501-548) or
 * compliant (compliance m/N: the position row only; negative or non-finite becomes 0).  Islands
 * holding one are solved by the workgroup-per-island tier.  Errors as rphys_world_add_ball_joint. */
RPhysResult rphys_world_add_distance_joint(RPhysWorld *world, uint32_t body_a, uint32_t body_b, float rest, float compliance,
                                           uint32_t *out_joint_id);
/* bodies.rs:333-338 add_hinge: all translational axes and the two rotational axes other than
 * `axis` locked, `axis` stored as the hinge axis.  Errors: axis above 2 -> ERR_BAD_ARG. */
RPhysResult rphys_world_add_hinge_joint(RPhysWorld *world, uint32_t body_a, uint32_t body_b, const RPhysIsometry *frame_a,
                                        const RPhysIsometry *frame_b, uint32_t axis, uint32_t *out_joint_id);
/* replaces World::set_d6_limit_r and JointSet::set_d6_limit_t (world.rs:968; bodies.rs:361-390):
 * on = RPHYS_FALSE removes axis `axis`'s limit; a limit [lo, hi] (m, or rad of the axis's twist
 * angle) constrains its axis, and on a rotational axis other than the stored hinge axis forgets
 * the stored axis (bodies.rs:878-880).  Errors: unknown joint -> ERR_INVALID_ID; a ball joint, an
 * axis above 2 or a non-finite bound -> ERR_BAD_ARG. */
RPhysResult rphys_world_set_joint_limit(RPhysWorld *world, uint32_t joint, RPhysBool angular, uint32_t axis, RPhysBool on, float lo, float hi);
/* replaces World::set_d6_drive_r / set_d6_drive_t (world.rs:966-967; bodies.rs:347-359): a drive
 * toward target constrains its axis; on = RPHYS_FALSE removes it.  A position drive (m or rad) is
 * a position row toward its target (rigid, or with the axis's compliance); a velocity drive (m/s or
 * rad/s) is a motor row whose impulse per step is capped at |max_impulse|.  kp and kd are stored as
 * the reference's Drive stores them (bodies.rs:65-71); neither solve reads them.
 * Errors as rphys_world_set_joint_limit (a non-finite drive value -> ERR_BAD_ARG). */
RPhysResult rphys_world_set_joint_drive(RPhysWorld *world, uint32_t joint, RPhysBool angular, uint32_t axis, RPhysBool on, float target,
                                        float kp, float kd, float max_impulse, RPhysBool velocity_mode);
/* replaces World::set_d6_damping (world.rs:971; bodies.rs:370-378): viscous damping (N s/m or
 * N m s/rad) and Coulomb friction (N or N m) on one axis; negative or non-finite values become 0;
 * neither constrains the axis.  Errors as rphys_world_set_joint_limit. */
RPhysResult rphys_world_set_joint_damping(RPhysWorld *world, uint32_t joint, RPhysBool angular, uint32_t axis, float damping, float friction);
/* D6Axis::compliance (bodies.rs:52-104) of one axis: 0 rigid, else XPBD compliance (m/N or
 * rad/(N m)) of its position row.  Errors as rphys_world_set_joint_limit. */
RPhysResult rphys_world_set_joint_compliance(RPhysWorld *world, uint32_t joint, RPhysBool angular, uint32_t axis, float compliance);
uint32_t rphys_world_num_joints(const RPhysWorld *world);
RPhysResult rphys_world_set_pair_props(RPhysWorld *world, uint32_t mat_a, uint32_t mat_b, float mu_s, float mu_k,
                                       float restitution, float stribeck_v);
/* World::set_pair_props' contact model (materials.rs:75-90 MatPairProps stiffness, dissipation,
 * patch_friction; the pair's friction and restitution stay): stiffness > 0 (N/m per contact point)
 * makes the pair compliant, an implicit Hunt-Crossley row with that dissipation (s/m) in place of
 * the rigid normal row, with no restitution, shock, projection or warm start (solver.rs:545-567,
 * 848-856); patch_friction solves each manifold's friction as one patch at the points' centroid,
 * sliding and torsion limited by the manifold's normal impulse (solver.rs:592-619, 882-915).  Like
 * the other pair properties, a batch uses its first world's.  0 and RPHYS_FALSE restore the rigid
 * model.  Errors: ids above 15 or a negative or non-finite stiffness, non-finite dissipation ->
 * ERR_BAD_ARG. */
RPhysResult rphys_world_set_pair_contact_model(RPhysWorld *world, uint32_t mat_a, uint32_t mat_b, float stiffness, float dissipation,
                                               RPhysBool patch_friction);
/* World::set_compliant_damping_at_start (world.rs:992): the compliant rows' damping linearised at the
 * step-start penetration instead of the predicted mid-step one (default off; a batch uses its
 * first world's). */
RPhysResult rphys_world_set_compliant_damping_at_start(RPhysWorld *world, RPhysBool on);
/* World::set_solver_iterations (world.rs:994-996), (velocity, position), default (20, 20) */
RPhysResult rphys_world_set_solver_iterations(RPhysWorld *world, uint32_t velocity, uint32_t position);
/* Residual mode (build addition; the reference runs a fixed count): each island (component of the
 * bodies the solve writes) runs PGS sweeps until its largest row change
 * max(|dJn|/m_n/(1+|bias|), |dJt1|/m_t1, |dJt2|/m_t2) <= tol (m/s; the sweep criterion of
 * solver.rs:1149), at most max_iters.  on = 0 returns to the fixed velocity iterations. */
RPhysResult rphys_world_set_residual_mode(RPhysWorld *world, uint32_t on, float tol, uint32_t max_iters);
RPhysResult rphys_world_set_gravity(RPhysWorld *world, RPhysVec3 g);          /* World::set_gravity (world.rs:560-563) */
RPhysResult rphys_body_set_sleep_allowed(RPhysWorld *world, uint32_t body_id, RPhysBool allowed);  /* world.rs:565-570 */
/* World::add_collider's material (world.rs:614-622, materials.rs:11-12 core ids 0..15): the pair
 * table entry (friction, restitution) the body's contacts take.  The reference's FFI gives every
 * body MaterialId::Default (lib.rs:544-557); this sets the id a World-level host would pass. */
RPhysResult rphys_body_set_material_id(RPhysWorld *world, uint32_t body_id, uint32_t material_id);
/* n ticks of dt enqueued back to back with one host synchronisation at the end */
RPhysResult rphys_world_step_n(RPhysWorld *world, float dt, uint32_t n_ticks);

typedef struct RPhysStepStats {
    uint32_t pairs;              /* broadphase pairs (StepStats.pairs_tested) */
    uint32_t contacts;           /* contacts after cull (StepStats.contacts) */
    uint32_t terrain_candidates; /* (collider, triangle) tests */
    uint32_t island_contacts;    /* contacts writing at least one body */
    uint32_t islands;
    uint32_t warm_entries;       /* warm-start cache entries after the tick */
    uint32_t pgs_iters;          /* PGS sweeps run by the island that ran the most */
    uint32_t pgs_iters_sum;      /* sum over islands of (sweeps x the island's contacts) */
    /* W6 restitution over the ticks of the last step call (solver.rs:1009-1245): time-of-impact
     * bounce rows, their support rows, pseudo rows, and components swept instead of solved
     * directly */
    uint32_t bounce_rows, bounce_support_rows, bounce_pseudo_rows, bounce_fallbacks;
    uint32_t relax_rows;         /* contacts on jointed bodies bounced by the relax pass (solver.rs:1001-1008) */
} RPhysStepStats;
RPhysResult rphys_world_get_stats(const RPhysWorld *world, RPhysStepStats *out);

/* Profiled stepping (build addition; the reference has no counterpart): runs n ticks on the
 * library's stream with HIP events at the stage boundaries of every tick and returns the stage
 * times and the tick counters summed over the ticks.  Stage k's name comes from
 * rphys_profile_stage_name(k).  Same results as rphys_world_step_n. */
#define RPHYS_PROFILE_MAX_STAGES 24
typedef struct RPhysProfile {
    uint32_t ticks;
    uint32_t n_stages;
    /* summed per stage; the small-island stages (velocity and projection) run alongside the
     * larger islands' stages on a stream of their own, so the stages can add up to more than the
     * ticks took */
    double stage_ms[RPHYS_PROFILE_MAX_STAGES];
    double pairs, terrain_candidates, contacts, island_contacts, islands;
    double island_contacts_global; /* contacts of the islands solved by the 1024-thread LDS-flag tier */
    double pgs_iters_max, pgs_iters_sum;  /* per-tick RPhysStepStats.pgs_iters / pgs_iters_sum, summed */
    /* Inside the sweeps of the islands larger than 4,096 contacts (summed over their workgroups):
     * ms in [0] velocity level order, [1] warm start, [2] PGS iterations, [3] velocity shock order,
     * [4] velocity shock pass, [5] ballistic pass, [6] projection iterations, [7] projection shock
     * order, [8] projection shock pass; [9] / [10] levels of the list / shock orders, summed. */
    double sweep_phase[11];
    /* Residual mode: islands, and their contacts, that reached the tolerance before the sweep
     * limit (summed per tick).  Learned warm start: contacts that passed the model's gate. */
    double islands_converged, island_contacts_converged, warm_start_learned;
    /* Leading levels of at least 1,024 rows of the largest island, list order and shock order,
     * summed per tick: the levels that run as GPU-wide launches (fixed iteration counts) */
    double wide_levels, wide_levels_shock;
    /* Residual mode: sweeps x entries of the islands that reached the tolerance, summed per tick
     * (over island_contacts_converged: the mean sweeps of the islands that converge) */
    double pgs_iters_sum_converged;
} RPhysProfile;
RPhysResult rphys_world_step_profiled(RPhysWorld *world, float dt, uint32_t n_ticks, RPhysProfile *out);
/* Bandwidth probe (build addition): copies `bytes` between two fresh device buffers `reps` times
 * with a plain copy kernel and returns read + write bytes per second in GB/s.  A reference for
 * the step's roofline figures and for calibrating HBM traffic counters. */
RPhysResult rphys_debug_copy_probe(uint64_t bytes, uint32_t reps, double *gbps);
const char *rphys_profile_stage_name(uint32_t stage);

/* One contact with its build-defined feature id (kind<<24 | which<<16 | k, bit 23 = swapped
 * shapes) and terrain cell ((gz*(nx-1)+gx)*2+tri; 0xFFFFFFFF for body pairs).  Colliders are
 * 0-based collider indices (the reference's Contact.a_collider / b_collider). */
typedef struct RPhysContactEx {
    uint32_t a_collider, b_collider;
    float point[3];
    float normal[3];
    float depth;
    float impulse;
    uint32_t feature;
    uint32_t cell;
} RPhysContactEx;
RPhysResult rphys_debug_get_contacts_ex(const RPhysWorld *world, RPhysContactEx *out, uint32_t cap, uint32_t *out_count);
/* Broadphase pairs (collider indices i < j) of the last tick, in the reference's sorted order. */
RPhysResult rphys_debug_get_pairs(const RPhysWorld *world, uint32_t *out_ij, uint32_t cap_pairs, uint32_t *out_count);

/* The bytes the world's step hash digests (its BLAKE3 input, world.rs:882-909): up to cap of them
 * into out; *out_len receives their count.  For locating where two hashes of one state differ.
 * Errors: null world or out_len -> ERR_NULL. */
RPhysResult rphys_debug_get_hash_input(const RPhysWorld *world, uint8_t *out, uint64_t cap, uint64_t *out_len);

/* Joint `joint`'s stored velocity rows (what next step's warm start reads, world.rs:896-900): up to
 * cap (tag, impulse) pairs in row order, tag = angular << 5 | axis << 3 | kind (0 locked, 1 motor,
 * 2 / 3 lower / upper limit, 4 damper, 5 friction); *out_n receives the row count (0: none stored).
 * Errors: null world or out_n -> ERR_NULL; unknown joint -> ERR_INVALID_ID. */
RPhysResult rphys_debug_get_joint_rows(const RPhysWorld *world, uint32_t joint, uint8_t *tags, float *impulses, uint32_t cap, uint32_t *out_n);

/* Full carried state of a world: per body pos/rot/lin/ang/last linear velocity/awake/sleep
 * ticks, and the warm-start cache (keys (body_a << 32 | body_b or 0xFFFFFFFF), body-A-frame
 * anchors, normal impulses, world friction impulses), for checkpointing and replay tools.
 * Arrays are per body (3 or 4 floats each); warm arrays hold up to warm_cap entries. */
RPhysResult rphys_debug_export_state(const RPhysWorld *world, float *pos3, float *rot4, float *lin3, float *ang3, float *last_lin3,
                                     uint32_t *awake, uint32_t *sleep_ticks, uint64_t *warm_keys, float *warm_anchor3,
                                     float *warm_jn, float *warm_jt3, uint32_t warm_cap, uint32_t *warm_count);

/* ---- batches: many independent worlds stepped by one launch sequence ---- */
typedef struct RPhysBatch RPhysBatch;
/* Builds a batch from n_worlds existing worlds (their bodies, heightfields, settings and carried
 * state are copied; the worlds stay usable).  Settings (iterations, gravity, pair props) are
 * taken from world 0. */
RPhysBatch *rphys_batch_create(RPhysWorld *const *worlds, uint32_t n_worlds);
void rphys_batch_destroy(RPhysBatch *batch);
RPhysResult rphys_batch_step_n(RPhysBatch *batch, float dt, uint32_t n_ticks);
/* Copies world wi of the batch (bodies, settings and carried state as the batch has stepped them,
 * tick included) into dst, replacing dst's contents; dst then steps on its own from there. */
RPhysResult rphys_batch_get_world(RPhysBatch *batch, uint32_t world_index, RPhysWorld *dst);
/* out: 32 * n_worlds bytes */
RPhysResult rphys_batch_get_hashes(const RPhysBatch *batch, uint8_t *out);
RPhysResult rphys_batch_get_state(const RPhysBatch *batch, uint32_t world_index, float *pos3, float *rot4, float *lin3, float *ang3);
RPhysResult rphys_batch_snapshot(RPhysBatch *batch, uint32_t slot);   /* device-side copy of all worlds' state */
RPhysResult rphys_batch_restore(RPhysBatch *batch, uint32_t slot);
/* Launch-shape tuning (build addition; the reference has no GPU launches to tune).  The island
 * sweeps run one workgroup per island, taken grid-stride, in tiers by island size; the grids
 * only change speed, never results.  rphys_batch_autotune grid-searches grid_small in
 * {1024..32768}, grid_large in {128..2048}, grid_huge in {64..256} (powers of two): each candidate steps the same
 * `ticks` graph-captured ticks from a device snapshot, per-world hashes must agree across all
 * candidates (else RPHYS_ERR_INTERNAL); the four fastest and the default shape run three more
 * times, the best of each shape's runs decides, the winner is installed and the batch is left at
 * the snapshot.  Caching results per (config, size) is the caller's (rphys.py keeps a JSON file). */
typedef struct RPhysTuning {
    uint32_t size_bytes;  /* sizeof(RPhysTuning) */
    uint32_t grid_small;  /* islands of <= 256 entries: one-wave workgroups (default 8192); islands of
                           * 257..4,096 entries use grid_small / 4 four-wave workgroups */
    uint32_t grid_large;  /* <= 131,072 contacts: 1024 threads, LDS flags (default 256) */
    uint32_t grid_huge;   /* larger islands: 1024 threads, HBM flags (default 64) */
} RPhysTuning;
RPhysResult rphys_batch_set_tuning(RPhysBatch *batch, const RPhysTuning *tuning);
RPhysResult rphys_batch_get_tuning(const RPhysBatch *batch, RPhysTuning *tuning);
RPhysResult rphys_batch_autotune(RPhysBatch *batch, float dt, uint32_t ticks, RPhysTuning *best, double *best_ms_per_tick);
RPhysResult rphys_batch_total_bodies(const RPhysBatch *batch, uint32_t *out);
/* Record the next tick of the batch into a HIP graph and replay it for n ticks (same results as
 * rphys_batch_step_n; launch overhead paid once). */
RPhysResult rphys_batch_step_profiled(RPhysBatch *batch, float dt, uint32_t n_ticks, RPhysProfile *out); /* as rphys_world_step_profiled */
RPhysResult rphys_batch_step_graph(RPhysBatch *batch, float dt, uint32_t n_ticks);

/* ---- Learned PGS warm start (build addition; the reference has no counterpart) ----
 * A 12 -> 64 -> 64 -> 3 ReLU MLP (BF16 weights and activations, FP32 accumulation on MFMA) maps
 * per-contact features (normal/tangential relative velocity, penetration, inverse masses, four
 * angular effective-mass terms, the previous frame's matched normal impulse, mu_s) to a correction
 * of the previous-frame warm start: converged (normal, tangent 1, tangent 2) impulses minus (matched
 * jn, matched jt projected on the contact's tangents; zeros without a match).  Parameter vector (f32, rphys_mlp_param_count()
 * entries): W1[64][12] b1[64] W2[64][64] b2[64] W3[3][64] b3[3] x_mean[12] x_std[12] y_scale[3];
 * inputs are normalised (x - x_mean) / x_std, outputs scaled by y_scale.
 * Warm-start mode: 0 the previous frame's impulses (the reference, default), 1 zero, 2 the
 * previous frame's plus the model's correction (jn clamped at 0, jt to the friction cone as a
 * cached impulse would be) for every solved contact whose normalised inputs and outputs all lie
 * within +-8; other contacts get mode 0's warm start exactly, so a zero model is mode 0.  The
 * solver's iterations are unchanged: with residual mode on, only the number of sweeps to the
 * residual depends on the warm start. */
uint32_t rphys_mlp_param_count(void);
uint32_t rphys_mlp_gradient_rows(void);  /* rows of one training step (rphys_mlp_gradient) */
RPhysResult rphys_world_set_warm_start(RPhysWorld *world, uint32_t mode, const float *params);
RPhysResult rphys_batch_set_warm_start(RPhysBatch *batch, uint32_t mode, const float *params);
RPhysResult rphys_batch_set_residual_mode(RPhysBatch *batch, uint32_t on, float tol, uint32_t max_iters);  /* as the world call, every world */
/* Contact log: while enabled, every tick appends its solved contacts (island-list order) as
 * features [12] and targets [3]: converged impulses (normal before the shock pass, friction) minus
 * the previous-frame warm start as defined above, until full. */
RPhysResult rphys_batch_set_contact_log(RPhysBatch *batch, uint32_t capacity_rows);  /* 0 disables */
RPhysResult rphys_batch_contact_log_rows(RPhysBatch *batch, uint32_t *rows, uint32_t *dropped);
RPhysResult rphys_batch_read_contact_log(RPhysBatch *batch, uint32_t first, uint32_t count, float *feat12, float *targ3);
typedef struct RPhysMlpTrainConfig {
    uint32_t size_bytes;  /* sizeof(RPhysMlpTrainConfig) */
    uint32_t epochs;
    float lr;             /* Adam (beta1 0.9, beta2 0.999, eps 1e-8) */
    uint32_t reserved;
    uint64_t seed;        /* weight initialisation and the per-epoch shuffles */
} RPhysMlpTrainConfig;
/* Trains on the batch's contact log on the device (mean squared error in units of y_scale, steps
 * of rphys_mlp_gradient_rows() contacts); writes the parameters and each epoch's mean loss. */
RPhysResult rphys_batch_train_warm_start(RPhysBatch *batch, const RPhysMlpTrainConfig *cfg, float *params_out, float *epoch_loss);
/* Device inference on host features (n rows of 12) -> n rows of 3 (checking hook). */
RPhysResult rphys_mlp_forward(const float *params, const float *feat12, uint32_t n, float *out3);
/* No reference counterpart: rphys_mlp_forward for callers that cannot allocate device memory
 * (code under HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).  params
 * (RP_MLP_NPARAM floats, a valid model), feat12 (n rows of 12) and out4 (n rows of 4: the three
 * predictions, then 1 where the row passes the model's gate, else 0) are device pointers on the
 * current device; n <= 2^30 (else RPHYS_ERR_BAD_ARG).  The kernels run on `stream` (a
 * hipStream_t; NULL: the null stream) and the call returns once they are enqueued.  The bf16
 * weight tiles go to a fixed 26,624-byte device array the library holds per GPU until the process
 * ends: no call allocates.  Calls that overlap on one GPU share that array and race; serialise them
 * (one stream, or events between them). */
RPhysResult rphys_mlp_forward_noalloc(const float *params, const float *feat12, uint32_t n, float *out4, void *stream);
/* Test plumbing (no reference counterpart): device buffers and blocking copies (to_host 0:
 * host -> device, 1: device -> host) for the device-pointer entry points above. */
RPhysResult rphys_debug_device_alloc(uint64_t bytes, void **out);
RPhysResult rphys_debug_device_free(void *ptr);
RPhysResult rphys_debug_device_copy(void *dst, const void *src, uint64_t bytes, uint32_t to_host);
/* One training step's loss and gradient (the first 5,187 parameters: weights and biases) over
 * rphys_mlp_gradient_rows() given rows, without updating (checking hook). */
RPhysResult rphys_mlp_gradient(const float *params, const float *feat12, const float *targ3, float *loss, float *grad);

/* Test hooks (no reference counterpart), process-wide, 0 = default.
 * rphys_debug_set_sweep_tiers: the largest island solved by sweep tiers 0 (one wave, default 256
 * entries), 3 (one wave, 4,096) and 1 (1024 threads with LDS flags, 131,072), larger islands by
 * tier 2 (flags in HBM); and the width from which an island's leading level runs as a GPU-wide
 * launch (default 1,024).  Bounds are clamped to each tier's flag storage, so islands only move
 * to wider tiers; every setting gives the same results.
 * rphys_debug_set_initial_capacity: pair, terrain-candidate, contact and warm-cache capacities
 * of scenes built afterwards (the warm cache still holds at least its entries + 1,024), so the
 * overflow-and-replay path runs on small scenes. */
RPhysResult rphys_debug_set_sweep_tiers(uint32_t max_tier0, uint32_t max_tier3, uint32_t max_tier1, uint32_t wide_min);
RPhysResult rphys_debug_set_initial_capacity(uint32_t pairs, uint32_t cands, uint32_t contacts, uint32_t warm);
/* Test hook (no reference counterpart), process-wide: the next n device allocations the library
 * makes for scenes and snapshot slots fail as out of memory (0 clears), so a host can force the
 * device-failure path: the call that needed them returns RPHYS_ERR_INTERNAL with a message, and
 * the world or batch keeps its last good state.  Returns the allocations still to fail through
 * out_pending (may be null). */
RPhysResult rphys_debug_fail_device_allocs(uint32_t n, uint32_t *out_pending);
/* Test hook (no reference counterpart), process-wide: how sweep tier 0 solves its islands'
 * velocities.  Islands of up to lane_max entries one island per lane (0 = default, 150;
 * RPHYS_LANE_NONE: none, and tier 0 solves every island a wave each), then up to pair_max by a
 * pair of lanes (at most 128), up to quad_max by a group of four lanes, the larger ones a wave
 * each (0 = default: 150 for both, that is none); but at least the wave_min largest islands a wave
 * each (0 = default, 2,048; RPHYS_LANE_NONE: none).  Bounds are clamped to tier 0's and kept in
 * order; every setting gives the same results. */
#define RPHYS_LANE_NONE 0xFFFFFFFFu
RPhysResult rphys_debug_set_lane_tier(uint32_t lane_max, uint32_t pair_max, uint32_t quad_max, uint32_t wave_min);

#ifdef __cplusplus
}
#endif
#endif
