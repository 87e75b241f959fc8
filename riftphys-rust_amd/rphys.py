"""ctypes binding of librphys_amd.so (include/riftphys_c.h).

Mirrors the reference's C ABI (crates/riftphys-ffi/src/riftphys_c.h): World wraps an RPhysWorld
handle, with the same 1-based body ids, result codes and error messages, plus the library's
additive exports (density bodies, pair props, iteration counts, debug views, batches).
The library must be built (riftphys-rust_amd/csrc/Makefile); there is no fallback: every call
runs the device step.
"""
import ctypes as C
import json
import math
import os
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "librphys_amd.so")

OK, ERR_NULL, ERR_INVALID_ID, ERR_CAPACITY, ERR_BAD_ARG, ERR_INTERNAL = 0, 1, 2, 3, 4, 0x80000000
BODY_STATIC, BODY_DYNAMIC = 0, 1
SHAPE_SPHERE, SHAPE_BOX, SHAPE_CAPSULE = 0, 1, 2
TERRAIN_BODY_ID = 0xFFFFFFFF
LANE_NONE = 0xFFFFFFFF  # rphys_debug_set_lane_tier: no lane tier / no wave islands
STEP_FLAG_SOLVER_ITERS = 1


class Vec3(C.Structure):
    _fields_ = [("x", C.c_float), ("y", C.c_float), ("z", C.c_float)]


class Quat(C.Structure):
    _fields_ = [("x", C.c_float), ("y", C.c_float), ("z", C.c_float), ("w", C.c_float)]


class Isometry(C.Structure):
    _fields_ = [("pos", Vec3), ("rot", Quat)]


class Velocity(C.Structure):
    _fields_ = [("lin", Vec3), ("ang", Vec3)]


class Material(C.Structure):
    _fields_ = [("mu_static", C.c_float), ("mu_dynamic", C.c_float), ("restitution", C.c_float)]


class CapsuleBodyDesc(C.Structure):
    _fields_ = [("pose", Isometry), ("vel", Velocity), ("radius", C.c_float), ("half_height", C.c_float), ("mass", C.c_float),
                ("body_type", C.c_uint32), ("material", Material), ("user_tag", C.c_uint32)]


class BoxBodyDesc(C.Structure):
    _fields_ = [("pose", Isometry), ("vel", Velocity), ("hx", C.c_float), ("hy", C.c_float), ("hz", C.c_float), ("mass", C.c_float),
                ("body_type", C.c_uint32), ("material", Material), ("user_tag", C.c_uint32)]


class SphereBodyDesc(C.Structure):
    _fields_ = [("pose", Isometry), ("vel", Velocity), ("radius", C.c_float), ("mass", C.c_float),
                ("body_type", C.c_uint32), ("material", Material), ("user_tag", C.c_uint32)]


class Hash32(C.Structure):
    _fields_ = [("bytes", C.c_uint8 * 32)]


class StepParams(C.Structure):
    _fields_ = [("size_bytes", C.c_uint32), ("flags", C.c_uint32), ("substeps", C.c_uint32), ("solver_iters", C.c_uint32), ("dt", C.c_float)]


class HeightfieldDesc(C.Structure):
    _fields_ = [("size_bytes", C.c_uint32), ("flags", C.c_uint32), ("width", C.c_uint32), ("height", C.c_uint32), ("row_stride", C.c_uint32),
                ("cell_size_x", C.c_float), ("cell_size_z", C.c_float), ("height_scale", C.c_float), ("height_offset", C.c_float),
                ("heights", C.POINTER(C.c_int16))]


class ContactEvent(C.Structure):
    _fields_ = [("body_id_a", C.c_uint32), ("body_id_b", C.c_uint32), ("position", Vec3), ("normal", Vec3), ("impulse", C.c_float)]


class PlayerDesc(C.Structure):
    _fields_ = [("start_pose", Isometry), ("radius", C.c_float), ("height", C.c_float), ("speed", C.c_float)]


class RayQuery(C.Structure):
    _fields_ = [("origin", Vec3), ("dir", Vec3), ("max_distance", C.c_float), ("ignore_body", C.c_uint32)]


class RayHit(C.Structure):
    _fields_ = [("hit", C.c_uint32), ("body_id", C.c_uint32), ("fraction", C.c_float), ("point", Vec3), ("normal", Vec3)]


class CapsuleSweepQuery(C.Structure):
    _fields_ = [("from_", Vec3), ("to", Vec3), ("radius", C.c_float), ("half_height", C.c_float), ("ignore_body", C.c_uint32)]


class CapsuleSweepHit(C.Structure):
    _fields_ = [("hit", C.c_uint32), ("body_id", C.c_uint32), ("fraction", C.c_float), ("started_overlapping", C.c_uint32),
                ("point", Vec3), ("normal", Vec3)]


class StepStats(C.Structure):
    _fields_ = [("pairs", C.c_uint32), ("contacts", C.c_uint32), ("terrain_candidates", C.c_uint32), ("island_contacts", C.c_uint32),
                ("islands", C.c_uint32), ("warm_entries", C.c_uint32), ("pgs_iters", C.c_uint32), ("pgs_iters_sum", C.c_uint32),
                ("bounce_rows", C.c_uint32), ("bounce_support_rows", C.c_uint32), ("bounce_pseudo_rows", C.c_uint32),
                ("bounce_fallbacks", C.c_uint32), ("relax_rows", C.c_uint32)]


class ContactEx(C.Structure):
    _fields_ = [("a_collider", C.c_uint32), ("b_collider", C.c_uint32), ("point", C.c_float * 3), ("normal", C.c_float * 3),
                ("depth", C.c_float), ("impulse", C.c_float), ("feature", C.c_uint32), ("cell", C.c_uint32)]


class Profile(C.Structure):
    _fields_ = [("ticks", C.c_uint32), ("n_stages", C.c_uint32), ("stage_ms", C.c_double * 24), ("pairs", C.c_double),
                ("terrain_candidates", C.c_double), ("contacts", C.c_double), ("island_contacts", C.c_double), ("islands", C.c_double),
                ("island_contacts_global", C.c_double), ("pgs_iters_max", C.c_double), ("pgs_iters_sum", C.c_double),
                ("sweep_phase", C.c_double * 11), ("islands_converged", C.c_double), ("island_contacts_converged", C.c_double),
                ("warm_start_learned", C.c_double), ("wide_levels", C.c_double), ("wide_levels_shock", C.c_double),
                ("pgs_iters_sum_converged", C.c_double)]


class MlpTrainConfig(C.Structure):
    _fields_ = [("size_bytes", C.c_uint32), ("epochs", C.c_uint32), ("lr", C.c_float), ("reserved", C.c_uint32), ("seed", C.c_uint64)]


class Tuning(C.Structure):
    _fields_ = [("size_bytes", C.c_uint32), ("grid_small", C.c_uint32), ("grid_large", C.c_uint32), ("grid_huge", C.c_uint32)]


# (name, restype, argtypes) of every export in include/riftphys_c.h
P = C.c_void_p
U32, U64, F32 = C.c_uint32, C.c_uint64, C.c_float
PU32 = C.POINTER(C.c_uint32)
PF32 = C.POINTER(C.c_float)
EXPORTS = [
    ("rphys_get_abi_version", U32, []),
    ("rphys_get_build_id", C.c_char_p, []),
    ("rphys_last_error_message", C.c_char_p, []),
    ("rphys_last_error_copy", U32, [C.c_char_p, U32, PU32]),
    ("rphys_world_create", P, [U32, U32]),
    ("rphys_world_destroy", None, [P]),
    ("rphys_world_set_epoch", U32, [P, U64]),
    ("rphys_world_set_rng_seed", U32, [P, U64]),
    ("rphys_add_capsule_body", U32, [P, C.POINTER(CapsuleBodyDesc), PU32]),
    ("rphys_add_box_body", U32, [P, C.POINTER(BoxBodyDesc), PU32]),
    ("rphys_add_sphere_body", U32, [P, C.POINTER(SphereBodyDesc), PU32]),
    ("rphys_body_remove", U32, [P, U32]),
    ("rphys_body_get_pose", U32, [P, U32, C.POINTER(Isometry)]),
    ("rphys_body_set_pose", U32, [P, U32, C.POINTER(Isometry)]),
    ("rphys_body_get_velocity", U32, [P, U32, C.POINTER(Velocity)]),
    ("rphys_body_set_velocity", U32, [P, U32, C.POINTER(Velocity)]),
    ("rphys_body_apply_impulse", U32, [P, U32, Vec3]),
    ("rphys_world_step_ex", U32, [P, C.POINTER(StepParams)]),
    ("rphys_world_step", U32, [P, F32]),
    ("rphys_world_tick_index", U32, [P, C.POINTER(U64)]),
    ("rphys_world_num_bodies", U32, [P, PU32]),
    ("rphys_world_get_step_hash", U32, [P, C.POINTER(Hash32)]),
    ("rphys_world_step_and_hash", U32, [P, F32, C.POINTER(Hash32)]),
    ("rphys_world_raycast", U32, [P, P, P]),
    ("rphys_world_capsule_sweep", U32, [P, P, P]),
    ("rphys_world_raycast_n", U32, [P, P, U32, P]),
    ("rphys_world_capsule_sweep_n", U32, [P, P, U32, P]),
    ("rphys_world_capsule_sweep", U32, [P, P, P]),
    ("rphys_world_set_heightfield_i16", U32, [P, C.POINTER(HeightfieldDesc), F32, F32, F32]),
    ("rphys_world_set_heightfield_raw32_square", U32, [P, P, U32, F32, F32, F32, F32, F32]),
    ("rphys_world_clear_heightfield", U32, [P]),
    ("rphys_world_get_contacts", U32, [P, C.POINTER(ContactEvent), U32, PU32]),
    ("rphys_add_player", U32, [P, C.POINTER(PlayerDesc), PU32]),
    ("rphys_player_set_input", U32, [P, U32, Vec3]),
    ("rphys_player_get_pose", U32, [P, U32, C.POINTER(Isometry), PU32]),
    ("rphys_world_snapshot_create", P, [P]),
    ("rphys_world_snapshot_restore", U32, [P, P]),
    ("rphys_world_snapshot_free", None, [P]),
    ("rphys_add_body_density", U32, [P, U32, PF32, C.POINTER(Isometry), C.POINTER(Velocity), U32, F32, PU32]),
    ("rphys_world_set_pair_props", U32, [P, U32, U32, F32, F32, F32, F32]),
    ("rphys_world_set_pair_contact_model", U32, [P, U32, U32, F32, F32, U32]),
    ("rphys_world_set_compliant_damping_at_start", U32, [P, U32]),
    ("rphys_world_add_ball_joint", U32, [P, U32, U32, C.POINTER(Isometry), C.POINTER(Isometry), PU32]),
    ("rphys_world_add_d6_joint", U32, [P, U32, U32, C.POINTER(Isometry), C.POINTER(Isometry), U32, U32, C.c_int32, PU32]),
    ("rphys_world_add_distance_joint", U32, [P, U32, U32, C.c_float, C.c_float, PU32]),
    ("rphys_world_add_hinge_joint", U32, [P, U32, U32, C.POINTER(Isometry), C.POINTER(Isometry), U32, PU32]),
    ("rphys_world_set_joint_limit", U32, [P, U32, U32, U32, U32, C.c_float, C.c_float]),
    ("rphys_world_set_joint_drive", U32, [P, U32, U32, U32, U32, C.c_float, C.c_float, C.c_float, C.c_float, U32]),
    ("rphys_world_set_joint_damping", U32, [P, U32, U32, U32, C.c_float, C.c_float]),
    ("rphys_world_set_joint_compliance", U32, [P, U32, U32, U32, C.c_float]),
    ("rphys_world_num_joints", U32, [P]),
    ("rphys_world_set_solver_iterations", U32, [P, U32, U32]),
    ("rphys_world_set_gravity", U32, [P, Vec3]),
    ("rphys_world_set_residual_mode", U32, [P, U32, F32, U32]),
    ("rphys_body_set_sleep_allowed", U32, [P, U32, U32]),
    ("rphys_body_set_material_id", U32, [P, U32, U32]),
    ("rphys_world_step_n", U32, [P, F32, U32]),
    ("rphys_world_get_stats", U32, [P, C.POINTER(StepStats)]),
    ("rphys_world_step_profiled", U32, [P, F32, U32, C.POINTER(Profile)]),
    ("rphys_profile_stage_name", C.c_char_p, [U32]),
    ("rphys_debug_copy_probe", U32, [U64, U32, C.POINTER(C.c_double)]),
    ("rphys_debug_get_contacts_ex", U32, [P, C.POINTER(ContactEx), U32, PU32]),
    ("rphys_debug_get_pairs", U32, [P, PU32, U32, PU32]),
    ("rphys_debug_get_joint_rows", U32, [P, U32, C.POINTER(C.c_uint8), PF32, U32, PU32]),
    ("rphys_debug_get_hash_input", U32, [P, C.POINTER(C.c_uint8), U64, C.POINTER(U64)]),
    ("rphys_debug_export_state", U32, [P, PF32, PF32, PF32, PF32, PF32, PU32, PU32, C.POINTER(C.c_uint64), PF32, PF32, PF32, U32, PU32]),
    ("rphys_batch_create", P, [C.POINTER(P), U32]),
    ("rphys_batch_destroy", None, [P]),
    ("rphys_batch_step_n", U32, [P, F32, U32]),
    ("rphys_batch_get_hashes", U32, [P, C.POINTER(C.c_uint8)]),
    ("rphys_batch_get_state", U32, [P, U32, PF32, PF32, PF32, PF32]),
    ("rphys_batch_snapshot", U32, [P, U32]),
    ("rphys_batch_get_world", U32, [P, U32, P]),
    ("rphys_batch_restore", U32, [P, U32]),
    ("rphys_batch_set_tuning", U32, [P, C.POINTER(Tuning)]),
    ("rphys_batch_get_tuning", U32, [P, C.POINTER(Tuning)]),
    ("rphys_batch_autotune", U32, [P, F32, U32, C.POINTER(Tuning), C.POINTER(C.c_double)]),
    ("rphys_batch_total_bodies", U32, [P, PU32]),
    ("rphys_batch_step_graph", U32, [P, F32, U32]),
    ("rphys_batch_step_profiled", U32, [P, F32, U32, C.POINTER(Profile)]),
    ("rphys_mlp_param_count", U32, []),
    ("rphys_mlp_gradient_rows", U32, []),
    ("rphys_world_set_warm_start", U32, [P, U32, PF32]),
    ("rphys_batch_set_warm_start", U32, [P, U32, PF32]),
    ("rphys_batch_set_contact_log", U32, [P, U32]),
    ("rphys_batch_set_residual_mode", U32, [P, U32, F32, U32]),
    ("rphys_batch_contact_log_rows", U32, [P, PU32, PU32]),
    ("rphys_batch_read_contact_log", U32, [P, U32, U32, PF32, PF32]),
    ("rphys_batch_train_warm_start", U32, [P, C.POINTER(MlpTrainConfig), PF32, PF32]),
    ("rphys_mlp_forward", U32, [PF32, PF32, U32, PF32]),
    ("rphys_mlp_gradient", U32, [PF32, PF32, PF32, PF32, PF32]),
    ("rphys_mlp_forward_noalloc", U32, [P, P, U32, P, P]),
    ("rphys_debug_device_alloc", U32, [U64, C.POINTER(P)]),
    ("rphys_debug_device_free", U32, [P]),
    ("rphys_debug_device_copy", U32, [P, P, U64, U32]),
    ("rphys_debug_set_sweep_tiers", U32, [U32, U32, U32, U32]),
    ("rphys_debug_set_lane_tier", U32, [U32, U32, U32, U32]),
    ("rphys_debug_set_initial_capacity", U32, [U32, U32, U32, U32]),
    ("rphys_debug_fail_device_allocs", U32, [U32, PU32]),
]

_lib = None


def lib():
    """Loads the library (raises if it is not built: the step has no other implementation)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("librphys_amd.so is not built (run __graft_entry__.build() or make -C riftphys-rust_amd/csrc)")
        L = C.CDLL(LIB_PATH)
        for name, res, args in EXPORTS:
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


class RPhysError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__("rphys error %#x: %s" % (code, msg))
        self.code = code


def check(code):
    if code != OK:
        m = lib().rphys_last_error_message()
        raise RPhysError(code, m.decode() if m else "")
    return code


def iso(p, q):
    return Isometry(Vec3(*[float(x) for x in p]), Quat(*[float(x) for x in q]))


def vel(l, a):
    return Velocity(Vec3(*[float(x) for x in l]), Vec3(*[float(x) for x in a]))


def _profile_dicts(pr):
    """(stage name -> ms summed, counters) of an RPhysProfile."""
    stages = {lib().rphys_profile_stage_name(k).decode(): pr.stage_ms[k] for k in range(pr.n_stages)}
    counters = dict(ticks=pr.ticks, pairs=pr.pairs, terrain_candidates=pr.terrain_candidates, contacts=pr.contacts,
                    island_contacts=pr.island_contacts, islands=pr.islands, island_contacts_global=pr.island_contacts_global,
                    pgs_iters_max=pr.pgs_iters_max, pgs_iters_sum=pr.pgs_iters_sum, islands_converged=pr.islands_converged,
                    island_contacts_converged=pr.island_contacts_converged, warm_start_learned=pr.warm_start_learned,
                    wide_levels=pr.wide_levels, wide_levels_shock=pr.wide_levels_shock,
                    pgs_iters_sum_converged=pr.pgs_iters_sum_converged)
    names = ("vel_order", "warm", "pgs", "vel_shock_order", "vel_shock", "ballistic", "projection", "proj_shock_order",
             "proj_shock", "levels_list", "levels_shock")
    counters["sweep_phase"] = {k: pr.sweep_phase[i] for i, k in enumerate(names)}
    return stages, counters


def _f32(a, shape=None):
    a = np.ascontiguousarray(a, np.float32)
    if shape is not None:
        assert a.shape == shape, (a.shape, shape)
    return a


def mlp_forward(params, feat):
    """Device inference of the warm-start model on host features [n][12] -> [n][3]."""
    p, f = _f32(params), _f32(feat)
    out = np.zeros((len(f), 3), np.float32)
    check(lib().rphys_mlp_forward(p.ctypes.data_as(PF32), f.ctypes.data_as(PF32), len(f), out.ctypes.data_as(PF32)))
    return out


def mlp_gradient(params, feat, targ):
    """(loss, gradient of the weights and biases) of one training step over the given rows."""
    p, f, t = _f32(params), _f32(feat), _f32(targ)
    n = lib().rphys_mlp_gradient_rows()
    assert len(f) == n and len(t) == n
    loss = np.zeros(1, np.float32)
    grad = np.zeros(lib().rphys_mlp_param_count() - 27, np.float32)
    check(lib().rphys_mlp_gradient(p.ctypes.data_as(PF32), f.ctypes.data_as(PF32), t.ctypes.data_as(PF32), loss.ctypes.data_as(PF32),
                                   grad.ctypes.data_as(PF32)))
    return float(loss[0]), grad


def copy_probe(nbytes, reps=20):
    """GB/s (read + write) of the library's plain copy kernel over nbytes (rphys_debug_copy_probe)."""
    g = C.c_double(0.0)
    check(lib().rphys_debug_copy_probe(nbytes, reps, C.byref(g)))
    return g.value


class World:
    """An RPhysWorld, optionally loaded with a scene dict from scenes.py."""

    def __init__(self, scene=None, max_bodies=None, max_colliders=None):
        L = lib()
        n = 0 if scene is None else len(scene["shape"])
        mb = max_bodies if max_bodies is not None else max(n, 1)
        mc = max_colliders if max_colliders is not None else mb
        self.h = L.rphys_world_create(mb, mc)
        if not self.h:
            raise RPhysError(ERR_INTERNAL, "rphys_world_create failed")
        self.dt = 1.0 / 60.0
        if scene is not None:
            self.load(scene)

    def close(self):
        if getattr(self, "h", None):
            lib().rphys_world_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def load(self, s):
        L = lib()
        for a, b, ms, mk, e, vs in s.get("pairs", []):
            check(L.rphys_world_set_pair_props(self.h, a, b, ms, mk, e, vs))
        for a, b, k, d, patch in s.get("contact_models", []):
            check(L.rphys_world_set_pair_contact_model(self.h, a, b, k, d, 1 if patch else 0))
        if s.get("damping_at_start"):
            check(L.rphys_world_set_compliant_damping_at_start(self.h, 1))
        vi, pi = s.get("iters", (20, 20))
        check(L.rphys_world_set_solver_iterations(self.h, vi, pi))
        if s.get("residual"):  # (tol, max_iters): each island iterates to the residual
            tol, mx = s["residual"]
            check(L.rphys_world_set_residual_mode(self.h, 1, tol, mx))
        if s.get("gravity") is not None:
            check(L.rphys_world_set_gravity(self.h, Vec3(*[float(x) for x in s["gravity"]])))
        ids = [self.add_body(int(s["shape"][i]), s["dims"][i], s["pose"][i], s["vel"][i], int(s["dynamic"][i]), float(s["density"][i]))
               for i in range(len(s["shape"]))]
        for i, b in enumerate(ids):
            if s.get("material") is not None:
                check(L.rphys_body_set_material_id(self.h, b, int(s["material"][i])))
            if s.get("sleep_allowed") is not None and not s["sleep_allowed"][i]:
                check(L.rphys_body_set_sleep_allowed(self.h, b, 0))
        for a, b, fa, fb in s.get("joints", []):  # scene joints name bodies by scene index
            self.add_ball_joint(ids[a], ids[b], fa, fb)
        for a, b, rest, comp in s.get("dist", []):  # after the ball joints, before the D6 joints
            self.add_distance_joint(ids[int(a)], ids[int(b)], rest, comp)
        for j in s.get("d6", []):  # after the ball and distance joints: joint indices in scene order
            q = self.add_d6_joint(ids[int(j["a"])], ids[int(j["b"])], j["fa"], j["fb"], j.get("t_mask", 7), j.get("r_mask", 0), j.get("hinge", -1))
            for ang, ax, lo, hi in j.get("limits", []):
                check(L.rphys_world_set_joint_limit(self.h, q, int(ang), ax, 1, lo, hi))
            for ang, ax, target, kp, kd, mx, vel in j.get("drives", []):
                check(L.rphys_world_set_joint_drive(self.h, q, int(ang), ax, 1, target, kp, kd, mx, int(vel)))
            for ang, ax, c, fr in j.get("damping", []):
                check(L.rphys_world_set_joint_damping(self.h, q, int(ang), ax, c, fr))
            for ang, ax, c in j.get("compliance", []):
                check(L.rphys_world_set_joint_compliance(self.h, q, int(ang), ax, c))
        hf = s.get("hf")
        if hf is not None:
            self.set_heightfield(hf)
        self.dt = float(s.get("dt", 1.0 / 60.0))

    def add_body(self, shape, dims, pose, v, dynamic=1, density=0.0):
        d = np.ascontiguousarray(dims, np.float32)
        out = C.c_uint32(0)
        check(lib().rphys_add_body_density(self.h, shape, d.ctypes.data_as(PF32), C.byref(iso(pose[:3], pose[3:])),
                                           C.byref(vel(v[:3], v[3:])), BODY_DYNAMIC if dynamic else BODY_STATIC, density, C.byref(out)))
        return out.value

    def add_player(self, pose, radius, height, speed):
        """rphys_add_player: returns the player index."""
        d = PlayerDesc()
        d.start_pose = iso(pose[:3], pose[3:])
        d.radius, d.height, d.speed = float(radius), float(height), float(speed)
        out = C.c_uint32()
        check(lib().rphys_add_player(self.h, C.byref(d), C.byref(out)))
        return out.value

    def player_set_input(self, idx, d):
        check(lib().rphys_player_set_input(self.h, idx, Vec3(*[float(x) for x in d])))

    def player(self, idx):
        """rphys_player_get_pose: (pose7, grounded)."""
        p = Isometry()
        g = C.c_uint32()
        check(lib().rphys_player_get_pose(self.h, idx, C.byref(p), C.byref(g)))
        return np.array([p.pos.x, p.pos.y, p.pos.z, p.rot.x, p.rot.y, p.rot.z, p.rot.w], np.float32), bool(g.value)

    def raycast(self, origin, direction, max_distance, ignore_body=0):
        """rphys_world_raycast: (hit, body_id, fraction, point, normal)."""
        q = RayQuery(Vec3(*[float(x) for x in origin]), Vec3(*[float(x) for x in direction]), float(max_distance), int(ignore_body))
        h = RayHit()
        check(lib().rphys_world_raycast(self.h, C.byref(q), C.byref(h)))
        return h

    def raycast_n(self, queries):
        """rphys_world_raycast_n over [(origin, direction, max_distance, ignore_body)]."""
        n = len(queries)
        qa = (RayQuery * n)(*[RayQuery(Vec3(*map(float, o)), Vec3(*map(float, d)), float(m), int(i)) for o, d, m, i in queries])
        ha = (RayHit * n)()
        check(lib().rphys_world_raycast_n(self.h, qa, n, ha))
        return list(ha)

    def capsule_sweep(self, frm, to, radius, half_height, ignore_body=0):
        """rphys_world_capsule_sweep: (hit, body_id, fraction, started_overlapping, point, normal)."""
        q = CapsuleSweepQuery(Vec3(*[float(x) for x in frm]), Vec3(*[float(x) for x in to]), float(radius), float(half_height), int(ignore_body))
        h = CapsuleSweepHit()
        check(lib().rphys_world_capsule_sweep(self.h, C.byref(q), C.byref(h)))
        return h

    def capsule_sweep_n(self, queries):
        n = len(queries)
        qa = (CapsuleSweepQuery * n)(*[CapsuleSweepQuery(Vec3(*map(float, f)), Vec3(*map(float, t)), float(r), float(hh), int(i))
                                       for f, t, r, hh, i in queries])
        ha = (CapsuleSweepHit * n)()
        check(lib().rphys_world_capsule_sweep_n(self.h, qa, n, ha))
        return list(ha)

    def add_body_spec(self, b):
        """One body from a dict of the scene fields (shape, dims, pose, vel, dynamic, density,
        optional material and sleep_allowed); returns its id."""
        d = np.zeros(3, np.float32)
        d[:len(b["dims"])] = b["dims"]
        i = self.add_body(int(b["shape"]), d, b["pose"], b["vel"], int(b["dynamic"]), float(b["density"]))
        if "material" in b:
            check(lib().rphys_body_set_material_id(self.h, i, int(b["material"])))
        if not b.get("sleep_allowed", 1):
            check(lib().rphys_body_set_sleep_allowed(self.h, i, 0))
        return i

    def add_ball_joint(self, body_a, body_b, frame_a, frame_b):
        """Ball joint between two body ids, frames (px, py, pz, qx, qy, qz, qw) in each body's space;
        returns the joint index."""
        out = C.c_uint32(0)
        check(lib().rphys_world_add_ball_joint(self.h, body_a, body_b, C.byref(iso(frame_a[:3], frame_a[3:])),
                                               C.byref(iso(frame_b[:3], frame_b[3:])), C.byref(out)))
        return out.value

    def add_d6_joint(self, body_a, body_b, frame_a, frame_b, t_mask=7, r_mask=0, hinge=-1):
        """D6 joint constraining the translational axes of t_mask and rotational axes of r_mask
        (bit i: axis i of frame A), with stored hinge axis `hinge` (-1: none); returns its index."""
        out = C.c_uint32(0)
        check(lib().rphys_world_add_d6_joint(self.h, body_a, body_b, C.byref(iso(frame_a[:3], frame_a[3:])),
                                             C.byref(iso(frame_b[:3], frame_b[3:])), t_mask, r_mask, hinge, C.byref(out)))
        return out.value

    def add_distance_joint(self, body_a, body_b, rest, compliance=0.0):
        out = C.c_uint32(0)
        check(lib().rphys_world_add_distance_joint(self.h, body_a, body_b, float(rest), float(compliance), C.byref(out)))
        return out.value

    def add_hinge_joint(self, body_a, body_b, frame_a, frame_b, axis):
        out = C.c_uint32(0)
        check(lib().rphys_world_add_hinge_joint(self.h, body_a, body_b, C.byref(iso(frame_a[:3], frame_a[3:])),
                                                C.byref(iso(frame_b[:3], frame_b[3:])), axis, C.byref(out)))
        return out.value

    def set_heightfield(self, hf):
        """terrain.rs via rphys_world_set_heightfield_i16 is lossy for float heights, so the
        raw32 square entry is used when the grid is square and the cell sizes agree."""
        hts = np.ascontiguousarray(hf["heights"], np.float32)
        nx, nz = int(hf["nx"]), int(hf["nz"])
        if nx != nz:
            raise ValueError("scene heightfields are square")
        sx = float(hf["sx"]) * (nx - 1)
        sz = float(hf["sz"]) * (nz - 1)
        check(lib().rphys_world_set_heightfield_raw32_square(self.h, hts.ctypes.data, hts.nbytes, sx, sz, float(hf["ox"]), float(hf["oz"]), 0.0))

    def step(self, dt=None):
        check(lib().rphys_world_step(self.h, self.dt if dt is None else dt))

    def step_n(self, n, dt=None):
        check(lib().rphys_world_step_n(self.h, self.dt if dt is None else dt, n))

    def step_profiled(self, n, dt=None):
        """Runs n ticks with per-stage HIP event timing; returns (stage name -> ms summed, counters)."""
        pr = Profile()
        check(lib().rphys_world_step_profiled(self.h, self.dt if dt is None else dt, n, C.byref(pr)))
        return _profile_dicts(pr)

    def hash(self):
        h = Hash32()
        check(lib().rphys_world_get_step_hash(self.h, C.byref(h)))
        return bytes(h.bytes)

    def tick(self):
        t = C.c_uint64(0)
        check(lib().rphys_world_tick_index(self.h, C.byref(t)))
        return t.value

    def num_bodies(self):
        n = C.c_uint32(0)
        check(lib().rphys_world_num_bodies(self.h, C.byref(n)))
        return n.value

    def state(self):
        n = self.num_bodies()
        pos = np.zeros((n, 3), np.float32); rot = np.zeros((n, 4), np.float32)
        lin = np.zeros((n, 3), np.float32); ang = np.zeros((n, 3), np.float32)
        p = Isometry(); v = Velocity()
        for i in range(n):
            if lib().rphys_body_get_pose(self.h, i + 1, C.byref(p)) != OK:
                continue
            check(lib().rphys_body_get_velocity(self.h, i + 1, C.byref(v)))
            pos[i] = (p.pos.x, p.pos.y, p.pos.z); rot[i] = (p.rot.x, p.rot.y, p.rot.z, p.rot.w)
            lin[i] = (v.lin.x, v.lin.y, v.lin.z); ang[i] = (v.ang.x, v.ang.y, v.ang.z)
        return pos, rot, lin, ang

    def stats(self):
        s = StepStats()
        check(lib().rphys_world_get_stats(self.h, C.byref(s)))
        return s

    def pairs(self):
        n = C.c_uint32(0)
        check(lib().rphys_debug_get_pairs(self.h, None, 0, C.byref(n)))
        out = np.zeros((max(n.value, 1), 2), np.uint32)
        check(lib().rphys_debug_get_pairs(self.h, out.ctypes.data_as(PU32), n.value, C.byref(n)))
        return out[: n.value]

    def contacts(self):
        n = C.c_uint32(0)
        check(lib().rphys_debug_get_contacts_ex(self.h, None, 0, C.byref(n)))
        arr = (ContactEx * max(n.value, 1))()
        check(lib().rphys_debug_get_contacts_ex(self.h, arr, n.value, C.byref(n)))
        return np.ctypeslib.as_array(arr)[: n.value].copy()

    def hash_input(self):
        """The bytes the step hash digests (rphys_debug_get_hash_input)."""
        n = C.c_uint64(0)
        check(lib().rphys_debug_get_hash_input(self.h, None, 0, C.byref(n)))
        out = np.zeros(max(n.value, 1), np.uint8)
        check(lib().rphys_debug_get_hash_input(self.h, out.ctypes.data_as(C.POINTER(C.c_uint8)), n.value, C.byref(n)))
        return bytes(out[:n.value])

    def joint_rows(self, q):
        """Joint q's stored rows as (tags, impulses) (rphys_debug_get_joint_rows)."""
        tags = (C.c_uint8 * 32)()
        lams = np.zeros(32, np.float32)
        n = C.c_uint32(0)
        check(lib().rphys_debug_get_joint_rows(self.h, q, tags, lams.ctypes.data_as(PF32), 32, C.byref(n)))
        return list(tags)[:n.value], lams[:n.value].copy()

    def export_state(self):
        """Bodies, sleep state, last velocities and the warm cache (rphys_debug_export_state)."""
        n = self.num_bodies()
        cnt = C.c_uint32(0)
        check(lib().rphys_debug_export_state(self.h, None, None, None, None, None, None, None, None, None, None, None, 0, C.byref(cnt)))
        m = cnt.value
        st = dict(pos=np.zeros((n, 3), np.float32), rot=np.zeros((n, 4), np.float32), lin=np.zeros((n, 3), np.float32),
                  ang=np.zeros((n, 3), np.float32), last_lin=np.zeros((n, 3), np.float32), awake=np.zeros(n, np.uint32),
                  sticks=np.zeros(n, np.uint32), wkeys=np.zeros(max(m, 1), np.uint64), wla=np.zeros((max(m, 1), 3), np.float32),
                  wjn=np.zeros(max(m, 1), np.float32), wjt=np.zeros((max(m, 1), 3), np.float32))
        f = lambda a: a.ctypes.data_as(PF32)
        u = lambda a: a.ctypes.data_as(PU32)
        check(lib().rphys_debug_export_state(self.h, f(st["pos"]), f(st["rot"]), f(st["lin"]), f(st["ang"]), f(st["last_lin"]),
                                             u(st["awake"]), u(st["sticks"]), st["wkeys"].ctypes.data_as(C.POINTER(C.c_uint64)),
                                             f(st["wla"]), f(st["wjn"]), f(st["wjt"]), m, C.byref(cnt)))
        for k in ("wkeys", "wla", "wjn", "wjt"):
            st[k] = st[k][:m]
        st["tick"] = self.tick()
        return st

    def set_warm_start(self, mode, params=None):
        p = None if params is None else _f32(params)
        check(lib().rphys_world_set_warm_start(self.h, mode, None if p is None else p.ctypes.data_as(PF32)))

    def set_residual_mode(self, on, tol=1.0e-4, max_iters=100):
        check(lib().rphys_world_set_residual_mode(self.h, 1 if on else 0, tol, max_iters))

    def snapshot(self):
        s = lib().rphys_world_snapshot_create(self.h)
        if not s:
            raise RPhysError(ERR_INTERNAL, "snapshot failed")
        return s

    def restore(self, snap):
        check(lib().rphys_world_snapshot_restore(self.h, snap))

    def free_snapshot(self, snap):
        lib().rphys_world_snapshot_free(snap)


class Batch:
    """Many worlds stepped together (rphys_batch_*)."""

    def __init__(self, worlds):
        arr = (P * len(worlds))(*[w.h for w in worlds])
        self.h = lib().rphys_batch_create(arr, len(worlds))
        if not self.h:
            m = lib().rphys_last_error_message()
            raise RPhysError(ERR_INTERNAL, m.decode() if m else "batch create failed")
        self.n = len(worlds)

    def close(self):
        if getattr(self, "h", None):
            lib().rphys_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def step_n(self, n, dt=1.0 / 60.0, graph=False):
        if graph:
            check(lib().rphys_batch_step_graph(self.h, dt, n))
        else:
            check(lib().rphys_batch_step_n(self.h, dt, n))

    def step_profiled(self, n, dt=1.0 / 60.0):
        """n ticks with per-stage HIP event timing (rphys_batch_step_profiled)."""
        pr = Profile()
        check(lib().rphys_batch_step_profiled(self.h, dt, n, C.byref(pr)))
        return _profile_dicts(pr)

    def hashes(self):
        out = np.zeros(32 * self.n, np.uint8)
        check(lib().rphys_batch_get_hashes(self.h, out.ctypes.data_as(C.POINTER(C.c_uint8))))
        return [bytes(out[32 * i: 32 * i + 32]) for i in range(self.n)]

    def total_bodies(self):
        n = C.c_uint32(0)
        check(lib().rphys_batch_total_bodies(self.h, C.byref(n)))
        return n.value

    def state(self, wi, n):
        pos = np.zeros((n, 3), np.float32); rot = np.zeros((n, 4), np.float32)
        lin = np.zeros((n, 3), np.float32); ang = np.zeros((n, 3), np.float32)
        check(lib().rphys_batch_get_state(self.h, wi, pos.ctypes.data_as(PF32), rot.ctypes.data_as(PF32), lin.ctypes.data_as(PF32), ang.ctypes.data_as(PF32)))
        return pos, rot, lin, ang

    def set_warm_start(self, mode, params=None):
        p = None if params is None else _f32(params)
        check(lib().rphys_batch_set_warm_start(self.h, mode, None if p is None else p.ctypes.data_as(PF32)))

    def set_residual_mode(self, on, tol=1.0e-4, max_iters=100):
        check(lib().rphys_batch_set_residual_mode(self.h, 1 if on else 0, tol, max_iters))

    def set_contact_log(self, rows):
        check(lib().rphys_batch_set_contact_log(self.h, rows))

    def contact_log_rows(self):
        n, dropped = C.c_uint32(0), C.c_uint32(0)
        check(lib().rphys_batch_contact_log_rows(self.h, C.byref(n), C.byref(dropped)))
        return n.value, dropped.value

    def read_contact_log(self, first, count):
        f = np.zeros((count, 12), np.float32); t = np.zeros((count, 3), np.float32)
        check(lib().rphys_batch_read_contact_log(self.h, first, count, f.ctypes.data_as(PF32), t.ctypes.data_as(PF32)))
        return f, t

    def train_warm_start(self, epochs=20, lr=1.0e-3, seed=1):
        cfg = MlpTrainConfig(C.sizeof(MlpTrainConfig), epochs, lr, 0, seed)
        params = np.zeros(lib().rphys_mlp_param_count(), np.float32)
        loss = np.zeros(epochs, np.float32)
        check(lib().rphys_batch_train_warm_start(self.h, C.byref(cfg), params.ctypes.data_as(PF32), loss.ctypes.data_as(PF32)))
        return params, loss

    def world(self, wi):
        """A new World holding world wi of the batch as the batch has stepped it."""
        w = World()
        check(lib().rphys_batch_get_world(self.h, wi, w.h))
        return w

    def snapshot(self, slot=0):
        check(lib().rphys_batch_snapshot(self.h, slot))

    def restore(self, slot=0):
        check(lib().rphys_batch_restore(self.h, slot))

    def set_tuning(self, grid_small, grid_large, grid_huge):
        t = Tuning(C.sizeof(Tuning), grid_small, grid_large, grid_huge)
        check(lib().rphys_batch_set_tuning(self.h, C.byref(t)))

    def get_tuning(self):
        t = Tuning()
        check(lib().rphys_batch_get_tuning(self.h, C.byref(t)))
        return dict(grid_small=t.grid_small, grid_large=t.grid_large, grid_huge=t.grid_huge)

    def autotune(self, config, ticks=2, dt=1.0 / 60.0, cache=None):
        """Installs the launch shape cached for (config, body-count bucket), or grid-searches one on
        the device (rphys_batch_autotune: same ticks from a snapshot per candidate, results checked
        equal) and caches it.  Returns (entry, found_in_cache); the batch state is unchanged."""
        path = cache or default_tuning_cache()
        key = tuning_key(config, self.total_bodies())
        data = load_tuning_cache(path)
        if key in data:
            e = data[key]
            self.set_tuning(e["grid_small"], e["grid_large"], e["grid_huge"])
            return e, True
        t, ms = Tuning(), C.c_double(0.0)
        check(lib().rphys_batch_autotune(self.h, dt, ticks, C.byref(t), C.byref(ms)))
        e = dict(grid_small=t.grid_small, grid_large=t.grid_large, grid_huge=t.grid_huge, ms_per_tick=ms.value, ticks=ticks)
        data[key] = e
        store_tuning_cache(path, data)
        return e, False


def tuning_key(config, bodies):
    """Cache key of a launch shape: the workload's name and its body count's power-of-two bucket."""
    return "%s/n2^%d" % (config, int(math.floor(math.log2(max(1, int(bodies))))))


def default_tuning_cache():
    base = os.environ.get("RPHYS_TUNING_CACHE")
    if base:
        return base
    return os.path.join(os.path.expanduser("~"), ".cache", "rphys_amd", "tuning.json")


def load_tuning_cache(path):
    try:
        with open(path) as f:
            data = json.load(f)
        return data if isinstance(data, dict) else {}
    except (OSError, ValueError):
        return {}


def store_tuning_cache(path, data):
    """Writes the cache atomically (temporary file, then rename); an unwritable location is skipped."""
    try:
        d = os.path.dirname(path) or "."
        os.makedirs(d, exist_ok=True)
        fd, tmp = tempfile.mkstemp(dir=d, prefix=".tuning", suffix=".json")
        with os.fdopen(fd, "w") as f:
            json.dump(data, f, indent=1, sort_keys=True)
        os.replace(tmp, path)
    except OSError:
        pass
