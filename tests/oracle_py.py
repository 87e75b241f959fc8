"""ctypes view of the CPU oracle (oracle/liboracle.so).  Test infrastructure only: imported by
tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg, never by the product."""
import ctypes as C
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.path.join(ROOT, "oracle", "liboracle.so")
OR_SPHERE, OR_BOX, OR_CAPSULE = 0, 1, 2


class OrStats(C.Structure):
    _fields_ = [("pairs", C.c_uint32), ("contacts", C.c_uint32), ("islands", C.c_uint32),
                ("terrain_candidates", C.c_uint32), ("levels", C.c_uint32), ("warm_matched", C.c_uint32),
                ("iters_used", C.c_uint32), ("residual", C.c_float), ("iters_sum", C.c_uint32),
                ("toi_rows", C.c_uint32), ("support_rows", C.c_uint32), ("pseudo_rows", C.c_uint32),
                ("bounce_components", C.c_uint32), ("bounce_fallbacks", C.c_uint32), ("bounce_solves", C.c_uint32),
                ("relax_rows", C.c_uint32), ("patches", C.c_uint32)]


class OrContact(C.Structure):
    _fields_ = [("a_col", C.c_uint32), ("b_col", C.c_uint32), ("point", C.c_float * 3), ("normal", C.c_float * 3),
                ("depth", C.c_float), ("impulse", C.c_float), ("feat", C.c_uint32), ("cell", C.c_uint32)]


class OrQueryHit(C.Structure):
    _fields_ = [("hit", C.c_int), ("body", C.c_uint32), ("fraction", C.c_float), ("point", C.c_float * 3), ("normal", C.c_float * 3),
                ("started", C.c_int)]


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("oracle not built: run `make -C oracle` (or __graft_entry__.build())")
        L = C.CDLL(LIB_PATH)
        P = C.c_void_p
        FP = C.POINTER(C.c_float)
        P32_ = C.POINTER(C.c_uint32)
        L.or_world_create.restype = P
        L.or_world_create.argtypes = [C.c_uint32]
        L.or_world_destroy.argtypes = [P]
        L.or_set_epoch.argtypes = [P, C.c_uint64]
        L.or_set_rng_seed.argtypes = [P, C.c_uint64]
        L.or_set_gravity.argtypes = [P, C.c_float, C.c_float, C.c_float]
        L.or_set_iterations.argtypes = [P, C.c_uint32, C.c_uint32]
        L.or_set_residual_mode.argtypes = [P, C.c_int, C.c_float, C.c_uint32]
        L.or_set_pair_props.argtypes = [P, C.c_uint32, C.c_uint32, C.c_float, C.c_float, C.c_float, C.c_float]
        L.or_add_body.restype = C.c_int32
        L.or_add_body.argtypes = [P, C.c_int, FP, FP, FP, C.c_int, C.c_float]
        L.or_add_body_nc.restype = C.c_int32
        L.or_add_body_nc.argtypes = [P, C.c_int, FP, FP, FP, C.c_int, C.c_float]
        L.or_add_distance_joint.restype = C.c_int32
        L.or_add_distance_joint.argtypes = [P, C.c_uint32, C.c_uint32, C.c_float, C.c_float]
        L.or_remove_body.argtypes = [P, C.c_uint32]
        L.or_add_ball_joint.argtypes = [P, C.c_uint32, C.c_uint32, FP, FP]
        L.or_add_ball_joint.restype = C.c_int32
        L.or_num_joints.argtypes = [P]
        L.or_num_joints.restype = C.c_uint32
        L.or_get_joint_warm.argtypes = [P, C.c_uint32, C.POINTER(C.c_uint8), FP]
        L.or_get_joint_warm.restype = C.c_uint32
        L.or_shake_ok.argtypes = [C.c_float]
        L.or_shake_ok.restype = C.c_int
        L.or_set_heightfield.argtypes = [P, C.c_uint32, C.c_uint32, C.c_float, C.c_float, C.c_float, C.c_float, FP, C.c_uint32]
        L.or_set_body_pose.argtypes = [P, C.c_uint32, FP]
        L.or_set_body_vel.argtypes = [P, C.c_uint32, FP]
        L.or_set_sleep_allowed.argtypes = [P, C.c_uint32, C.c_int]
        L.or_set_body_material.argtypes = [P, C.c_uint32, C.c_uint32]
        L.or_add_player.argtypes = [P, FP, C.c_float, C.c_float, C.c_float]
        L.or_add_player.restype = C.c_int32
        L.or_player_set_input.argtypes = [P, C.c_uint32, FP]
        L.or_player_get.argtypes = [P, C.c_uint32, FP, C.POINTER(C.c_int), FP]
        L.or_player_get.restype = C.c_uint32
        L.or_raycast.argtypes = [P, FP, FP, C.c_float, C.c_int32, C.POINTER(OrQueryHit)]
        L.or_raycast.restype = C.c_int
        L.or_capsule_sweep.argtypes = [P, FP, FP, C.c_float, C.c_float, C.c_int32, C.POINTER(OrQueryHit)]
        L.or_capsule_sweep.restype = C.c_int
        DP = C.POINTER(C.c_double)
        L.or_bn_min_degree.restype = C.c_uint32
        L.or_bn_min_degree.argtypes = [C.c_uint32, P32_, P32_, P32_]
        L.or_bn_pseudo_solve.restype = C.c_int
        L.or_bn_pseudo_solve.argtypes = [C.c_uint32, P32_, P32_, DP, DP, DP]
        L.or_bn_bounce_qp.restype = C.c_int
        L.or_bn_bounce_qp.argtypes = [C.c_uint32, P32_, P32_, DP, DP, DP, DP, P32_]
        L.or_toi_bounce.restype = C.c_float
        L.or_toi_bounce.argtypes = [C.c_float, C.c_float, C.c_float, C.c_float, C.c_float, FP, C.POINTER(C.c_int)]
        L.or_step.argtypes = [P, C.c_float]
        L.or_set_threads.argtypes = [C.c_int]
        P8, P16, P32 = C.POINTER(C.c_uint8), C.POINTER(C.c_uint16), C.POINTER(C.c_uint32)
        L.or_collide_shape_triangle.restype = C.c_int
        L.or_collide_shape_triangle.argtypes = [C.c_int, FP, FP, FP, C.c_float, FP]
        L.or_warm_pick_sequence.argtypes = [P32, FP, C.c_uint32, P32, FP, C.c_uint32, C.POINTER(C.c_int32)]
        L.or_export_carried.restype = C.c_uint32
        L.or_export_carried.argtypes = [P, FP, P8, P16, P32, FP, FP, FP, C.c_uint32, P32, P8, P8, FP, P32]
        L.or_capacity.restype = C.c_uint32
        L.or_capacity.argtypes = [P]
        L.or_step_hash.argtypes = [P, C.c_char_p]
        L.or_step_hash_preimage.restype = C.c_size_t
        L.or_step_hash_preimage.argtypes = [P, C.c_void_p, C.c_size_t]
        L.or_num_bodies.restype = C.c_uint32
        L.or_num_bodies.argtypes = [P]
        L.or_tick.restype = C.c_uint64
        L.or_tick.argtypes = [P]
        L.or_get_state.argtypes = [P, FP, FP, FP, FP]
        L.or_get_stats.argtypes = [P, C.POINTER(OrStats)]
        L.or_get_pairs.restype = C.c_uint32
        L.or_get_pairs.argtypes = [P, C.POINTER(C.c_uint32), C.c_uint32]
        L.or_get_contacts.restype = C.c_uint32
        L.or_get_contacts.argtypes = [P, C.POINTER(OrContact), C.c_uint32]
        L.or_get_awake.restype = C.c_uint32
        L.or_get_awake.argtypes = [P, C.POINTER(C.c_uint8), C.c_uint32]
        L.or_snapshot.restype = P
        L.or_snapshot.argtypes = [P]
        L.or_restore.argtypes = [P, P]
        L.or_snapshot_free.argtypes = [P]
        L.or_blake3.argtypes = [C.c_char_p, C.c_size_t, C.c_char_p]
        L.or_collide.restype = C.c_int
        L.or_collide.argtypes = [C.c_int, FP, FP, C.c_int, FP, FP, C.c_float, FP]
        L.or_collide_box_triangle.restype = C.c_int
        L.or_collide_box_triangle.argtypes = [FP, FP, FP, C.c_float, FP]
        L.or_collide_sphere_triangle.restype = C.c_int
        L.or_collide_sphere_triangle.argtypes = [C.c_float, FP, FP, C.c_float, FP]
        L.or_sincosf.argtypes = [C.c_float, FP, FP]
        L.or_set_pair_contact_model.argtypes = [P, C.c_uint32, C.c_uint32, C.c_float, C.c_float, C.c_int]
        L.or_set_compliant_damping_at_start.argtypes = [P, C.c_int]
        L.or_add_d6_joint.restype = C.c_int32
        L.or_add_d6_joint.argtypes = [P, C.c_uint32, C.c_uint32, FP, FP, C.c_uint32, C.c_uint32, C.c_int32]
        L.or_set_d6_limit.argtypes = [P, C.c_uint32, C.c_int, C.c_uint32, C.c_int, C.c_float, C.c_float]
        L.or_set_d6_drive.argtypes = [P, C.c_uint32, C.c_int, C.c_uint32, C.c_int, C.c_float, C.c_float, C.c_float, C.c_float, C.c_int]
        L.or_set_d6_damping.argtypes = [P, C.c_uint32, C.c_int, C.c_uint32, C.c_float, C.c_float]
        L.or_set_d6_compliance.argtypes = [P, C.c_uint32, C.c_int, C.c_uint32, C.c_float]
        L.or_patch_radius.restype = C.c_float
        L.or_patch_radius.argtypes = [FP, C.c_uint32]
        L.or_atan2f_unit.restype = C.c_float
        L.or_atan2f_unit.argtypes = [C.c_float, C.c_float]
        U = C.POINTER(C.c_uint32)
        L.or_load_state.argtypes = [P, FP, FP, FP, FP, FP, U, U, C.c_uint64, C.POINTER(C.c_uint64), FP, FP, FP, C.c_uint32]
        L.or_broadphase.restype = C.c_uint32
        L.or_broadphase.argtypes = [FP, C.c_uint32, U, C.c_uint32]
        _lib = L
    return _lib


def fptr(a):
    return a.ctypes.data_as(C.POINTER(C.c_float))


def blake3(data: bytes) -> bytes:
    out = C.create_string_buffer(32)
    lib().or_blake3(data, len(data), out)
    return out.raw


class OracleWorld:
    """The oracle world, loaded with a scene dict from scenes.py."""

    def __init__(self, scene=None, capacity=None):
        L = lib()
        n = 0 if scene is None else len(scene["shape"])
        self.cap = capacity if capacity is not None else max(n, 1)
        self.h = L.or_world_create(self.cap)
        if scene is not None:
            self.load(scene)

    def load(self, s):
        L = lib()
        for a, b, ms, mk, e, vs in s.get("pairs", []):
            L.or_set_pair_props(self.h, a, b, ms, mk, e, vs)
        for a, b, k, d, patch in s.get("contact_models", []):
            L.or_set_pair_contact_model(self.h, a, b, k, d, int(patch))
        if s.get("damping_at_start"):
            L.or_set_compliant_damping_at_start(self.h, 1)
        vi, pi = s.get("iters", (20, 20))
        L.or_set_iterations(self.h, vi, pi)
        if s.get("residual"):
            tol, mx = s["residual"]
            L.or_set_residual_mode(self.h, 1, tol, mx)
        if s.get("gravity") is not None:
            L.or_set_gravity(self.h, *[float(x) for x in s["gravity"]])
        for i in range(len(s["shape"])):
            d = np.ascontiguousarray(s["dims"][i], np.float32)
            p = np.ascontiguousarray(s["pose"][i], np.float32)
            v = np.ascontiguousarray(s["vel"][i], np.float32)
            b = L.or_add_body(self.h, int(s["shape"][i]), fptr(d), fptr(p), fptr(v), int(s["dynamic"][i]), float(s["density"][i]))
            if s.get("material") is not None:
                L.or_set_body_material(self.h, b, int(s["material"][i]))
            if s.get("sleep_allowed") is not None and not s["sleep_allowed"][i]:
                L.or_set_sleep_allowed(self.h, b, 0)
        for a, b, fa, fb in s.get("joints", []):
            fa = np.ascontiguousarray(fa, np.float32)
            fb = np.ascontiguousarray(fb, np.float32)
            assert L.or_add_ball_joint(self.h, int(a), int(b), fptr(fa), fptr(fb)) >= 0
        for a, b, rest, comp in s.get("dist", []):  # after the ball joints, before the D6 joints
            assert L.or_add_distance_joint(self.h, int(a), int(b), float(rest), float(comp)) >= 0
        for j in s.get("d6", []):  # after the ball and distance joints: joint indices in scene order
            q = self.add_d6_joint(int(j["a"]), int(j["b"]), j["fa"], j["fb"], j.get("t_mask", 7), j.get("r_mask", 0), j.get("hinge", -1))
            for ang, ax, lo, hi in j.get("limits", []):
                L.or_set_d6_limit(self.h, q, int(ang), ax, 1, lo, hi)
            for ang, ax, target, kp, kd, mx, vel in j.get("drives", []):
                L.or_set_d6_drive(self.h, q, int(ang), ax, 1, target, kp, kd, mx, int(vel))
            for ang, ax, c, fr in j.get("damping", []):
                L.or_set_d6_damping(self.h, q, int(ang), ax, c, fr)
            for ang, ax, c in j.get("compliance", []):
                L.or_set_d6_compliance(self.h, q, int(ang), ax, c)
        hf = s.get("hf")
        if hf is not None:
            hts = np.ascontiguousarray(hf["heights"], np.float32)
            # the cell size the reference's raw32 loader derives (lib.rs:1516-1517): size / (dim - 1)
            n1 = np.float32(hf["nx"] - 1)
            sx = float(np.float32(np.float32(np.float32(hf["sx"]) * n1) / n1))
            sz = float(np.float32(np.float32(np.float32(hf["sz"]) * n1) / n1))
            L.or_set_heightfield(self.h, hf["nx"], hf["nz"], sx, sz, float(hf["ox"]), float(hf["oz"]), fptr(hts), 4)
        self.dt = float(s.get("dt", 1.0 / 60.0))

    def __del__(self):
        try:
            lib().or_world_destroy(self.h)
        except Exception:
            pass

    def add_body_spec(self, b):
        """One body from a dict of the scene fields (shape, dims, pose, vel, dynamic, density,
        optional material and sleep_allowed); returns its index."""
        L = lib()
        d = np.zeros(3, np.float32)
        d[:len(b["dims"])] = b["dims"]
        p = np.ascontiguousarray(b["pose"], np.float32)
        v = np.ascontiguousarray(b["vel"], np.float32)
        i = L.or_add_body(self.h, int(b["shape"]), fptr(d), fptr(p), fptr(v), int(b["dynamic"]), float(b["density"]))
        if "material" in b:
            L.or_set_body_material(self.h, i, int(b["material"]))
        if not b.get("sleep_allowed", 1):
            L.or_set_sleep_allowed(self.h, i, 0)
        return i

    def add_d6_joint(self, a, b, fa, fb, t_mask=7, r_mask=0, hinge=-1):
        """A D6 joint between bodies a and b (0-based), frames {p, q} in each body's space."""
        fa = np.ascontiguousarray(fa, np.float32)
        fb = np.ascontiguousarray(fb, np.float32)
        q = lib().or_add_d6_joint(self.h, a, b, fptr(fa), fptr(fb), t_mask, r_mask, hinge)
        assert q >= 0
        return q

    def add_hinge_joint(self, a, b, fa, fb, axis):
        return self.add_d6_joint(a, b, fa, fb, 7, 7 & ~(1 << axis), axis)

    def add_player(self, pose, radius, height, speed):
        return lib().or_add_player(self.h, fptr(np.asarray(pose, np.float32)), float(radius), float(height), float(speed))

    def player_set_input(self, idx, d):
        lib().or_player_set_input(self.h, idx, fptr(np.asarray(d, np.float32)))

    def player(self, idx):
        """(pose7, grounded, vertical velocity, body index)"""
        pose = np.zeros(7, np.float32)
        g = C.c_int()
        vv = np.zeros(1, np.float32)
        b = lib().or_player_get(self.h, idx, fptr(pose), C.byref(g), fptr(vv))
        return pose, bool(g.value), float(vv[0]), b

    def raycast(self, origin, direction, max_distance, ignore=-1):
        """World::raycast with the direction normalised as rphys_world_raycast does (ignore: a body
        index or -1)."""
        d = np.asarray(direction, np.float32)
        l2 = np.float32(np.float32(d[0] * d[0] + d[1] * d[1]) + d[2] * d[2])
        h = OrQueryHit()
        if not max_distance > 0 or not l2 > np.float32(1e-20):
            return h
        d = d / np.sqrt(l2)
        lib().or_raycast(self.h, fptr(np.asarray(origin, np.float32)), fptr(np.ascontiguousarray(d, np.float32)), float(max_distance),
                         int(ignore), C.byref(h))
        return h

    def capsule_sweep(self, frm, to, radius, half_height, ignore=-1):
        h = OrQueryHit()
        lib().or_capsule_sweep(self.h, fptr(np.asarray(frm, np.float32)), fptr(np.asarray(to, np.float32)), float(radius), float(half_height),
                               int(ignore), C.byref(h))
        return h

    def step(self, dt=None):
        lib().or_step(self.h, self.dt if dt is None else dt)

    def hash(self):
        out = C.create_string_buffer(32)
        lib().or_step_hash(self.h, out)
        return out.raw

    def hash_input(self):
        """The bytes the step hash digests."""
        n = lib().or_step_hash_preimage(self.h, None, 0)
        out = C.create_string_buffer(max(n, 1))
        lib().or_step_hash_preimage(self.h, out, n)
        return out.raw[:n]

    def state(self):
        n = lib().or_num_bodies(self.h)
        pos = np.zeros((n, 3), np.float32); rot = np.zeros((n, 4), np.float32)
        lin = np.zeros((n, 3), np.float32); ang = np.zeros((n, 3), np.float32)
        lib().or_get_state(self.h, fptr(pos), fptr(rot), fptr(lin), fptr(ang))
        return pos, rot, lin, ang

    def stats(self):
        s = OrStats()
        lib().or_get_stats(self.h, C.byref(s))
        return s

    def pairs(self):
        n = lib().or_get_pairs(self.h, None, 0)
        out = np.zeros((max(n, 1), 2), np.uint32)
        lib().or_get_pairs(self.h, out.ctypes.data_as(C.POINTER(C.c_uint32)), n)
        return out[:n]

    def contacts(self):
        n = lib().or_get_contacts(self.h, None, 0)
        arr = (OrContact * max(n, 1))()
        lib().or_get_contacts(self.h, arr, n)
        return np.ctypeslib.as_array(arr)[:n].copy() if n else np.zeros(0, dtype=np.ctypeslib.as_array(arr).dtype)

    def load_state(self, st):
        """Continues from a state exported by the device library (rphys.World.export_state)."""
        c = lambda a, t=np.float32: np.ascontiguousarray(a, t)
        pos, rot, lin, ang, last = c(st["pos"]), c(st["rot"]), c(st["lin"]), c(st["ang"]), c(st["last_lin"])
        awake, sticks = c(st["awake"], np.uint32), c(st["sticks"], np.uint32)
        wk, wla, wjn, wjt = c(st["wkeys"], np.uint64), c(st["wla"]), c(st["wjn"]), c(st["wjt"])
        U = C.POINTER(C.c_uint32)
        lib().or_load_state(self.h, fptr(pos), fptr(rot), fptr(lin), fptr(ang), fptr(last), awake.ctypes.data_as(U),
                            sticks.ctypes.data_as(U), int(st["tick"]), wk.ctypes.data_as(C.POINTER(C.c_uint64)), fptr(wla),
                            fptr(wjn), fptr(wjt), len(wk))

    def joint_warm(self, q):
        """Joint q's stored rows as (tags, impulses)."""
        tags = (C.c_uint8 * 32)()  # a D6 joint holds up to 30 rows
        lams = np.zeros(32, np.float32)
        n = lib().or_get_joint_warm(self.h, q, tags, fptr(lams))
        return list(tags)[:n], lams[:n].copy()

    def snapshot(self):
        return lib().or_snapshot(self.h)

    def restore(self, snap):
        lib().or_restore(self.h, snap)

    def free_snapshot(self, snap):
        lib().or_snapshot_free(snap)
