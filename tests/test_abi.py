"""The C ABI without a GPU: every symbol include/riftphys_c.h declares is exported, the v1 struct
layouts match the reference (ffi/src/tests.rs:1416-1432), and the host-side error conventions
(lib.rs:96-447, 1103-1141) hold.  None of these calls reach the device."""
import ctypes as C
import os
import re
import subprocess
import tempfile

import numpy as np
import pytest

import rphys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "riftphys_c.h")


def declared_functions():
    src = open(HEADER).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    names = re.findall(r"\b(rphys_[a-z0-9_]+)\s*\(", src)
    return sorted(set(names))


def test_every_declared_symbol_is_exported():
    L = rphys.lib()
    names = declared_functions()
    assert len(names) >= 45
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing
    bound = {e[0] for e in rphys.EXPORTS}
    assert set(names) == bound, set(names) ^ bound


def test_v1_layouts_match_the_reference():
    prog = r"""
#include <stdio.h>
#include <stddef.h>
#include "riftphys_c.h"
int main(void) {
  printf("%zu %zu %zu %zu %zu %zu %zu %zu %zu %zu %zu\n", sizeof(RPhysVec3), sizeof(RPhysQuat), sizeof(RPhysIsometry),
         sizeof(RPhysVelocity), sizeof(RPhysHash32), sizeof(RPhysStepParams), sizeof(RPhysHeightfieldDesc),
         sizeof(RPhysContactEvent), sizeof(RPhysBoxBodyDesc), offsetof(RPhysHeightfieldDesc, heights), sizeof(RPhysMaterial));
  return 0;
}
"""
    with tempfile.TemporaryDirectory() as d:
        c = os.path.join(d, "l.c")
        open(c, "w").write(prog)
        exe = os.path.join(d, "l")
        subprocess.run(["gcc", "-I", os.path.join(ROOT, "include"), c, "-o", exe], check=True)
        out = subprocess.run([exe], check=True, capture_output=True, text=True).stdout.split()
    vec3, quat, iso, vel, h32, sp, hfd, ce, box, hoff, mat = map(int, out)
    assert (vec3, quat, iso, vel, h32) == (12, 16, 28, 24, 32)
    assert sp == 20 and hfd == 48 and hoff == 40 and ce == 36 and mat == 12
    assert box == 28 + 24 + 16 + 4 + 12 + 4
    assert C.sizeof(rphys.StepParams) == sp and C.sizeof(rphys.HeightfieldDesc) == hfd and C.sizeof(rphys.ContactEvent) == ce
    assert C.sizeof(rphys.BoxBodyDesc) == box


def test_player_controller_entries_validate_as_the_reference():
    """riftphys_c.h:227-243 / lib.rs:1580-1693: argument checks, the player index counting from 0,
    one body and one collider consumed (capacity), get_pose answering from the host mirror before
    any step.  (No call here reaches the device.)"""
    L = rphys.lib()
    assert C.sizeof(rphys.PlayerDesc) == 40
    w = L.rphys_world_create(2, 2)
    try:
        d = rphys.PlayerDesc()
        d.start_pose = rphys.iso([0, 1, 0], [0, 0, 0, 1])
        d.radius, d.height, d.speed = 0.3, 1.8, 4.0
        idx = C.c_uint32(7)
        assert L.rphys_add_player(w, None, C.byref(idx)) == rphys.ERR_NULL and idx.value == 0xFFFFFFFF
        d.radius = 0.0
        assert L.rphys_add_player(w, C.byref(d), C.byref(idx)) == rphys.ERR_BAD_ARG
        d.radius = 0.3
        d.start_pose.pos.x = float("nan")
        assert L.rphys_add_player(w, C.byref(d), C.byref(idx)) == rphys.ERR_BAD_ARG
        d.start_pose.pos.x = 0.0
        assert L.rphys_player_set_input(w, 0, rphys.Vec3(1, 0, 0)) == rphys.ERR_BAD_ARG  # no player yet
        assert L.rphys_add_player(w, C.byref(d), C.byref(idx)) == rphys.OK and idx.value == 0
        d.start_pose.pos.x = 3.0
        assert L.rphys_add_player(w, C.byref(d), C.byref(idx)) == rphys.OK and idx.value == 1
        assert L.rphys_add_player(w, C.byref(d), C.byref(idx)) == rphys.ERR_CAPACITY  # two bodies at most
        n = C.c_uint32()
        assert L.rphys_world_num_bodies(w, C.byref(n)) == rphys.OK and n.value == 2
        assert L.rphys_player_set_input(w, 1, rphys.Vec3(1, 0, 0)) == rphys.OK
        assert L.rphys_player_set_input(w, 2, rphys.Vec3(1, 0, 0)) == rphys.ERR_BAD_ARG
        assert L.rphys_player_set_input(w, 0, rphys.Vec3(float("inf"), 0, 0)) == rphys.ERR_BAD_ARG
        assert L.rphys_player_set_input(None, 0, rphys.Vec3(1, 0, 0)) == rphys.ERR_NULL
        pose = rphys.Isometry()
        g = C.c_uint32(9)
        assert L.rphys_player_get_pose(w, 1, C.byref(pose), C.byref(g)) == rphys.OK
        assert (pose.pos.x, pose.pos.y, pose.rot.w, g.value) == (3.0, 1.0, 1.0, 0)
        pose.pos.x = 5.0
        assert L.rphys_player_get_pose(w, 2, C.byref(pose), C.byref(g)) == rphys.ERR_BAD_ARG
        assert (pose.pos.x, pose.rot.w, g.value) == (0.0, 1.0, 0)
        assert L.rphys_player_get_pose(w, 0, None, None) == rphys.ERR_NULL
    finally:
        L.rphys_world_destroy(w)


def box_desc(**kw):
    d = rphys.BoxBodyDesc()
    d.pose = rphys.iso([0, 1, 0], [0, 0, 0, 1])
    d.vel = rphys.vel([0, 0, 0], [0, 0, 0])
    d.hx = d.hy = d.hz = 0.5
    d.mass = 1.0
    d.body_type = rphys.BODY_DYNAMIC
    for k, v in kw.items():
        setattr(d, k, v)
    return d


def last_error():
    m = rphys.lib().rphys_last_error_message()
    return m.decode() if m else ""


def test_abi_version_and_null_handling():
    L = rphys.lib()
    assert L.rphys_get_abi_version() == 1
    out = C.c_uint32(7)
    assert L.rphys_add_box_body(None, C.byref(box_desc()), C.byref(out)) == rphys.ERR_NULL
    assert out.value == 0
    assert L.rphys_world_step_ex(None, None) == rphys.ERR_NULL
    assert L.rphys_world_step(None, 0.016) == rphys.ERR_NULL
    h = rphys.Hash32()
    assert L.rphys_world_get_step_hash(None, C.byref(h)) == rphys.ERR_NULL
    t = C.c_uint64(5)
    assert L.rphys_world_tick_index(None, C.byref(t)) == rphys.ERR_NULL and t.value == 0


def test_body_ids_validation_and_capacity():
    L = rphys.lib()
    w = L.rphys_world_create(2, 2)
    try:
        a, b = C.c_uint32(), C.c_uint32()
        assert L.rphys_add_box_body(w, C.byref(box_desc()), C.byref(a)) == rphys.OK and a.value == 1
        assert L.rphys_add_box_body(w, C.byref(box_desc()), C.byref(b)) == rphys.OK and b.value == 2
        c = C.c_uint32(9)
        assert L.rphys_add_box_body(w, C.byref(box_desc()), C.byref(c)) == rphys.ERR_CAPACITY and c.value == 0
        assert "capacity" in last_error()
        p = rphys.Isometry()
        assert L.rphys_body_get_pose(w, 0, C.byref(p)) == rphys.ERR_INVALID_ID
        assert p.rot.w == 1.0  # identity written on failure
        assert L.rphys_body_get_pose(w, rphys.TERRAIN_BODY_ID, C.byref(p)) == rphys.ERR_INVALID_ID
        assert L.rphys_body_get_pose(w, 3, C.byref(p)) == rphys.ERR_INVALID_ID
        assert L.rphys_body_get_pose(w, 1, C.byref(p)) == rphys.OK and abs(p.pos.y - 1.0) < 1e-7
        assert L.rphys_body_remove(w, 1) == rphys.OK
        assert L.rphys_body_get_pose(w, 1, C.byref(p)) == rphys.ERR_INVALID_ID and "removed" in last_error()
        n = C.c_uint32()
        assert L.rphys_world_num_bodies(w, C.byref(n)) == rphys.OK and n.value == 2
    finally:
        L.rphys_world_destroy(w)


def test_bad_arguments_are_rejected():
    L = rphys.lib()
    w = L.rphys_world_create(8, 8)
    try:
        out = C.c_uint32()
        assert L.rphys_add_box_body(w, C.byref(box_desc(hx=-1.0)), C.byref(out)) == rphys.ERR_BAD_ARG
        assert L.rphys_add_box_body(w, C.byref(box_desc(mass=0.0)), C.byref(out)) == rphys.ERR_BAD_ARG
        assert L.rphys_add_box_body(w, C.byref(box_desc(body_type=7)), C.byref(out)) == rphys.ERR_BAD_ARG
        d = box_desc()
        d.pose.pos.x = float("nan")
        assert L.rphys_add_box_body(w, C.byref(d), C.byref(out)) == rphys.ERR_BAD_ARG
        sp = rphys.StepParams(8, 0, 1, 0, 0.016)
        assert L.rphys_world_step_ex(w, C.byref(sp)) == rphys.ERR_BAD_ARG and "size_bytes" in last_error()
        sp = rphys.StepParams(20, 0, 0, 0, 0.016)
        assert L.rphys_world_step_ex(w, C.byref(sp)) == rphys.ERR_BAD_ARG
        sp = rphys.StepParams(20, 0, 1, 0, -1.0)
        assert L.rphys_world_step_ex(w, C.byref(sp)) == rphys.ERR_BAD_ARG
        sp = rphys.StepParams(20, 0, 1, 0, float("inf"))
        assert L.rphys_world_step_ex(w, C.byref(sp)) == rphys.ERR_BAD_ARG
        # a static body needs no mass (lib.rs:304-312)
        assert L.rphys_add_box_body(w, C.byref(box_desc(mass=0.0, body_type=rphys.BODY_STATIC)), C.byref(out)) == rphys.OK
        bad = np.zeros(3, np.float32)
        assert L.rphys_world_set_heightfield_raw32_square(w, bad.ctypes.data, 12, 1, 1, 0, 0, 0) == rphys.ERR_BAD_ARG
        # contact models: material ids within the table, a finite non-negative stiffness, a finite dissipation
        assert L.rphys_world_set_pair_contact_model(None, 0, 0, 1.0, 0.0, 0) == rphys.ERR_NULL
        assert L.rphys_world_set_pair_contact_model(w, 16, 0, 1.0, 0.0, 0) == rphys.ERR_BAD_ARG
        assert L.rphys_world_set_pair_contact_model(w, 0, 0, -1.0, 0.0, 0) == rphys.ERR_BAD_ARG
        assert L.rphys_world_set_pair_contact_model(w, 0, 0, float("nan"), 0.0, 0) == rphys.ERR_BAD_ARG
        assert L.rphys_world_set_pair_contact_model(w, 0, 0, 1.0e4, float("inf"), 0) == rphys.ERR_BAD_ARG
        assert L.rphys_world_set_pair_contact_model(w, 3, 6, 1.0e4, 0.5, 1) == rphys.OK
        assert L.rphys_world_set_compliant_damping_at_start(None, 1) == rphys.ERR_NULL
        assert L.rphys_world_set_compliant_damping_at_start(w, 1) == rphys.OK
        # stepping an empty-of-dynamics world is allowed without a device only when it has no bodies
        L2 = L.rphys_world_create(4, 4)
        assert L.rphys_world_step(L2, 0.016) == rphys.OK
        t = C.c_uint64()
        assert L.rphys_world_tick_index(L2, C.byref(t)) == rphys.OK and t.value == 1
        L.rphys_world_destroy(L2)
    finally:
        L.rphys_world_destroy(w)


def test_last_error_is_thread_local():
    import threading

    L = rphys.lib()
    L.rphys_world_step(None, 0.016)
    seen = []

    def other():
        seen.append(L.rphys_last_error_message())

    th = threading.Thread(target=other)
    th.start(); th.join()
    assert seen == [None]
    assert last_error() != ""
    buf = C.create_string_buffer(8)
    n = C.c_uint32()
    assert L.rphys_last_error_copy(buf, 8, C.byref(n)) == rphys.OK and n.value == len(last_error())
    assert buf.value == last_error().encode()[:7]


def test_query_argument_checks():
    """lib.rs:1270-1410: null -> ERR_NULL, non-finite input -> ERR_BAD_ARG, a dead or sentinel
    ignore_body -> ERR_INVALID_ID, max_distance <= 0 or a zero direction -> a miss; outputs are the
    miss record on every failure.  (None of these reach the device.)"""
    L = rphys.lib()
    w = L.rphys_world_create(4, 4)
    try:
        nan = float("nan")
        hit = rphys.RayHit()
        q = rphys.RayQuery(rphys.Vec3(0, 1, 0), rphys.Vec3(0, -1, 0), 5.0, 0)
        assert L.rphys_world_raycast(None, C.byref(q), C.byref(hit)) == rphys.ERR_NULL
        assert L.rphys_world_raycast(w, None, C.byref(hit)) == rphys.ERR_NULL
        assert (hit.hit, hit.body_id, hit.fraction, hit.normal.y) == (0, 0, 1.0, 1.0)
        for bad in (rphys.RayQuery(rphys.Vec3(nan, 1, 0), rphys.Vec3(0, -1, 0), 5.0, 0),
                    rphys.RayQuery(rphys.Vec3(0, 1, 0), rphys.Vec3(0, float("inf"), 0), 5.0, 0),
                    rphys.RayQuery(rphys.Vec3(0, 1, 0), rphys.Vec3(0, -1, 0), nan, 0)):
            assert L.rphys_world_raycast(w, C.byref(bad), C.byref(hit)) == rphys.ERR_BAD_ARG
            assert hit.hit == 0 and hit.fraction == 1.0
        for ig in (7, rphys.TERRAIN_BODY_ID):
            q.ignore_body = ig
            assert L.rphys_world_raycast(w, C.byref(q), C.byref(hit)) == rphys.ERR_INVALID_ID
        q.ignore_body = 0
        q.max_distance = 0.0
        assert L.rphys_world_raycast(w, C.byref(q), C.byref(hit)) == rphys.OK and hit.hit == 0
        q.max_distance, q.dir = 5.0, rphys.Vec3(0, 0, 0)
        assert L.rphys_world_raycast(w, C.byref(q), C.byref(hit)) == rphys.OK and hit.hit == 0
        sh = rphys.CapsuleSweepHit()
        sq = rphys.CapsuleSweepQuery(rphys.Vec3(0, 1, 0), rphys.Vec3(0, 0, 0), 0.3, 0.5, 0)
        assert L.rphys_world_capsule_sweep(w, None, C.byref(sh)) == rphys.ERR_NULL
        for r, hh in ((0.0, 0.5), (-1.0, 0.5), (nan, 0.5), (0.3, -0.1), (0.3, nan)):
            sq.radius, sq.half_height = r, hh
            assert L.rphys_world_capsule_sweep(w, C.byref(sq), C.byref(sh)) == rphys.ERR_BAD_ARG
            assert (sh.hit, sh.fraction, sh.started_overlapping) == (0, 1.0, 0)
        sq.radius, sq.half_height, sq.ignore_body = 0.3, 0.5, 3
        assert L.rphys_world_capsule_sweep(w, C.byref(sq), C.byref(sh)) == rphys.ERR_INVALID_ID
        sq.ignore_body = 0
        # an empty world with no terrain: nothing to hit, answered without the device
        assert L.rphys_world_capsule_sweep(w, C.byref(sq), C.byref(sh)) == rphys.OK and sh.hit == 0
        qs = (rphys.RayQuery * 2)(q, rphys.RayQuery(rphys.Vec3(0, 1, 0), rphys.Vec3(nan, 0, 0), 1.0, 0))
        hs = (rphys.RayHit * 2)()
        assert L.rphys_world_raycast_n(w, qs, 2, hs) == rphys.ERR_BAD_ARG
        assert hs[0].hit == 0 and hs[1].hit == 0
    finally:
        L.rphys_world_destroy(w)


def test_ball_joint_export_validates_its_arguments():
    L = rphys.lib()
    w = L.rphys_world_create(4, 4)
    try:
        a, b = C.c_uint32(), C.c_uint32()
        assert L.rphys_add_box_body(w, C.byref(box_desc()), C.byref(a)) == rphys.OK
        assert L.rphys_add_box_body(w, C.byref(box_desc()), C.byref(b)) == rphys.OK
        fa, fb = rphys.iso((0, 0.5, 0), (0, 0, 0, 1)), rphys.iso((0, -0.5, 0), (0, 0, 0, 1))
        j = C.c_uint32(7)
        assert L.rphys_world_add_ball_joint(None, a, b, C.byref(fa), C.byref(fb), C.byref(j)) == rphys.ERR_NULL
        assert L.rphys_world_add_ball_joint(w, a, b, None, C.byref(fb), C.byref(j)) == rphys.ERR_NULL
        assert L.rphys_world_add_ball_joint(w, 0, b, C.byref(fa), C.byref(fb), C.byref(j)) == rphys.ERR_INVALID_ID
        assert j.value == 0xFFFFFFFF
        assert L.rphys_world_add_ball_joint(w, rphys.TERRAIN_BODY_ID, b, C.byref(fa), C.byref(fb), C.byref(j)) == rphys.ERR_INVALID_ID
        assert L.rphys_world_add_ball_joint(w, a, a, C.byref(fa), C.byref(fb), C.byref(j)) == rphys.ERR_BAD_ARG
        bad = rphys.iso((0, float("nan"), 0), (0, 0, 0, 1))
        assert L.rphys_world_add_ball_joint(w, a, b, C.byref(bad), C.byref(fb), C.byref(j)) == rphys.ERR_BAD_ARG
        assert L.rphys_world_num_joints(w) == 0
        assert L.rphys_world_add_ball_joint(w, a, b, C.byref(fa), C.byref(fb), C.byref(j)) == rphys.OK and j.value == 0
        assert L.rphys_world_add_ball_joint(w, b, a, C.byref(fb), C.byref(fa), C.byref(j)) == rphys.OK and j.value == 1
        assert L.rphys_world_num_joints(w) == 2
        assert L.rphys_body_remove(w, a.value) == rphys.OK
        assert L.rphys_world_add_ball_joint(w, a, b, C.byref(fa), C.byref(fb), C.byref(j)) == rphys.ERR_INVALID_ID
    finally:
        L.rphys_world_destroy(w)


def test_d6_joint_exports_validate_their_arguments():
    L = rphys.lib()
    w = L.rphys_world_create(4, 4)
    try:
        a, b = C.c_uint32(), C.c_uint32()
        assert L.rphys_add_box_body(w, C.byref(box_desc()), C.byref(a)) == rphys.OK
        assert L.rphys_add_box_body(w, C.byref(box_desc()), C.byref(b)) == rphys.OK
        fa, fb = rphys.iso((0, 0.5, 0), (0, 0, 0, 1)), rphys.iso((0, -0.5, 0), (0, 0, 0, 1))
        j = C.c_uint32(7)
        assert L.rphys_world_add_d6_joint(w, a, b, C.byref(fa), C.byref(fb), 8, 0, -1, C.byref(j)) == rphys.ERR_BAD_ARG
        assert j.value == 0xFFFFFFFF
        assert L.rphys_world_add_d6_joint(w, a, b, C.byref(fa), C.byref(fb), 7, 0, 3, C.byref(j)) == rphys.ERR_BAD_ARG
        assert L.rphys_world_add_d6_joint(w, a, a, C.byref(fa), C.byref(fb), 7, 0, -1, C.byref(j)) == rphys.ERR_BAD_ARG
        assert L.rphys_world_add_d6_joint(None, a, b, C.byref(fa), C.byref(fb), 7, 0, -1, C.byref(j)) == rphys.ERR_NULL
        assert L.rphys_world_add_hinge_joint(w, a, b, C.byref(fa), C.byref(fb), 3, C.byref(j)) == rphys.ERR_BAD_ARG
        assert L.rphys_world_num_joints(w) == 0
        assert L.rphys_world_add_ball_joint(w, a, b, C.byref(fa), C.byref(fb), C.byref(j)) == rphys.OK and j.value == 0
        assert L.rphys_world_add_hinge_joint(w, a, b, C.byref(fa), C.byref(fb), 2, C.byref(j)) == rphys.OK and j.value == 1
        assert L.rphys_world_add_d6_joint(w, b, a, C.byref(fb), C.byref(fa), 7, 5, -1, C.byref(j)) == rphys.OK and j.value == 2
        assert L.rphys_world_num_joints(w) == 3
        # setters: a ball joint has no axes; unknown joints; axes 0..2; finite values
        assert L.rphys_world_set_joint_limit(w, 0, 1, 2, 1, -1.0, 1.0) == rphys.ERR_BAD_ARG and "D6" in last_error()
        assert L.rphys_world_set_joint_limit(w, 9, 1, 2, 1, -1.0, 1.0) == rphys.ERR_INVALID_ID
        assert L.rphys_world_set_joint_limit(w, 1, 1, 3, 1, -1.0, 1.0) == rphys.ERR_BAD_ARG
        assert L.rphys_world_set_joint_limit(w, 1, 1, 2, 1, float("nan"), 1.0) == rphys.ERR_BAD_ARG
        assert L.rphys_world_set_joint_limit(w, 1, 1, 2, 0, float("nan"), 1.0) == rphys.OK  # removing needs no bounds
        assert L.rphys_world_set_joint_limit(w, 1, 1, 2, 1, -1.0, 1.0) == rphys.OK
        assert L.rphys_world_set_joint_drive(w, 1, 1, 2, 1, float("inf"), 0.0, 0.0, 1.0, 1) == rphys.ERR_BAD_ARG
        assert L.rphys_world_set_joint_drive(w, 2, 0, 1, 1, 0.5, 0.0, 0.0, 1.0, 0) == rphys.OK
        assert L.rphys_world_set_joint_drive(None, 2, 0, 1, 1, 0.5, 0.0, 0.0, 1.0, 0) == rphys.ERR_NULL
        assert L.rphys_world_set_joint_damping(w, 1, 1, 2, -3.0, float("nan")) == rphys.OK  # clamped to 0, as the reference does
        assert L.rphys_world_set_joint_damping(w, 0, 1, 2, 1.0, 1.0) == rphys.ERR_BAD_ARG
        assert L.rphys_world_set_joint_compliance(w, 1, 1, 2, float("inf")) == rphys.ERR_BAD_ARG
        assert L.rphys_world_set_joint_compliance(w, 1, 1, 2, 0.01) == rphys.OK
        assert L.rphys_world_set_joint_compliance(w, 7, 1, 2, 0.01) == rphys.ERR_INVALID_ID
        # distance joints: the ball joint's body checks; a distance joint has no D6 axes
        assert L.rphys_world_add_distance_joint(w, a, a, 1.0, 0.0, C.byref(j)) == rphys.ERR_BAD_ARG and j.value == 0xFFFFFFFF
        assert L.rphys_world_add_distance_joint(w, 9, b, 1.0, 0.0, C.byref(j)) == rphys.ERR_INVALID_ID
        assert L.rphys_world_add_distance_joint(None, a, b, 1.0, 0.0, C.byref(j)) == rphys.ERR_NULL
        assert L.rphys_world_add_distance_joint(w, a, b, float("nan"), -1.0, C.byref(j)) == rphys.OK and j.value == 3
        assert L.rphys_world_set_joint_limit(w, 3, 0, 0, 1, -1.0, 1.0) == rphys.ERR_BAD_ARG
        assert L.rphys_world_num_joints(w) == 4
    finally:
        L.rphys_world_destroy(w)
