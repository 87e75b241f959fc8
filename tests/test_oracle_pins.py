"""More pins of the oracle (CPU only): sleep, the capsule narrowphase cases in closed form,
warm-start matching, multi-chunk BLAKE3 against an independent restatement, and the step hash's
byte stream restated from provenance.rs:6-24 / world.rs:882-909."""
import ctypes as C
import math
import struct

import numpy as np
import pytest

import blake3_py
import oracle_py as O
import scenes

f32 = lambda a: np.ascontiguousarray(a, np.float32)
IDQ = [0.0, 0.0, 0.0, 1.0]
DT = float(np.float32(1.0 / 60.0))


# ---------------------------------------------------------------- sleep (sleep.rs:59-80)
def _free_world(vels, allow=None):
    L = O.lib()
    h = L.or_world_create(8)
    L.or_set_gravity(h, 0.0, 0.0, 0.0)
    for k, v in enumerate(vels):
        L.or_add_body_nc(h, O.OR_SPHERE, O.fptr(f32([0.5, 0, 0])), O.fptr(f32([3.0 * k, 0, 0] + IDQ)), O.fptr(f32(v)), 1, 1000.0)
        if allow is not None and not allow[k]:
            L.or_set_sleep_allowed(h, k, 0)
    return h


def _awake(h, n):
    out = (C.c_uint8 * n)()
    O.lib().or_get_awake(h, out, n)
    return list(out)


def test_body_falls_asleep_after_31_quiet_ticks():
    """|v|^2 < 0.01 and |w|^2 < 0.01 for more than 30 consecutive ticks puts a body to sleep;
    a body just above either threshold, or one not allowed to sleep, stays awake; a moving tick
    restarts the count."""
    L = O.lib()
    vels = [[0.09, 0, 0, 0, 0, 0], [0.1001, 0, 0, 0, 0, 0], [0, 0, 0, 0, 0.1001, 0], [0.09, 0, 0, 0, 0, 0], [0, 0, 0, 0.099, 0, 0]]
    h = _free_world(vels, allow=[1, 1, 1, 0, 1])
    try:
        for t in range(30):
            L.or_step(h, DT)
        assert _awake(h, 5) == [1, 1, 1, 1, 1]
        L.or_set_body_vel(h, 4, O.fptr(f32([0.2, 0, 0, 0, 0, 0])))  # body 4 moves for one tick
        L.or_step(h, DT)                                    # tick 31
        assert _awake(h, 5) == [0, 1, 1, 1, 1]
        L.or_set_body_vel(h, 4, O.fptr(f32([0.0, 0, 0, 0, 0, 0])))
        for t in range(30):                                          # ticks 32..61: body 4 quiet
            L.or_step(h, DT)
        assert _awake(h, 5) == [0, 1, 1, 1, 1]
        pos0 = np.zeros((5, 3), np.float32)
        L.or_get_state(h, O.fptr(pos0), None, None, None)
        L.or_step(h, DT)                                    # tick 62: its 31st quiet tick
        assert _awake(h, 5) == [0, 1, 1, 1, 0]
        pos1 = np.zeros((5, 3), np.float32)
        L.or_get_state(h, O.fptr(pos1), None, None, None)
        assert (pos1[0] == pos0[0]).all() and pos1[1, 0] > pos0[1, 0]  # a sleeping body is not integrated
    finally:
        L.or_world_destroy(h)


# ---------------------------------------------------------------- capsule narrowphase in closed form
def collide(sa, da, pa, sb, db, pb, cd=0.0):
    out = np.zeros(28, np.float32)
    n = O.lib().or_collide(sa, O.fptr(f32(da)), O.fptr(f32(pa)), sb, O.fptr(f32(db)), O.fptr(f32(pb)), cd, O.fptr(out))
    return out[: 7 * n].reshape(n, 7)


def qz(angle):
    return [0.0, 0.0, math.sin(angle / 2), math.cos(angle / 2)]


def test_capsule_sphere_side_and_cap_regions():
    """narrowphase.rs:511-522: the closest point on the capsule's segment, then sphere-sphere:
    normal from the segment point to the sphere centre, penetration r_a + r_b - distance, point
    halfway between the two surface points."""
    cap, sph = [0.3, 0.5, 0.0], [0.25, 0, 0]
    p = collide(O.OR_CAPSULE, cap, [0, 0, 0] + IDQ, O.OR_SPHERE, sph, [0.5, 0.2, 0] + IDQ)
    assert len(p) == 1
    np.testing.assert_allclose(p[0, 3:6], [1, 0, 0], atol=1e-6)
    assert abs(p[0, 6] - 0.05) < 1e-6
    np.testing.assert_allclose(p[0, :3], [0.275, 0.2, 0.0], atol=1e-6)
    # beyond the segment's end: the end point is the closest
    p = collide(O.OR_CAPSULE, cap, [0, 0, 0] + IDQ, O.OR_SPHERE, sph, [0.1, 1.0, 0] + IDQ)
    d = math.hypot(0.1, 0.5)
    np.testing.assert_allclose(p[0, 3:6], [0.1 / d, 0.5 / d, 0], atol=1e-6)
    assert abs(p[0, 6] - (0.55 - d)) < 1e-6
    # the capsule turned onto the x axis, and the order swapped (normal from A to B)
    p = collide(O.OR_SPHERE, sph, [0.2, 0.5, 0] + IDQ, O.OR_CAPSULE, cap, [0, 0, 0] + qz(-math.pi / 2))
    np.testing.assert_allclose(p[0, 3:6], [0, -1, 0], atol=1e-6)
    assert abs(p[0, 6] - 0.05) < 1e-6
    assert not len(collide(O.OR_CAPSULE, cap, [0, 0, 0] + IDQ, O.OR_SPHERE, sph, [0.56, 0.2, 0] + IDQ))


def test_box_capsule_deepest_test_point_wins():
    """narrowphase.rs:595-622: box-sphere at the capsule's two end points and at the segment point
    nearest the box centre; the deepest wins, the earlier one on ties."""
    box, cap, r, hh = [1.0, 0.5, 1.0], [0.3, 0.5, 0.0], 0.3, 0.5
    # level, 4 cm into the top face: all three test points tie, the first end point (pos - axis h) wins
    p = collide(O.OR_BOX, box, [0, 0, 0] + IDQ, O.OR_CAPSULE, cap, [0, 0.5 + r - 0.04, 0] + qz(-math.pi / 2))
    assert len(p) == 1
    np.testing.assert_allclose(p[0, 3:6], [0, 1, 0], atol=1e-6)
    assert abs(p[0, 6] - 0.04) < 1e-5
    axis = np.array([math.cos(-math.pi / 2 + math.pi / 2), math.sin(-math.pi / 2 + math.pi / 2), 0])  # R(-90 deg) y = +x
    first_end_x = -hh * axis[0]
    assert abs(p[0, 0] - first_end_x) < 1e-5
    # tilted: the lower end point is the deepest
    tilt = 0.2
    q = qz(-math.pi / 2 + tilt)
    ax = np.array([math.sin(math.pi / 2 - tilt), math.cos(math.pi / 2 - tilt), 0.0])  # rot * +y
    cy = 0.5 + r - 0.02
    ends = [np.array([0, cy, 0]) - hh * ax, np.array([0, cy, 0]) + hh * ax]
    low = min(ends, key=lambda e: e[1])
    p = collide(O.OR_BOX, box, [0, 0, 0] + IDQ, O.OR_CAPSULE, cap, [0, cy, 0] + q)
    assert abs(p[0, 6] - (r - (low[1] - 0.5))) < 1e-5
    assert abs(p[0, 0] - low[0]) < 1e-5  # the point lies under that end


def test_capsule_triangle_deepest_of_three_points():
    """narrowphase.rs:709-733: sphere-triangle at the end points and the midpoint; the deepest wins.
    The normal runs from the triangle to the sphere centre (narrowphase.rs:690-697)."""
    tri = f32([-20, 0, -20, -20, 0, 40, 40, 0, -20])  # x + z <= 20, normal (b-a)x(c-a) = +y
    r, hh, tilt = 0.3, 0.5, 0.3
    q = qz(tilt)  # the axis leans towards -x: the +y end moves to -x, the -y end to +x and down
    ax = np.array([-math.sin(tilt), math.cos(tilt), 0.0])
    c = np.array([1.0, r + hh * math.cos(tilt) - 0.03, 2.0])
    low = c - hh * ax
    out = np.zeros(28, np.float32)
    n = O.lib().or_collide_shape_triangle(O.OR_CAPSULE, O.fptr(f32([r, hh, 0])), O.fptr(f32(list(c) + q)), O.fptr(tri), 0.0, O.fptr(out))
    assert n == 1
    np.testing.assert_allclose(out[3:6], [0, 1, 0], atol=2e-5)  # closest-point rounding on a 60 m triangle
    assert abs(out[6] - (r - low[1])) < 1e-5
    np.testing.assert_allclose(out[:3], [low[0], 0.0, low[2]], atol=1e-5)


# ---------------------------------------------------------------- warm-start matching (solver.rs:717-762)
def test_warm_match_nearest_unused_within_1cm_lowest_index_on_ties():
    key = [5, 0xFFFFFFFF, 0, 0]
    other = [5, 7, 0, 0]
    cache_keys = [other, key, key, key, key, key]  # sorted: (5, 7) < (5, 0xFFFFFFFF)
    cache_keys = sorted(cache_keys)
    la = [[0.0, 0, 0], [0.008, 0, 0], [0.003, 0, 0], [0.0, 0.003, 0], [0.0201, 0, 0], [0.5, 0, 0]]
    queries = [key, key, key, key, other, [6, 0xFFFFFFFF, 0, 0]]
    anchors = [[0.0, 0, 0], [0.0, 0, 0], [0.02, 0, 0], [0.0, 0, 0], [0.0, 0, 0], [0.0, 0, 0]]
    out = np.zeros(len(queries), np.int32)
    O.lib().or_warm_pick_sequence(np.ascontiguousarray(cache_keys, np.uint32).ctypes.data_as(C.POINTER(C.c_uint32)), O.fptr(f32(la)), len(la),
                                  np.ascontiguousarray(queries, np.uint32).ctypes.data_as(C.POINTER(C.c_uint32)), O.fptr(f32(anchors)), len(queries),
                                  out.ctypes.data_as(C.POINTER(C.c_int32)))
    # index 0 is the (5, 7) entry; entries 2 and 3 tie at 3 mm: the first query takes 2, the
    # second the tie's other one, the third the 0.1 mm one at 20.1 mm; the fourth finds only the
    # 8 mm one left; (5, 7)'s entry matches its query; an unknown key matches nothing
    assert list(out) == [2, 3, 4, 1, 0, -1]


# ---------------------------------------------------------------- BLAKE3 (multi-block, multi-chunk)
def test_blake3_matches_an_independent_restatement():
    assert blake3_py.blake3(b"").hex() == "af1349b9f5f9a1a6a0404dea36dcc9499bcb25c9adc112b7cc9a93cae41f3262"
    assert blake3_py.blake3(b"abc").hex() == "6437b3ac38465133ffb63b75273a8db548c558465d79db03fd359c6cd5bd9d85"
    data = bytes(i % 251 for i in range(40000))
    for n in (1, 63, 64, 65, 127, 128, 1023, 1024, 1025, 2047, 2048, 2049, 3072, 3073, 4096, 4097, 5121, 8193, 16384, 31744, 40000):
        assert O.blake3(data[:n]) == blake3_py.blake3(data[:n]), n


# ---------------------------------------------------------------- the step hash's byte stream
def _canon(x):
    x = np.float32(x)
    if not np.isfinite(x):
        return 0x7FC00000
    b = int(x.view(np.uint32))
    return 0 if b == 0x80000000 else b


def _vec(vals):
    return b"".join(struct.pack("<I", _canon(v)) for v in vals)


def restated_step_hash(h, epoch, rng):
    """World::step_hash (provenance.rs:6-24) with hash_carried (world.rs:884-909), from the
    oracle's exported state: epoch, RNG state, the empty schedule's digest, each body's index,
    pose and velocity, then the warm cache by key (BTreeMap order) with per-key entry counts, no
    patches, the stored joint rows by joint index, no rotors, last velocities, sleep flags and
    counters over the sleep slots; every collection prefixed with its length."""
    L = O.lib()
    n, cap = L.or_num_bodies(h), L.or_capacity(h)
    pos = np.zeros((n, 3), np.float32); rot = np.zeros((n, 4), np.float32); lin = np.zeros((n, 3), np.float32); ang = np.zeros((n, 3), np.float32)
    L.or_get_state(h, O.fptr(pos), O.fptr(rot), O.fptr(lin), O.fptr(ang))
    nw = L.or_export_carried(h, None, None, None, None, None, None, None, 0, None, None, None, None, None)
    sl = max(cap, n)
    last = np.zeros((n, 3), np.float32); awake = np.zeros(sl, np.uint8); sticks = np.zeros(sl, np.uint16)
    keys = np.zeros((max(nw, 1), 4), np.uint32); wla = np.zeros((max(nw, 1), 3), np.float32)
    wjn = np.zeros(max(nw, 1), np.float32); wjt = np.zeros((max(nw, 1), 3), np.float32)
    nj = L.or_num_joints(h)
    jq = np.zeros(max(nj, 1), np.uint32); jrows = np.zeros(max(nj, 1), np.uint8); jtags = np.zeros((max(nj, 1), 3), np.uint8)
    jlam = np.zeros((max(nj, 1), 3), np.float32); njw = C.c_uint32(0)
    P8, P16, P32 = C.POINTER(C.c_uint8), C.POINTER(C.c_uint16), C.POINTER(C.c_uint32)
    L.or_export_carried(h, O.fptr(last), awake.ctypes.data_as(P8), sticks.ctypes.data_as(P16), keys.ctypes.data_as(P32), O.fptr(wla),
                        O.fptr(wjn), O.fptr(wjt), nw, jq.ctypes.data_as(P32), jrows.ctypes.data_as(P8), jtags.ctypes.data_as(P8),
                        O.fptr(jlam), C.byref(njw))
    out = struct.pack("<QQ", epoch, rng) + blake3_py.blake3(b"")
    for i in range(n):
        out += struct.pack("<I", i) + _vec(pos[i]) + _vec(rot[i]) + _vec(lin[i]) + _vec(ang[i])
    groups = {}
    for e in range(nw):
        groups.setdefault(tuple(int(k) for k in keys[e]), []).append(e)
    out += struct.pack("<Q", len(groups))
    for k in sorted(groups):
        out += struct.pack("<4I", *k) + struct.pack("<I", len(groups[k]))
        for e in groups[k]:
            out += _vec(wla[e]) + _vec([wjn[e]]) + _vec(wjt[e])
    out += struct.pack("<Q", 0)  # warm patches
    out += struct.pack("<Q", njw.value)
    for k in range(njw.value):
        out += struct.pack("<II", int(jq[k]), int(jrows[k]))
        for r in range(int(jrows[k])):
            out += struct.pack("<B", int(jtags[k, r])) + _vec([jlam[k, r]])
    out += struct.pack("<Q", 0)  # rotors
    out += struct.pack("<Q", n) + b"".join(_vec(v) for v in last)
    out += struct.pack("<Q", sl) + bytes(awake)
    out += struct.pack("<Q", sl) + b"".join(struct.pack("<H", int(t)) for t in sticks)
    return blake3_py.blake3(out)


@pytest.mark.parametrize("kind", ["mixed", "joints"])
def test_step_hash_byte_stream_restated(kind):
    L = O.lib()
    s = scenes.small_mixed(n=48, seed=5) if kind == "mixed" else scenes.ragdolls_small(n_towers=2, tower_h=3, n_ragdolls=2, chain=5)
    o = O.OracleWorld(s, capacity=len(s["shape"]) + 3)  # sleep slots beyond the bodies are hashed too
    L.or_set_epoch(o.h, 0x1122334455667788)
    L.or_set_rng_seed(o.h, 0xABCDEF)
    for t in range(60):
        o.step()
        if t in (0, 29, 59):
            assert o.hash() == restated_step_hash(o.h, 0x1122334455667788, 0xABCDEF | 1), "tick %d" % (t + 1)


def _on_flat_terrain(shape, dims, y, n=17, cell=0.5):
    s = scenes._empty(1)
    s["shape"][0] = shape
    s["dims"][0] = dims
    s["pose"][0] = (4.0, y, 4.0, 0.0, 0.0, 0.0, 1.0)
    s["hf"] = dict(nx=n, nz=n, sx=np.float32(cell), sz=np.float32(cell), ox=0.0, oz=0.0, heights=np.zeros(n * n, np.float32))
    s["pairs"] = list(scenes.CONFIG_PAIRS)
    return s


@pytest.mark.parametrize("name, shape, dims, y, ny", [("sphere", O.OR_SPHERE, (0.3, 0.0, 0.0), 0.31, 1.0),
                                                       ("capsule", O.OR_CAPSULE, (0.2, 0.3, 0.0), 0.51, 1.0),
                                                       ("box", O.OR_BOX, (0.3, 0.3, 0.3), 0.31, -1.0)])
def test_terrain_contact_normal_conventions(name, shape, dims, y, ny):
    """The reference orients a sphere's (and so a capsule's) terrain contact from the triangle to
    the shape (narrowphase.rs:683-693: normal = (center - closest) / dist), a box's from the box to
    the terrain (narrowphase.rs:402), and the solver takes every terrain contact's normal the same
    way (body a the shape, b the terrain).  So a resting sphere's or capsule's terrain rows push it
    toward the terrain in the velocity solve, and the position projection lifts it back: in islands
    where such a row must also hold up other bodies the rows have no common solution, which is the
    residual floor of the capsule-chain worlds.  This pins the two conventions the oracle and the
    device reproduce (a flipped sign fails it)."""
    o = O.OracleWorld(_on_flat_terrain(shape, dims, y))
    for _ in range(20):
        o.step()
    c = o.contacts()
    assert len(c) > 0
    assert all(x["b_col"] == 0xFFFFFFFF for x in c)
    assert np.all(c["normal"][:, 1] == np.float32(ny)), (name, c["normal"])
