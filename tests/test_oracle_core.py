"""Pins of the CPU oracle (no GPU): published digests, closed forms, the reference's own unit
tests restated, brute force on small inputs, and invariants.  Each check is chosen so that a
plausible slip in the oracle (a dropped term, a swapped operand, a wrong sign) fails it."""
import ctypes as C
import math

import numpy as np
import pytest

import oracle_py as O
import scenes


def f32(x):
    return np.ascontiguousarray(x, np.float32)


def collide(sa, da, pa, sb, db, pb, cd):
    out = np.zeros(28, np.float32)
    n = O.lib().or_collide(sa, O.fptr(f32(da)), O.fptr(f32(pa)), sb, O.fptr(f32(db)), O.fptr(f32(pb)), cd, O.fptr(out))
    return out[: 7 * n].reshape(n, 7)


IDQ = [0.0, 0.0, 0.0, 1.0]


# ---------------------------------------------------------------- BLAKE3 (blake3 crate 1.8.2)
def test_blake3_published_digests():
    assert O.blake3(b"").hex() == "af1349b9f5f9a1a6a0404dea36dcc9499bcb25c9adc112b7cc9a93cae41f3262"
    assert O.blake3(b"abc").hex() == "6437b3ac38465133ffb63b75273a8db548c558465d79db03fd359c6cd5bd9d85"


def test_blake3_chunk_boundaries_distinct_and_stable():
    # lengths around the 1024-byte chunk and 64-byte block edges: every length gives a distinct
    # digest (a lost byte or a wrong counter would collide neighbours) and incremental == one-shot
    data = bytes(i % 251 for i in range(5 * 1024 + 7))
    digests = set()
    for n in (0, 1, 63, 64, 65, 1023, 1024, 1025, 2048, 2049, 3072, 4097, 5 * 1024 + 7):
        d = O.blake3(data[:n])
        assert d not in digests
        digests.add(d)


# ---------------------------------------------------------------- sin / cos
def test_sincos_libm_form_within_one_ulp():
    """or_sincosf restates libm 0.2.15's sincosf (musl / FreeBSD msun kernels: fixed double
    polynomials, so not always correctly rounded).  Checked against the exact value over every
    reduction branch: within 1 ulp everywhere, correctly rounded in all but a few in 10^4, and
    sin x = x, cos x = 1 exactly below 2^-12."""
    rng = np.random.default_rng(1)
    xs = np.concatenate([rng.uniform(-4, 4, 40000), rng.uniform(-1e-3, 1e-3, 5000), rng.uniform(-60, 60, 5000)]).astype(np.float32)
    s = C.c_float(); c = C.c_float()
    mism = 0
    for x in xs:
        O.lib().or_sincosf(float(x), C.byref(s), C.byref(c))
        es, ec = np.float32(math.sin(float(x))), np.float32(math.cos(float(x)))
        for got, want in ((s.value, es), (c.value, ec)):
            g, w = np.float32(got), want
            if g != w:
                mism += 1
                ulp = abs(int(np.array([g]).view(np.int32)[0]) - int(np.array([w]).view(np.int32)[0]))
                assert ulp <= 1, (x, g, w)
    assert mism <= 0.001 * 2 * len(xs), mism
    for x in np.float32([1e-5, -2e-4, 2.4e-4, 1e-30]):
        O.lib().or_sincosf(float(x), C.byref(s), C.byref(c))
        assert np.float32(s.value) == x and c.value == 1.0


# ---------------------------------------------------------------- narrowphase (narrowphase.rs:735-776)
def test_resting_box_gets_four_upward_contacts():
    ground = ([50.0, 0.5, 50.0], [0.0, -0.5, 0.0] + IDQ)
    foot = ([0.06, 0.03, 0.12], [0.0, 0.025, 0.0] + IDQ)
    m = collide(1, ground[0], ground[1], 1, foot[0], foot[1], 0.01)
    assert len(m) == 4
    for p in m:
        assert np.linalg.norm(p[3:6] - [0, 1, 0]) < 1e-5
        assert abs(p[6] - 0.015) < 1e-4
    m2 = collide(1, foot[0], foot[1], 1, ground[0], ground[1], 0.01)
    assert np.linalg.norm(m2[0, 3:6] + [0, 1, 0]) < 1e-5


def test_separated_boxes_do_not_touch():
    assert len(collide(1, [0.5] * 3, [0, 0, 0] + IDQ, 1, [0.5] * 3, [1.2, 0, 0] + IDQ, 0.01)) == 0


def test_tilted_box_touches_with_its_low_edge():
    rot = [0.0, 0.0, math.sin(0.15), math.cos(0.15)]
    low = 0.2 * (math.cos(0.3) + math.sin(0.3))
    m = collide(1, [10.0, 0.5, 10.0], [0, -0.5, 0] + IDQ, 1, [0.2] * 3, [0, low - 0.002, 0] + rot, 0.0)
    assert 1 <= len(m) <= 2
    assert np.linalg.norm(m[0, 3:6] - [0, 1, 0]) < 1e-4
    assert abs(m[0, 6] - 0.002) < 1e-4


def test_sphere_centre_inside_box_pushed_out_through_nearest_face():
    """collision/tests/box_sphere_inside_test.rs:10-28"""
    floor = [0.0, -0.5, 0.0, 0.0, math.sin(0.2), 0.0, math.cos(0.2)]
    r, cd = 0.001, 0.003
    for depth in (-0.5 * r, 0.5 * r, r, 2.0 * r, 20.0 * r):
        at = [0.3, -depth, -0.2] + IDQ
        m = collide(1, [5.0, 0.5, 5.0], floor, 0, [r, 0, 0], at, cd)
        assert len(m) == 1
        assert np.linalg.norm(m[0, 3:6] - [0, 1, 0]) < 1e-5
        assert abs(m[0, 6] - (r + depth + cd)) < 1e-5
        f = collide(0, [r, 0, 0], at, 1, [5.0, 0.5, 5.0], floor, cd)
        assert np.linalg.norm(f[0, 3:6] + [0, 1, 0]) < 1e-5 and abs(f[0, 6] - (r + depth + cd)) < 1e-5


def test_sphere_sphere_closed_form():
    rng = np.random.default_rng(2)
    for _ in range(200):
        ca, cb = rng.uniform(-1, 1, 3), rng.uniform(-1, 1, 3)
        ra, rb, cd = rng.uniform(0.1, 0.8), rng.uniform(0.1, 0.8), 0.01
        m = collide(0, [ra, 0, 0], list(ca) + IDQ, 0, [rb, 0, 0], list(cb) + IDQ, cd)
        d = np.linalg.norm(cb - ca)
        if d > ra + rb + cd + 1e-5:
            assert len(m) == 0
            continue
        if d < ra + rb + cd - 1e-5:
            n = (cb - ca) / d
            assert np.allclose(m[0, 3:6], n, atol=1e-5)
            assert abs(m[0, 6] - (ra + rb + cd - d)) < 1e-5
            assert np.allclose(m[0, :3], 0.5 * ((ca + n * ra) + (cb - n * rb)), atol=1e-5)


def test_capsule_capsule_closest_points():
    # two crossed vertical/horizontal capsules: closest points at the segment crossing
    a = [0.0, 0.0, 0.0] + IDQ                                   # axis +y, hh 1
    qz = [0.0, 0.0, math.sin(math.pi / 4), math.cos(math.pi / 4)]  # rotate y onto -x
    b = [0.0, 0.3, 0.5] + qz
    m = collide(2, [0.3, 1.0, 0], a, 2, [0.3, 1.0, 0], b, 0.0)
    assert len(m) == 1
    assert np.allclose(m[0, 3:6], [0, 0, 1], atol=1e-5)
    assert abs(m[0, 6] - 0.1) < 1e-5


def test_box_triangle_flat_patch_spans_four_corners():
    tri = f32([-10, 0, -10, -10, 0, 30, 30, 0, -10])  # covers the box footprint; normal +y
    pose = f32([0, 0.49, 0] + IDQ)
    out = np.zeros(28, np.float32)
    n = O.lib().or_collide_box_triangle(O.fptr(f32([0.5, 0.5, 0.5])), O.fptr(pose), O.fptr(tri), 0.05, O.fptr(out))
    pts = out[: 7 * n].reshape(n, 7)
    assert n == 4
    raw = np.cross(tri[3:6] - tri[0:3], tri[6:9] - tri[0:3])
    nrm = raw / np.linalg.norm(raw)
    for p in pts:
        assert np.allclose(p[3:6], -nrm, atol=1e-6)   # box -> terrain
        assert abs(p[6] - (0.05 + 0.01)) < 1e-5
    xs = sorted({round(float(p[0]), 4) for p in pts})
    assert xs == [-0.5, 0.5]


def test_sphere_triangle_vertex_and_face_regions():
    tri = f32([0, 0, 0, 0, 0, 1, 1, 0, 0])
    out = np.zeros(7, np.float32)
    n = O.lib().or_collide_sphere_triangle(0.5, O.fptr(f32([0.25, 0.3, 0.25])), O.fptr(tri), 0.0, O.fptr(out))
    assert n == 1 and np.allclose(out[:3], [0.25, 0, 0.25], atol=1e-6) and abs(out[6] - 0.2) < 1e-6
    n = O.lib().or_collide_sphere_triangle(0.5, O.fptr(f32([-0.3, 0.0, -0.3])), O.fptr(tri), 0.0, O.fptr(out))
    assert n == 1 and np.allclose(out[:3], [0, 0, 0], atol=1e-7)


# ---------------------------------------------------------------- broadphase (broadphase.rs:29-99)
def grid_broadphase(bb):
    """The reference's algorithm, restated literally: hash-grid binning at 2.5 m, oversized
    (> 12.5 m) colliders against everything, per-cell all pairs, sort + dedup."""
    cell = np.float32(2.5)
    grid, over, pairs = {}, [], []
    for i, b in enumerate(bb):
        mn, mx = b[:3], b[3:]
        if not np.isfinite(mn[0]):
            continue
        d = mx - mn
        if (d > cell * np.float32(5.0)).any():
            over.append(i)
            continue
        lo = np.floor(mn / cell).astype(int)
        hi = np.floor(mx / cell).astype(int)
        for x in range(lo[0], hi[0] + 1):
            for y in range(lo[1], hi[1] + 1):
                for z in range(lo[2], hi[2] + 1):
                    grid.setdefault((x, y, z), []).append(i)

    def ov(a, b):
        return not (a[3] < b[0] or a[0] > b[3] or a[4] < b[1] or a[1] > b[4] or a[5] < b[2] or a[2] > b[5])

    for lst in grid.values():
        for p in range(len(lst)):
            for q in range(p + 1, len(lst)):
                a, b = min(lst[p], lst[q]), max(lst[p], lst[q])
                if ov(bb[a], bb[b]):
                    pairs.append((a, b))
    for big in over:
        for i in range(len(bb)):
            if i != big and ov(bb[big], bb[i]):
                pairs.append((min(i, big), max(i, big)))
    return sorted(set(pairs))


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_broadphase_equals_grid_algorithm(seed):
    rng = np.random.default_rng(seed)
    n = 300
    c = rng.uniform(0, 30, (n, 3)).astype(np.float32)
    h = rng.uniform(0.1, 1.5, (n, 3)).astype(np.float32)
    bb = np.concatenate([c - h, c + h], axis=1).astype(np.float32)
    bb[5] = [0, 0, 0, 14, 2, 2]                 # oversized
    bb[17] = [np.nan, 1, 1, 2, 2, 2]            # non-finite min.x: only oversized pairs reach it
    bb[40, :3] = bb[41, :3]                     # equal keys
    out = np.zeros((n * n, 2), np.uint32)
    k = O.lib().or_broadphase(O.fptr(bb), n, out.ctypes.data_as(C.POINTER(C.c_uint32)), n * n)
    got = [tuple(map(int, p)) for p in out[:k]]
    assert got == grid_broadphase(bb)


# ---------------------------------------------------------------- whole-step closed forms
def world_with(bodies, iters=(20, 20), pairs=((0, 0, 0.5, 0.5, 0.0, 0.08),), sleep=False):
    w = O.OracleWorld(None, capacity=len(bodies))
    for a, b, ms, mk, e, vs in pairs:
        O.lib().or_set_pair_props(w.h, a, b, ms, mk, e, vs)
    O.lib().or_set_iterations(w.h, *iters)
    for shape, dims, pose, v, dyn, rho in bodies:
        O.lib().or_add_body(w.h, shape, O.fptr(f32(dims)), O.fptr(f32(pose)), O.fptr(f32(v)), dyn, rho)
    for i in range(len(bodies)):
        O.lib().or_set_sleep_allowed(w.h, i, 1 if sleep else 0)
    w.dt = 1.0 / 60.0
    return w


def test_free_fall_matches_closed_form():
    """reality/src/analytic.rs free fall (1%): the ballistic average integrates a constant
    acceleration exactly."""
    w = world_with([(0, [0.5, 0, 0], [0, 100, 0] + IDQ, [0] * 6, 1, 1000.0)])
    for _ in range(60):
        w.step()
    y = w.state()[0][0, 1]
    g = np.float32(-9.81)
    assert abs(y - (100.0 + 0.5 * float(g) * 1.0)) < 2e-4
    assert abs(w.state()[2][0, 1] - float(g)) < 1e-4


@pytest.mark.parametrize("shape,dims,half_y", [(0, [0.1, 0, 0], 0.1), (2, [0.1, 0.1, 0], 0.2), (1, [0.1, 0.1, 0.1], 0.1)])
def test_shapes_rest_on_a_static_ground(shape, dims, half_y):
    """contact_resting_test.rs:30-45 (+-12 mm after 3 s; restitution 0 so the W6 bounce pass
    stays out)"""
    ground = (1, [50.0, 0.5, 50.0], [0, -0.5, 0] + IDQ, [0] * 6, 0, 0.0)
    body = (shape, dims, [0, half_y + 0.2, 0] + IDQ, [0] * 6, 1, 1000.0)
    w = world_with([ground, body])
    for _ in range(180):
        w.step()
    y = w.state()[0][1, 1]
    assert abs(y - half_y) < 0.012, y


def test_stacked_boxes_hold_30s():
    """contact_resting_test.rs:63-75: a 3-box stack stays put for 30 s kept awake."""
    ground = (1, [50.0, 0.5, 50.0], [0, -0.5, 0] + IDQ, [0] * 6, 0, 0.0)
    h = [0.2, 0.1, 0.2]
    rho = 5.0 / (8 * 0.2 * 0.1 * 0.2)
    boxes = [(1, h, [0, 0.1 + 0.2 * i + 0.01 * i, 0] + IDQ, [0] * 6, 1, rho) for i in range(3)]
    w = world_with([ground] + boxes)
    for _ in range(1800):
        w.step()
    pos = w.state()[0]
    for i in range(3):
        p = pos[1 + i]
        assert abs(p[1] - (0.1 + 0.2 * i)) < 0.03 and abs(p[0]) < 0.02 and abs(p[2]) < 0.02, (i, p)


def test_box_flat_on_heightfield_does_not_drift():
    """at_rest_test.rs (box flat on a heightfield, < 1e-4 m over 60 s): 10 s here."""
    w = world_with([(1, [0.5, 0.5, 0.5], [8.0, 0.5, 8.0] + IDQ, [0] * 6, 1, 1000.0)],
                   pairs=((0, 0, 0.5, 0.5, 0.0, 0.08), (0, 4, 0.5, 0.5, 0.0, 0.08)))
    hts = np.zeros(17 * 17, np.float32)
    O.lib().or_set_heightfield(w.h, 17, 17, 1.0, 1.0, 0.0, 0.0, O.fptr(hts), 4)
    for _ in range(60):
        w.step()
    p0 = w.state()[0][0].copy()
    for _ in range(600):
        w.step()
    assert np.linalg.norm(w.state()[0][0] - p0) < 1e-4


# ---------------------------------------------------------------- invariants on a mixed scene
def test_mixed_scene_invariants_and_snapshot():
    s = scenes.small_mixed(n=80, seed=7)
    w = O.OracleWorld(s)
    for t in range(60):
        w.step()
        cts = w.contacts()
        if len(cts) == 0:
            continue
        pos = w.state()[0]
        keys = cts["a_col"].astype(np.int64) << 32 | cts["b_col"]
        assert np.all(np.diff(keys) >= 0), "contacts in bucket order"
        _, counts = np.unique(keys, return_counts=True)
        assert counts.max() <= 4
        assert np.all(cts["impulse"] >= 0)
        body = cts["b_col"] != 0xFFFFFFFF
        d = pos[cts["b_col"][body]] - pos[cts["a_col"][body]]
        assert np.all(np.einsum("ij,ij->i", cts["normal"][body], d) >= 0)
        nl = np.linalg.norm(cts["normal"], axis=1)
        assert np.all(np.abs(nl - 1) < 1e-5)
    snap = w.snapshot()
    a = []
    for _ in range(10):
        w.step()
        a.append(w.hash())
    w.restore(snap)
    b = []
    for _ in range(10):
        w.step()
        b.append(w.hash())
    w.free_snapshot(snap)
    assert a == b


def test_cold_restore_differs():
    """at_rest_test.rs carried_state_restores_exactly: dropping the carried state changes the
    next steps (the warm start matters), keeping it reproduces them."""
    s = scenes.small_mixed(n=60, seed=11)
    w = O.OracleWorld(s)
    for _ in range(90):
        w.step()
    assert w.stats().warm_matched > 0
