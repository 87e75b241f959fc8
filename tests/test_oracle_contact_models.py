"""The oracle's contact models (CPU only): compliant Hunt-Crossley contact (solver.rs:545-567,
848-856; step.rs:340, 433) and patch friction (solver.rs:219-260, 592-619, 763-782, 882-915,
927-944) against the reference's own tests of them, restated with their closed forms
(compliant_contact_test.rs, patch_friction_test.rs), and the carried patch impulses in the step
hash and the snapshots.

Compliant: a box on four compliant corners sinks m g / (4 k) and stays; a sphere on one rests at
F / k within 1 % at 60 Hz, 240 Hz and 2 kHz; an elastic impact returns nearly all its speed and
dissipation takes more from faster impacts; the bounce at 2 kHz matches Hunt-Crossley's own ODE
(integrated here by RK4 at 1 us) within 5 %; friction reads the compliant normal force (Coulomb
sliding distance) and holds a box on a 20 degree slope.  Patch friction: a cube spinning flat stops
after I w0^2 / (2 mu m g r) with r the face's mean radius A (sqrt 2 + ln(1 + sqrt 2)) / 3, where
per-point friction gives r = sqrt 2 A (the corners); sliding distance stays Coulomb.
"""
import math

import numpy as np
import pytest

import oracle_py as O
import scenes

G = 9.81
f = np.float32


def run(s, ticks, every=None):
    w = O.OracleWorld(s)
    out = []
    for _ in range(ticks):
        w.step()
        if every is not None:
            out.append(every(w))
    return w, out


def body_y(w):
    return float(w.state()[0][1][1])


def test_resting_penetration_matches_the_spring():
    k = 20000.0
    s = scenes.contact_model_scene(k, 1.0, 0.5, 1, (0.1, 0.1, 0.1), (0.0, 0.1, 0.0), hz=1000.0)
    w, _ = run(s, 2000)
    sink = 0.1 - body_y(w)
    want = G / (4.0 * k)
    assert abs(sink - want) < 0.15 * want, (sink, want)
    assert np.linalg.norm(w.state()[2][1]) < 1e-3
    # no rigid rows: the compliant pair is left out of the warm cache and the position passes
    assert w.stats().warm_matched == 0


def max_up_speed(k, d, v0, hz=4000.0, ticks=400):
    s = scenes.contact_model_scene(k, d, 0.0, 0, (0.05,), (0.0, 0.06, 0.0), vel=(0.0, -v0, 0.0), hz=hz)
    _, vy = run(s, ticks, lambda w: float(w.state()[2][1][1]))
    return max(max(vy), 0.0) / v0


def test_dissipation_sets_the_bounce():
    elastic = max_up_speed(1.0e5, 0.0, 2.0)
    damped = max_up_speed(1.0e5, 0.5, 2.0)
    damped_fast = max_up_speed(1.0e5, 0.5, 4.0)
    assert elastic > 0.9, elastic
    assert damped < elastic - 0.2, (damped, elastic)
    assert damped_fast < damped, (damped_fast, damped)


def test_friction_uses_the_compliant_normal_force():
    mu, v0, k = 0.3, 2.0, 50000.0
    s = scenes.contact_model_scene(k, 2.0, mu, 1, (0.1, 0.1, 0.1), (0.0, 0.1 - G / (4.0 * k), 0.0), vel=(v0, 0.0, 0.0), hz=1000.0)
    w, _ = run(s, 3000)
    dist = float(w.state()[0][1][0])
    want = v0 * v0 / (2.0 * mu * G)
    assert abs(dist - want) < 0.05 * want, (dist, want)


@pytest.mark.parametrize("hz", [60.0, 240.0, 2000.0])
def test_implicit_contact_rests_at_force_over_stiffness(hz):
    k, r = 20000.0, 0.05
    s = scenes.contact_model_scene(k, 1.0, 0.5, 0, (r,), (0.0, r, 0.0), hz=hz)
    w, _ = run(s, int(3.0 * hz))
    sink = r - body_y(w)
    want = G / k
    assert abs(sink / want - 1.0) < 0.01, (hz, sink, want)


def hunt_crossley_e(k, d, v_in):
    """Restitution of m x'' = -k x (1 + d x') (never pulling) for 1 kg hitting at v_in, by RK4 at
    1 us (compliant_contact_test.rs:102-116)."""
    def acc(x, v):
        return 0.0 if x <= 0.0 else -max(k * x * (1.0 + d * v), 0.0)
    x, v, h = 0.0, v_in, 1.0e-6
    while True:
        k1x, k1v = v, acc(x, v)
        k2x, k2v = v + 0.5 * h * k1v, acc(x + 0.5 * h * k1x, v + 0.5 * h * k1v)
        k3x, k3v = v + 0.5 * h * k2v, acc(x + 0.5 * h * k2x, v + 0.5 * h * k2v)
        k4x, k4v = v + h * k3v, acc(x + h * k3x, v + h * k3v)
        x += h / 6.0 * (k1x + 2.0 * k2x + 2.0 * k3x + k4x)
        v += h / 6.0 * (k1v + 2.0 * k2v + 2.0 * k3v + k4v)
        if x < 0.0:
            return -v / v_in


def hc_bounce(k, d, hz, at_start=False):
    r, v_in = 0.05, 1.0
    s = scenes.contact_model_scene(k, d, 0.0, 0, (r,), (0.0, r + 0.002, 0.0), vel=(0.0, -v_in, 0.0), hz=hz, gravity=False)
    s["damping_at_start"] = at_start
    w, _ = run(s, int(0.5 * hz))
    return float(w.state()[2][1][1]) / v_in


@pytest.mark.parametrize("d", [0.0, 0.1, 0.3])
def test_implicit_contact_bounces_as_hunt_crossley(d):
    k = 1000.0
    want = hunt_crossley_e(k, d, 1.0)
    mid = hc_bounce(k, d, 2000.0)
    assert abs(mid / want - 1.0) < 0.05, (d, mid, want)
    # damping linearised at the step-start penetration: a different, still dissipative bounce
    start = hc_bounce(k, d, 2000.0, at_start=True)
    assert 0.0 < start <= 1.0 and (d == 0.0) == (start == mid), (start, mid)


def test_a_box_holds_on_a_compliant_slope():
    k, tilt = 20000.0, math.radians(20.0)
    c, s_ = math.cos(tilt), math.sin(tilt)
    h = 0.1 - G * c / (4.0 * k)
    start = (-s_ * h, c * h, 0.0)  # the tilt applied to (0, h, 0)
    rot = (0.0, 0.0, math.sin(tilt / 2.0), math.cos(tilt / 2.0))
    s = scenes.contact_model_scene(k, 1.0, 0.6, 1, (0.1, 0.1, 0.1), start, hz=60.0, ground_tilt=tilt, rot=rot)
    w = O.OracleWorld(s)
    for _ in range(60):
        w.step()
    settled = w.state()[0][1].copy()
    for _ in range(300):
        w.step()
    creep = float(np.linalg.norm(w.state()[0][1] - settled))
    assert creep < 1e-3, creep


# ---- patch friction (patch_friction_test.rs) ----
A, M, MU = 0.0524, 0.37, 0.2


def spin(patch, k, w0, v0, ticks=4000, hz=1480.0):
    """A cube resting on a face, spinning at w0 about the vertical and sliding at v0; returns the
    angle turned (the sum of w_y / hz over the ticks) and the distance slid."""
    sink = G * M / (4.0 * k) if k > 0.0 else 0.0
    s = scenes.contact_model_scene(k, 3.0 if k > 0.0 else 0.0, MU, 1, (A, A, A), (0.0, A - sink, 0.0), vel=(v0, 0.0, 0.0),
                                   ang=(0.0, w0, 0.0), hz=hz, patch=patch, mass=M, mats=(scenes.MAT_PMMA, scenes.MAT_WOOD))
    w = O.OracleWorld(s)
    angle = 0.0
    for _ in range(ticks):
        w.step()
        angle += float(w.state()[3][1][1]) / hz
    return angle, float(w.state()[0][1][0]), w


def face_radius():
    return A * (math.sqrt(2.0) + math.log(1.0 + math.sqrt(2.0))) / 3.0


@pytest.mark.parametrize("k", [0.0, 13591.0])
def test_a_spinning_face_stops_as_a_uniformly_pressed_face(k):
    w0 = 6.0
    inertia = M * (2.0 * A) ** 2 / 6.0  # the cube's own (the reference's test sets 8.1e-4)
    theta = lambda r: inertia * w0 * w0 / (2.0 * MU * M * G * r)
    patch, _, wp = spin(True, k, w0, 0.0)
    corners, _, _ = spin(False, k, w0, 0.0)
    assert wp.stats().patches == 1
    want_patch, want_corners = theta(face_radius()), theta(A * math.sqrt(2.0))
    assert abs(patch - want_patch) < 0.03 * want_patch, (k, patch, want_patch)
    assert abs(corners - want_corners) < 0.03 * want_corners, (k, corners, want_corners)


@pytest.mark.parametrize("k", [0.0, 13591.0])
def test_patch_sliding_distance_is_coulomb(k):
    v0 = 1.5
    want = v0 * v0 / (2.0 * MU * G)
    _, slid, _ = spin(True, k, 0.0, v0)
    assert abs(slid - want) < 0.03 * want, (k, slid, want)


def test_patch_impulses_are_hashed_and_carried_by_snapshots():
    # rigid patches store their impulses (warm_patches): the hash differs from per-point friction
    # even where the motion is the same, and a rollback replays the carried patch state
    a = spin(True, 0.0, 4.0, 1.0, ticks=40)[2]
    b = spin(False, 0.0, 4.0, 1.0, ticks=40)[2]
    assert a.hash() != b.hash()
    snap = O.lib().or_snapshot(a.h)
    try:
        h1 = []
        for _ in range(30):
            a.step()
            h1.append(a.hash())
        O.lib().or_restore(a.h, snap)
        h2 = []
        for _ in range(30):
            a.step()
            h2.append(a.hash())
        assert h1 == h2
    finally:
        O.lib().or_snapshot_free(snap)


def test_patch_radius_closed_forms():
    """The mean radius the patch integrates (6 x 6 centroid rule per fan triangle) against the
    exact mean distance over a square and a rectangle, and a segment's quarter length."""
    pts = [(A, A), (-A, A), (-A, -A), (A, -A)]
    r = patch_radius(pts)
    assert abs(r - face_radius()) < 0.01 * face_radius(), (r, face_radius())
    a, b = 0.3, 0.1  # rectangle of half sides a, b: mean of sqrt(x^2 + y^2) by quadrature
    xs = (np.arange(2000) + 0.5) / 2000
    X, Y = np.meshgrid(xs * a, xs * b)
    want = float(np.mean(np.sqrt(X * X + Y * Y)))
    r = patch_radius([(a, b), (-a, b), (-a, -b), (a, -b)])
    assert abs(r - want) < 0.01 * want, (r, want)
    assert abs(patch_radius([(-0.2, 0.0), (0.2, 0.0)]) - 0.1) < 1e-7


def patch_radius(pts):
    """The oracle's patch radius of plane points (offsets from the centroid, in the patch plane)."""
    xy = np.array(pts, np.float32).reshape(-1)
    return float(O.lib().or_patch_radius(xy.ctypes.data_as(O.C.POINTER(O.C.c_float)), len(pts)))


def test_patch_order_atan2_against_numpy():
    """The angle that orders a patch's points: within one f32 ulp of numpy's arctan2 (and equal to
    the correctly rounded value for nearly every input), signs and quadrants included."""
    rng = np.random.default_rng(4)
    ys = np.concatenate([rng.normal(size=4000), [0.0, -0.0, 1.0, -1.0, 0.0, 2.0]]).astype(np.float32)
    xs = np.concatenate([rng.normal(size=4000), [1.0, -1.0, 0.0, 0.0, -0.0, -2.0]]).astype(np.float32)
    got = np.array([O.lib().or_atan2f_unit(float(y), float(x)) for y, x in zip(ys, xs)], np.float32)
    want = np.arctan2(ys.astype(np.float64), xs.astype(np.float64)).astype(np.float32)
    ulps = np.abs(got.view(np.int32).astype(np.int64) - want.view(np.int32).astype(np.int64))
    assert ulps.max() <= 1 and (ulps == 0).mean() > 0.999, (ulps.max(), (ulps == 0).mean())
    pi = np.float32(math.pi)
    assert got[-6] == 0.0 and got[-5] == -pi and got[-4] == np.float32(math.pi / 2) and got[-3] == -np.float32(math.pi / 2)
    assert got[-2] == pi and got[-1] == np.float32(3.0 * math.pi / 4.0)
