"""The oracle against closed-form mechanics (CPU only).

Restated from the reference's own closed-form benchmarks, crates/riftphys-reality/src/analytic.rs:
the same scenes (a 100 x 1 x 100 m static box as the ground, tilted about +z for slopes; 1 kg
boxes and balls with sleep disabled; analytic.rs:67-100), the same measured quantities and the same
tolerances.  They pin what no reference golden covers bit for bit: the friction rows' sign, scale
and Stribeck curve (sliding, static friction, incline), the friction rows' torque (rolling without
slip), and the gyroscopic turn (torque-free tumbling), including its subdivided form on a needle.
Restitution is 0 in every scene here (the W6 bounce is not part of the oracle).
"""
import math

import numpy as np
import pytest

import oracle_py as O

f = np.float32
G = 9.81
WOOD = (0.55, 0.45, 0.10)    # materials.rs:101 (mu_s, mu_k, stribeck_v)
RUBBER = (1.10, 0.95, 0.15)  # materials.rs:98


def _qz(angle):
    s, c = f(math.sin(angle / 2)), f(math.cos(angle / 2))
    return np.array([0, 0, s, c], f)


def _rot(q, v):
    x, y, z, w = [float(t) for t in q]
    R = np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - w * z), 2 * (x * z + w * y)],
                  [2 * (x * y + w * z), 1 - 2 * (x * x + z * z), 2 * (y * z - w * x)],
                  [2 * (x * z - w * y), 2 * (y * z + w * x), 1 - 2 * (x * x + y * y)]])
    return R @ np.asarray(v, np.float64)


class W:
    """A small oracle world (World::with_capacity(16, 16), analytic.rs:67-71)."""

    def __init__(self, gravity=True, mat=None):
        self.L = O.lib()
        self.h = self.L.or_world_create(16)
        self.L.or_set_gravity(self.h, 0.0, -G if gravity else 0.0, 0.0)
        if mat is not None:  # every body's collider uses this material pair, restitution 0
            self.L.or_set_pair_props(self.h, 0, 0, mat[0], mat[1], 0.0, mat[2])
            self.L.or_set_pair_props(self.h, 0, 4, mat[0], mat[1], 0.0, mat[2])

    def __del__(self):
        self.L.or_world_destroy(self.h)

    def add(self, shape, dims, pos, rot, vel, dynamic, density, collider=True):
        d = np.array(dims, f)
        p = np.array([*pos, *rot], f)
        v = np.array(vel, f)
        add = self.L.or_add_body if collider else self.L.or_add_body_nc
        i = add(self.h, shape, O.fptr(d), O.fptr(p), O.fptr(v), int(dynamic), float(density))
        if dynamic:
            self.L.or_set_sleep_allowed(self.h, i, 0)
        return i

    def ground(self, tilt):
        q = _qz(tilt)
        return self.add(O.OR_BOX, (50.0, 0.5, 50.0), _rot(q, (0, -0.5, 0)), q, [0] * 6, False, 1000.0)

    def cuboid(self, pos, rot, vel, half, mass, collider=True):
        return self.add(O.OR_BOX, half, pos, rot, vel, True, mass / (8.0 * half[0] * half[1] * half[2]), collider)

    def sphere(self, pos, r, mass):
        return self.add(O.OR_SPHERE, (r, 0, 0), pos, (0, 0, 0, 1), [0] * 6, True, mass / (4.0 / 3.0 * math.pi * r ** 3))

    def run(self, seconds, dt=1.0 / 60.0, each=None):
        for _ in range(int(round(seconds / dt))):
            self.L.or_step(self.h, f(dt))
            if each is not None:
                each()

    def state(self, i):
        n = self.L.or_num_bodies(self.h)
        pos = np.zeros((n, 3), f); rot = np.zeros((n, 4), f); lin = np.zeros((n, 3), f); ang = np.zeros((n, 3), f)
        self.L.or_get_state(self.h, O.fptr(pos), O.fptr(rot), O.fptr(lin), O.fptr(ang))
        return pos[i].astype(np.float64), rot[i].astype(np.float64), lin[i].astype(np.float64), ang[i].astype(np.float64)


def test_sliding_stopping_distance():
    """analytic.rs:198-211: a block sliding to rest from 3 m/s stops at v0^2 / (2 mu_k g), 5%."""
    w = W(mat=WOOD)
    w.ground(0.0)
    h = (0.1, 0.05, 0.1)
    b = w.cuboid((0.0, h[1], 0.0), (0, 0, 0, 1), (3.0, 0, 0, 0, 0, 0), h, 1.0)
    w.run(3.0)
    truth = 9.0 / (2 * WOOD[1] * G)
    x = w.state(b)[0][0]
    assert abs(x - truth) / truth <= 0.05, (x, truth)


def _incline(tilt, seconds):
    """analytic.rs:213-223: distance a block travels downhill from rest."""
    w = W(mat=WOOD)
    w.ground(tilt)
    h = (0.1, 0.05, 0.1)
    q = _qz(tilt)
    start = _rot(q, (0, h[1], 0))
    b = w.cuboid(start, q, [0] * 6, h, 1.0)
    w.run(seconds)
    return -float(np.dot(w.state(b)[0] - start, _rot(q, (1, 0, 0))))


def test_static_friction_holds_below_atan_mu_s():
    """analytic.rs:225-236: on a slope of 0.8 atan(mu_s) the block creeps under 5 mm in 2 s."""
    d = _incline(0.8 * math.atan(WOOD[0]), 2.0)
    assert abs(d) <= 0.005, d


def test_incline_slide_distance():
    """analytic.rs:225-236: from rest on atan(mu_s) + 0.1 rad, g (sin - mu_k cos) t^2 / 2 in 1 s, 5%."""
    tilt = math.atan(WOOD[0]) + 0.1
    truth = 0.5 * G * (math.sin(tilt) - WOOD[1] * math.cos(tilt))
    d = _incline(tilt, 1.0)
    assert abs(d - truth) / truth <= 0.05, (d, truth)


@pytest.mark.parametrize("ball_first", [False, True])
def test_rolling_without_slip(ball_first):
    """analytic.rs:261-281: a solid ball rolls down 0.3 rad at a = 5/7 g sin(theta) (5%), with
    spin x radius / speed within 0.05 of 1.  Added after the ground the ball is the contacts' body
    B, added first it is body A: both sides of the friction rows' torque are exercised."""
    tilt, r = 0.3, 0.1
    w = W(mat=RUBBER)
    q = _qz(tilt)
    start = _rot(q, (0, r, 0))
    if ball_first:
        b = w.sphere(start, r, 1.0)
        w.ground(tilt)
    else:
        w.ground(tilt)
        b = w.sphere(start, r, 1.0)
    w.run(1.0)
    pos, _, lin, ang = w.state(b)
    down = -_rot(q, (1, 0, 0))
    d = float(np.dot(pos - start, down))
    v = float(np.dot(lin, down))
    truth = 0.5 * 5.0 / 7.0 * G * math.sin(tilt)
    assert abs(d - truth) / truth <= 0.05, (d, truth)
    assert abs(np.linalg.norm(ang) * r / max(v, 1e-6) - 1.0) <= 0.05


def _tumble(half, w0, seconds, mass=1.0):
    """Torque-free box (no gravity, no collider): the largest change of world angular momentum and
    of rotational energy over the run, and the flips of the body-frame y rate (analytic.rs:284-313)."""
    w = W(gravity=False)
    b = w.cuboid((0, 0, 0), (0, 0, 0, 1), (0, 0, 0, *w0), half, mass, collider=False)
    hx, hy, hz = [2.0 * x for x in half]
    inertia = mass / 12.0 * np.array([hy * hy + hz * hz, hx * hx + hz * hz, hx * hx + hy * hy])

    def mom():
        _, q, _, a = w.state(b)
        R = np.array([_rot(q, e) for e in np.eye(3)]).T
        wb = R.T @ a
        return R @ (inertia * wb), 0.5 * float(np.dot(wb, inertia * wb)), wb

    l0, e0, _ = mom()
    out = {"dl": 0.0, "de": 0.0, "flips": 0, "sign": 1.0}

    def each():
        l, e, wb = mom()
        out["dl"] = max(out["dl"], np.linalg.norm(l - l0) / np.linalg.norm(l0))
        out["de"] = max(out["de"], abs(e - e0) / e0)
        if np.sign(wb[1]) != out["sign"] and abs(wb[1]) > 0.5:
            out["flips"] += 1
            out["sign"] = np.sign(wb[1])

    w.run(seconds, each=each)
    return out, inertia


def test_torque_free_tumbling_dzhanibekov():
    """analytic.rs:284-313: a box spun about its intermediate axis keeps its world angular momentum
    and rotational energy within 1% over 10 s and flips at least once."""
    out, _ = _tumble((0.3, 0.2, 0.1), (0.02, 4.0, 0.02), 10.0)
    assert out["dl"] <= 0.01 and out["de"] <= 0.01, out
    assert out["flips"] >= 1, out


def test_needle_turn_is_subdivided_and_keeps_its_energy():
    """step.rs:216-270: on a thin link the turn gain dt |dev(I^-1)| |L| exceeds 0.5, so the turn
    runs in ceil(gain / 0.5) midpoint pieces.  A needle spun off its axes then keeps its rotational
    energy within 1% over 5 s (0.3% here); run as one undivided piece it drifts by about 10%."""
    half, w0, dt = (0.01, 0.5, 0.01), (3.0, 0.5, 0.2), 1.0 / 60.0
    out, inertia = _tumble(half, w0, 5.0)
    inv = 1.0 / inertia
    dev = np.abs(inv - inv.mean()).max()
    gain = dt * dev * np.linalg.norm(inertia * np.array(w0))
    assert gain > 1.0, gain  # several pieces from the first step
    assert out["de"] <= 0.01 and out["dl"] <= 0.01, out
