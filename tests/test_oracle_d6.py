"""The oracle's D6 joints (CPU only): velocity rows (solver.rs:336-424 d6_rows: dampers, Coulomb
friction, locked axes, velocity motors, speculative limits; warm-started by row tag) and the
position solve (bodies.rs:616-873 solve_d6: translational and rotational rows, hinge twist in
radians, limits, position drives, compliance), against the reference's own joint tests restated
with their closed forms (hinge_angle_test.rs, joint_damping_test.rs, joint_energy_test.rs).

A rod hinged at its end falls onto a -1 rad stop and rests there, and reaches a -2.2 rad stop past
the bottom of its swing; a spinning bar's damper decays its spin at the implicit damper's discrete
rate ln(1 + c dt / I) / dt, its friction stops it after I w0^2 / (2 tau_f) and holds it; the cube
pendulum's large-swing period matches the physical pendulum (RK4) within 1 % and its energy stays
within 2 %; a velocity motor spins a bar up at max_impulse / (I dt) per second to its target; a
position drive holds its angle; a sprung hinge sags to its static angle.
"""
import math

import numpy as np
import pytest

import oracle_py as O
import scenes

G = 9.81
f = np.float32


def twist(w, b):
    """The body's rotation angle about z (quaternions here turn about z only)."""
    q = w.state()[1][b]
    return 2.0 * math.atan2(float(q[2]), float(q[3]))


def test_hinge_rests_on_a_limit_in_radians():
    w = O.OracleWorld(scenes.hinge_rod(limit=(-1.0, 1.0)))
    furthest = 0.0
    for _ in range(5 * 240):
        w.step()
        a = twist(w, 1)
        if abs(a) > abs(furthest):
            furthest = a
    rest, stop = math.degrees(twist(w, 1)), math.degrees(-1.0)
    assert abs(rest - stop) < 0.5, (rest, stop)
    assert math.degrees(furthest) > stop - 1.0, math.degrees(furthest)


def test_hinge_limit_beyond_two_radians_holds():
    w = O.OracleWorld(scenes.hinge_rod(limit=(-2.2, 0.1)))
    furthest = 0.0
    for _ in range(3 * 240):
        w.step()
        a = twist(w, 1)
        if abs(a) > abs(furthest):
            furthest = a
    stop = math.degrees(-2.2)
    assert stop - 1.0 < math.degrees(furthest) < stop + 1.0, (math.degrees(furthest), stop)


def izz():
    hx, hy = scenes.SPIN_HALF[0], scenes.SPIN_HALF[1]
    return scenes.SPIN_MASS / 3.0 * (hx * hx + hy * hy)


@pytest.mark.parametrize("hz", [60.0, 240.0])
@pytest.mark.parametrize("rate", [0.5, 2.0])
def test_damped_hinge_decays_at_the_discrete_rate(hz, rate):
    w = O.OracleWorld(scenes.hinge_spinner(3.0, damping=rate * izz(), hz=hz))
    t = 2.0
    for _ in range(int(t * hz)):
        w.step()
    ang, lin = w.state()[3][1], w.state()[2][1]
    measured = math.log(3.0 / float(ang[2])) / t
    discrete = math.log(1.0 + rate / hz) * hz
    assert abs(measured - discrete) < 0.005 * discrete, (measured, discrete)
    assert abs(measured - rate) < 0.02 * rate
    assert abs(ang[0]) < 1e-4 and abs(ang[1]) < 1e-4 and np.linalg.norm(lin) < 1e-4


@pytest.mark.parametrize("hz", [60.0, 240.0])
def test_friction_hinge_stops_at_the_closed_form_angle(hz):
    w0, tau = 3.0, 0.2
    want, t_stop = izz() * w0 * w0 / (2.0 * tau), izz() * w0 / tau
    w = O.OracleWorld(scenes.hinge_spinner(w0, friction=tau, hz=hz))
    total, last = 0.0, 0.0
    for _ in range(int((t_stop + 2.0) * hz)):
        w.step()
        a = twist(w, 1)
        d = a - last
        d -= 2 * math.pi * round(d / (2 * math.pi))
        total += d
        last = a
    assert abs(total - want) < 0.02 * want, (total, want)
    assert abs(float(w.state()[3][1][2])) < 1e-4


def test_cube_pendulum_period_matches_the_physical_pendulum():
    m, arm = 80.0, 5.0
    i_pivot = m * (2.0 * 2.0) / 6.0 + m * arm * arm
    k = m * G * arm / i_pivot
    th, om, t, h = math.pi / 2, 0.0, 0.0, 1e-5
    fn = lambda x: -k * math.sin(x)
    while True:
        k1t, k1o = om, fn(th)
        k2t, k2o = om + 0.5 * h * k1o, fn(th + 0.5 * h * k1t)
        k3t, k3o = om + 0.5 * h * k2o, fn(th + 0.5 * h * k2t)
        k4t, k4o = om + h * k3o, fn(th + h * k3t)
        om_new = om + h / 6.0 * (k1o + 2 * k2o + 2 * k3o + k4o)
        th += h / 6.0 * (k1t + 2 * k2t + 2 * k3t + k4t)
        t += h
        if t > 1.0 and om < 0.0 and om_new >= 0.0:
            break
        om = om_new
    period = 2.0 * t
    w = O.OracleWorld(scenes.cube_pendulum())
    dt, prev, turns, t_engine = 1.0 / 62.5, 0.0, 0, 0.0
    for n in range(1, 2000):
        w.step()
        wz = float(w.state()[3][1][2])
        if n > 5 and prev != 0.0 and (prev > 0.0) != (wz > 0.0):
            turns += 1
        if turns == 2:
            t_engine = n * dt
            break
        prev = wz
    assert abs((t_engine - period) / period) < 0.01, (t_engine, period)


def test_cube_pendulum_keeps_its_energy():
    m = 80.0
    inertia = m * 4.0 / 6.0
    w = O.OracleWorld(scenes.cube_pendulum())

    def energy():
        p, _, lin, ang = w.state()
        return m * G * (float(p[1][1]) - 15.0) + 0.5 * m * float(np.dot(lin[1], lin[1])) + 0.5 * inertia * float(np.dot(ang[1], ang[1]))
    w.step()
    e0, lo, hi = energy(), 1.0, 1.0
    for _ in range(int(5 * 62.5)):
        w.step()
        r = energy() / e0
        lo, hi = min(lo, r), max(hi, r)
    assert lo > 0.98 and hi < 1.02, (lo, hi)


def test_velocity_motor_spins_up_at_its_impulse_cap():
    """A bar (no gravity) on a motor to 2 rad/s capped at 0.01 N s a step: dw = cap / I a step
    until the target is reached, then held."""
    hz, cap, target = 60.0, 0.01, 2.0
    s = scenes.hinge_spinner(0.0, hz=hz)
    s["d6"][0]["drives"] = [(True, 2, target, 0.0, 0.0, cap, True)]
    w = O.OracleWorld(s)
    w.step()
    w1 = float(w.state()[3][1][2])
    assert abs(w1 - cap / izz()) < 1e-4 * (cap / izz()), (w1, cap / izz())
    for _ in range(int(3 * hz)):
        w.step()
    assert abs(float(w.state()[3][1][2]) - target) < 1e-4


def test_position_drive_holds_its_angle():
    s = scenes.hinge_spinner(0.0, hz=120.0)
    s["d6"][0]["drives"] = [(True, 2, 0.5, 0.0, 0.0, 0.0, False)]
    w = O.OracleWorld(s)
    for _ in range(60):
        w.step()
    assert abs(twist(w, 1) - 0.5) < 1e-4, twist(w, 1)


def test_sprung_hinge_sags_to_its_static_angle():
    """A rod on a hinge with compliance c on its twist: the static sag theta solves
    theta = -c m g (L / 2) cos(theta)."""
    c, hz = 0.01, 240.0
    s = scenes.hinge_rod(compliance=c, hz=hz)
    s["d6"][0]["damping"] = [(True, 2, 2.0, 0.0)]
    w = O.OracleWorld(s)
    for _ in range(int(8 * hz)):
        w.step()
    m = 8.0 * 0.5 * 0.025 * 0.025 * 1000.0
    th = 0.0
    for _ in range(100):
        th = -c * m * G * 0.5 * math.cos(th)
    assert abs(twist(w, 1) - th) < 0.01 * abs(th), (twist(w, 1), th)


def test_hinge_chain_replays_from_a_snapshot():
    w = O.OracleWorld(scenes.hinge_chain())
    for _ in range(60):
        w.step()
    snap = O.lib().or_snapshot(w.h)
    try:
        a = []
        for _ in range(60):
            w.step()
            a.append(w.hash())
        O.lib().or_restore(w.h, snap)
        b = []
        for _ in range(60):
            w.step()
            b.append(w.hash())
        assert a == b
    finally:
        O.lib().or_snapshot_free(snap)
