"""D6 joints on the device (velocity rows of dampers, friction, locked axes, velocity motors and
speculative limits; the position solve with hinge twist, limits, position drives and compliance)
against the CPU oracle, bit for bit: the step hash every tick (it carries each joint's stored rows
by tag), the final state.

The reference's joint scenes (hinge_angle_test.rs, joint_damping_test.rs, joint_energy_test.rs,
joint_bits_probe.rs) at reduced lengths, a sprung damped hinge (the position-level damper), ragdolls
whose chains mix hinges, swing limits and ball joints among contacts, a batch with graph-captured
ticks and rollback slots, and a world snapshot.
"""
import numpy as np
import pytest

import oracle_py
import rphys
import scenes

pytestmark = pytest.mark.gpu


def run_pair(s, ticks, every=1):
    o = oracle_py.OracleWorld(s)
    g = rphys.World(s)
    for t in range(ticks):
        o.step()
        g.step()
        if (t + 1) % every == 0 or t + 1 == ticks:
            assert o.hash() == g.hash(), "%s: hash differs at tick %d" % (s.get("name"), t + 1)
    so, sg = o.state(), g.state()
    for a, b in zip(so, sg):
        np.testing.assert_array_equal(np.asarray(a, np.float32).view(np.uint32), np.asarray(b, np.float32)[:, :a.shape[1]].view(np.uint32))
    return o, g


def spinner(w0, hz=60.0, **kw):
    s = scenes.hinge_spinner(w0, hz=hz, **kw)
    return s


def motor_spinner():
    s = scenes.hinge_spinner(0.0, hz=60.0)
    s["d6"][0]["drives"] = [(True, 2, 2.0, 0.0, 0.0, 0.01, True)]
    return s


def held_spinner():
    s = scenes.hinge_spinner(0.0, hz=120.0)
    s["d6"][0]["drives"] = [(True, 2, 0.5, 0.0, 0.0, 0.0, False)]
    return s


def sprung_rod():
    s = scenes.hinge_rod(compliance=0.01, hz=240.0)
    s["d6"][0]["damping"] = [(True, 2, 2.0, 0.0)]
    return s


def limited_translation():
    """A box on a prismatic-like D6 joint: x free within [-0.2, 0.3] m, y on a compliant spring,
    z locked, rotations locked, sliding under gravity tilted by the frame."""
    s = scenes.hinge_rod(hz=120.0)
    j = s["d6"][0]
    j["t_mask"], j["r_mask"], j["hinge"] = 6, 7, -1
    j["fa"] = np.array([0, 0, 0, 0, 0, np.sin(0.3), np.cos(0.3)], np.float32)
    j["limits"] = [(False, 0, -0.2, 0.3)]
    j["compliance"] = [(False, 1, 0.002)]
    j["damping"] = [(False, 0, 0.0, 0.2)]
    return s


CASES = {
    "hinge_rod_limit_1rad": lambda: (scenes.hinge_rod(limit=(-1.0, 1.0)), 360),
    "hinge_rod_limit_2p2rad": lambda: (scenes.hinge_rod(limit=(-2.2, 0.1)), 360),
    "spinner_damped_60hz": lambda: (spinner(3.0, damping=1.0 * scenes.SPIN_MASS / 3.0 * 0.26), 120),
    "spinner_friction_240hz": lambda: (spinner(3.0, friction=0.2, hz=240.0), 360),
    "spinner_motor": lambda: (motor_spinner(), 200),
    "spinner_position_drive": lambda: (held_spinner(), 120),
    "sprung_damped_rod": lambda: (sprung_rod(), 480),
    "cube_pendulum": lambda: (scenes.cube_pendulum(), 200),
    "hinge_chain": lambda: (scenes.hinge_chain(), 240),
    "limited_translation": lambda: (limited_translation(), 240),
}


@pytest.mark.parametrize("name", sorted(CASES))
def test_reference_joint_scenes_match_the_oracle(name):
    s, ticks = CASES[name]()
    s["name"] = name
    run_pair(s, ticks)


def test_ragdolls_with_d6_joints_match_the_oracle():
    s = scenes.d6_ragdolls()
    s["name"] = "d6_ragdolls"
    o, g = run_pair(s, 120, every=2)
    assert len(s["d6"]) > 10


def test_d6_joint_rows_are_carried_with_their_tags():
    """The hinge chain's limited and motor joints store rows tagged angular << 5 | axis << 3 | kind;
    the device's stored rows equal the oracle's, tag for tag and bit for bit."""
    s = scenes.hinge_chain()
    o, g = run_pair(s, 60)
    for q in range(4):
        to, lo = o.joint_warm(q)
        tg, lg = g.joint_rows(q)
        assert to == tg and lo.view(np.uint32).tolist() == lg.view(np.uint32).tolist(), q
    tags, _ = g.joint_rows(1)
    assert (32 | 2 << 3 | 1) in tags  # the motor row of axis r2


def test_mixed_ball_and_d6_joints_hash_their_stored_rows():
    """One D6 joint among ball joints: every joint's stored rows and the step hash's input bytes
    agree with the oracle's (the stored-joint count covers both kinds)."""
    s = scenes.ragdolls_small(n_towers=3, tower_h=4, n_ragdolls=4)
    a, b, fa, fb = s["joints"].pop()
    s["d6"] = [dict(a=a, b=b, fa=np.asarray(fa, np.float32), fb=np.asarray(fb, np.float32), t_mask=7, r_mask=0, hinge=-1)]
    o, g = run_pair(s, 20)
    assert o.hash_input() == g.hash_input()
    nj = len(s["joints"]) + 1
    for q in range(nj):
        to, lo = o.joint_warm(q)
        tg, lg = g.joint_rows(q)
        assert to == tg and lo.view(np.uint32).tolist() == lg.view(np.uint32).tolist(), q


def test_batch_with_d6_joints_matches_single_worlds_and_replays():
    worlds_s = [scenes.d6_ragdolls(seed=30 + k, n_towers=2, tower_h=3, n_ragdolls=2 + k) for k in range(3)]
    singles = [rphys.World(s) for s in worlds_s]
    for w in singles:
        w.step_n(30)
    b = rphys.Batch([rphys.World(s) for s in worlds_s])
    b.step_n(30, graph=True)
    assert b.hashes() == [w.hash() for w in singles]
    o = oracle_py.OracleWorld(worlds_s[2])
    for _ in range(30):
        o.step()
    assert o.hash() == b.hashes()[2]
    b.snapshot(5)
    b.step_n(12)
    want = b.hashes()
    b.restore(5)
    b.step_n(12, graph=True)
    assert b.hashes() == want


def test_world_snapshot_carries_the_d6_rows():
    g = rphys.World(scenes.hinge_chain())
    g.step_n(40)
    snap = g.snapshot()
    a = []
    for _ in range(20):
        g.step()
        a.append(g.hash())
    g.restore(snap)
    b = []
    for _ in range(20):
        g.step()
        b.append(g.hash())
    g.free_snapshot(snap)
    assert a == b


@pytest.fixture
def caps():
    L = rphys.lib()
    yield L
    L.rphys_debug_set_initial_capacity(0, 0, 0, 0)


def test_overflow_replay_with_d6_joints_bit_exact(caps):
    """Tiny initial capacities on the D6 ragdolls: the first ticks with contacts overflow and replay,
    the D6 joints' carried rows restored from their backup; bit-exact with the oracle every tick."""
    caps.rphys_debug_set_initial_capacity(64, 256, 64, 16)
    s = scenes.d6_ragdolls(seed=13, n_towers=4, tower_h=5, n_ragdolls=6)
    run_pair(s, 60)


def test_residual_mode_with_d6_joints_matches_the_oracle():
    s = scenes.d6_ragdolls(seed=17)
    s["residual"] = (1.0e-4, 60)
    o = oracle_py.OracleWorld(s)
    g = rphys.World(s)
    for t in range(80):
        o.step()
        g.step()
        assert o.hash() == g.hash(), "tick %d" % (t + 1)
        so, sg = o.stats(), g.stats()
        assert (so.iters_used, so.iters_sum) == (sg.pgs_iters, sg.pgs_iters_sum), (t + 1, so.iters_used, sg.pgs_iters)


@pytest.mark.parametrize("hz,ticks", [(60.0, 180), (240.0, 480)])
def test_jointed_bodies_in_a_restitution_pile_match_the_oracle(hz, ticks):
    """The W6 pile with a hinged box falling onto it and a box on a distance joint swinging in:
    restitution's time-of-impact rows and LCP treat both jointed boxes as immovable."""
    s = scenes.jointed_bounce_pile(hz=hz)
    s["name"] = "jointed_bounce_pile_%g" % hz
    o, g = run_pair(s, ticks)
    assert o.stats().contacts > 0
