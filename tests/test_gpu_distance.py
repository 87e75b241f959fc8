"""Distance joints on the device against the reference's own golden bits and the CPU oracle.

G1 (crates/riftphys-world/tests/distance_bits_test.rs:25-63): the device's state after 3 s of a box
swinging on a rigid, a compliant and a guard-controller distance joint, at 60 and 240 Hz, hashes
(FNV-1a-64 of the 13 f32 state words) to the six values the reference records; the step hash
agrees with the oracle's every tick.  Also: ragdolls whose chains mix rigid and compliant distance
joints with ball joints among contacts, a D6 joint with no velocity rows (all its axes compliant:
its position row alone holds it), and a batch with graph-captured ticks and rollback slots.
"""
import numpy as np
import pytest

import oracle_py
import rphys
import scenes
from test_oracle_goldens import GOLDEN, _qmul, _sincos, fnv_state

pytestmark = pytest.mark.gpu
f = np.float32


def g1_scene(kind, hz):
    """distance_bits_test.rs:25-44 as a scene: a static pivot at (0, 5, 0) and a 0.2 x 0.1 x 0.16 m
    box of density 500, turned and moving, on a distance joint to the pivot."""
    sy, cy = _sincos(f(0.3) * f(0.5))
    sx, cx = _sincos(f(0.2) * f(0.5))
    q = _qmul((f(0), sy, f(0), cy), (sx, f(0), f(0), cx))
    start = np.array([0.7, 4.3, 0.2], f)
    d = start - np.array([0, 5, 0], f)
    rest = np.sqrt(f(f(d[0] * d[0] + d[1] * d[1]) + d[2] * d[2]))
    joint = [(0, 1, rest, 0.0), (0, 1, f(rest * f(0.8)), 2.0e-3), (0, 1, f(rest * f(0.9)), 1.0e-3)][kind]
    s = dict(shape=np.array([1, 1], np.int32), dims=np.array([[0.5, 0.5, 0.5], [0.1, 0.05, 0.08]], f),
             pose=np.array([[0, 5, 0, 0, 0, 0, 1], [*start, *q]], f), vel=np.array([np.zeros(6), [0.0, 0.5, 1.5, 0.4, -1.0, 2.0]], f),
             dynamic=np.array([0, 1], np.int32), density=np.array([1000.0, 500.0], f), sleep_allowed=np.array([1, 0], np.int32),
             gravity=(0.0, -9.81, 0.0), hf=None, dt=f(1.0) / f(hz), dist=[joint], name="g1_%d_%g" % (kind, hz))
    return s


@pytest.mark.parametrize("kind,hz,want", GOLDEN)
def test_distance_joint_goldens_g1_on_the_device(kind, hz, want):
    s = g1_scene(kind, hz)
    g = rphys.World(s)
    o = oracle_py.OracleWorld(s)
    for t in range(int(3.0 * hz)):
        g.step()
        o.step()
        assert g.hash() == o.hash(), "%s: hash differs at tick %d" % (s["name"], t + 1)
    pos, rot, lin, ang = g.state()
    got = fnv_state(pos[1], rot[1], lin[1], ang[1])
    assert got == want, "kind %d at %g Hz: %016x, reference golden %016x" % (kind, hz, got, want)


def distance_ragdolls(seed=11, n_towers=3, tower_h=4, n_ragdolls=4):
    """ragdolls_small with one chain joint in three a rigid distance joint between the two capsules'
    centres (its rest their current distance) and one in three a compliant one (10 % longer)."""
    s = scenes.ragdolls_small(seed=seed, n_towers=n_towers, tower_h=tower_h, n_ragdolls=n_ragdolls)
    balls, dist = [], []
    for k, (a, b, fa, fb) in enumerate(s["joints"]):
        r = np.linalg.norm(np.asarray(s["pose"][b][:3], np.float64) - np.asarray(s["pose"][a][:3], np.float64))
        if k % 3 == 0:
            dist.append((a, b, float(r), 0.0))
        elif k % 3 == 1:
            dist.append((a, b, float(r) * 1.1, 1.0e-4))
        else:
            balls.append((a, b, fa, fb))
    s["joints"], s["dist"] = balls, dist
    return s


def test_ragdolls_with_distance_joints_match_the_oracle():
    s = distance_ragdolls()
    o = oracle_py.OracleWorld(s)
    g = rphys.World(s)
    for t in range(120):
        o.step()
        g.step()
        if t % 2 == 1:
            assert o.hash() == g.hash(), "hash differs at tick %d" % (t + 1)
    for a, b in zip(o.state(), g.state()):
        np.testing.assert_array_equal(np.asarray(a, f).view(np.uint32), np.asarray(b, f)[:, :a.shape[1]].view(np.uint32))


def test_d6_joint_without_velocity_rows_matches_the_oracle():
    """A rod on a D6 joint whose three translational axes are compliant and nothing else is
    constrained: no velocity rows, so only its position rows hold the rod."""
    s = scenes.hinge_rod(hz=120.0)
    j = s["d6"][0]
    j["t_mask"], j["r_mask"], j["hinge"] = 7, 0, -1
    j["compliance"] = [(False, 0, 1.0e-5), (False, 1, 1.0e-5), (False, 2, 1.0e-5)]
    o = oracle_py.OracleWorld(s)
    g = rphys.World(s)
    for t in range(240):
        o.step()
        g.step()
        assert o.hash() == g.hash(), "hash differs at tick %d" % (t + 1)
    p = g.state()[0][1]
    assert abs(np.linalg.norm(p - np.array([0, 10, 0], f)) - 0.5) < 0.05  # held near its rest length


def test_batch_with_distance_joints_matches_single_worlds_and_replays():
    worlds_s = [distance_ragdolls(seed=40 + k, n_towers=2, tower_h=3, n_ragdolls=2 + k) for k in range(3)]
    singles = [rphys.World(s) for s in worlds_s]
    for w in singles:
        w.step_n(30)
    b = rphys.Batch([rphys.World(s) for s in worlds_s])
    b.step_n(30, graph=True)
    assert b.hashes() == [w.hash() for w in singles]
    o = oracle_py.OracleWorld(worlds_s[1])
    for _ in range(30):
        o.step()
    assert o.hash() == b.hashes()[1]
    b.snapshot(2)
    b.step_n(10)
    want = b.hashes()
    b.restore(2)
    b.step_n(10, graph=True)
    assert b.hashes() == want
