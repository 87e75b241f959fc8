"""Contact models on the device (compliant Hunt-Crossley rows and friction patches, solved by the
contact-model tier) against the CPU oracle, bit for bit: the step hash every tick (it carries the
patch impulses), the final state.

The reference's compliant and patch-friction scenes (compliant_contact_test.rs,
patch_friction_test.rs) at reduced lengths, a settling pile and a ragdoll scene whose every pair
uses patch friction (joints, patches and contacts in one island; patches between sleeping bodies
in no island), compliant terrain, residual mode, a batch with contact models, graph-captured ticks
and rollback slots, and a world snapshot.
"""
import math

import numpy as np
import pytest

import oracle_py
import rphys
import scenes

pytestmark = pytest.mark.gpu
G = 9.81


def run_pair(s, ticks, cap=None, every=1):
    o = oracle_py.OracleWorld(s) if cap is None else oracle_py.OracleWorld(s, capacity=cap)
    g = rphys.World(s) if cap is None else rphys.World(s, max_bodies=cap)
    for t in range(ticks):
        o.step()
        g.step()
        if (t + 1) % every == 0 or t + 1 == ticks:
            assert o.hash() == g.hash(), "%s: hash differs at tick %d" % (s.get("name"), t + 1)
    so, sg = o.state(), g.state()
    for a, b in zip(so, sg):
        np.testing.assert_array_equal(np.asarray(a, np.float32).view(np.uint32), np.asarray(b, np.float32)[:, :a.shape[1]].view(np.uint32))
    return o, g


CASES = {
    "compliant_box_rest": lambda: (scenes.contact_model_scene(20000.0, 1.0, 0.5, 1, (0.1, 0.1, 0.1), (0.0, 0.1, 0.0), hz=1000.0), 300),
    "compliant_sphere_bounce": lambda: (scenes.contact_model_scene(1000.0, 0.3, 0.0, 0, (0.05,), (0.0, 0.052, 0.0), vel=(0.0, -1.0, 0.0),
                                                                   hz=2000.0, gravity=False), 400),
    "compliant_slide": lambda: (scenes.contact_model_scene(50000.0, 2.0, 0.3, 1, (0.1, 0.1, 0.1), (0.0, 0.1 - G / 200000.0, 0.0),
                                                           vel=(2.0, 0.0, 0.0), hz=1000.0), 300),
    "patch_spin_rigid": lambda: (scenes.contact_model_scene(0.0, 0.0, 0.2, 1, (0.0524,) * 3, (0.0, 0.0524, 0.0), ang=(0.0, 6.0, 0.0),
                                                            hz=1480.0, patch=True, mass=0.37, mats=(scenes.MAT_PMMA, scenes.MAT_WOOD)), 400),
    "patch_spin_compliant": lambda: (scenes.contact_model_scene(13591.0, 3.0, 0.2, 1, (0.0524,) * 3,
                                                                (0.0, 0.0524 - G * 0.37 / (4 * 13591.0), 0.0), vel=(1.0, 0.0, 0.0),
                                                                ang=(0.0, 4.0, 0.0), hz=1480.0, patch=True, mass=0.37,
                                                                mats=(scenes.MAT_PMMA, scenes.MAT_WOOD)), 400),
}


@pytest.mark.parametrize("name", sorted(CASES))
def test_reference_contact_model_scenes_match_the_oracle(name):
    s, ticks = CASES[name]()
    s["name"] = name
    run_pair(s, ticks)


def test_damping_at_the_step_start_matches_the_oracle():
    s = scenes.contact_model_scene(1000.0, 0.3, 0.0, 0, (0.05,), (0.0, 0.052, 0.0), vel=(0.0, -1.0, 0.0), hz=2000.0, gravity=False)
    s["damping_at_start"] = True
    o, g = run_pair(s, 300)
    s2 = dict(s, damping_at_start=False)
    o2 = oracle_py.OracleWorld(s2)
    for _ in range(300):
        o2.step()
    assert o2.hash() != o.hash()


def test_compliant_slope_hold_matches_the_oracle():
    k, tilt = 20000.0, math.radians(20.0)
    c, s_ = math.cos(tilt), math.sin(tilt)
    h = 0.1 - G * c / (4.0 * k)
    rot = (0.0, 0.0, math.sin(tilt / 2.0), math.cos(tilt / 2.0))
    s = scenes.contact_model_scene(k, 1.0, 0.6, 1, (0.1, 0.1, 0.1), (-s_ * h, c * h, 0.0), hz=60.0, ground_tilt=tilt, rot=rot)
    run_pair(s, 240)


def patch_pile(seed=3, n=64, compliant_terrain=True):
    """A settling pile over a heightfield: every body pair solves friction by patches; terrain
    contacts compliant (with patches) or rigid."""
    s = scenes.small_mixed(n=n, seed=seed)
    s["contact_models"] = [(0, 0, 0.0, 0.0, True), (0, 4, 2.0e5 if compliant_terrain else 0.0, 1.0, True)]
    return s


def test_patch_pile_matches_the_oracle():
    o, g = run_pair(patch_pile(), 150)
    assert o.stats().patches > 5


def test_ragdolls_with_patches_and_joints_match_the_oracle():
    s = scenes.ragdolls_small(n_towers=3, tower_h=4, n_ragdolls=4)
    s["contact_models"] = [(0, 0, 0.0, 0.0, True), (0, 4, 0.0, 0.0, True)]
    run_pair(s, 90)


def test_patches_between_sleeping_bodies_match_the_oracle():
    """A pile allowed to sleep: once asleep its patches write no body (no island) and keep their
    carried impulses; hashes agree through the fall, the settling and the sleep."""
    s = patch_pile(seed=5, n=24, compliant_terrain=False)
    o, g = run_pair(s, 260, every=2)
    n = len(s["shape"])
    aw = np.zeros(n, np.uint8)
    oracle_py.lib().or_get_awake(o.h, aw.ctypes.data_as(oracle_py.C.POINTER(oracle_py.C.c_uint8)), n)
    assert int(np.sum(aw == 0)) > 0


def test_residual_mode_in_the_contact_model_tier_matches_the_oracle():
    s = patch_pile(seed=7, n=48)
    s["residual"] = (1.0e-4, 60)
    o = oracle_py.OracleWorld(s)
    g = rphys.World(s)
    for t in range(80):
        o.step()
        g.step()
        assert o.hash() == g.hash(), "tick %d" % (t + 1)
        so, sg = o.stats(), g.stats()
        assert (so.iters_used, so.iters_sum) == (sg.pgs_iters, sg.pgs_iters_sum), (t + 1, so.iters_used, sg.pgs_iters)


def test_batch_with_contact_models_matches_single_worlds_and_replays():
    # the batch takes its first world's pair table: all three worlds share it
    models = [(0, 0, 0.0, 0.0, True), (0, 4, 2.0e5, 1.0, True)]
    worlds_s = []
    for k in range(3):
        s = scenes.small_mixed(n=40 + 8 * k, seed=20 + k)
        s["contact_models"] = models
        worlds_s.append(s)
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
    b.snapshot(3)
    b.step_n(12)
    want = b.hashes()
    b.restore(3)
    b.step_n(12, graph=True)
    assert b.hashes() == want


def test_world_snapshot_carries_the_patch_impulses():
    g = rphys.World(patch_pile(seed=9, n=32))
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


def test_patch_world_with_spare_capacity_matches_the_oracle():
    """A world whose capacity exceeds its bodies hashes sleep slots past the live bodies after the
    carried patch entries: hashes agree while patches are carried."""
    s = patch_pile(seed=11, n=24)
    n = len(s["shape"])
    o, g = run_pair(s, 60, cap=n + 9, every=3)
    assert o.stats().patches > 0
