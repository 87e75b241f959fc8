"""Kinematic players on the device (rphys_add_player, rphys_player_set_input,
rphys_player_get_pose; World::update_players at the start of each tick) against the CPU oracle
and the reference's own player tests (ffi/src/tests.rs:1515-1605).

Players walk through a mixed scene over terrain, bumping and kicking bodies: the step hash, the
players' poses and grounded flags equal the oracle's every tick.  A batch of worlds with players
equals the single worlds (stream and graph launches); a world snapshot replays them.
"""
import numpy as np
import pytest

import oracle_py
import rphys
import scenes

pytestmark = pytest.mark.gpu
f = np.float32


def test_reference_player_tests_on_the_device():
    s = scenes.bounce_scene_of([scenes.bounce_body(1, (0.2, 0.2, 0.2), (50.0, 5.0, 50.0), 1.0, 0)])
    g = rphys.World(s, max_bodies=4)
    p = g.add_player([0.0, 10.0, 0.0, 0, 0, 0, 1], 0.4, 1.8, 5.0)
    g.player_set_input(p, [1.0, 0.0, 0.0])
    for _ in range(10):
        g.step()
    pose, grounded = g.player(p)
    assert pose[1] < 10.0 and pose[0] > 0.0 and not grounded
    s = scenes.bounce_scene_of([scenes.bounce_body(1, (0.2, 0.2, 0.2), (50.0, 5.0, 50.0), 1.0, 0)])
    s["hf"] = dict(nx=2, nz=2, sx=f(10.0), sz=f(10.0), ox=f(0.0), oz=f(0.0), heights=np.zeros(4, f))
    g = rphys.World(s, max_bodies=4)
    p = g.add_player([5.0, 3.0, 5.0, 0, 0, 0, 1], 0.5, 2.0, 5.0)
    for _ in range(60):
        g.step()
    pose, grounded = g.player(p)
    assert grounded and 1.4 < pose[1] < 1.6


def players_scene(seed=3, n=64):
    s = scenes.small_mixed(n=n, seed=seed)
    hf = s["hf"]
    cx = float(hf["ox"]) + float(hf["sx"]) * (hf["nx"] - 1) * 0.5
    cz = float(hf["oz"]) + float(hf["sz"]) * (hf["nz"] - 1) * 0.5
    top = float(np.max(hf["heights"])) + 2.5
    starts = [(cx - 4.0, top, cz), (cx + 3.0, top + 1.0, cz - 2.0), (cx, top + 0.5, cz + 4.0)]
    dims = [(0.4, 1.8, 4.0), (0.6, 1.6, 6.0), (0.3, 2.0, 3.0)]
    inputs = [(1.0, 0.0, 0.2), (-0.5, 0.0, 1.0), (0.0, 0.0, -1.0)]
    return s, starts, dims, inputs


def add_players(w, starts, dims):
    return [w.add_player([*p, 0, 0, 0, 1], *d) for p, d in zip(starts, dims)]


def test_players_walk_through_a_mixed_scene_like_the_oracle():
    s, starts, dims, inputs = players_scene()
    cap = len(s["shape"]) + len(starts)
    o, g = oracle_py.OracleWorld(s, capacity=cap), rphys.World(s, max_bodies=cap)
    po, pg = add_players(o, starts, dims), add_players(g, starts, dims)
    assert po == pg == [0, 1, 2]
    grounded = 0
    for t in range(150):
        if t == 20:
            for i, d in enumerate(inputs):
                o.player_set_input(po[i], d)
                g.player_set_input(pg[i], d)
        if t == 100:
            o.player_set_input(po[0], (0.0, 0.0, 0.0))
            g.player_set_input(pg[0], (0.0, 0.0, 0.0))
        o.step()
        g.step()
        assert o.hash() == g.hash(), "hash differs at tick %d" % (t + 1)
        for i in range(len(po)):
            pose_o, gr_o, _, _ = o.player(po[i])
            pose_g, gr_g = g.player(pg[i])
            np.testing.assert_array_equal(pose_o.view(np.uint32), pose_g.view(np.uint32))
            assert gr_o == gr_g
            grounded += gr_g
    assert grounded > 100


def test_batch_with_players_matches_single_worlds_and_replays():
    s, starts, dims, inputs = players_scene(seed=4, n=40)
    cap = len(s["shape"]) + len(starts)
    singles, fresh = [], []
    for k in range(2):
        for lst in (singles, fresh):
            w = rphys.World(s, max_bodies=cap)
            ids = add_players(w, starts[:2 + k], dims[:2 + k])
            for i in ids:
                w.player_set_input(i, inputs[i])
            lst.append(w)
    b = rphys.Batch(fresh)
    for _ in range(40):
        for w in singles:
            w.step()
    b.step_n(40)
    assert b.hashes() == [w.hash() for w in singles]
    # graph-captured ticks from the same start
    fresh2 = []
    for k in range(2):
        w = rphys.World(s, max_bodies=cap)
        for i in add_players(w, starts[:2 + k], dims[:2 + k]):
            w.player_set_input(i, inputs[i])
        fresh2.append(w)
    b2 = rphys.Batch(fresh2)
    b2.step_n(40, graph=True)
    assert b2.hashes() == b.hashes()
    # a world snapshot carries the players
    w = singles[1]
    snap = w.snapshot()
    a = []
    for _ in range(20):
        w.step()
        a.append((w.hash(), tuple(w.player(0)[0]), w.player(2)[1]))
    w.restore(snap)
    c = []
    for _ in range(20):
        w.step()
        c.append((w.hash(), tuple(w.player(0)[0]), w.player(2)[1]))
    w.free_snapshot(snap)
    assert a == c
