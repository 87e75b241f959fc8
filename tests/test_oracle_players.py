"""The oracle's kinematic players (World::update_players, step.rs:544-639; add_player,
world.rs:466-475; the FFI's player entries, lib.rs:1580-1693) against closed forms (CPU only).

A player dropped over flat terrain falls under g until the capsule (radius, height / 2) meets
the ground within the 5 cm margin, then stands grounded; it walks at its speed along the input
direction; a dynamic box it runs into is woken and kicked with -10 N s
along the contact normal (which points from the player to the box); a snapshot carries the controller state, so a rollback replays.
"""
import numpy as np
import pytest

import oracle_py as O
import scenes

f = np.float32
DT = f(1.0) / f(60.0)


def flat_world(bodies=(), h=0.5, n=33, cell=1.0):
    hts = np.full(n * n, h, f)
    hf = dict(nx=n, nz=n, sx=f(cell), sz=f(cell), ox=f(-16.0), oz=f(-16.0), heights=hts)
    s = scenes.bounce_scene_of(list(bodies) or [scenes.bounce_body(1, (0.2, 0.2, 0.2), (10.0, 5.0, 10.0), 1.0, 0)])
    s["hf"] = hf
    s["dt"] = DT
    return O.OracleWorld(s, capacity=16)


def test_player_falls_then_stands_grounded_on_flat_terrain():
    w = flat_world()
    r, height = 0.3, 1.6
    p = w.add_player([0.0, 3.0, 0.0, 0, 0, 0, 1], r, height, 4.0)
    assert p == 0
    ys, vv = [], []
    for t in range(180):
        w.step()
        pose, grounded, v, _ = w.player(p)
        ys.append(pose[1]); vv.append(v)
    # free fall at the start: vertical velocity -g t
    assert abs(vv[0] - (-9.81 * DT)) < 1e-6 and abs(vv[9] - (-9.81 * DT * 10)) < 1e-5
    pose, grounded, v, _ = w.player(p)
    rest = 0.5 + height * 0.5 + r
    assert grounded and v == 0.0
    assert rest - 0.06 < pose[1] < rest + 0.06
    assert max(ys[-30:]) - min(ys[-30:]) < 1e-3, "standing still"


def test_reference_player_tests():
    """ffi/src/tests.rs:1515-1605: a player started in the air falls and moves along +x with its
    input, and is not grounded after 10 ticks; one dropped on a flat 2 x 2 heightfield of 10 m cells
    settles grounded with y in (1.4, 1.6) (half height 1 + radius 0.5)."""
    s = scenes.bounce_scene_of([scenes.bounce_body(1, (0.2, 0.2, 0.2), (50.0, 5.0, 50.0), 1.0, 0)])
    s["hf"] = None
    w = O.OracleWorld(s, capacity=16)
    p = w.add_player([0.0, 10.0, 0.0, 0, 0, 0, 1], 0.4, 1.8, 5.0)
    w.player_set_input(p, [1.0, 0.0, 0.0])
    for _ in range(10):
        w.step(f(1.0) / f(60.0))
    pose, grounded, _, _ = w.player(p)
    assert pose[1] < 10.0 and pose[0] > 0.0 and not grounded
    s = scenes.bounce_scene_of([scenes.bounce_body(1, (0.2, 0.2, 0.2), (50.0, 5.0, 50.0), 1.0, 0)])
    s["hf"] = dict(nx=2, nz=2, sx=f(10.0), sz=f(10.0), ox=f(0.0), oz=f(0.0), heights=np.zeros(4, f))
    w = O.OracleWorld(s, capacity=16)
    p = w.add_player([5.0, 3.0, 5.0, 0, 0, 0, 1], 0.5, 2.0, 5.0)
    for _ in range(60):
        w.step(f(1.0) / f(60.0))
    pose, grounded, _, _ = w.player(p)
    assert grounded and 1.4 < pose[1] < 1.6


def test_player_walks_at_its_speed():
    w = flat_world()
    speed = 3.0
    p = w.add_player([0.0, 1.9, 0.0, 0, 0, 0, 1], 0.3, 1.6, speed)
    for _ in range(60):
        w.step()
    x0 = w.player(p)[0][0]
    w.player_set_input(p, [2.0, 0.0, 0.0])  # normalised: the speed alone sets the pace
    for _ in range(30):
        w.step()
    x1 = w.player(p)[0][0]
    assert abs((x1 - x0) - speed * DT * 30) < 0.02


def test_player_kicks_a_dynamic_box_it_runs_into():
    box = scenes.bounce_body(1, (0.3, 0.3, 0.3), (2.0, 0.8, 0.0), 2.0, 0)
    w = flat_world([box])
    for _ in range(40):
        w.step()
    # the controller's capsule (r 0.6) is wider than the player's collider (world.rs:466-475 gives
    # it the default r 0.4), so the controller reaches the box before the contact solve does
    p = w.add_player([0.0, 2.3, 0.0, 0, 0, 0, 1], 0.6, 1.6, 4.0)
    for _ in range(20):
        w.step()
    w.player_set_input(p, [1.0, 0.0, 0.0])
    kicked = False
    for _ in range(60):
        w.step()
        lin = w.state()[2][0]
        if abs(lin[0]) > 1.0:
            kicked = True
            break
    # the kick is -10 N s along the contact normal, which points from the player to the box
    # (collide_shapes' A -> B): the box is thrown back toward the player (5 m/s for 2 kg before
    # the step's own contact solve takes some of it)
    assert kicked and -5.0 < lin[0] < -1.0, lin


def test_rollback_replays_a_moving_player():
    w = flat_world()
    p = w.add_player([0.0, 2.5, 0.0, 0, 0, 0, 1], 0.3, 1.6, 2.0)
    w.player_set_input(p, [0.0, 0.0, 1.0])
    for _ in range(10):
        w.step()
    snap = O.lib().or_snapshot(w.h)
    try:
        a = []
        for _ in range(40):
            w.step()
            a.append((w.hash(), w.player(p)[1], w.player(p)[2]))
        O.lib().or_restore(w.h, snap)
        b = []
        for _ in range(40):
            w.step()
            b.append((w.hash(), w.player(p)[1], w.player(p)[2]))
        assert a == b
    finally:
        O.lib().or_snapshot_free(snap)
