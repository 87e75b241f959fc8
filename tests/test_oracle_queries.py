"""The oracle's spatial queries (World::raycast, World::sweep_capsule; queries.rs:50-261) against
closed forms (CPU only).

The ray tests collider bounding boxes (entry distance, the box face's normal, the lowest collider
index on ties, ignore_body) and terrain triangles (height and the surface normal of a flat and a
sloped field); a collider beats the terrain at equal distance.  The capsule sweep reports a start
overlap against the true shapes (fraction 0), then the first hit of its tip or base sphere
against the boxes grown by the radius, and misses when it stays in place.
"""
import math

import numpy as np
import pytest

import oracle_py as O
import scenes

f = np.float32


def world(bodies, hf=None, cap=16):
    s = scenes.bounce_scene_of(bodies) if bodies else scenes.bounce_scene_of([scenes.bounce_floor(0)])
    if not bodies:
        for k in ("shape", "dims", "pose", "vel", "dynamic", "density", "material", "sleep_allowed"):
            s[k] = s[k][:0]
    s["hf"] = hf
    return O.OracleWorld(s, capacity=cap)


def box_at(x, y, z, he=(0.5, 0.5, 0.5)):
    return scenes.bounce_body(1, he, (x, y, z), 1.0, 0)


def flat_hf(n=9, h=0.0, cell=1.0, slope=0.0):
    hts = np.array([[h + slope * (i * cell) for i in range(n)] for _ in range(n)], f).reshape(-1)
    return dict(nx=n, nz=n, sx=f(cell), sz=f(cell), ox=f(-4.0), oz=f(-4.0), heights=hts)


def test_ray_down_onto_a_box_hits_its_top_face():
    w = world([box_at(0.0, 1.0, 0.0)])
    h = w.raycast((0.1, 5.0, 0.2), (0, -1, 0), 10.0)
    assert h.hit and h.body == 0
    assert abs(h.fraction - 3.5 / 10.0) < 1e-6
    assert tuple(h.normal) == (0.0, 1.0, 0.0)
    assert abs(h.point[1] - 1.5) < 1e-6
    # the ray stops short of the box: no hit
    assert not w.raycast((0.1, 5.0, 0.2), (0, -1, 0), 3.4).hit
    # a direction that is not unit length is normalised first
    h2 = w.raycast((0.1, 5.0, 0.2), (0, -7, 0), 10.0)
    assert h2.hit and abs(h2.fraction - 0.35) < 1e-6


def test_ray_ties_go_to_the_lowest_collider_and_ignore_skips_a_body():
    # two boxes side by side, the ray grazes both top faces at the same height
    w = world([box_at(1.0, 0.0, 0.0), box_at(0.0, 0.0, 0.0)])
    h = w.raycast((0.5, 3.0, 0.0), (0, -1, 0), 10.0)
    assert h.hit and h.body == 0
    h = w.raycast((0.5, 3.0, 0.0), (0, -1, 0), 10.0, ignore=0)
    assert h.hit and h.body == 1
    # the nearer box wins regardless of order
    w = world([box_at(0.0, 0.0, 0.0), box_at(0.0, 2.0, 0.0)])
    assert w.raycast((0.0, 6.0, 0.0), (0, -1, 0), 10.0).body == 1


def test_ray_hits_flat_and_sloped_terrain_with_the_surface_normal():
    w = world([], hf=flat_hf(h=0.25))
    h = w.raycast((0.3, 4.0, -0.6), (0, -1, 0), 10.0)
    assert h.hit and h.body == 0xFFFFFFFF
    assert abs(h.point[1] - 0.25) < 1e-6 and abs(h.fraction - 0.375) < 1e-6
    assert np.allclose(h.normal, (0, 1, 0))
    s = 0.5
    w = world([], hf=flat_hf(slope=s))
    x0 = 1.3
    h = w.raycast((x0, 10.0, 0.4), (0, -1, 0), 20.0)
    want_y = s * (x0 + 4.0)
    assert h.hit and abs(h.point[1] - want_y) < 1e-5
    n = np.array([-s, 1.0, 0.0]) / math.sqrt(1 + s * s)
    assert np.allclose(h.normal, n, atol=1e-6)
    # a horizontal ray above the field misses it
    assert not w.raycast((-3.0, 9.0, 0.0), (1, 0, 0), 5.0).hit


def test_collider_wins_a_tie_with_the_terrain():
    # a box whose top face is level with the terrain: equal distance, the collider stays best
    w = world([box_at(0.0, -0.5, 0.0)], hf=flat_hf(h=0.0))
    h = w.raycast((0.2, 2.0, 0.2), (0, -1, 0), 5.0)
    assert h.hit and h.body == 0


def test_capsule_starting_inside_a_box_reports_the_overlap():
    w = world([box_at(0.0, 0.0, 0.0)])
    h = w.capsule_sweep((0.0, 0.9, 0.0), (0.0, 5.0, 0.0), 0.3, 0.5)
    assert h.hit and h.started and h.body == 0 and h.fraction == 0.0
    assert abs(np.linalg.norm(h.normal) - 1.0) < 1e-5


def test_capsule_swept_down_onto_a_box_hits_with_its_base_sphere():
    w = world([box_at(0.0, 0.0, 0.0)])
    r, hh = 0.3, 0.5
    frm, to = (0.0, 4.0, 0.0), (0.0, 0.0, 0.0)
    h = w.capsule_sweep(frm, to, r, hh)
    # the base sphere's centre starts at 3.5 and meets the box top (0.5) grown by r at 0.8
    want = (3.5 - 0.8) / 4.0
    assert h.hit and not h.started and h.body == 0
    assert abs(h.fraction - want) < 1e-6
    assert tuple(h.normal) == (0.0, 1.0, 0.0)
    assert abs(h.point[1] - 0.5) < 1e-5
    # sideways into a box: the x face
    h = w.capsule_sweep((-4.0, 0.0, 0.0), (0.0, 0.0, 0.0), r, hh)
    assert h.hit and abs(h.fraction - (4.0 - 0.8) / 4.0) < 1e-6 and tuple(h.normal) == (-1.0, 0.0, 0.0)


def test_capsule_sweep_in_place_misses_and_sweeps_onto_terrain():
    w = world([box_at(5.0, 5.0, 5.0)], hf=flat_hf(h=0.0))
    assert not w.capsule_sweep((0.0, 3.0, 0.0), (0.0, 3.0, 0.0), 0.2, 0.4).hit
    h = w.capsule_sweep((0.0, 3.0, 0.0), (0.0, -1.0, 0.0), 0.2, 0.4)
    # base sphere centre 2.6 meets the triangles' zero-thickness boxes grown by 0.2 at 0.2
    assert h.hit and h.body == 0xFFFFFFFF and abs(h.fraction - 2.4 / 4.0) < 1e-6
    # a capsule resting in the terrain starts overlapped
    h = w.capsule_sweep((0.0, 0.5, 0.0), (0.0, 3.0, 0.0), 0.2, 0.4)
    assert h.hit and h.started and h.body == 0xFFFFFFFF
