"""Spatial queries on the device (rphys_world_raycast, rphys_world_capsule_sweep and their batched
forms) against the CPU oracle, bit for bit: hit flag, body id, fraction, point and normal.

A mixed scene over a heightfield is stepped on both sides (hashes equal), then seeded rays and
capsule sweeps are cast from above, sideways and from inside bodies; the batched entry points
must answer as the single ones; a world of terrain alone is queried as well.
"""
import numpy as np
import pytest

import oracle_py
import rphys
import scenes

pytestmark = pytest.mark.gpu
TERRAIN = 0xFFFFFFFF


def same_hit(o, g, sweep):
    assert bool(o.hit) == bool(g.hit), (o.hit, g.hit)
    if not o.hit:
        assert g.body_id == 0 and g.fraction == 1.0
        return
    want_id = TERRAIN if o.body == TERRAIN else o.body + 1
    assert g.body_id == want_id, (g.body_id, want_id)
    assert np.float32(o.fraction).view(np.uint32) == np.float32(g.fraction).view(np.uint32), (o.fraction, g.fraction)
    gp = np.array([g.point.x, g.point.y, g.point.z], np.float32)
    gn = np.array([g.normal.x, g.normal.y, g.normal.z], np.float32)
    np.testing.assert_array_equal(np.array(o.point[:], np.float32).view(np.uint32), gp.view(np.uint32))
    np.testing.assert_array_equal(np.array(o.normal[:], np.float32).view(np.uint32), gn.view(np.uint32))
    if sweep:
        assert bool(o.started) == bool(g.started_overlapping)


def stepped_pair(ticks=30):
    s = scenes.small_mixed(n=96, seed=13)
    o, g = oracle_py.OracleWorld(s), rphys.World(s)
    for _ in range(ticks):
        o.step()
        g.step()
    assert o.hash() == g.hash()
    return s, o, g


def random_queries(rng, n, pos):
    lo, hi = pos.min(axis=0) - 2.0, pos.max(axis=0) + 2.0
    out = []
    for k in range(n):
        kind = k % 3
        if kind == 0:      # from above, downwards with a tilt
            o = rng.uniform(lo, hi).astype(np.float32); o[1] = hi[1] + 3.0
            d = np.array([rng.uniform(-0.3, 0.3), -1.0, rng.uniform(-0.3, 0.3)], np.float32)
        elif kind == 1:    # sideways through the pile
            o = rng.uniform(lo, hi).astype(np.float32)
            a = rng.uniform(0, 2 * np.pi)
            d = np.array([np.cos(a), rng.uniform(-0.2, 0.2), np.sin(a)], np.float32)
        else:              # from a body's centre outwards
            o = pos[rng.integers(len(pos))].astype(np.float32)
            d = rng.normal(size=3).astype(np.float32)
        out.append((o, d, np.float32(rng.uniform(1.0, 40.0))))
    return out


def test_raycasts_match_the_oracle():
    s, o, g = stepped_pair()
    pos = o.state()[0]
    rng = np.random.default_rng(5)
    qs = random_queries(rng, 240, pos)
    hits = 0
    for k, (org, d, m) in enumerate(qs):
        ig = (k % 7 == 0) and k // 7 % len(pos) or 0
        ho = o.raycast(org, d, m, ignore=(ig - 1) if ig else -1)
        hg = g.raycast(org, d, m, ignore_body=ig)
        same_hit(ho, hg, False)
        hits += bool(ho.hit)
    assert hits > 60
    # the batched form answers as the single calls
    many = g.raycast_n([(org, d, m, 0) for org, d, m in qs])
    for (org, d, m), hg in zip(qs, many):
        same_hit(o.raycast(org, d, m), hg, False)


def test_capsule_sweeps_match_the_oracle():
    s, o, g = stepped_pair()
    pos = o.state()[0]
    rng = np.random.default_rng(8)
    started = moved = 0
    qs = []
    for k in range(150):
        frm = rng.uniform(pos.min(axis=0) - 1.0, pos.max(axis=0) + 1.0).astype(np.float32)
        if k % 5 == 0:
            frm = pos[rng.integers(len(pos))].astype(np.float32)
        to = (frm + rng.normal(size=3) * rng.uniform(0.0, 6.0)).astype(np.float32)
        if k % 11 == 0:
            to = frm.copy()
        r, hh = np.float32(rng.uniform(0.1, 0.5)), np.float32(rng.uniform(0.0, 0.8))
        qs.append((frm, to, r, hh))
        ho = o.capsule_sweep(frm, to, r, hh)
        hg = g.capsule_sweep(frm, to, r, hh)
        same_hit(ho, hg, True)
        started += bool(ho.hit and ho.started)
        moved += bool(ho.hit and not ho.started)
    assert started > 5 and moved > 5
    many = g.capsule_sweep_n([(frm, to, r, hh, 0) for frm, to, r, hh in qs])
    for (frm, to, r, hh), hg in zip(qs, many):
        same_hit(o.capsule_sweep(frm, to, r, hh), hg, True)


def test_queries_on_terrain_alone():
    s = scenes.small_mixed(n=8, seed=2)
    for k in ("shape", "dims", "pose", "vel", "dynamic", "density"):
        s[k] = s[k][:0]
    o, g = oracle_py.OracleWorld(s, capacity=4), rphys.World(s, max_bodies=4)
    hf = s["hf"]
    cx = float(hf["ox"]) + float(hf["sx"]) * (hf["nx"] - 1) * 0.5
    cz = float(hf["oz"]) + float(hf["sz"]) * (hf["nz"] - 1) * 0.5
    for dx, dz in [(0.0, 0.0), (3.3, -2.1), (-7.7, 5.5)]:
        org, d = (cx + dx, 60.0, cz + dz), (0.05, -1.0, 0.02)
        ho, hg = o.raycast(org, d, 100.0), g.raycast(org, d, 100.0)
        assert ho.hit and ho.body == TERRAIN
        same_hit(ho, hg, False)
        ho, hg = o.capsule_sweep(org, (cx + dx, -20.0, cz + dz), 0.4, 0.3), g.capsule_sweep(org, (cx + dx, -20.0, cz + dz), 0.4, 0.3)
        assert ho.hit
        same_hit(ho, hg, True)
