"""Ball joints in the oracle (CPU): pins from closed forms and invariants.

The reference's own ball-joint tests run the Rust world; none of its golden vectors cover a ball
joint, so the oracle's joint rows are pinned here by what the mechanics fixes: a rigid pendulum's
half period, the anchor staying put, a hanging body's joint impulse equal to its weight times dt,
the jointed-pair exclusion (step.rs:85-93,102), and replay after restore.
"""
import math
import os
import re
import sys

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)
sys.path.insert(0, os.path.join(HERE, "..", "riftphys-rust_amd"))

import oracle_py  # noqa: E402
import scenes  # noqa: E402


def qrot(q, v):
    x, y, z, w = (float(c) for c in q)
    b = np.array([x, y, z])
    v = np.asarray(v, np.float64)
    return v * (w * w - b @ b) + b * (2.0 * (v @ b)) + np.cross(b, v) * 2.0 * w


def anchor_gaps(world, scene):
    pos, rot = world.state()[0], world.state()[1]
    out = []
    for a, b, fa, fb in scene["joints"]:
        xa = pos[a] + qrot(rot[a], fa[:3])
        xb = pos[b] + qrot(rot[b], fb[:3])
        out.append(np.linalg.norm(xb - xa))
    return np.array(out)


def two_body(anchor_pose, body_pose, fa, fb, body_shape=2, body_dims=(0.1, 0.15, 0.0), iters=(10, 20), dyn_a=0):
    s = scenes._empty(2)
    s["shape"][:] = [1, body_shape]
    s["dims"][0] = (0.2, 0.2, 0.2)
    s["dims"][1] = body_dims
    s["dynamic"][:] = [dyn_a, 1]
    s["pose"][0] = anchor_pose
    s["pose"][1] = body_pose
    s["joints"] = [(0, 1, fa, fb)]
    s["hf"] = None
    s["pairs"] = list(scenes.CONFIG_PAIRS)
    s["iters"] = iters
    s["dt"] = np.float32(1.0 / 60.0)
    return s


def test_pendulum_holds_its_anchor_and_swings_with_the_physical_period():
    # capsule (r 0.1, hh 0.15: length 0.5) hanging from its end at a static box's centre, released
    # horizontally; a physical pendulum of pivot distance 0.25 and I = m (L^2/12 + 0.25^2) swings to
    # the other horizontal side in half a period: T = 2 pi sqrt(I / (m g d)) K(90 deg) / (pi / 2)
    q = scenes.quat_yaw_pitch(0.0, math.pi / 2)
    s = two_body((0, 5, 0, 0, 0, 0, 1), (0, 5, 0.25, *q), (0, 0, 0, 0, 0, 0, 1), (0, -0.25, 0, 0, 0, 0, 1))
    w = oracle_py.OracleWorld(s)
    zs, gaps = [], []
    for _ in range(120):
        w.step()
        zs.append(float(w.state()[0][1][2]))
        gaps.append(anchor_gaps(w, s).max())
    assert max(gaps) < 1e-5
    # the far apex: the first swing's lowest z
    t_far = int(np.argmin(zs[:70])) + 1
    r, hh = 0.1, 0.15
    # the engine's capsule inertia about a transverse axis, m (r^2 / 4 + h^2 / 12) with h = 2 hh
    # (geom/src/mass.rs:92-119), plus the pivot offset 0.25
    I_over_m = r * r / 4.0 + (2 * hh) ** 2 / 12.0 + 0.25 ** 2
    T_small = 2 * math.pi * math.sqrt(I_over_m / (9.81 * 0.25))
    K90 = 1.8540746773013719  # complete elliptic integral K(sin 45 deg)
    half = 0.5 * T_small * K90 / (math.pi / 2) * 60.0
    assert abs(t_far - half) < 0.08 * half, (t_far, half)
    # energy: the far apex is back at the release height
    assert abs(min(zs[:70]) + 0.25) < 1e-3


def test_hanging_capsule_joint_carries_its_weight():
    # capsule hanging straight down from a static box; frame A = identity, so rows 0, 8, 16 are
    # world x, y, z and the joint impulse on the capsule is +lambda_y = m g dt per step
    s = two_body((0, 5, 0, 0, 0, 0, 1), (0, 5 - 0.2 - 0.25, 0, 0, 0, 0, 1), (0, -0.2, 0, 0, 0, 0, 1), (0, 0.25, 0, 0, 0, 0, 1))
    w = oracle_py.OracleWorld(s)
    for _ in range(20):
        w.step()
    tags, lams = w.joint_warm(0)
    assert tags == [0, 8, 16]
    r, hh = 0.1, 0.15
    m = 1000.0 * (math.pi * r * r * 2 * hh + 4.0 / 3.0 * math.pi * r ** 3)
    expect = m * 9.81 / 60.0
    assert abs(lams[1] - expect) < 1e-3 * expect, (lams, expect)
    assert abs(lams[0]) < 1e-4 and abs(lams[2]) < 1e-4
    assert anchor_gaps(w, s).max() < 1e-5


def test_jointed_pair_makes_no_contacts():
    # two overlapping boxes: contacts without the joint, none with it (the pair itself still counts)
    def scene(joined):
        s = two_body((0, 3, 0, 0, 0, 0, 1), (0.3, 3, 0, 0, 0, 0, 1), (0.15, 0, 0, 0, 0, 0, 1), (-0.15, 0, 0, 0, 0, 0, 1),
                     body_shape=1, body_dims=(0.2, 0.2, 0.2))
        if not joined:
            s["joints"] = []
        return s
    for joined in (False, True):
        w = oracle_py.OracleWorld(scene(joined))
        w.step()
        st = w.stats()
        assert st.pairs == 1
        assert (st.contacts == 0) == joined


def test_snapshot_restore_replays_joint_worlds():
    # restore sets poses through set_body_pose, which normalises each rotation again (lib.rs:
    # 1778-1784, dynamics.rs:130-147); a joint's orientation update need not be a fixed point of
    # that, so two replays from one restore agree with each other, bit for bit
    s = scenes.ragdolls_small()
    w = oracle_py.OracleWorld(s)
    for _ in range(30):
        w.step()
    snap = w.snapshot()
    runs = []
    for _ in range(2):
        w.restore(snap)
        hs = []
        for _ in range(10):
            w.step()
            hs.append(w.hash())
        runs.append(hs)
    w.free_snapshot(snap)
    assert runs[0] == runs[1]
    tags, lams = w.joint_warm(0)
    assert len(tags) == 3 and np.any(lams != 0)


def test_shake_test_matches_device_threshold():
    # the device compares min(|w|, 1) against the first float whose 2 acos < 0.5 rad
    hdr = open(os.path.join(HERE, "..", "riftphys-rust_amd", "csrc", "rp_dev.h")).read()
    bits = int(re.search(r"#define RP_SHAKE_COS_BITS (0x[0-9a-fA-F]+)u", hdr).group(1), 16)
    L = oracle_py.lib()
    lo = bits - 200000
    cand = np.arange(lo, bits + 200000, dtype=np.uint32)
    rng = np.random.default_rng(3)
    cand = np.concatenate([cand, rng.integers(0, 0x3F800000, 20000, dtype=np.uint32).astype(np.uint32)])
    vals = cand.view(np.float32)
    thr = np.array([bits], np.uint32).view(np.float32)[0]
    for v in vals:
        assert bool(L.or_shake_ok(float(v))) == bool(min(abs(v), np.float32(1.0)) >= thr), v
    assert L.or_shake_ok(float("nan")) == 1


def test_towers_and_chains_workload_shape():
    s = scenes.towers_ragdolls()
    assert len(s["shape"]) == 16384
    assert len(s["joints"]) == 200 * 14
    assert int((s["shape"] == 2).sum()) == 3000
    w = oracle_py.OracleWorld(s)
    assert anchor_gaps(w, s).max() < 1e-5
