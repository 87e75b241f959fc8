"""The oracle against the reference's own golden bits (CPU only).

G1: crates/riftphys-world/tests/distance_bits_test.rs:25-63 records six FNV-1a-64 hashes of one
body's 13 f32 state words after 3 s of a box swinging on a distance joint from a static pivot:
a rigid rod, a compliant rod (started stretched) and the guard controller's compliant joint
(world.rs:527-531: a distance joint at rest_guard, k_guard; GuardCtrl is never stepped by the
world), each at 60 and 240 Hz.  The scene has no colliders, so the hashes pin every non-contact
stage of World::step bit for bit: S1 gravity, the legacy distance velocity row
(solver.rs:795-820), S4 integration with the subdivided gyroscopic midpoint turn
(step.rs:207-270: the box's principal moments differ and its spin is not principal), the XPBD
distance row with SHAKE directions (bodies.rs:501-548), S6's velocity from corrections
(step.rs:312-379), glam's Quat / Mat3A operation order and libm's sincosf.
"""
import ctypes as C

import numpy as np
import pytest

import oracle_py as O

f = np.float32
GOLDEN = [(0, 60.0, 0x87f81a83f9615530), (0, 240.0, 0xe7411745726c1915),
          (1, 60.0, 0x6f111d612da3e10f), (1, 240.0, 0xed532e5e87913e18),
          (2, 60.0, 0xbb57c95b1772e5eb), (2, 240.0, 0xc1a92484be836750)]


def _sincos(x):
    s, c = C.c_float(), C.c_float()
    O.lib().or_sincosf(f(x), C.byref(s), C.byref(c))
    return f(s.value), f(c.value)


def _qmul(l, r):
    """glam's SSE2 Quat * Quat, term by term in f32."""
    lx, ly, lz, lw = l
    rx, ry, rz, rw = r
    ax, ay, az, aw = lw * rx, lw * ry, lw * rz, lw * rw
    bx, by, bz, bw = lx * rw, -(lx * rz), lx * ry, -(lx * rx)
    cx, cy, cz, cw = ly * rz, ly * rw, -(ly * rx), -(ly * ry)
    dx, dy, dz, dw = -(lz * ry), lz * rx, lz * rw, -(lz * rz)
    return ((ax + bx) + (cx + dx), (ay + by) + (cy + dy), (az + bz) + (cz + dz), (aw + bw) + (cw + dw))


def fnv_state(pos, rot, lin, ang):
    """distance_bits_test.rs:11-20: FNV-1a-64 over the f32 bit patterns, xor then multiply."""
    h = 0xcbf29ce484222325
    for x in [*pos, *rot, *lin, *ang]:
        h = ((h ^ int(f(x).view(np.uint32))) * 0x100000001b3) & 0xFFFFFFFFFFFFFFFF
    return h


def distance_scene(kind, hz):
    """distance_bits_test.rs:25-44, built through the oracle's C entry points."""
    L = O.lib()
    h = L.or_world_create(8)
    try:
        L.or_set_gravity(h, 0.0, -9.81, 0.0)
        z = np.zeros(6, f)
        pivot = np.array([0, 5, 0, 0, 0, 0, 1], f)
        unit = np.array([0.5, 0.5, 0.5], f)  # no mass: a static body ignores its shape
        a = L.or_add_body_nc(h, O.OR_BOX, O.fptr(unit), O.fptr(pivot), O.fptr(z), 0, 1000.0)
        sy, cy = _sincos(f(0.3) * f(0.5))  # Quat::from_rotation_y(0.3)
        sx, cx = _sincos(f(0.2) * f(0.5))  # Quat::from_rotation_x(0.2)
        q = _qmul((f(0), sy, f(0), cy), (sx, f(0), f(0), cx))
        start = np.array([0.7, 4.3, 0.2], f)
        pose = np.array([*start, *q], f)
        vel = np.array([0.0, 0.5, 1.5, 0.4, -1.0, 2.0], f)
        he = np.array([0.1, 0.05, 0.08], f)  # MassProps::from_box_density(he, 500)
        b = L.or_add_body_nc(h, O.OR_BOX, O.fptr(he), O.fptr(pose), O.fptr(vel), 1, 500.0)
        L.or_set_sleep_allowed(h, b, 0)
        d = start - np.array([0, 5, 0], f)
        rest = np.sqrt(f(f(d[0] * d[0] + d[1] * d[1]) + d[2] * d[2]))
        if kind == 0:
            L.or_add_distance_joint(h, a, b, rest, 0.0)
        elif kind == 1:
            L.or_add_distance_joint(h, a, b, f(rest * f(0.8)), 2.0e-3)
        else:
            L.or_add_distance_joint(h, a, b, f(rest * f(0.9)), 1.0e-3)
        dt = f(1.0) / f(hz)
        for _ in range(int(3.0 * hz)):
            L.or_step(h, dt)
        n = 2
        pos = np.zeros((n, 3), f); rot = np.zeros((n, 4), f); lin = np.zeros((n, 3), f); ang = np.zeros((n, 3), f)
        L.or_get_state(h, O.fptr(pos), O.fptr(rot), O.fptr(lin), O.fptr(ang))
        p = pos[1]
        assert np.all(np.isfinite(p)) and np.linalg.norm(p - np.array([0, 5, 0], f)) < 3.0
        return fnv_state(pos[1], rot[1], lin[1], ang[1])
    finally:
        L.or_world_destroy(h)


@pytest.mark.parametrize("kind,hz,want", GOLDEN)
def test_distance_joint_goldens_g1(kind, hz, want):
    got = distance_scene(kind, hz)
    assert got == want, "kind %d at %g Hz: %016x, reference golden %016x" % (kind, hz, got, want)


# G2/G3: restitution_toi_test.rs:596-619 (`w6_scene_hashes`, recorded in docs/PARITY_SPEC.md
# :2120-2126) prints the first 8 bytes of World::step_hash after three W6 bounce scenes at 60 and
# 240 Hz.  They pin, bit for bit: box-box SAT contacts and manifolds, the warm start and its cache,
# the shock pass, the W6 time-of-impact bounce rows, the interior-point LCP on the pile's
# support component, the pseudo pass, the integration with pseudo-velocities, sleep bookkeeping
# and the step-hash byte stream (World::with_capacity(32|16) fixes the hash's sleep-slot count).
W6_GOLDEN = [("pile", 60.0, "b6be817421a5936a"), ("pile", 240.0, "1a15ce00411301fa"),
             ("stack", 60.0, "56a889bb6e73d42c"), ("stack", 240.0, "90eac03965c85039"),
             ("drone", 60.0, "75f5412b43fa9cc0"), ("drone", 240.0, "a88414ba7682a6ed")]


def w6_golden_run(kind, hz, world_cls, **kw):
    """Runs one golden scene in world_cls (the oracle or the device binding); returns the world
    and the largest (toi rows, support rows, pseudo rows) seen by its stats, if it has them."""
    import scenes
    if kind == "pile":
        s = scenes.bounce_pile(hz)
        w = world_cls(s, **kw)
        for _ in range(int(hz)):
            w.step()
        w.add_body_spec(scenes.bounce_pile_drop())
        ticks = int(hz)
    elif kind == "stack":
        w = world_cls(scenes.bounce_stack(hz), **kw)
        ticks = int(3 * hz)
    else:
        w = world_cls(scenes.bounce_drone(hz), **kw)
        ticks = int(3 * hz)
    seen = [0, 0, 0, 0]
    for _ in range(ticks):
        w.step()
        st = w.stats() if hasattr(w, "stats") else None
        if st is not None and hasattr(st, "toi_rows"):
            seen = [max(seen[0], st.toi_rows), max(seen[1], st.support_rows), max(seen[2], st.pseudo_rows),
                    seen[3] + st.bounce_fallbacks]
    return w, seen


@pytest.mark.parametrize("kind,hz,want", W6_GOLDEN)
def test_w6_scene_goldens_g2_g3(kind, hz, want):
    cap = 16 if kind == "drone" else 32
    w, seen = w6_golden_run(kind, hz, lambda s: O.OracleWorld(s, capacity=cap))
    got = w.hash()[:8].hex()
    assert got == want, "%s at %g Hz: %s, reference golden %s" % (kind, hz, got, want)
    assert seen[0] > 0, "no time-of-impact row fired"
    assert seen[3] == 0, "a bounce component fell back to sweeping"
    if kind == "pile":
        assert seen[1] > 100, "the pile's support component is the large interior-point system"
