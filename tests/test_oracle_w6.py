"""The oracle's W6 restitution (time-of-impact bounce rows, interior-point LCP, pseudo pass)
against closed forms and the reference's own restitution tests (CPU only).

Restated from crates/riftphys-world/tests/restitution_toi_test.rs: ball drops reach e^2 h and
e^4 h (:66-96), the drone box bounces flat without spin (:104-138), a bounce on a resting box
leaves it in place (:160-205), a zero-g head-on collision matches the 1-D closed form
(:207-232), a rollback across a bounce replays (:234-250), a side hit shares momentum through a
frictionless floor (:252-286).  The LCP numerics (bounce_lcp.rs) are checked against dense
linear algebra and a brute-force LCP: the minimum-degree order of a chain fills nothing
(bounce_lcp.rs:383-391), the sparse factor solves (K + CFM diag K) x = u (:371-381), and the
interior-point solution meets the LCP's complementarity on random systems.
"""
import ctypes as C
import itertools
import math

import numpy as np
import pytest

import oracle_py as O
import scenes

f = np.float32
G = 9.81


def world(bodies, gravity=True, hz=60.0, cap=16, pairs=()):
    s = scenes.bounce_scene_of(bodies, gravity=gravity, hz=hz)
    s["pairs"] = list(pairs)
    return O.OracleWorld(s, capacity=cap)


def pair_e(a, b):
    """materials.rs mix_pair: max restitution (no override applies to these pairs)."""
    tab = {scenes.MAT_RUBBER: 0.80, scenes.MAT_ALUMINUM: 0.40, scenes.MAT_CONCRETE: 0.0, scenes.MAT_ICE: 0.02}
    return max(tab[a], tab[b])


def body_state(w, i):
    pos, rot, lin, ang = w.state()
    return pos[i], rot[i], lin[i], ang[i]


def apexes(w, b, hz, secs, n, bottom):
    out, rising, apex = [], False, 0.0
    for _ in range(int(secs * hz)):
        w.step()
        y, vy = bottom(w), body_state(w, b)[2][1]
        if not rising and vy > 0.0:
            rising, apex = True, y
        if rising:
            apex = max(apex, y)
            if vy < 0.0:
                out.append(apex)
                rising = False
                if len(out) == n:
                    break
    return out


@pytest.mark.parametrize("hz", [60.0, 240.0])
@pytest.mark.parametrize("mat", [scenes.MAT_RUBBER, scenes.MAT_ALUMINUM])
def test_ball_drops_reach_e2_h_and_e4_h(hz, mat):
    e = pair_e(mat, mat)
    r = 0.1
    for h in (0.3, 1.0, 3.0):
        w = world([scenes.bounce_floor(mat), scenes.bounce_body(0, (r,), (0.0, r + h, 0.0), 0.1, mat)], hz=hz)
        a = apexes(w, 1, hz, 6.0, 2, lambda w: body_state(w, 1)[0][1] - r)
        w1, w2 = e * e * h, e ** 4 * h
        assert abs(a[0] - w1) < 0.02 * w1, (hz, mat, h, a, w1)
        if math.sqrt(2.0 * G * w1) > 1.3:
            assert len(a) > 1 and abs(a[1] - w2) < 0.03 * w2, (hz, mat, h, a, w2)


def test_flat_drone_bounces_to_e2_h_without_spin():
    hz, h = 60.0, 2.0
    half = (0.25, 0.05, 0.25)
    e = pair_e(scenes.MAT_ALUMINUM, scenes.MAT_CONCRETE)
    w = world([scenes.bounce_floor(scenes.MAT_CONCRETE), scenes.bounce_body(1, half, (0.0, h + half[1], 0.0), 1.2, scenes.MAT_ALUMINUM)], hz=hz)
    rising, apex, spin, toi = False, 0.0, 0.0, 0
    for _ in range(int(3 * hz)):
        w.step()
        toi = max(toi, w.stats().toi_rows)
        p, q, v, om = body_state(w, 1)
        y = p[1] - half[1]
        if not rising and v[1] > 0.0:
            rising = True
        if rising:
            spin = max(spin, float(np.linalg.norm(om)))
            apex = max(apex, y)
            if v[1] < 0.0:
                break
    want = e * e * h
    assert toi >= 1, "no time-of-impact row fired"
    assert abs(apex - want) < 0.03 * want, (apex, want)
    assert spin < 1e-3, spin


@pytest.mark.parametrize("hz,m_up", [(60.0, 2.0), (240.0, 2.0), (60.0, 0.2), (240.0, 0.2)])
def test_bounce_on_a_resting_box_does_not_lift_it(hz, m_up):
    e = pair_e(scenes.MAT_ALUMINUM, scenes.MAT_ALUMINUM)
    half = (0.2, 0.1, 0.2)
    w = world([scenes.bounce_floor(scenes.MAT_CONCRETE), scenes.bounce_body(1, half, (0.0, 0.1, 0.0), 2.0, scenes.MAT_ALUMINUM),
               scenes.bounce_body(1, half, (0.0, 1.3, 0.0), m_up, scenes.MAT_ALUMINUM)], hz=hz)
    lo, hi = 1e9, -1e9
    rising, apex, done = False, 0.0, False
    for _ in range(int(1.5 * hz)):
        w.step()
        y = body_state(w, 1)[0][1]
        lo, hi = min(lo, y), max(hi, y)
        pu, _, vu, _ = body_state(w, 2)
        yu = pu[1] - 0.3
        if not done and not rising and vu[1] > 0.0:
            rising = True
        if not done and rising:
            apex = max(apex, yu)
            if vu[1] < 0.0:
                done = True
    want = e * e * 1.0
    assert abs(apex - want) < 0.03 * want, (apex, want)
    assert hi - lo < 1e-3, (lo, hi)


@pytest.mark.parametrize("hz", [60.0, 240.0])
def test_head_on_in_zero_g_matches_the_closed_form(hz):
    e = pair_e(scenes.MAT_ALUMINUM, scenes.MAT_ALUMINUM)
    m1, m2, v1, v2, r = 1.0, 2.0, 3.0, -1.0, 0.1
    w = world([scenes.bounce_body(0, (r,), (0.0, 1.0, 0.0), m1, scenes.MAT_ALUMINUM, vel=(v1, 0, 0, 0, 0, 0)),
               scenes.bounce_body(0, (r,), (f(2.0 * r) + f(0.83), 1.0, 0.0), m2, scenes.MAT_ALUMINUM, vel=(v2, 0, 0, 0, 0, 0))],
              gravity=False, hz=hz)
    t_end = 0.5
    for _ in range(int(round(t_end * hz))):
        w.step()
    p = m1 * v1 + m2 * v2
    want1 = (p + m2 * e * (v2 - v1)) / (m1 + m2)
    want2 = (p + m1 * e * (v1 - v2)) / (m1 + m2)
    got1, got2 = body_state(w, 0)[2][0], body_state(w, 1)[2][0]
    gap = body_state(w, 1)[0][0] - body_state(w, 0)[0][0] - 2.0 * r
    want_gap = e * (v1 - v2) * (t_end - 0.83 / (v1 - v2))
    assert abs(got1 - want1) < 0.01 * abs(want1), (got1, want1)
    assert abs(got2 - want2) < 0.01 * abs(want2), (got2, want2)
    assert abs(gap - want_gap) < 2e-3, (gap, want_gap)


def test_rollback_across_a_bounce_replays():
    hz = 60.0
    s, c = C.c_float(), C.c_float()
    O.lib().or_sincosf(f(0.2) * f(0.5), C.byref(s), C.byref(c))  # Quat::from_rotation_z(0.2)
    box = scenes.bounce_body(1, (0.25, 0.05, 0.25), (0.0, 1.05, 0.0), 1.2, scenes.MAT_ALUMINUM, rot=(0.0, 0.0, s.value, c.value))
    w = world([scenes.bounce_floor(scenes.MAT_CONCRETE), box], hz=hz)
    for _ in range(24):
        w.step()
    snap = O.lib().or_snapshot(w.h)
    try:
        first, toi = [], 0
        for _ in range(30):
            w.step()
            toi = max(toi, w.stats().toi_rows)
            first.append(w.hash())
        assert toi > 0, "no bounce in the window"
        O.lib().or_restore(w.h, snap)
        again = []
        for _ in range(30):
            w.step()
            again.append(w.hash())
        assert first == again
    finally:
        O.lib().or_snapshot_free(snap)


@pytest.mark.parametrize("hz", [60.0, 240.0])
def test_side_hit_on_a_resting_box_shares_momentum(hz):
    e = pair_e(scenes.MAT_RUBBER, scenes.MAT_RUBBER)
    m, v = 1.0, 3.0
    # pair_props(Rubber, Ice) is the curated override (materials.rs:260-263) with friction removed
    slick = (scenes.MAT_ICE, scenes.MAT_RUBBER, 0.0, 0.0, 0.05, 0.08)
    w = world([scenes.bounce_floor(scenes.MAT_ICE), scenes.bounce_body(1, (0.15, 0.15, 0.15), (0.0, 0.15, 0.0), m, scenes.MAT_RUBBER)],
              hz=hz, pairs=[slick])
    for _ in range(30):
        w.step()
    w.add_body_spec(scenes.bounce_body(0, (0.1,), (f(-0.15) - f(0.1) - f(0.023), 0.15, 0.0), m, scenes.MAT_RUBBER, vel=(v, 0, 0, 0, 0, 0)))
    hit = False
    for _ in range(int(0.2 * hz)):
        w.step()
        if body_state(w, 2)[2][0] < 0.5 * v:
            hit = True
            break
    assert hit
    vb, vx = body_state(w, 2)[2][0], body_state(w, 1)[2][0]
    want_ball, want_box = (1.0 - e) * 0.5 * v, (1.0 + e) * 0.5 * v
    assert abs(vx - want_box) < 0.02 * want_box, (vx, want_box)
    assert abs(vb - want_ball) < 0.02 * want_box, (vb, want_ball)
    assert abs(m * (vb + vx) - m * v) < 0.01 * m * v


# ---- the LCP numerics (bounce_lcp.rs) ----

def csr(dense):
    n = dense.shape[0]
    ptr, idx, val = [0], [], []
    for i in range(n):
        for j in range(n):
            if dense[i, j] != 0.0 or i == j:
                idx.append(j)
                val.append(dense[i, j])
        ptr.append(len(idx))
    return np.array(ptr, np.uint32), np.array(idx, np.uint32), np.array(val, np.float64)


def chain(n):
    """bounce_lcp.rs:346-355: a chain of rows coupled to neighbours."""
    k = np.zeros((n, n))
    for i in range(n):
        if i > 0:
            k[i, i - 1] = -0.5
        k[i, i] = 2.0 + 0.1 * i
        if i + 1 < n:
            k[i, i + 1] = -0.5
    return k


def u32p(a):
    return a.ctypes.data_as(C.POINTER(C.c_uint32))


def dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def test_min_degree_is_a_permutation_and_fills_a_chain_not_at_all():
    ptr, idx, _ = csr(chain(25))
    order = np.zeros(25, np.uint32)
    l_nnz = O.lib().or_bn_min_degree(25, u32p(ptr), u32p(idx), u32p(order))
    assert sorted(order.tolist()) == list(range(25))
    assert l_nnz == 24
    # an arrow (row 0 coupled to all): the spokes (degree 1) go first, lowest index first; with one
    # spoke left the hub also has degree 1 and the tie goes to the lower index; no fill
    k = np.eye(12) * 4.0
    k[0, 1:] = k[1:, 0] = 1.0
    ptr, idx, _ = csr(k)
    order = np.zeros(12, np.uint32)
    assert O.lib().or_bn_min_degree(12, u32p(ptr), u32p(idx), u32p(order)) == 11
    assert order.tolist() == list(range(1, 11)) + [0, 11]


def test_sparse_factor_solves_the_compliant_system():
    rng = np.random.default_rng(3)
    for k in (chain(40), chain(7)):
        n = k.shape[0]
        ptr, idx, val = csr(k)
        u = np.sin(np.arange(n) * 0.37)
        x = np.zeros(n)
        assert O.lib().or_bn_pseudo_solve(n, u32p(ptr), u32p(idx), dp(val), dp(u), dp(x)) == 0
        want = np.linalg.solve(k + np.diag(1e-12 * np.diag(k)), u)
        assert np.max(np.abs(x - want)) <= 1e-12 * (1.0 + np.max(np.abs(want)))
    # a random sparse SPD matrix from rows sharing bodies
    J = rng.normal(size=(30, 18)) * (rng.random((30, 18)) < 0.15)
    k = J @ J.T + np.eye(30) * 0.5
    ptr, idx, val = csr(k)
    u = rng.normal(size=30)
    x = np.zeros(30)
    O.lib().or_bn_pseudo_solve(30, u32p(ptr), u32p(idx), dp(val), dp(u), dp(x))
    want = np.linalg.solve(k + np.diag(1e-12 * np.diag(k)), u)
    assert np.max(np.abs(x - want)) <= 1e-10 * (1.0 + np.max(np.abs(want)))


def brute_lcp(k, q):
    """y >= 0, w = K y + q >= 0, y w = 0, by enumerating the active set (K positive definite)."""
    n = len(q)
    for act in itertools.product([0, 1], repeat=n):
        a = [i for i in range(n) if act[i]]
        y = np.zeros(n)
        if a:
            y[a] = np.linalg.solve(k[np.ix_(a, a)], -q[a])
        w = k @ y + q
        if np.all(y >= -1e-12) and np.all(w >= -1e-10):
            return y
    raise AssertionError("no solution")


def test_bounce_qp_meets_complementarity_on_random_systems():
    rng = np.random.default_rng(11)
    for trial in range(40):
        n = int(rng.integers(1, 7))
        J = rng.normal(size=(n, 6))
        k = J @ J.T + np.eye(n) * rng.uniform(0.05, 1.0)
        b = rng.normal(size=n)
        acc0 = np.abs(rng.normal(size=n)) * (rng.random(n) < 0.5)
        ptr, idx, val = csr(k)
        x = np.zeros(n)
        solves = np.zeros(1, np.uint32)
        conv = O.lib().or_bn_bounce_qp(n, u32p(ptr), u32p(idx), dp(val), dp(b), dp(acc0), dp(x), u32p(solves))
        assert conv == 1, trial
        y = acc0 + x
        want = brute_lcp(k, -(k @ acc0) - b)
        assert np.max(np.abs(y - want)) < 1e-6 * (1.0 + np.max(np.abs(want))), (trial, y, want)
        assert 1 <= solves[0] <= 64


def test_toi_bounce_closed_forms():
    """solver.rs:197-212: a pair closing at v0 from gap s0 under relative acceleration a meets at
    the smallest root t_c of s0 + v0 t + a t^2 / 2, leaves at -e (v0 + a t_c), and the pseudo
    velocity puts it where the bounce leaves it at the step's end."""
    L = O.lib()
    dt, e = f(1.0 / 60.0), f(0.5)
    u, hp = C.c_float(), C.c_int()
    # no relative acceleration: meets at s0 / |v0|
    s0, v0 = f(0.02), f(-3.0)
    v_end = L.or_toi_bounce(s0, v0, 0.0, e, dt, C.byref(u), C.byref(hp))
    t_c = s0 / -v0
    assert hp.value == 1
    assert abs(v_end - e * 3.0) < 1e-6
    s_end = e * 3.0 * (dt - t_c)
    assert abs(u.value - ((s_end - s0) / dt - v_end)) < 1e-5
    # already touching (s0 <= 0): bounces at once, no pseudo row
    v_end = L.or_toi_bounce(-0.001, -2.0, -9.81, 0.8, dt, C.byref(u), C.byref(hp))
    assert hp.value == 0 and abs(v_end - (0.8 * 2.0 - 9.81 * dt)) < 1e-5
    # too weak to leave within the step: ends at rest on the contact
    v_end = L.or_toi_bounce(-0.001, -1.01, -900.0, 0.1, dt, C.byref(u), C.byref(hp))
    assert v_end == 0.0
    # no root inside the step: the pair bounces at the step's end
    v_end = L.or_toi_bounce(0.5, -1.5, 0.0, 0.5, dt, C.byref(u), C.byref(hp))
    assert abs(v_end - 0.0) < 1e-6 or abs(v_end - 0.75) < 1e-6
