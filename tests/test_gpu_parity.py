"""Device step vs the CPU oracle on the same seeded scenes (the parity gate).

Integer/index outputs are compared bit for bit (broadphase pairs, contact colliders, feature ids,
terrain cells); the f32 state is compared bit for bit as well (both sides evaluate the same
operation sequence), with the 1e-5 relative bound checked separately as the stated tolerance.
All sizes here run in seconds on the oracle.
"""
import numpy as np
import pytest

import oracle_py
import rphys
import scenes

pytestmark = pytest.mark.gpu

REL_TOL = 1e-5  # positions / velocities after one tick


def pair_worlds(scene):
    o = oracle_py.OracleWorld(scene)
    g = rphys.World(scene)
    return o, g


def assert_state_close(o, g, rel=REL_TOL, exact=True):
    so, sg = o.state(), g.state()
    for name, a, b in zip(("pos", "rot", "lin", "ang"), so, sg):
        if exact:
            bad = np.nonzero(a.view(np.uint32) != b.view(np.uint32))
            assert bad[0].size == 0, "%s differs at bodies %s: oracle %s device %s" % (name, bad[0][:5], a[bad[0][:3]], b[bad[0][:3]])
        scale = np.maximum(np.abs(a), 1.0)
        assert np.all(np.abs(a - b) <= rel * scale), name


def contacts_equal(co, cg):
    assert len(co) == len(cg), (len(co), len(cg))
    if len(co) == 0:
        return
    np.testing.assert_array_equal(co["a_col"], cg["a_collider"])
    np.testing.assert_array_equal(co["b_col"], cg["b_collider"])
    np.testing.assert_array_equal(co["feat"], cg["feature"])
    np.testing.assert_array_equal(co["cell"], cg["cell"])
    np.testing.assert_array_equal(co["point"].view(np.uint32), cg["point"].view(np.uint32))
    np.testing.assert_array_equal(co["normal"].view(np.uint32), cg["normal"].view(np.uint32))
    np.testing.assert_array_equal(np.asarray(co["depth"]).view(np.uint32), np.asarray(cg["depth"]).view(np.uint32))


@pytest.mark.parametrize("seed", [7, 11])
def test_small_mixed_first_ticks_bit_exact(seed):
    s = scenes.small_mixed(n=96, seed=seed)
    o, g = pair_worlds(s)
    assert o.hash() == g.hash(), "initial state hash"
    for t in range(40):
        o.step()
        g.step()
        po, pg = o.pairs(), g.pairs()
        np.testing.assert_array_equal(po, pg, err_msg="pairs at tick %d" % (t + 1))
        contacts_equal(o.contacts(), g.contacts())
        assert_state_close(o, g)
        assert o.hash() == g.hash(), "hash at tick %d" % (t + 1)


def test_box_lattice_ticks_bit_exact():
    s = scenes.box_lattice()
    o, g = pair_worlds(s)
    for t in range(30):
        o.step()
        g.step()
        if t in (0, 9, 29):
            np.testing.assert_array_equal(o.pairs(), g.pairs())
            contacts_equal(o.contacts(), g.contacts())
        assert_state_close(o, g)
        assert o.hash() == g.hash(), "hash at tick %d" % (t + 1)


def test_run_to_run_hashes_identical():
    s = scenes.small_mixed(n=128, seed=3)
    hs = []
    for _ in range(2):
        g = rphys.World(s)
        seq = []
        for _ in range(30):
            g.step()
            seq.append(g.hash())
        hs.append(seq)
        g.close()
    assert hs[0] == hs[1]


def test_step_n_equals_single_steps():
    s = scenes.small_mixed(n=64, seed=5)
    a, b = rphys.World(s), rphys.World(s)
    for _ in range(20):
        a.step()
    b.step_n(20)
    assert a.hash() == b.hash()


def test_snapshot_rollback_replays_hashes():
    """ffi/src/tests.rs:1616-1684: snapshot -> steps -> restore -> steps, every hash equal."""
    s = scenes.small_mixed(n=64, seed=9)
    g = rphys.World(s)
    for _ in range(25):
        g.step()
    snap = g.snapshot()
    first = []
    for _ in range(15):
        g.step()
        first.append(g.hash())
    g.restore(snap)
    second = []
    for _ in range(15):
        g.step()
        second.append(g.hash())
    g.free_snapshot(snap)
    assert first == second


def test_batch_matches_single_worlds():
    ss = [scenes.small_mixed(n=n, seed=20 + i) for i, n in enumerate([40, 90, 17])]
    singles = [rphys.World(s) for s in ss]
    fresh = [rphys.World(s) for s in ss]
    batch = rphys.Batch(fresh)
    for _ in range(12):
        for w in singles:
            w.step()
    batch.step_n(12)
    assert batch.hashes() == [w.hash() for w in singles]


def test_batch_graph_matches_stream():
    ss = [scenes.small_mixed(n=n, seed=40 + i) for i, n in enumerate([33, 70])]
    b1 = rphys.Batch([rphys.World(s) for s in ss])
    b2 = rphys.Batch([rphys.World(s) for s in ss])
    b1.step_n(10)
    b2.step_n(10, graph=True)
    assert b1.hashes() == b2.hashes()


def test_batch_rollback_window():
    """8-tick rollback window replayed from a device snapshot gives the same hashes."""
    ss = [scenes.small_mixed(n=n, seed=60 + i) for i, n in enumerate([50, 80])]
    b = rphys.Batch([rphys.World(s) for s in ss])
    b.step_n(20)
    b.snapshot(0)
    b.step_n(8)
    h1 = b.hashes()
    b.restore(0)
    b.step_n(8)
    assert b.hashes() == h1


def test_batch_profiled_matches_step_n():
    ss = [scenes.small_mixed(n=n, seed=80 + i) for i, n in enumerate([45, 75, 30])]
    b1 = rphys.Batch([rphys.World(s) for s in ss])
    b2 = rphys.Batch([rphys.World(s) for s in ss])
    b1.step_n(14)
    b2.step_n(4)
    stages, ctr = b2.step_profiled(10)
    assert ctr["ticks"] == 10 and set(stages) >= {"broadphase", "vel_solve_small", "projection_shock"}
    assert b1.hashes() == b2.hashes()


@pytest.mark.parametrize("make", [lambda: scenes.small_mixed(n=96, seed=7), lambda: scenes.box_lattice()], ids=["mixed96", "boxes512"])
def test_residual_mode_bit_exact_with_equal_iteration_counts(make):
    """Each island iterates until its sweep residual <= 1e-4 (at most 100 sweeps); the device and the
    oracle must stop every island at the same sweep and agree bit for bit."""
    s = make()
    s["residual"] = (1.0e-4, 100)
    o, g = pair_worlds(s)
    for t in range(25):
        o.step()
        g.step()
        assert_state_close(o, g)
        assert o.hash() == g.hash(), "hash at tick %d" % (t + 1)
        so, sg = o.stats(), g.stats()
        assert (so.iters_used, so.iters_sum) == (sg.pgs_iters, sg.pgs_iters_sum), (t + 1, so.iters_used, sg.pgs_iters)


def test_graph_recaptured_when_iterations_change():
    s = scenes.small_mixed(n=64, seed=13)
    a, b = rphys.World(s), rphys.World(s)
    a.step_n(6)
    for _ in range(6):
        b.step()
    for w in (a, b):
        rphys.check(rphys.lib().rphys_world_set_solver_iterations(w.h, 4, 7))
    a.step_n(6)
    for _ in range(6):
        b.step()
    assert a.hash() == b.hash()
