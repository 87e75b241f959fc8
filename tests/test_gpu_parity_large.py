"""Device step vs the oracle at the full 16,384-body mixed workload, where the settled pile forms
one island of tens of thousands of contacts (the 1024-thread, level-ordered sweep tier).

The device runs the scene from spawn; its state (poses, velocities, sleep counters, last
velocities, warm cache) is then handed to the oracle, and both continue for a few ticks, compared
bit for bit (pairs, contacts with feature ids, state, step hash).  The oracle's cost per settled
tick (about half a second) bounds the number of compared ticks.
"""
import numpy as np
import pytest

import oracle_py
import rphys
import scenes

pytestmark = pytest.mark.gpu


def handover(scene, ticks):
    g = rphys.World(scene)
    g.step_n(ticks)
    st = g.export_state()
    o = oracle_py.OracleWorld(scene)
    o.load_state(st)
    assert o.hash() == g.hash(), "state handed to the oracle hashes differently"
    return g, o


def compare_tick(g, o, t):
    """Steps both one tick (the device through the profiled step) and compares; returns the device
    tick's profile counters."""
    _, ctr = g.step_profiled(1)
    o.step()
    np.testing.assert_array_equal(o.pairs(), g.pairs(), err_msg="pairs at tick %d" % t)
    co, cg = o.contacts(), g.contacts()
    assert len(co) == len(cg), (len(co), len(cg))
    np.testing.assert_array_equal(co["feat"], cg["feature"])
    np.testing.assert_array_equal(co["cell"], cg["cell"])
    so, sg = o.state(), g.state()
    for name, a, b in zip(("pos", "rot", "lin", "ang"), so, sg):
        bad = np.nonzero(a.view(np.uint32) != b.view(np.uint32))[0]
        assert bad.size == 0, "%s differs at %d bodies, first %s" % (name, bad.size, bad[:5])
    assert o.hash() == g.hash(), "hash at tick %d" % t
    return ctr


def test_settled_pile_continues_bit_exact():
    """The compared ticks run the big island in split form: its wide list-order and shock-order
    levels as GPU-wide launches, the rest as one-workgroup tail sweeps."""
    s = scenes.mixed_shapes()
    g, o = handover(s, 150)
    for t in range(151, 153):
        ctr = compare_tick(g, o, t)
        assert ctr["wide_levels"] > 0 and ctr["wide_levels_shock"] > 0, ctr
    st = g.stats()
    assert st.islands > 0 and st.island_contacts > 20000, (st.islands, st.island_contacts)


def test_step_profiled_equals_step_n():
    s = scenes.mixed_shapes()
    a, b = rphys.World(s), rphys.World(s)
    a.step_n(160)
    b.step_n(120)
    stages, ctr = b.step_profiled(40)
    assert ctr["ticks"] == 40 and ctr["island_contacts_global"] > 0, ctr
    assert a.hash() == b.hash()


def test_world_of_65536_bodies_bit_exact():
    """A world of 65,536 mixed bodies (four times the broadphase's one-workgroup sort width):
    ordered by the radix-sort broadphase path, bit-exact with the oracle from spawn (pairs,
    contacts, state, hash) over its first ticks, where the dense spawn already overlaps."""
    s = scenes.mixed_shapes(n=65536, hf_n=512, extent=200.0)
    g, o = rphys.World(s), oracle_py.OracleWorld(s)
    oracle_py.lib().or_set_threads(8)
    try:
        for t in range(1, 4):
            ctr = compare_tick(g, o, t)
        assert ctr["pairs"] > 0, ctr
    finally:
        oracle_py.lib().or_set_threads(1)


def test_batch_with_one_world_above_the_sort_width_matches_single_worlds():
    """A batch holding one world above the one-workgroup sort width and two small ones (the whole
    scene then takes the radix-sort order): the same per-world hashes as the worlds stepped
    alone, every tick."""
    scs = [scenes.small_mixed(n=96, seed=3), scenes.mixed_shapes(n=20000, hf_n=256, extent=100.0), scenes.box_lattice()]
    b = rphys.Batch([rphys.World(x) for x in scs])
    singles = [rphys.World(x) for x in scs]
    for t in range(12):
        b.step_n(1)
        for w in singles:
            w.step()
        assert b.hashes() == [w.hash() for w in singles], "tick %d" % (t + 1)
