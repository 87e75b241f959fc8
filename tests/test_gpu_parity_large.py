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
