"""Device paths that the default sizes leave idle, forced on small scenes and checked against the
oracle bit for bit: every sweep tier (one island per lane, lane groups, one wave for islands of up
to 256 and up to 4,096 entries, 1024 threads with LDS flags, 1024 threads with flags in HBM), the
GPU-wide level launches of the split form, and the capacity
overflow that rolls a tick back, grows the buffers and replays it (with ball joints, whose warm
rows the failed tick has already rewritten), including rollback slots made before the growth.
"""
import numpy as np
import pytest

import oracle_py
import rphys
import scenes

pytestmark = pytest.mark.gpu


@pytest.fixture
def hooks():
    L = rphys.lib()
    yield L
    L.rphys_debug_set_sweep_tiers(0, 0, 0, 0)
    L.rphys_debug_set_lane_tier(0, 0, 0, 0)
    L.rphys_debug_set_initial_capacity(0, 0, 0, 0)


def run_pair(scene, ticks, check_every=1):
    o, g = oracle_py.OracleWorld(scene), rphys.World(scene)
    for t in range(ticks):
        o.step()
        g.step()
        if (t + 1) % check_every == 0 or t + 1 == ticks:
            assert o.hash() == g.hash(), "hash at tick %d" % (t + 1)
    return g


# (max tier 0, max tier 3, max tier 1, wide-level minimum; 0 = default): every island of two or
# more entries in tier 2 (flags in HBM); a spread over all four tiers; the split form with narrow
# "wide" levels in tier 1 (LDS flags) and in tier 2
TIERS = {"all_hbm_flags": (1, 1, 1, 0), "spread": (8, 40, 200, 0), "split_lds": (1, 1, 0, 8), "split_hbm": (1, 1, 1, 8)}


def dense_pile():
    """1,200 mixed bodies dropped into a 12 m square: piles of a few thousand contacts whose list
    orders open with wide levels (the 512-box lattice's chains leave its first level one rank wide)."""
    return scenes.mixed_shapes(n=1200, hf_n=48, extent=12.0, height=10.0, margin=4.0)


@pytest.mark.parametrize("tiers", ["all_hbm_flags", "spread"])
@pytest.mark.parametrize("scene", ["mixed", "lattice", "joints"])
def test_sweep_tiers_bit_exact(hooks, tiers, scene):
    hooks.rphys_debug_set_sweep_tiers(*TIERS[tiers])
    if scene == "mixed":
        s, ticks = scenes.small_mixed(n=96, seed=7), 40
    elif scene == "lattice":
        s, ticks = scenes.box_lattice(), 30
    else:
        s, ticks = scenes.ragdolls_small(), 40
    run_pair(s, ticks)


# (largest island solved one per lane, by a lane pair, by a group of four lanes; tier 0's larger
# ones a wave each; least number of the largest islands a wave each): every tier-0 island by
# lanes; none (a wave per island); every split, with lanes up to 4 entries, pairs up to 16, groups
# of four up to 64; lane pairs up to 128; groups of four for every island of two or more entries;
# the default bounds, with the 8 largest islands a wave each
N = rphys.LANE_NONE
LANES = {"all": (256, 0, 0, N), "off": (N, 0, 0, 0), "split": (4, 16, 64, N), "pairs": (1, 128, 128, N), "groups": (1, 1, 256, N),
         "wave8": (0, 0, 0, 8)}


@pytest.mark.parametrize("lanes", ["all", "off", "split", "pairs", "groups", "wave8"])
@pytest.mark.parametrize("scene", ["mixed", "joints"])
def test_lane_tier_bounds_bit_exact(hooks, lanes, scene):
    """Tier 0's velocity solve one island per lane, by lane pairs and groups of four, or a wave per
    island: the same bits as the oracle every tick."""
    hooks.rphys_debug_set_lane_tier(*LANES[lanes])
    s, ticks = (scenes.small_mixed(n=96, seed=7), 40) if scene == "mixed" else (scenes.ragdolls_small(), 40)
    run_pair(s, ticks)


@pytest.mark.parametrize("lanes", ["all", "split", "groups", "off"])
def test_residual_mode_in_the_lane_tier(hooks, lanes):
    """Residual mode by lanes, lane groups and a wave per island: each island stops at its own
    tolerance, with the oracle's sweep counts and hashes."""
    hooks.rphys_debug_set_lane_tier(*LANES[lanes])
    s = scenes.ragdolls_small()
    s["residual"] = (1.0e-4, 60)
    o, g = oracle_py.OracleWorld(s), rphys.World(s)
    for t in range(40):
        o.step()
        g.step()
        assert o.hash() == g.hash(), "hash at tick %d" % (t + 1)
        so, sg = o.stats(), g.stats()
        assert (so.iters_used, so.iters_sum) == (sg.pgs_iters, sg.pgs_iters_sum), (t + 1, so.iters_used, sg.pgs_iters)


@pytest.mark.parametrize("tiers", ["split_lds", "split_hbm"])
def test_split_form_bit_exact(hooks, tiers):
    """Big islands' leading wide levels as GPU-wide launches (fixed iterations), checked every
    tenth tick over the fall and the settling of the pile."""
    hooks.rphys_debug_set_sweep_tiers(*TIERS[tiers])
    g = run_pair(dense_pile(), 100, check_every=10)
    _, counters = g.step_profiled(1)
    assert counters["wide_levels"] > 0 and counters["wide_levels_shock"] > 0, "the split form did not run"


@pytest.mark.parametrize("tiers", ["all_hbm_flags", "split_lds"])
def test_residual_mode_in_the_big_tiers(hooks, tiers):
    """Residual mode in tiers 2 and 1 (one-kernel sweeps): the same per-island stop and sweep
    counts as the oracle.  Tier 2's only LDS word is the sweep residual, at LDS address 0."""
    hooks.rphys_debug_set_sweep_tiers(*TIERS[tiers])
    s = scenes.box_lattice()
    s["residual"] = (1.0e-4, 40)
    o, g = oracle_py.OracleWorld(s), rphys.World(s)
    for t in range(20):
        o.step()
        g.step()
        assert o.hash() == g.hash(), "hash at tick %d" % (t + 1)
        so, sg = o.stats(), g.stats()
        assert (so.iters_used, so.iters_sum) == (sg.pgs_iters, sg.pgs_iters_sum), (t + 1, so.iters_used, sg.pgs_iters)


def test_overflow_replay_with_joints_bit_exact(hooks):
    """Tiny initial capacities: the first ticks with contacts overflow the pair, contact and
    warm-cache buffers; the device rolls each failed tick back (joint warm rows included), grows,
    and replays, bit-exact with the oracle every tick."""
    hooks.rphys_debug_set_initial_capacity(64, 256, 64, 16)
    s = scenes.ragdolls_small(n_towers=6, tower_h=5, n_ragdolls=8)
    run_pair(s, 60)


def test_rollback_slot_made_before_growth(hooks):
    """A rollback slot taken while the warm-cache capacity is small, then re-taken after a rebuild
    grew it: the slot is reallocated, and the restored window replays its hashes."""
    hooks.rphys_debug_set_initial_capacity(0, 0, 0, 16)
    worlds = [rphys.World(scenes.box_lattice(seed=scenes.seed_for(1, k))) for k in range(2)]
    b = rphys.Batch(worlds)
    b.snapshot(0)  # empty warm cache, capacity ~1,024 entries
    b.step_n(40)   # the settled lattices hold several thousand warm entries: rebuilt larger
    b.snapshot(0)
    b.step_n(8)
    want = b.hashes()
    b.restore(0)
    b.step_n(8)
    assert b.hashes() == want
    o = oracle_py.OracleWorld(scenes.box_lattice(seed=scenes.seed_for(1, 1)))
    for _ in range(48):
        o.step()
    assert o.hash() == want[1]
