"""Ball joints: the device step vs the oracle on scenes with capsule chains (bit for bit), residual
mode with joints, the batch with joint worlds (graph, rollback) and the at-scale workload's shape."""
import ctypes as C

import numpy as np
import pytest

import oracle_py
import rphys
import scenes
from test_gpu_parity import assert_state_close, contacts_equal

pytestmark = pytest.mark.gpu


def chains(seed=11, **kw):
    return scenes.ragdolls_small(seed=seed, **kw)


@pytest.mark.parametrize("seed", [11, 12])
def test_chain_scene_bit_exact(seed):
    s = chains(seed)
    o, g = oracle_py.OracleWorld(s), rphys.World(s)
    assert o.hash() == g.hash()
    for t in range(150):
        o.step()
        g.step()
        if t < 40 or t % 10 == 9:
            np.testing.assert_array_equal(o.pairs(), g.pairs(), err_msg="pairs at tick %d" % (t + 1))
            contacts_equal(o.contacts(), g.contacts())
            assert_state_close(o, g)
            assert o.hash() == g.hash(), "hash at tick %d" % (t + 1)


def test_chain_scene_iters_20_20_bit_exact():
    s = chains(13)
    s["iters"] = (20, 20)
    o, g = oracle_py.OracleWorld(s), rphys.World(s)
    for t in range(80):
        o.step()
        g.step()
    assert_state_close(o, g)
    assert o.hash() == g.hash()


def test_chain_scene_residual_mode_bit_exact():
    s = chains(14)
    s["residual"] = (1.0e-4, 60)
    o, g = oracle_py.OracleWorld(s), rphys.World(s)
    for t in range(60):
        o.step()
        g.step()
        assert o.hash() == g.hash(), "hash at tick %d" % (t + 1)
        so, sg = o.stats(), g.stats()
        assert (so.iters_used, so.iters_sum) == (sg.pgs_iters, sg.pgs_iters_sum), (t + 1, so.iters_used, sg.pgs_iters)


def test_towers_and_chains_medium_bit_exact():
    s = scenes.towers_ragdolls(n_towers=36, tower_h=12, n_loose=4, n_ragdolls=24, seed=21, hf_n=96, extent=30.0, margin=8.0, drop=4.0)
    o, g = oracle_py.OracleWorld(s), rphys.World(s)
    for t in range(90):
        o.step()
        g.step()
        if t % 15 == 14:
            assert_state_close(o, g)
            assert o.hash() == g.hash(), "hash at tick %d" % (t + 1)


def test_joint_batch_matches_single_worlds_graph_and_rollback():
    ss = [chains(30 + i, n_ragdolls=3 + i) for i in range(3)]
    singles = [rphys.World(s) for s in ss]
    b = rphys.Batch([rphys.World(s) for s in ss])
    for _ in range(20):
        for w in singles:
            w.step()
    b.step_n(20, graph=True)
    assert b.hashes() == [w.hash() for w in singles]
    b.snapshot(0)
    b.step_n(8, graph=True)
    h1 = b.hashes()
    b.restore(0)
    b.step_n(8, graph=True)
    assert b.hashes() == h1


def test_world_snapshot_restore_with_joints_replays():
    s = chains(15)
    g = rphys.World(s)
    g.step_n(25)
    snap = g.snapshot()
    runs = []
    for _ in range(2):
        g.restore(snap)
        hs = []
        for _ in range(6):
            g.step()
            hs.append(g.hash())
        runs.append(hs)
    g.free_snapshot(snap)
    assert runs[0] == runs[1]


def test_contact_log_skips_joint_entries():
    """The contact log (training data) holds one row per solved contact, none for joint entries of
    the island lists; in residual mode only the contacts of islands that reached the tolerance
    (an island stopped at the sweep limit holds no converged impulses)."""
    ss = [chains(40 + i, n_ragdolls=3 + i) for i in range(2)]
    nj = sum(len(s["joints"]) for s in ss)
    b = rphys.Batch([rphys.World(s) for s in ss])
    b.step_n(10)
    b.set_contact_log(400000)
    expect = 0
    for _ in range(5):
        _, c = b.step_profiled(1)
        expect += c["island_contacts"] - nj  # every joint here is active with its rows
    rows, dropped = b.contact_log_rows()
    assert dropped == 0 and rows == expect and rows > 0
    b.set_residual_mode(True, 1.0e-2, 60)
    conv = 0
    for _ in range(5):
        _, c = b.step_profiled(1)
        conv += c["island_contacts_converged"]
    rows2, dropped = b.contact_log_rows()
    added = rows2 - rows
    assert dropped == 0 and 0 < added <= conv and added >= conv - 5 * nj, (added, conv)
    feat, targ = b.read_contact_log(0, rows2)
    assert np.isfinite(feat).all() and np.isfinite(targ).all()


def test_full_towers_and_chains_world_bit_exact():
    """One whole at-scale world (16,384 bodies, 2,800 ball joints), device and oracle from spawn."""
    s = scenes.towers_ragdolls(seed=scenes.seed_for(5, 0))
    o, g = oracle_py.OracleWorld(s), rphys.World(s)
    for t in range(90):
        o.step()
        g.step()
        if t % 30 == 29:
            assert_state_close(o, g)
            assert o.hash() == g.hash(), "hash at tick %d" % (t + 1)


def _two_body(anchor_pose, body_pose, fa, fb):
    s = scenes._empty(2)
    s["shape"][:] = [1, 2]
    s["dims"][0] = (0.2, 0.2, 0.2)
    s["dims"][1] = (0.1, 0.15, 0.0)
    s["dynamic"][:] = [0, 1]
    s["pose"][0] = anchor_pose
    s["pose"][1] = body_pose
    s["joints"] = [(0, 1, fa, fb)]
    s["hf"] = None
    s["pairs"] = list(scenes.CONFIG_PAIRS)
    s["iters"] = (10, 20)
    s["dt"] = np.float32(1.0 / 60.0)
    return s


def test_joint_to_a_static_body_bit_exact():
    """A capsule pendulum hung from a static box: the joint writes one body only."""
    q = scenes.quat_yaw_pitch(0.3, 1.2)
    s = _two_body((0, 5, 0, 0, 0, 0, 1), (0.05, 5.1, 0.24, *q), (0, 0, 0, 0, 0, 0, 1), (0, -0.25, 0, 0, 0, 0, 1))
    o, g = oracle_py.OracleWorld(s), rphys.World(s)
    for t in range(120):
        o.step()
        g.step()
        assert o.hash() == g.hash(), "hash at tick %d" % (t + 1)
    assert_state_close(o, g)


def test_removed_jointed_body_bit_exact():
    """A chain link removed mid-run: its joints stay, now to a static body, and lose their warm rows."""
    s = chains(16)
    o, g = oracle_py.OracleWorld(s), rphys.World(s)
    for _ in range(20):
        o.step()
        g.step()
    victim = s["joints"][3][1]  # a capsule in the middle of the first chain (scene index)
    assert oracle_py.lib().or_remove_body(o.h, victim) == 1
    assert rphys.lib().rphys_body_remove(g.h, victim + 1) == rphys.OK
    for t in range(40):
        o.step()
        g.step()
        assert o.hash() == g.hash(), "hash at tick %d after the removal" % (t + 1)


def test_set_pose_and_restore_on_joint_worlds_match_the_oracle():
    """Teleporting a jointed body drops its joints' warm rows (world.rs:818-826), and a world restore
    renormalises poses and restores the warm joint rows (lib.rs:1778-1790): device and oracle agree."""
    s = chains(17)
    o, g = oracle_py.OracleWorld(s), rphys.World(s)
    for _ in range(15):
        o.step()
        g.step()
    so, sg = o.snapshot(), g.snapshot()
    b = s["joints"][5][0]
    pose = np.array(o.state()[0][b].tolist() + o.state()[1][b].tolist(), np.float32)
    pose[1] += 0.3
    oracle_py.lib().or_set_body_pose(o.h, b, oracle_py.fptr(pose))
    assert rphys.lib().rphys_body_set_pose(g.h, b + 1, C.byref(rphys.iso(pose[:3], pose[3:]))) == rphys.OK
    for t in range(10):
        o.step()
        g.step()
        assert o.hash() == g.hash(), "hash at tick %d after the teleport" % (t + 1)
    o.restore(so)
    g.restore(sg)
    assert o.hash() == g.hash(), "hash after the restore"
    for t in range(10):
        o.step()
        g.step()
        assert o.hash() == g.hash(), "hash at tick %d after the restore" % (t + 1)
    o.free_snapshot(so)
    g.free_snapshot(sg)


def test_batch_mixing_worlds_with_and_without_joints():
    # a batch takes its solver settings from world 0: every world here runs (10, 20) iterations
    ss = [scenes.small_mixed(n=40, seed=50), chains(51), scenes.small_mixed(n=70, seed=52), chains(53, n_ragdolls=2)]
    for s in ss:
        s["iters"] = (10, 20)
    singles = [rphys.World(s) for s in ss]
    b = rphys.Batch([rphys.World(s) for s in ss])
    for _ in range(15):
        for w in singles:
            w.step()
    b.step_n(15, graph=True)
    assert b.hashes() == [w.hash() for w in singles]
    o = oracle_py.OracleWorld(ss[1])
    for _ in range(15):
        o.step()
    assert o.hash() == singles[1].hash()
