"""The configurations at their full sizes on the device.

c3: the batched workload, 1,024 worlds of log-uniform 32..2,048 bodies (605,152 bodies), graph-
captured ticks with an autotuned launch shape and 8-tick rollback windows whose replayed hashes
must match, and 9 sampled worlds (including the largest) bit-exact against single oracle worlds.

c1: the 512-box lattice for all of its 600 ticks, the step hash equal to the oracle's every tick,
and the gates on penetration and energy drift over the 600 ticks (within 1% of the oracle's;
with equal bits they agree exactly, and the gate quantities are reported).
"""
import numpy as np
import pytest

import oracle_py
import rphys
import scenes

pytestmark = pytest.mark.gpu


def test_c3_full_batch_graphs_autotune_rollback_and_oracle(tmp_path):
    worlds = [scenes.batch_world(w) for w in range(1024)]
    sizes = [len(s["shape"]) for s in worlds]
    assert sum(sizes) == 605152
    b = rphys.Batch([rphys.World(s) for s in worlds])
    b.step_n(4, graph=True)
    h4 = b.hashes()
    e, _ = b.autotune("test_c3_1024", ticks=1, cache=str(tmp_path / "tuning.json"))
    assert b.hashes() == h4  # the search leaves the batch where it was
    b.step_n(6, graph=True)
    b.snapshot(0)
    b.step_n(8, graph=True)
    first = b.hashes()
    b.restore(0)
    b.step_n(8, graph=True)
    replay = b.hashes()
    assert first == replay
    sample = sorted(set(list(range(0, 1024, 128)) + [int(np.argmax(sizes))]))
    for w in sample:
        o = oracle_py.OracleWorld(worlds[w])
        for _ in range(18):
            o.step()
        assert o.hash() == replay[w], "world %d (%d bodies) differs from the oracle after 18 ticks" % (w, sizes[w])


def _gates(pos, rot, lin, ang, masses, inertia_diag, heights_y0):
    """Penetration below the flat floor (largest, m) and total energy (J) of the box lattice."""
    lowest = []
    for p, q in zip(pos, rot):
        x, y, z, w = [float(t) for t in q]
        R = np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - w * z), 2 * (x * z + w * y)],
                      [2 * (x * y + w * z), 1 - 2 * (x * x + z * z), 2 * (y * z - w * x)],
                      [2 * (x * z - w * y), 2 * (y * z + w * x), 1 - 2 * (x * x + y * y)]])
        lowest.append(float(p[1]) - 0.5 * np.abs(R[1]).sum())
    pen = max(0.0, heights_y0 - min(lowest))
    ke = 0.5 * masses * (lin.astype(np.float64) ** 2).sum(axis=1)
    ke += 0.5 * inertia_diag * (ang.astype(np.float64) ** 2).sum(axis=1)  # cubes: isotropic inertia
    pe = masses * 9.81 * pos[:, 1].astype(np.float64)
    return pen, float((ke + pe).sum())


def test_c1_600_ticks_hash_every_tick_and_drift_gates():
    s = scenes.box_lattice()
    o, g = oracle_py.OracleWorld(s), rphys.World(s)
    m = 1000.0  # 1 m cubes at 1000 kg/m^3
    idiag = m / 6.0
    e0 = None
    worst_pen = {"oracle": 0.0, "device": 0.0}
    for t in range(600):
        o.step()
        g.step()
        assert o.hash() == g.hash(), "hash at tick %d" % (t + 1)
        if (t + 1) % 50 == 0:
            so, sg = o.state(), g.state()
            po, eo = _gates(*so, m, idiag, 0.0)
            pd, ed = _gates(*sg, m, idiag, 0.0)
            if e0 is None:
                e0 = eo
            worst_pen["oracle"] = max(worst_pen["oracle"], po)
            worst_pen["device"] = max(worst_pen["device"], pd)
            # the north star's gates: penetration and energy drift within 1% of the oracle's
            assert abs(pd - po) <= 0.01 * max(po, 1e-3)
            assert abs((ed - e0) - (eo - e0)) <= 0.01 * max(abs(eo - e0), 1.0)
    assert worst_pen["device"] < 0.1  # the settled pile rests on the floor
