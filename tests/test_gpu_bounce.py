"""W6 restitution on the device (time-of-impact bounce rows, the interior-point LCP, the pseudo
pass, the restitution relax pass) against the reference's golden bits and the CPU oracle.

G2/G3 (restitution_toi_test.rs:596-619, docs/PARITY_SPEC.md:2120-2126): the first 8 bytes of the
step hash after the 20-box pile, the box on a two-box stack and the drone box, at 60 and 240 Hz,
must be the reference's.  The other cases compare device and oracle bit for bit every tick:
ball drops, a zero-g head-on collision, bounces on terrain with the default material (e 0.1), a
jointed chain that takes the relax pass, and a batch mixing bounce worlds with a world that never
bounces.
"""
import time

import numpy as np
import pytest

import oracle_py
import rphys
import scenes
from test_oracle_goldens import W6_GOLDEN, w6_golden_run

pytestmark = pytest.mark.gpu


def assert_same_state(o, g, what):
    for name, a, b in zip(("pos", "rot", "lin", "ang"), o.state(), g.state()):
        bad = np.nonzero(a.view(np.uint32) != b.view(np.uint32))[0]
        assert bad.size == 0, "%s: %s differs at bodies %s (oracle %s, device %s)" % (what, name, bad[:5], a[bad[:2]], b[bad[:2]])


@pytest.mark.parametrize("kind,hz,want", W6_GOLDEN)
def test_w6_scene_goldens_on_the_device(kind, hz, want):
    cap = 16 if kind == "drone" else 32
    t0 = time.time()
    g, _ = w6_golden_run(kind, hz, lambda s: rphys.World(s, max_bodies=cap))
    dt = time.time() - t0
    got = g.hash()[:8].hex()
    print("%s at %g Hz: %s (reference %s), %.2f s" % (kind, hz, got, want, dt))
    assert got == want, "%s at %g Hz: device %s, reference golden %s" % (kind, hz, got, want)


def lockstep(scene, ticks, cap=16):
    """Steps the oracle and the device together; returns them and the bounce and relax rows seen
    (each tick's counts must agree)."""
    o = oracle_py.OracleWorld(scene, capacity=cap)
    g = rphys.World(scene, max_bodies=cap)
    rows = relax = 0
    for t in range(ticks):
        o.step()
        g.step()
        so, sg = o.stats(), g.stats()
        assert (so.toi_rows, so.relax_rows) == (sg.bounce_rows, sg.relax_rows), "%s: tick %d" % (scene["name"], t + 1)
        rows += sg.bounce_rows
        relax += sg.relax_rows
        assert o.hash() == g.hash(), "%s: hash differs at tick %d" % (scene["name"], t + 1)
    assert_same_state(o, g, scene["name"])
    return o, g, rows, relax


def test_pile_bounce_every_tick_against_the_oracle():
    """The 240 Hz pile: its drop's support component (about 170 rows) runs the interior point."""
    s = scenes.bounce_pile(240.0)
    o = oracle_py.OracleWorld(s, capacity=32)
    g = rphys.World(s, max_bodies=32)
    for _ in range(240):
        o.step()
        g.step()
    assert o.hash() == g.hash(), "settled pile"
    drop = scenes.bounce_pile_drop()
    o.add_body_spec(drop)
    g.add_body_spec(drop)
    support = 0
    for t in range(240):
        o.step()
        g.step()
        st = g.stats()
        support = max(support, st.bounce_support_rows)
        assert st.bounce_fallbacks == 0
        assert o.hash() == g.hash(), "pile: hash differs at tick %d after the drop" % (t + 1)
    assert support > 100
    assert_same_state(o, g, "pile")


@pytest.mark.parametrize("mat", [scenes.MAT_RUBBER, scenes.MAT_ALUMINUM])
@pytest.mark.parametrize("hz", [60.0, 240.0])
def test_ball_drop_against_the_oracle(mat, hz):
    r = 0.1
    s = scenes.bounce_scene_of([scenes.bounce_floor(mat), scenes.bounce_body(0, (r,), (0.0, r + 1.0, 0.0), 0.1, mat)], hz=hz)
    _, _, rows, _ = lockstep(s, int(2.0 * hz))
    assert rows > 0


def test_head_on_in_zero_g_against_the_oracle():
    m = scenes.MAT_ALUMINUM
    s = scenes.bounce_scene_of([scenes.bounce_body(0, (0.1,), (0.0, 1.0, 0.0), 1.0, m, vel=(3, 0, 0, 0, 0, 0)),
                                scenes.bounce_body(0, (0.1,), (np.float32(0.2) + np.float32(0.83), 1.0, 0.0), 2.0, m,
                                                   vel=(-1, 0, 0, 0, 0, 0))], gravity=False, hz=60.0)
    _, _, rows, _ = lockstep(s, 30)
    assert rows > 0


def test_default_material_bounces_on_terrain_against_the_oracle():
    """Without the configs' restitution-0 overrides every pair bounces (Default e 0.1): spheres,
    boxes and capsules dropped on a heightfield hit it faster than 1 m/s."""
    s = scenes.small_mixed(n=64, seed=5)
    s["pairs"] = []
    s["name"] = "mixed64_default_materials"
    _, g, rows, _ = lockstep(s, 90, cap=64)
    assert rows > 0


def test_jointed_chain_takes_the_relax_pass_against_the_oracle():
    """Ball-jointed capsule chains bounce in the relax pass (solver.rs:1001-1008), free bodies
    alongside them by time of impact."""
    s = scenes.ragdolls_small(n_towers=1, tower_h=3, n_ragdolls=2, chain=5)
    s["pairs"] = []
    s["name"] = "chains_default_materials"
    _, _, rows, relax = lockstep(s, 90, cap=len(s["shape"]))
    assert relax > 0 and rows > 0


def test_batch_of_bounce_worlds_matches_single_worlds():
    resting = scenes.bounce_scene_of([scenes.bounce_floor(scenes.MAT_CONCRETE),
                                      scenes.bounce_body(1, (0.2, 0.1, 0.2), (0.0, 0.1, 0.0), 2.0, scenes.MAT_ALUMINUM)])
    ss = [scenes.bounce_stack(60.0), resting, scenes.bounce_drone(60.0), scenes.bounce_stack(60.0)]
    singles = [rphys.World(s, max_bodies=16) for s in ss]
    batch = rphys.Batch([rphys.World(s, max_bodies=16) for s in ss])
    for _ in range(90):
        for w in singles:
            w.step()
    batch.step_n(90)
    assert batch.hashes() == [w.hash() for w in singles]
    batch2 = rphys.Batch([rphys.World(s, max_bodies=16) for s in ss])
    batch2.step_n(90, graph=True)
    assert batch2.hashes() == [w.hash() for w in singles]
