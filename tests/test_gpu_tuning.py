"""Launch-shape autotuner on the device: it leaves the batch at its starting state, installs one of
its candidates, and stepping with the tuned grids gives the same hashes as the default grids (the
grids only change how islands are spread over workgroups)."""
import pytest

import rphys
import scenes

pytestmark = pytest.mark.gpu

G0, G1, G2 = {1024, 2048, 4096, 8192, 16384, 32768}, {128, 256, 512, 1024, 2048}, {64, 128, 256}


def make_batch():
    return rphys.Batch([rphys.World(scenes.batch_world(w)) for w in range(6)])


def test_autotune_keeps_state_and_results(tmp_path):
    cache = str(tmp_path / "tuning.json")
    b = make_batch()
    b.step_n(40, graph=True)
    h0 = b.hashes()
    e, cached = b.autotune("test6", ticks=1, cache=cache)
    assert not cached
    assert e["grid_small"] in G0 and e["grid_large"] in G1 and e["grid_huge"] in G2 and e["ms_per_tick"] > 0.0
    assert b.get_tuning() == {k: e[k] for k in ("grid_small", "grid_large", "grid_huge")}
    assert b.hashes() == h0
    b.step_n(10, graph=True)
    tuned = b.hashes()

    ref = make_batch()
    assert ref.get_tuning() == {"grid_small": 8192, "grid_large": 256, "grid_huge": 64}
    ref.step_n(50, graph=True)
    assert ref.hashes() == tuned

    # the second call reads the cache; odd grids still give the same results
    e2, cached2 = ref.autotune("test6", ticks=1, cache=cache)
    assert cached2 and e2 == e
    ref.set_tuning(3, 5, 7)
    ref.step_n(5)
    b.step_n(5)
    assert ref.hashes() == b.hashes()


def test_set_tuning_rejects_zero_grids():
    b = make_batch()
    with pytest.raises(rphys.RPhysError):
        b.set_tuning(0, 256, 64)
