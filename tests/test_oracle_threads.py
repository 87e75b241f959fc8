"""The oracle's parallel regions (OpenMP where the reference uses rayon: narrowphase over pairs,
terrain over colliders, PGS level batches of at least 64 rows; step.rs:14,113-141,
solver.rs:18-23,789-880) give the single-threaded results bit for bit, as the reference's
level schedule does (tests/parallel_replay_test.rs:2-7)."""
import pytest

import oracle_py as O
import scenes


def hashes(scene, ticks, threads):
    L = O.lib()
    L.or_set_threads(threads)
    try:
        o = O.OracleWorld(scene)
        out = []
        for _ in range(ticks):
            o.step()
            out.append(o.hash())
        return out
    finally:
        L.or_set_threads(1)


@pytest.mark.parametrize("kind", ["mixed", "mixed_residual", "joints"])
def test_threads_give_identical_hashes(kind):
    if kind.startswith("mixed"):
        s = scenes.mixed_shapes(n=1200, hf_n=64, extent=24.0, height=8.0, margin=4.0)
        if kind == "mixed_residual":
            s["residual"] = (1.0e-4, 40)
        ticks = 50
    else:
        s, ticks = scenes.ragdolls_small(n_towers=6, n_ragdolls=10), 50
    assert hashes(s, ticks, 1) == hashes(s, ticks, 4)
