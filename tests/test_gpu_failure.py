"""Device failures forced from the host (rphys_debug_fail_device_allocs): the reference turns
any failure inside a call into RPHYS_ERR_INTERNAL with a thread-local message (lib.rs:96-106);
here a failed device allocation does the same, and the world or batch keeps its last good state.

A failed scene upload (first step, and the rebuild after a body is added), a failed rebuild
while growing the buffers after a capacity overflow, and a failed rollback-slot allocation each
return ERR_INTERNAL with the allocation message; the next call succeeds and the world goes on
bit-exact with the oracle, as if the failed call had not been made.  No kernel faults: the
failure is an allocation the host refuses.
"""
import ctypes as C

import pytest

import oracle_py
import rphys
import scenes

pytestmark = pytest.mark.gpu


@pytest.fixture
def L():
    L = rphys.lib()
    L.rphys_debug_fail_device_allocs(0, None)
    yield L
    L.rphys_debug_fail_device_allocs(0, None)
    L.rphys_debug_set_initial_capacity(0, 0, 0, 0)


def pending(L):
    p = C.c_uint32()
    L.rphys_debug_fail_device_allocs(0, C.byref(p))
    return p.value


def expect_internal(fn):
    with pytest.raises(rphys.RPhysError) as e:
        fn()
    assert e.value.code == rphys.ERR_INTERNAL
    assert "device allocation failed" in str(e.value), str(e.value)


def test_failed_scene_upload_keeps_the_world(L):
    s = scenes.small_mixed(n=48, seed=21)
    o, g = oracle_py.OracleWorld(s, capacity=64), rphys.World(s, max_bodies=64)
    L.rphys_debug_fail_device_allocs(1, None)
    expect_internal(g.step)
    assert pending(L) == 0 and g.tick() == 0
    for t in range(10):
        o.step()
        g.step()
        assert o.hash() == g.hash(), "tick %d" % (t + 1)
    extra = scenes.bounce_body(1, (0.3, 0.2, 0.25), tuple(float(x) for x in o.state()[0][3] + [0.0, 3.0, 0.0]), 1.0, 0)
    o.add_body_spec(extra)
    g.add_body_spec(extra)
    L.rphys_debug_fail_device_allocs(7, None)  # a few allocations in, part-way through the build
    expect_internal(g.step)
    L.rphys_debug_fail_device_allocs(0, None)
    assert g.tick() == 10
    for t in range(30):
        o.step()
        g.step()
        assert o.hash() == g.hash(), "tick %d" % (t + 11)


def test_failed_growth_after_an_overflow_keeps_the_world(L):
    """Tiny initial capacities: ticks overflow and rebuild larger.  Each rebuild is made to fail
    once: the failed tick is rolled back and not counted, and the retried tick grows and matches."""
    L.rphys_debug_set_initial_capacity(64, 256, 64, 16)
    s = scenes.small_mixed(n=64, seed=5)
    o, g = oracle_py.OracleWorld(s), rphys.World(s)
    g.step()  # builds the first (small) scene
    o.step()
    failures = 0
    for t in range(1, 60):
        if failures < 3:
            L.rphys_debug_fail_device_allocs(1, None)
            try:
                g.step()
            except rphys.RPhysError as e:
                assert e.code == rphys.ERR_INTERNAL and "device allocation failed" in str(e)
                assert g.tick() == t
                failures += 1
                L.rphys_debug_fail_device_allocs(0, None)
                g.step()
            else:
                assert pending(L) == 1  # no rebuild this tick: nothing allocated
        else:
            g.step()
        o.step()
        assert g.tick() == t + 1
        assert o.hash() == g.hash(), "tick %d" % (t + 1)
    assert failures >= 1


def test_failed_rollback_slot_keeps_the_batch(L):
    worlds = [rphys.World(scenes.small_mixed(n=40, seed=30 + k)) for k in range(3)]
    b = rphys.Batch(worlds)
    b.step_n(10)
    L.rphys_debug_fail_device_allocs(1, None)
    expect_internal(lambda: b.snapshot(2))
    assert pending(L) == 0
    b.snapshot(2)
    b.step_n(12)
    want = b.hashes()
    b.restore(2)
    b.step_n(12)
    assert b.hashes() == want
    o = oracle_py.OracleWorld(scenes.small_mixed(n=40, seed=31))
    for _ in range(22):
        o.step()
    assert o.hash() == want[1]
