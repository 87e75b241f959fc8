"""The reference's wire formats on the device path: rphys_world_get_contacts (lib.rs:1802-1878)
and the i16 heightfield setter (lib.rs:1440-1490), against the debug contact view and the oracle."""
import ctypes as C

import numpy as np
import pytest

import oracle_py
import rphys
import scenes

pytestmark = pytest.mark.gpu


def get_contacts(g, cap=None):
    L = rphys.lib()
    n = C.c_uint32(123)
    assert L.rphys_world_get_contacts(g.h, None, 0, C.byref(n)) == rphys.OK
    total = n.value
    cap = total if cap is None else cap
    arr = (rphys.ContactEvent * max(cap, 1))()
    for e in arr:
        e.body_id_a = 0xDEAD
    assert L.rphys_world_get_contacts(g.h, arr, cap, C.byref(n)) == rphys.OK
    assert n.value == total
    return total, arr


def test_get_contacts_matches_the_contact_view():
    s = scenes.small_mixed(n=96, seed=7)
    g = rphys.World(s)
    total, _ = get_contacts(g)
    assert total == 0  # no tick yet: no contacts
    g.step_n(90)  # settled: most contacts carry load
    ex = g.contacts()
    total, arr = get_contacts(g)
    assert total == len(ex) > 0
    for e, x in zip(arr, ex):
        # one collider per body, in body order: collider i is body id i + 1
        assert e.body_id_a == x["a_collider"] + 1
        assert e.body_id_b == (rphys.TERRAIN_BODY_ID if x["b_collider"] == 0xFFFFFFFF else x["b_collider"] + 1)
        assert np.float32(e.position.x).view(np.uint32) == np.float32(x["point"][0]).view(np.uint32)
        assert (e.position.y, e.position.z) == (x["point"][1], x["point"][2])
        assert (e.normal.x, e.normal.y, e.normal.z) == tuple(x["normal"])
        assert np.float32(e.impulse).view(np.uint32) == np.float32(x["impulse"]).view(np.uint32)
    assert any(x["impulse"] > 0 for x in ex)
    # a short buffer takes the first events; the count stays the tick's total
    t2, part = get_contacts(g, cap=3)
    assert t2 == total and [e.body_id_a for e in part] == [e.body_id_a for e in arr[:3]]


def test_get_contacts_after_a_removal_skips_the_removed_body():
    s = scenes.small_mixed(n=96, seed=7)
    g = rphys.World(s)
    g.step_n(90)
    total, arr = get_contacts(g)
    body_contacts = [e for e in arr if e.body_id_b != rphys.TERRAIN_BODY_ID]
    assert body_contacts
    victim = body_contacts[0].body_id_a
    keep = [e for e in arr if victim not in (e.body_id_a, e.body_id_b)]
    assert rphys.lib().rphys_body_remove(g.h, victim) == rphys.OK
    t2, arr2 = get_contacts(g)
    assert t2 == total  # the last tick's contacts, as the reference counts them
    got = [(e.body_id_a, e.body_id_b, e.impulse) for e in arr2[:len(keep)]]
    assert got == [(e.body_id_a, e.body_id_b, e.impulse) for e in keep]
    assert all(e.body_id_a == 0xDEAD for e in arr2[len(keep):])  # nothing written past the packed events
    g.step()
    _, arr3 = get_contacts(g)
    assert all(victim not in (e.body_id_a, e.body_id_b) for e in arr3)


def test_i16_heightfield_matches_the_oracle():
    """rphys_world_set_heightfield_i16 with a row stride wider than the grid: heights are
    h * scale + offset + y_offset in f32 (lib.rs:1476-1483), then q6 (terrain.rs:31-40)."""
    s = scenes.small_mixed(n=96, seed=11)
    hf = s.pop("hf")
    nx = nz = int(hf["nx"])
    stride = nx + 5
    rng = np.random.default_rng(5)
    raw = rng.integers(-300, 300, size=(nz, stride)).astype(np.int16)
    scale, offset, y_off = np.float32(0.01), np.float32(0.5), np.float32(0.25)
    cell = np.float32(hf["sx"])
    g = rphys.World(s)
    d = rphys.HeightfieldDesc()
    d.size_bytes = C.sizeof(rphys.HeightfieldDesc)
    d.width, d.height, d.row_stride = nx, nz, stride
    d.cell_size_x = d.cell_size_z = float(cell)
    d.height_scale, d.height_offset = float(scale), float(offset)
    d.heights = raw.ctypes.data_as(C.POINTER(C.c_int16))
    assert rphys.lib().rphys_world_set_heightfield_i16(g.h, C.byref(d), 0.0, 0.0, float(y_off)) == rphys.OK
    hts = ((raw[:, :nx].astype(np.float32) * scale + offset) + y_off).astype(np.float32).ravel()
    o = oracle_py.OracleWorld(s)
    oracle_py.lib().or_set_heightfield(o.h, nx, nz, float(cell), float(cell), 0.0, 0.0, oracle_py.fptr(np.ascontiguousarray(hts)), 4)
    assert o.hash() == g.hash()
    for t in range(40):
        o.step()
        g.step()
        assert o.hash() == g.hash(), "hash at tick %d" % (t + 1)
    assert len(g.contacts()) > 0 and any(c["b_collider"] == 0xFFFFFFFF for c in g.contacts())
