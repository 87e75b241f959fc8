"""rphys_mlp_forward_noalloc (the entry point that allocates nothing: its bf16 weight tiles live
in a file-scope device array) against rphys_mlp_forward on the same model and features, bit for
bit, on two calls in a row, the second with a smaller input; and its argument checks."""
import ctypes as C

import numpy as np
import pytest

import mlp_ref
import rphys

pytestmark = pytest.mark.gpu


class DevBuf:
    def __init__(self, arr=None, nbytes=None):
        self.L = rphys.lib()
        self.nbytes = arr.nbytes if arr is not None else nbytes
        self.ptr = C.c_void_p()
        rphys.check(self.L.rphys_debug_device_alloc(self.nbytes, C.byref(self.ptr)))
        if arr is not None:
            rphys.check(self.L.rphys_debug_device_copy(self.ptr, arr.ctypes.data, self.nbytes, 0))

    def get(self, shape):
        out = np.empty(shape, np.float32)
        rphys.check(self.L.rphys_debug_device_copy(out.ctypes.data, self.ptr, out.nbytes, 1))
        return out

    def free(self):
        rphys.check(self.L.rphys_debug_device_free(self.ptr))


def case(seed, n):
    rng = np.random.default_rng(seed)
    xm = rng.normal(size=12).astype(np.float32)
    xs = rng.uniform(0.5, 2.0, 12).astype(np.float32)
    p = mlp_ref.init(seed, xm=xm, xs=xs, ys=np.array([3.0, 0.5, 0.7], np.float32))
    feat = (xm + xs * rng.normal(size=(n, 12))).astype(np.float32)
    feat[::17] += 40.0 * xs  # some rows out of the gate
    return np.ascontiguousarray(p, np.float32), np.ascontiguousarray(feat, np.float32)


def run_noalloc(p, feat):
    L = rphys.lib()
    dp, df = DevBuf(p), DevBuf(feat)
    do = DevBuf(nbytes=16 * len(feat))
    try:
        rphys.check(L.rphys_mlp_forward_noalloc(dp.ptr, df.ptr, len(feat), do.ptr, None))
        return do.get((len(feat), 4))  # the blocking copy follows the null stream's kernels
    finally:
        for b in (dp, df, do):
            b.free()


def test_noalloc_forward_equals_forward_twice_second_smaller():
    for seed, n in ((21, 5000), (22, 777)):
        p, feat = case(seed, n)
        got = run_noalloc(p, feat)
        want = rphys.mlp_forward(p, feat)
        np.testing.assert_array_equal(got[:, :3].view(np.uint32), want.view(np.uint32))
        assert set(np.unique(got[:, 3])) <= {0.0, 1.0}
        assert (got[::17, 3] == 0.0).all() and got[:, 3].sum() > 0.5 * n


def test_noalloc_argument_checks():
    L = rphys.lib()
    p, feat = case(23, 8)
    dp = DevBuf(p)
    try:
        assert L.rphys_mlp_forward_noalloc(None, None, 8, None, None) == rphys.ERR_NULL
        assert L.rphys_mlp_forward_noalloc(dp.ptr, None, 8, None, None) == rphys.ERR_NULL
        assert L.rphys_mlp_forward_noalloc(dp.ptr, dp.ptr, (1 << 30) + 1, dp.ptr, None) == rphys.ERR_BAD_ARG
        assert b"2^30" in L.rphys_last_error_message()
        assert L.rphys_mlp_forward_noalloc(dp.ptr, None, 0, None, None) == rphys.OK
    finally:
        dp.free()
