"""Pins the CPU reference of the warm-start MLP (tests/mlp_ref.py) before it judges the device:
bfloat16 rounding on known values, and the analytic float64 gradient against central finite
differences of the float64 loss (a dropped term, wrong sign, index or transposed operand fails)."""
import numpy as np

import mlp_ref


def test_bf16_rounding_known_values():
    x = np.array([1.0, 1.0 + 2.0 ** -8, 1.0 + 3 * 2.0 ** -9, -2.5, 3.0e38, 1.0 + 2.0 ** -7 + 2.0 ** -8], np.float32)
    want = np.array([1.0, 1.0, 1.0 + 2.0 ** -7, -2.5, np.uint32(0x7F620000).view(np.float32), 1.0 + 2.0 ** -6], np.float32)
    got = mlp_ref.to_bf16(x)
    assert np.array_equal(got.view(np.uint32) & 0xFFFF, np.zeros(len(x), np.uint32))
    np.testing.assert_allclose(got, want, rtol=1e-7)


def test_float64_gradient_matches_finite_differences():
    rng = np.random.default_rng(5)
    feat = rng.normal(size=(64, 12)).astype(np.float32)
    targ = rng.normal(size=(64, 3)).astype(np.float32)
    p = mlp_ref.init(3, xm=np.zeros(12), xs=np.ones(12), ys=np.array([1.0, 2.0, 0.5]))
    p64 = p.astype(np.float64)
    _, g = mlp_ref.loss_grad(p64, feat, targ, bf16=False)
    idx = list(rng.choice(mlp_ref.NW, 60, replace=False)) + [mlp_ref.NW - 1, mlp_ref.NW - 2, mlp_ref.NW - 3, 0, 768, 768 + 64]
    for i in idx:
        h = 1e-6
        pp, pm = p64.copy(), p64.copy()
        pp[i] += h
        pm[i] -= h
        fd = (mlp_ref.loss_grad(pp, feat, targ, bf16=False)[0] - mlp_ref.loss_grad(pm, feat, targ, bf16=False)[0]) / (2 * h)
        assert abs(fd - g[i]) <= 1e-6 + 1e-4 * abs(fd), (i, fd, g[i])


def test_forward_shapes_and_output_scale():
    p = mlp_ref.init(1, ys=np.array([10.0, 1.0, 1.0]))
    feat = np.random.default_rng(2).normal(size=(5, 12)).astype(np.float32)
    y, out, _ = mlp_ref.forward(p, feat)
    assert y.shape == (5, 3) and np.allclose(out[:, 0], 10.0 * y[:, 0])
