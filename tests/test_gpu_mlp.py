"""Device MLP kernels of the learned warm start vs the CPU reference (tests/mlp_ref.py, which
emulates the same bfloat16 rounding points and accumulates in float64), and the learned warm start
in the solver.  The model predicts a correction of the previous-frame warm start, so a model that
outputs zeros, or one whose every prediction fails the gate, must reproduce the previous-frame
warm start bit for bit; a trained model must need no more sweeps to the residual than it."""
import numpy as np
import pytest

import mlp_ref
import rphys
import scenes

pytestmark = pytest.mark.gpu


def random_case(seed, n):
    rng = np.random.default_rng(seed)
    xm = rng.normal(size=12).astype(np.float32)
    xs = rng.uniform(0.5, 2.0, 12).astype(np.float32)
    p = mlp_ref.init(seed, xm=xm, xs=xs, ys=np.array([3.0, 0.5, 0.7], np.float32))
    feat = (xm + xs * rng.normal(size=(n, 12))).astype(np.float32)
    return p, feat, rng


def test_forward_matches_reference():
    p, feat, _ = random_case(11, 1000)  # not a multiple of the 64-contact block: tails masked
    got = rphys.mlp_forward(p, feat)
    _, want, _ = mlp_ref.forward(p, feat)
    scale = np.abs(want).max()
    err = np.abs(got - want)
    # accumulation order only; a bfloat16 activation may land one step apart (2^-8 relative)
    assert np.median(err) <= 1e-5 * scale and err.max() <= 2e-2 * scale, (np.median(err), err.max(), scale)


def test_gradient_matches_reference():
    p, feat, rng = random_case(12, rphys.lib().rphys_mlp_gradient_rows())
    targ = rng.normal(size=(len(feat), 3)).astype(np.float32)
    loss, g = rphys.mlp_gradient(p, feat, targ)
    lref, gref = mlp_ref.loss_grad(p, feat, targ)
    assert abs(loss - lref) <= 1e-4 * abs(lref)
    cos = float(np.dot(g, gref) / (np.linalg.norm(g) * np.linalg.norm(gref)))
    assert cos > 0.9999, cos
    np.testing.assert_allclose(g, gref, rtol=5e-2, atol=2e-3 * np.abs(gref).max())


def test_learned_warm_start_trains_and_cuts_sweeps():
    worlds = [scenes.box_lattice(seed=0x5EED0000 + 1000 + w) for w in range(4)]
    b = rphys.Batch([rphys.World(s) for s in worlds])
    b.step_n(60)  # the lattice collapses onto the floor
    b.set_contact_log(1000000)
    b.step_n(40)
    rows, dropped = b.contact_log_rows()
    assert rows > 50000 and dropped == 0, (rows, dropped)
    f, t = b.read_contact_log(0, 1000)
    # targets are the converged impulses less the previous-frame warm start, whose normal part is
    # feature 10 (the matched jn): their sum is the converged jn, never negative
    assert np.all(np.isfinite(f)) and np.all(np.isfinite(t))
    assert np.all(t[:, 0] + f[:, 10] >= -1e-6 * np.maximum(1.0, f[:, 10])), (t[:, 0] + f[:, 10]).min()
    assert np.any(f[:, 10] > 0.0)
    params, loss = b.train_warm_start(epochs=4, lr=1e-3, seed=7)
    assert np.all(np.isfinite(loss)) and loss[-1] < loss[0], loss
    # same start state, residual mode: learned vs the previous-frame warm start
    res = {}
    for mode in (0, 2):
        fresh = [scenes.box_lattice(seed=0x5EED0000 + 1100 + w) for w in range(2)]
        for s in fresh:
            s["residual"] = (1e-4, 60)
        bb = rphys.Batch([rphys.World(s) for s in fresh])
        bb.step_n(60)
        bb.set_warm_start(mode, params if mode == 2 else None)
        st, c = bb.step_profiled(20)
        res[mode] = c["pgs_iters_sum"] / c["island_contacts"]
    assert res[2] <= res[0] * 1.02, res


def constant_model(out_bias):
    """Zero weights, raw inputs (mean 0, std 1), unit output scale: every prediction is out_bias."""
    p = np.zeros(mlp_ref.NPARAM, np.float32)
    o = 64 * 12 + 64 + 64 * 64 + 64 + 3 * 64
    p[o:o + 3] = out_bias
    p[o + 3 + 12:o + 3 + 24] = 1.0
    p[o + 3 + 24:o + 3 + 27] = 1.0
    return p


@pytest.mark.parametrize("bias, gated", [(0.0, False), (100.0, True)])
def test_learned_fallback_is_previous_frame_bit_exact(bias, gated):
    """A zero correction, or predictions that all fail the gate (|y| > 8 normalised), leave the
    reference's warm start unchanged: same hashes as mode 0, tick by tick."""
    p = constant_model(bias)
    got = rphys.mlp_forward(p, np.zeros((3, 12), np.float32))
    assert np.all(got == bias)
    out = {}
    for mode in (0, 2):
        scs = [scenes.small_mixed(n=96, seed=5), scenes.box_lattice(seed=0x5EED0000 + 1200)]
        bb = rphys.Batch([rphys.World(s) for s in scs])
        bb.step_n(40)
        bb.set_residual_mode(True, 1e-4, 30)
        bb.set_warm_start(mode, p if mode == 2 else None)
        hs, learned = [], 0.0
        for _ in range(15):
            st, c = bb.step_profiled(1)
            hs.append(bb.hashes())
            learned += c["warm_start_learned"]
        out[mode] = (hs, learned, c["contacts"])
    assert out[2][0] == out[0][0]
    if gated:
        assert out[2][1] == 0.0
    else:
        assert out[2][1] > 0.0
