"""Device MLP kernels of the learned warm start vs the CPU reference (tests/mlp_ref.py, which
emulates the same bfloat16 rounding points and accumulates in float64), and the learned warm start
in the solver.  The model predicts a correction of the previous-frame warm start, so a model that
outputs zeros, or one whose every prediction fails the gate, must reproduce the previous-frame
warm start bit for bit; a trained model must start nearer the converged impulses than it on
held-out worlds, and need no more sweeps to the residual."""
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


def logged(make, seeds, warm=60, ticks=40, rows=1000000):
    """(features, targets) the solver logs over `ticks` ticks of fresh worlds after `warm` ticks."""
    b = rphys.Batch([rphys.World(make(s)) for s in seeds])
    b.step_n(warm)
    b.set_contact_log(rows)
    b.step_n(ticks)
    n, dropped = b.contact_log_rows()
    assert n > 0 and dropped == 0, (n, dropped)
    return b, n


# (scene generator, training seeds, held-out seeds, largest ratio of the learned start's squared
# error to the previous frame's on the held-out contacts; measured 0.72 and 0.49)
GENERATORS = {"lattice": (lambda s: scenes.box_lattice(seed=s), range(1000, 1004), range(1300, 1302), 0.85),
              "towers_chains": (lambda s: scenes.ragdolls_small(seed=s), range(1000, 1008), range(1300, 1304), 0.7)}


@pytest.mark.parametrize("gen", list(GENERATORS))
def test_learned_warm_start_is_nearer_the_converged_impulses_on_held_out_worlds(gen):
    """Trained on the converged impulses the solver logs in some worlds, the model's corrected
    warm start starts nearer the converged impulses than the previous-frame warm start on worlds
    it never saw: the solver's starting error, which it reaches even where a residual floor keeps
    islands at the sweep limit.  And in residual mode it needs no more sweeps than that start."""
    make, train_seeds, held_seeds, bound = GENERATORS[gen]
    b, rows = logged(make, [0x5EED0000 + k for k in train_seeds])
    f, t = b.read_contact_log(0, min(rows, 1000))
    # targets are the converged impulses less the previous-frame warm start, whose normal part is
    # feature 10 (the matched jn): their sum is the converged jn, never negative
    assert np.all(np.isfinite(f)) and np.all(np.isfinite(t))
    assert np.all(t[:, 0] + f[:, 10] >= -1e-6 * np.maximum(1.0, f[:, 10])), (t[:, 0] + f[:, 10]).min()
    assert np.any(f[:, 10] > 0.0)
    params, loss = b.train_warm_start(epochs=4, lr=1e-3, seed=7)
    assert np.all(np.isfinite(loss)) and loss[-1] < loss[0], loss
    hb, n = logged(make, [0x5EED0000 + k for k in held_seeds], ticks=20, rows=400000)
    hf, ht = hb.read_contact_log(0, n)
    pred = rphys.mlp_forward(params, hf)
    err_prev = float(np.mean(np.sum(ht * ht, 1)))  # the previous frame's start misses by the target itself
    err_learned = float(np.mean(np.sum((ht - pred) ** 2, 1)))
    assert err_learned < bound * err_prev, (err_learned, err_prev)
    # same start state, residual mode: learned vs the previous-frame warm start
    res = {}
    for mode in (0, 2):
        fresh = [make(0x5EED0000 + k) for k in held_seeds]
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
