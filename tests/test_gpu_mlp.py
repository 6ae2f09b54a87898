"""Learned-prefetch MLP kernels (MFMA GEMM, epilogues, BCE, Adam) against the
float64 oracle, through the C ABI. Tolerances are FP32 against float64:
relative 2e-3 on gradients, scaled by each gradient block's magnitude."""
import numpy as np
import pytest

import fcs
from oracle import mlp_oracle as M

pytestmark = pytest.mark.gpu


def _rand(dims, batch, seed):
    rng = np.random.default_rng(seed)
    n_in, h, n_out = dims
    x = rng.normal(0, 1, (batch, n_in)).astype(np.float32)
    y = (rng.random((batch, n_out)) < 0.3).astype(np.float32)
    return x, y


def _close(a, b, rel=2e-3):
    scale = max(float(np.max(np.abs(b))), 1e-30)
    return np.max(np.abs(a - b)) <= rel * scale


@pytest.mark.parametrize("dims,batch", [((37, 24, 19), 5), ((600, 70, 130), 7), ((64, 64, 64), 64)])
def test_forward_backward_match_oracle(gpu_ctx, dims, batch):
    print(f"[gpu] mlp fwd/bwd {dims} batch {batch}", flush=True)
    m = fcs.MLP(gpu_ctx, *dims, max_batch=64, seed=11)
    p = m.params().astype(np.float64)
    x, y = _rand(dims, batch, 1)
    z_ref, _ = M.forward(p, x, *dims)
    z = m.forward(x)
    assert _close(z, z_ref, 1e-4)
    loss = m.backward(x, y)
    L_ref, g_ref = M.loss_grad(p, x, y, *dims)
    assert abs(loss - L_ref) <= 1e-5 * max(1.0, abs(L_ref))
    g = m.grads().astype(np.float64)
    o, _ = M.layout(*dims)
    for name, (off, size) in o.items():
        assert _close(g[off:off + size], g_ref[off:off + size]), name
    m.close()


def test_adam_step_matches_oracle(gpu_ctx):
    dims = (37, 24, 19)
    m = fcs.MLP(gpu_ctx, *dims, max_batch=8, seed=3)
    p0 = m.params().astype(np.float64)
    x, y = _rand(dims, 6, 2)
    m.backward(x, y)
    g = m.grads().astype(np.float64)
    m.adam(1e-2)
    p_ref, mm, vv = M.adam(p0, g, np.zeros_like(p0), np.zeros_like(p0), 1, 1e-2)
    assert np.allclose(m.params(), p_ref, rtol=1e-5, atol=1e-6)
    # second step uses the carried moments and the step count
    m.backward(x, y)
    g2 = m.grads().astype(np.float64)
    m.adam(1e-2)
    p_ref2, _, _ = M.adam(p_ref, g2, mm, vv, 2, 1e-2)
    assert np.allclose(m.params(), p_ref2, rtol=1e-5, atol=1e-6)
    m.close()


def test_train_masks_learns_a_shift(gpu_ctx):
    """Targets = the input mask shifted by one position: the loss must drop."""
    rng = np.random.default_rng(4)
    n, n_pose, nbits = 512, 3, 61
    poses = rng.normal(0, 1, (n, n_pose)).astype(np.float32)
    mask = rng.random((n, nbits)) < 0.2
    tgt = np.roll(mask, 1, axis=1)
    m = fcs.MLP(gpu_ctx, n_pose + nbits, 64, nbits, max_batch=32, seed=5)
    losses = m.train_masks(poses, fcs.pack_bits(mask), fcs.pack_bits(tgt), epochs=30, batch=32, lr=3e-3, seed=6)
    assert np.all(np.isfinite(losses))
    assert losses[-1] < 0.35 * losses[0], losses
    # the trained network predicts the shift on held-out masks
    x = np.concatenate([rng.normal(0, 1, (16, n_pose)), rng.random((16, nbits)) < 0.2], 1).astype(np.float32)
    pred = 1 / (1 + np.exp(-m.forward(x))) > 0.5
    want = np.roll(x[:, n_pose:] > 0.5, 1, axis=1)
    assert (pred == want).mean() > 0.95
    m.close()
