"""The float64 MLP oracle pinned by central finite differences and the
closed-form loss at zero logits; planted mistakes must fail the check."""
import numpy as np
import pytest

from oracle import mlp_oracle as M

DIMS = (7, 5, 4)


def _problem(seed=0, batch=3):
    rng = np.random.default_rng(seed)
    n_in, h, n_out = DIMS
    _, npar = M.layout(*DIMS)
    p = rng.normal(0, 0.5, npar)
    x = rng.normal(0, 1, (batch, n_in))
    y = (rng.random((batch, n_out)) < 0.4).astype(np.float64)
    return p, x, y


def _fd(p, x, y, eps=1e-6):
    g = np.zeros_like(p)
    for i in range(p.size):
        q = p.copy(); q[i] += eps
        lp = M.loss(q, x, y, *DIMS)
        q[i] -= 2 * eps
        lm = M.loss(q, x, y, *DIMS)
        g[i] = (lp - lm) / (2 * eps)
    return g


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_gradient_matches_finite_differences(seed):
    p, x, y = _problem(seed)
    L, g = M.loss_grad(p, x, y, *DIMS)
    assert np.isclose(L, M.loss(p, x, y, *DIMS))
    assert np.allclose(g, _fd(p, x, y), rtol=1e-5, atol=1e-8)


@pytest.mark.parametrize("mutant", [1, 2])
def test_planted_mistakes_fail_the_check(mutant):
    p, x, y = _problem(5)
    _, g = M.loss_grad(p, x, y, *DIMS, mutant=mutant)
    assert not np.allclose(g, _fd(p, x, y), rtol=1e-5, atol=1e-8)


def test_zero_logits_loss_is_log2():
    n_in, h, n_out = DIMS
    _, npar = M.layout(*DIMS)
    p = np.zeros(npar)
    x = np.ones((2, n_in))
    y = np.array([[0, 1, 0, 1], [1, 1, 0, 0]], np.float64)
    assert np.isclose(M.loss(p, x, y, *DIMS), np.log(2.0))


def test_adam_first_step_moves_by_lr_times_sign():
    # closed form: with m0 = v0 = 0, step 1 moves every coordinate by lr * g / (|g| + eps')
    g = np.array([0.5, -2.0, 1e-3])
    p, m, v = M.adam(np.zeros(3), g, np.zeros(3), np.zeros(3), 1, lr=0.1, eps=1e-12)
    assert np.allclose(p, -0.1 * np.sign(g), rtol=1e-6)


def test_bit_packing_round_trip():
    import fcs
    rng = np.random.default_rng(3)
    m = rng.random((4, 77)) < 0.3
    w = fcs.pack_bits(m)
    assert w.shape == (4, 3)
    assert np.array_equal(M.unpack_bits(w, 77), m.astype(np.float64))
