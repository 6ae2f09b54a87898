"""Test infrastructure only (imported by tests/ and bench.py's CPU baseline,
never by the product path): a float64 numpy restatement of the learned
prefetch MLP of include/fcs/fcs_mlp.h -- forward pass, binary cross-entropy
with logits averaged over batch x outputs, backward pass and one Adam step.

The network has no counterpart in the reference, so nothing of the
reference pins it; the gradients are pinned by central finite differences
(tests/test_mlp_oracle.py) and the loss by its closed form at z = 0
(log 2 for every element)."""
import numpy as np


def layout(n_in, h, n_out):
    """Offsets of W1, b1, W2, b2, W3, b3 in the flat parameter array."""
    o = {}
    off = 0
    for name, size in (("W1", n_in * h), ("b1", h), ("W2", h * h), ("b2", h), ("W3", h * n_out), ("b3", n_out)):
        o[name] = (off, size)
        off += size
    return o, off


def unpack(p, n_in, h, n_out):
    o, _ = layout(n_in, h, n_out)
    g = lambda k: p[o[k][0]:o[k][0] + o[k][1]]
    return (g("W1").reshape(n_in, h), g("b1"), g("W2").reshape(h, h), g("b2"), g("W3").reshape(h, n_out), g("b3"))


def forward(p, x, n_in, h, n_out, relu_grad_bug=False):
    W1, b1, W2, b2, W3, b3 = unpack(np.asarray(p, np.float64), n_in, h, n_out)
    x = np.asarray(x, np.float64)
    h1 = np.maximum(x @ W1 + b1, 0.0)
    h2 = np.maximum(h1 @ W2 + b2, 0.0)
    z = h2 @ W3 + b3
    return z, (x, h1, h2)


def loss(p, x, y, n_in, h, n_out):
    z, _ = forward(p, x, n_in, h, n_out)
    y = np.asarray(y, np.float64)
    return float(np.mean(np.maximum(z, 0) - z * y + np.log1p(np.exp(-np.abs(z)))))


def loss_grad(p, x, y, n_in, h, n_out, mutant=0):
    """Loss and gradient (flat, parameter layout). `mutant` plants a mistake
    (1: ReLU derivative taken as 1, 2: bias gradient of the last layer
    dropped) so tests can show the finite-difference check catches it."""
    p = np.asarray(p, np.float64)
    W1, b1, W2, b2, W3, b3 = unpack(p, n_in, h, n_out)
    z, (x, h1, h2) = forward(p, x, n_in, h, n_out)
    y = np.asarray(y, np.float64)
    n = z.size
    L = float(np.mean(np.maximum(z, 0) - z * y + np.log1p(np.exp(-np.abs(z)))))
    dz = (1.0 / (1.0 + np.exp(-z)) - y) / n
    gW3 = h2.T @ dz
    gb3 = dz.sum(0) if mutant != 2 else np.zeros_like(b3)
    dh2 = dz @ W3.T
    if mutant != 1:
        dh2 = dh2 * (h2 > 0)
    gW2 = h1.T @ dh2
    gb2 = dh2.sum(0)
    dh1 = (dh2 @ W2.T) * (h1 > 0)
    gW1 = x.T @ dh1
    gb1 = dh1.sum(0)
    g = np.concatenate([gW1.ravel(), gb1, gW2.ravel(), gb2, gW3.ravel(), gb3])
    return L, g


def adam(p, g, m, v, step, lr, b1=0.9, b2=0.999, eps=1e-8):
    """One Adam step (step counts from 1); returns new (p, m, v)."""
    m = b1 * m + (1 - b1) * g
    v = b2 * v + (1 - b2) * g * g
    p = p - lr * (m / (1 - b1 ** step)) / (np.sqrt(v / (1 - b2 ** step)) + eps)
    return p, m, v


def unpack_bits(words, nbits):
    """LSB-first bit rows -> 0/1 float64 matrix."""
    w = np.asarray(words, np.uint32)
    b = ((w[..., :, None] >> np.arange(32, dtype=np.uint32)) & 1).reshape(w.shape[:-1] + (-1,))
    return b[..., :nbits].astype(np.float64)
