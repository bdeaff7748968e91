"""CPU reference (float64 numpy) of the cost model's MLP, its masked MSE loss, the loss
gradient and Adam -- the checker for the library's GPU kernels (xc_mlp_*).

Layout: parameters are one flat array, layer by layer W_l (dims[l+1] x dims[l], row-major)
then b_l (dims[l+1]); tanh on hidden layers, linear output.  The gradient here is pinned by
central finite differences in tests/test_mlp_ref.py, so it is not trusted by construction.
"""
import numpy as np


def split_params(dims, p):
    out, off = [], 0
    for l in range(len(dims) - 1):
        nw = dims[l + 1] * dims[l]
        W = p[off:off + nw].reshape(dims[l + 1], dims[l])
        off += nw
        b = p[off:off + dims[l + 1]]
        off += dims[l + 1]
        out.append((W, b))
    assert off == p.size
    return out


def num_params(dims):
    return sum(dims[l + 1] * dims[l] + dims[l + 1] for l in range(len(dims) - 1))


def forward(dims, p, x):
    """Returns (output, list of layer outputs)."""
    a = np.asarray(x, np.float64)
    acts = []
    layers = split_params(dims, np.asarray(p, np.float64))
    for l, (W, b) in enumerate(layers):
        z = a @ W.T + b
        a = np.tanh(z) if l + 1 < len(layers) else z
        acts.append(a)
    return a, acts


def loss(dims, p, x, t, mask):
    y, _ = forward(dims, p, x)
    ms = mask.sum()
    return float((mask * (y - t) ** 2).sum() / ms) if ms > 0 else 0.0


def loss_grad(dims, p, x, t, mask):
    """(loss, gradient) by backpropagation."""
    p = np.asarray(p, np.float64)
    x = np.asarray(x, np.float64)
    layers = split_params(dims, p)
    y, acts = forward(dims, p, x)
    ms = mask.sum()
    inv = 1.0 / ms if ms > 0 else 0.0
    L = float((mask * (y - t) ** 2).sum() * inv)
    dA = 2.0 * mask * (y - t) * inv
    grads = []
    for l in range(len(layers) - 1, -1, -1):
        W, _ = layers[l]
        A = acts[l]
        dZ = dA * (1.0 - A * A) if l + 1 < len(layers) else dA
        X = acts[l - 1] if l > 0 else x
        grads.append((dZ.T @ X, dZ.sum(axis=0)))
        dA = dZ @ W
    flat = []
    for gW, gb in reversed(grads):
        flat.append(gW.ravel())
        flat.append(gb)
    return L, np.concatenate(flat)


def adam(p, g, m, v, step, lr, b1=0.9, b2=0.999, eps=1e-8, wd=0.0):
    """One Adam step with decoupled weight decay; returns new (p, m, v)."""
    m = b1 * m + (1 - b1) * g
    v = b2 * v + (1 - b2) * g * g
    mh = m / (1 - b1 ** step)
    vh = v / (1 - b2 ** step)
    return p - lr * (mh / (np.sqrt(vh) + eps) + wd * p), m, v
