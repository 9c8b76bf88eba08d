"""ORACLE — FP64 numpy restatement of the surrogate MLP (test
infrastructure only).  Same architecture and flat parameter layout as the
GPU kernels: for each layer W[in][out] row-major then b[out]; SiLU on the
hidden layers, identity output; loss = mean over batch and outputs of the
squared error.  Its gradients are pinned by central finite differences
(tests/test_mlp_oracle.py)."""
import numpy as np


def shapes(n_in, hidden, depth, n_out):
    dims = [n_in] + [hidden] * depth + [n_out]
    return [(dims[i], dims[i + 1]) for i in range(len(dims) - 1)]


def unpack(flat, n_in, hidden, depth, n_out):
    out, o = [], 0
    for k, n in shapes(n_in, hidden, depth, n_out):
        W = flat[o:o + k * n].reshape(k, n)
        o += k * n
        b = flat[o:o + n]
        o += n
        out.append((W, b))
    assert o == flat.size
    return out


def n_params(n_in, hidden, depth, n_out):
    return sum(k * n + n for k, n in shapes(n_in, hidden, depth, n_out))


def silu(z):
    return z / (1.0 + np.exp(-z))


def forward(flat, x, arch):
    layers = unpack(np.asarray(flat, dtype=np.float64), *arch)
    a = np.asarray(x, dtype=np.float64)
    cache = [(a, None)]
    for i, (W, b) in enumerate(layers):
        z = a @ W + b
        if i < len(layers) - 1:
            a = silu(z)
        else:
            a = z
        cache.append((a, z))
    return a, cache


def loss_and_grad(flat, x, y, arch):
    layers = unpack(np.asarray(flat, dtype=np.float64), *arch)
    yp, cache = forward(flat, x, arch)
    B, n_out = yp.shape
    d = yp - y
    loss = float(np.mean(d * d))
    dz = 2.0 * d / (B * n_out)
    grads = [None] * len(layers)
    for i in range(len(layers) - 1, -1, -1):
        W, b = layers[i]
        a_prev = cache[i][0]
        grads[i] = (a_prev.T @ dz, dz.sum(axis=0))
        if i > 0:
            z_prev = cache[i][1]
            s = 1.0 / (1.0 + np.exp(-z_prev))
            dz = (dz @ W.T) * s * (1.0 + z_prev * (1.0 - s))
    flat_g = np.concatenate([np.concatenate([gW.ravel(), gb]) for gW, gb in grads])
    return loss, flat_g


def init_params(arch, seed=0):
    """He-style init: W ~ N(0, 2/fan_in) for SiLU layers, 1/fan_in for the output."""
    rng = np.random.default_rng(seed)
    parts = []
    sh = shapes(*arch)
    for i, (k, n) in enumerate(sh):
        scale = np.sqrt((2.0 if i < len(sh) - 1 else 1.0) / k)
        parts.append(rng.normal(0, scale, k * n))
        parts.append(np.zeros(n))
    return np.concatenate(parts)
