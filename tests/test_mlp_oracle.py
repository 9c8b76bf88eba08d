"""The FP64 MLP oracle's gradient against central finite differences (CPU)."""
import numpy as np

from oracle import mlp_ref


def test_oracle_gradient_matches_finite_differences():
    arch = (5, 8, 3, 4)
    rng = np.random.default_rng(0)
    w = mlp_ref.init_params(arch, seed=1) + rng.normal(0, 0.05, mlp_ref.n_params(*arch))
    x = rng.normal(0, 1, (16, 5))
    y = rng.normal(0, 1, (16, 4))
    loss, g = mlp_ref.loss_and_grad(w, x, y, arch)
    h = 1e-6
    idx = rng.choice(w.size, 60, replace=False)
    for i in idx:
        e = np.zeros_like(w)
        e[i] = h
        fd = (mlp_ref.loss_and_grad(w + e, x, y, arch)[0] - mlp_ref.loss_and_grad(w - e, x, y, arch)[0]) / (2 * h)
        assert abs(fd - g[i]) <= 1e-7 + 1e-5 * abs(fd)
    assert mlp_ref.n_params(5, 256, 4, 4) == 199940
