"""Surrogate MLP kernels (BF16 MFMA, FP32 accumulate) against the FP64
oracle of the same network, and end-to-end training on MC labels."""
import numpy as np
import pytest

from ismc_amd import surrogate as S
from oracle import mlp_ref

pytestmark = pytest.mark.gpu
ARCH = (5, 256, 4, 4)


@pytest.fixture(scope="module")
def mlp():
    m = S.MLP(*ARCH, max_batch=4096)
    yield m
    m.close()


def test_param_roundtrip(mlp):
    assert mlp.n_params == 199940 == mlp_ref.n_params(*ARCH)
    w = S.init_params(ARCH, seed=3)
    mlp.set_params(w)
    assert np.array_equal(mlp.get_params(), w)


def test_forward_matches_fp64_oracle(mlp):
    w = S.init_params(ARCH, seed=5)
    mlp.set_params(w)
    rng = np.random.default_rng(0)
    x = rng.uniform(-1, 1, (1000, 5)).astype(np.float32)
    y = mlp.forward(x)
    yr, _ = mlp_ref.forward(w.astype(np.float64), x.astype(np.float64), ARCH)
    err = np.abs(y - yr)
    scale = np.abs(yr).mean()
    print("forward max abs err %.3e (mean |y| %.3e)" % (err.max(), scale))
    assert err.max() < 0.03 * scale + 1e-3


def test_gradients_match_fp64_oracle(mlp):
    w = S.init_params(ARCH, seed=7)
    mlp.set_params(w)
    rng = np.random.default_rng(1)
    x = rng.uniform(-1, 1, (4096, 5)).astype(np.float32)
    t = rng.normal(0, 1, (4096, 4)).astype(np.float32)
    loss, g = mlp.gradients(x, t)
    lr, gr = mlp_ref.loss_and_grad(w.astype(np.float64), x.astype(np.float64), t.astype(np.float64), ARCH)
    cos = float(np.dot(g, gr) / (np.linalg.norm(g) * np.linalg.norm(gr)))
    rel = float(np.linalg.norm(g - gr) / np.linalg.norm(gr))
    print("loss gpu %.6f oracle %.6f, grad cosine %.6f, relative L2 %.4f" % (loss, lr, cos, rel))
    assert loss == pytest.approx(lr, rel=0.02)
    assert cos > 0.999 and rel < 0.05
    # per layer (weights and biases separately)
    o = 0
    for k, n in mlp_ref.shapes(*ARCH):
        for size in (k * n, n):
            a, b = g[o:o + size], gr[o:o + size]
            if np.linalg.norm(b) > 0:
                assert np.linalg.norm(a - b) / np.linalg.norm(b) < 0.08
            o += size


def test_train_step_is_one_adam_step_on_the_gradients():
    """ismc_mlp_train_step on a host batch equals ismc_mlp_gradients on the
    same batch followed by Adam's first step written out in numpy (zero
    moments: m = (1-b1) g, v = (1-b2) g^2, step size lr sqrt(1-b2)/(1-b1));
    its loss is the batch's loss before the step."""
    m = S.MLP(*ARCH, max_batch=4096)
    try:
        w0 = S.init_params(ARCH, seed=9)
        m.set_params(w0)
        rng = np.random.default_rng(4)
        x = rng.uniform(-1, 1, (1024, 5)).astype(np.float32)
        t = rng.normal(0, 1, (1024, 4)).astype(np.float32)
        loss_g, g = m.gradients(x, t)
        lr, b1, b2, eps = 1e-3, 0.9, 0.999, 1e-8
        loss = m.train_step(x, t, lr=lr, beta1=b1, beta2=b2, eps=eps)
        w1 = m.get_params()
        f = np.float32
        mm = f(1 - b1) * g
        vv = f(1 - b2) * g * g
        lr_t = f(lr * np.sqrt(1 - b2) / (1 - b1))
        expect = w0 - lr_t * mm / (np.sqrt(vv) + f(eps))
        assert loss == pytest.approx(loss_g, rel=1e-6)
        assert np.allclose(w1, expect, rtol=1e-5, atol=1e-7)
        assert np.abs(w1 - w0).max() > 0.5 * lr
    finally:
        m.close()


def test_training_fits_a_smooth_function(mlp):
    rng = np.random.default_rng(2)
    x = rng.uniform(-1, 1, (65536, 5)).astype(np.float32)
    y = np.stack([np.sin(2 * x[:, 0]) + x[:, 1] * x[:, 2], np.tanh(x[:, 3] - x[:, 4]),
                  x[:, 0] ** 2 - x[:, 1], np.cos(x[:, 2] + x[:, 4])], 1).astype(np.float32)
    mlp.set_params(S.init_params(ARCH, seed=1))
    mlp.load_dataset(x, y)
    losses, ms = mlp.train(15, batch=4096, lr=2e-3, seed=0)
    print("losses", np.round(losses, 5), "ms", ms)
    assert losses[-1] < 0.05 * losses[0] and losses[-1] < 5e-3


def test_surrogate_pipeline_small(engine):
    inp, tal, secs, packets = S.generate_dataset(8192, 4000, seed=11, engine=engine)
    assert tal.shape == (8192, 4) and np.all((tal >= 0) & (tal <= 1))
    sur, info = S.train_surrogate(inp, tal, epochs=20, batch=1024, lr=2e-3, lr_decay_every=10)
    print("holdout", info["holdout_rmse"], "train ms", info["train_ms"])
    assert info["losses"][-1] < info["losses"][0]
    assert info["holdout_rmse"]["R_d"] < 0.05 and info["holdout_rmse"]["T_d"] < 0.05


def test_labels_equal_simulate_batch_results(engine):
    """generate_dataset (tuned launch shape, tallies read as arrays) gives
    the same labels as simulate_batch's per-point Results of a default-shaped
    launch on the same samples, seed and streams: tallies do not depend on
    the launch shape."""
    inp, tal, se, secs, npk = S.generate_dataset(600, 3000, seed=21, engine=engine, return_se=True)
    cfgs = [S.sample_config(inp["mu_a"][i], inp["mu_s"][i], inp["g"][i], inp["n"][i], inp["t"][i])
            for i in range(600)]
    res = engine.simulate_batch(cfgs, 3000, seed=21 + 7, streams=np.arange(600, dtype=np.uint32))
    for i, r in enumerate(res):
        m = r.mean
        assert tal[i].tolist() == [m["R_d"], m["T_d"], m["T_u"], m["R_d"] + m["R_s"]], i
        np.testing.assert_allclose(se[i, :3], [r.se["R_d"], r.se["T_d"], r.se["T_u"]], rtol=1e-12)


def test_train_rejects_indices_outside_the_dataset():
    """ismc_mlp_train checks every batch index against the resident dataset
    on the host (an index outside it would make the gather read beyond the
    device buffers): negative or too-large indices give ISMC_EINVAL before
    any launch, and the network trains normally afterwards."""
    import ctypes
    m = S.MLP(max_batch=128)
    try:
        m.set_params(S.init_params(m.arch, 3))
        rng = np.random.default_rng(0)
        x = rng.uniform(-1, 1, (256, 5)).astype(np.float32)
        y = rng.normal(0, 1, (256, 4)).astype(np.float32)
        m.load_dataset(x, y)
        L = m._L
        losses = np.zeros(1, dtype=np.float32)
        ms = ctypes.c_float()
        for bad in (-1, 256, 1 << 30):
            perm = np.arange(256, dtype=np.int32)
            perm[100] = bad
            rc = L.ismc_mlp_train(m._h, perm.ctypes.data, 1, 2, 128, 1e-3, 0.9, 0.999, 1e-8,
                                  losses.ctypes.data, ctypes.byref(ms))
            assert rc == -1 and b"index" in L.ismc_mlp_last_error(m._h)
        before = m.get_params()
        l, _ = m.train(2, batch=128, lr=1e-3)
        assert np.all(np.isfinite(l)) and not np.array_equal(before, m.get_params())
    finally:
        m.close()


def test_packaged_network_against_fresh_labels(engine):
    """The packaged network against 2^14 fresh samples of the training box,
    labelled by 1e5 MC packets each (a seed the training never used): its
    RMSE on the sphere readings R_d and T_d stays within 1.5 times the
    labels' median MC standard error, and with the labels' own noise
    removed the network's error is within 1.25 times that standard error."""
    sur = S.Surrogate.load(S.packaged_weights())
    rep = S.holdout_report(sur, n_samples=1 << 14, packets=100_000, engine=engine)
    o = rep["outputs"]
    print({k: {q: round(v, 6) for q, v in o[k].items() if v is not None} for k in o},
          "%.3g label packets/s" % rep["label_packets_per_s"])
    for k in ("R_d", "T_d"):
        assert o[k]["rmse"] <= 1.5 * o[k]["median_label_se"], (k, o[k])
        assert o[k]["network_rmse_noise_removed"] <= 1.25 * o[k]["median_label_se"], (k, o[k])
