"""Learned surrogate of the transport: an MLP mapping (log mu_a, log mu_s,
g, n, thickness) to logit-transformed sphere readings (R_d, T_d, T_u and
total reflectance R_d + R_s), trained only on tallies of this Monte Carlo.
It seeds the inverse solver; MC re-simulation at the solution judges it.

The network runs on libismc's BF16-MFMA kernels (ismc_mlp_* in
include/ismc.h): 5 -> [256 SiLU] x 4 -> 4, FP32 master weights and Adam.
"""
import ctypes
import os
import time
from dataclasses import dataclass
from typing import Optional

import numpy as np

from . import _lib
from .config import Layer, SampleConfig
from .engine import Engine, config_array, default_engine
from .workloads import standard_sphere

# input box (uniform in these coordinates): log10 mu_a, log10 mu_s, g, n, t
RANGES = np.array([[-2.0, 1.0],       # mu_a 0.01 .. 10 /cm
                   [0.0, 3.0],        # mu_s 1 .. 1000 /cm
                   [0.60, 0.95],      # g
                   [1.33, 1.50],      # n
                   [0.05, 0.50]])     # t / cm
OUTPUTS = ("R_d", "T_d", "T_u", "R_tot")
CLAMP = 1e-6


def features(mua, mus, g, n, t):
    """Inputs scaled to [-1, 1] over the training box."""
    raw = np.stack([np.log10(np.asarray(mua, dtype=np.float64)),
                    np.log10(np.asarray(mus, dtype=np.float64)),
                    np.asarray(g, dtype=np.float64) + 0 * np.asarray(mua, dtype=np.float64),
                    np.asarray(n, dtype=np.float64) + 0 * np.asarray(mua, dtype=np.float64),
                    np.asarray(t, dtype=np.float64) + 0 * np.asarray(mua, dtype=np.float64)], axis=-1)
    lo, hi = RANGES[:, 0], RANGES[:, 1]
    return (2.0 * (raw - lo) / (hi - lo) - 1.0).astype(np.float32)


def logit(p):
    p = np.clip(np.asarray(p, dtype=np.float64), CLAMP, 1 - CLAMP)
    return np.log(p / (1 - p))


def expit(z):
    return 1.0 / (1.0 + np.exp(-np.asarray(z, dtype=np.float64)))


def sample_inputs(n, seed=0):
    """Uniform in (log mu_a, log mu_s, g, n, t) over the training box."""
    rng = np.random.default_rng(seed)
    u = rng.random((n, 5))
    v = RANGES[:, 0] + u * (RANGES[:, 1] - RANGES[:, 0])
    return {"mu_a": 10 ** v[:, 0], "mu_s": 10 ** v[:, 1], "g": v[:, 2], "n": v[:, 3], "t": v[:, 4]}


def sample_config(mu_a, mu_s, g, n, t):
    return SampleConfig(layers=[Layer(n=float(n), g=float(g), t=float(t), mu_a=float(mu_a),
                                      mu_s=float(mu_s))],
                        R=standard_sphere(), T=standard_sphere(), roulette_threshold=1e-4)


def generate_dataset(n_samples, packets, seed=0, engine: Optional[Engine] = None,
                     points_per_launch=16384, progress=None, autotune=True, return_se=False, stats=None):
    """MC labels for n_samples inputs: one batched launch per
    points_per_launch samples (one work-queue chunk per sample unless
    `autotune` picks the launch shape; labels are identical either way).
    Returns (inputs dict, tallies [n, 4] in OUTPUTS order, seconds, packets),
    with return_se the labels' MC standard errors [n, 4] before the seconds
    (R_tot's from E[R_d^2] + E[R_s^2]: a packet leaves its weight in one
    exit tally).  A `stats` dict, when given, receives the launches' kernel
    time ("kernel_ms"), summed event counts ("events") and the seconds the
    autotuner spent timing launch shapes on first use of this workload
    class ("tuning_s", part of the returned seconds)."""
    eng = engine or default_engine()
    inp = sample_inputs(n_samples, seed)
    tal = np.zeros((n_samples, 4))
    se = np.zeros((n_samples, 4))
    t0 = time.perf_counter()
    for s in range(0, n_samples, points_per_launch):
        e = min(n_samples, s + points_per_launch)
        cfgs = [sample_config(inp["mu_a"][i], inp["mu_s"][i], inp["g"][i], inp["n"][i], inp["t"][i])
                for i in range(s, e)]
        kw = {"chunk_packets": min(int(packets), 1 << 20)}
        if autotune:
            t1 = time.perf_counter()
            kw = eng.tuner.params(cfgs, packets, seed=seed + 7)
            if stats is not None:
                stats["tuning_s"] = stats.get("tuning_s", 0.0) + time.perf_counter() - t1
        sums, sumsq, npk = eng.simulate_batch_array(config_array(cfgs), packets, seed=seed + 7,
                                                    streams=np.arange(s, e, dtype=np.uint32), **kw)
        if stats is not None:
            stats["kernel_ms"] = stats.get("kernel_ms", 0.0) + eng.last_launch()["kernel_ms"]
            ev = stats.setdefault("events", {})
            for name, v in eng.last_events().items():
                ev[name] = ev.get(name, 0) + v
        # columns of sums in TALLIES order: A, R_d, R_s, T_d, T_u
        N = np.maximum(npk.astype(np.float64), 1.0)[:, None]
        m = sums / N
        tal[s:e] = np.stack([m[:, 1], m[:, 3], m[:, 4], m[:, 1] + m[:, 2]], axis=1)
        if return_se:
            rt = m[:, 1] + m[:, 2]
            v = np.maximum(sumsq / N - m * m, 0.0)
            v_tot = np.maximum((sumsq[:, 1] + sumsq[:, 2]) / N[:, 0] - rt * rt, 0.0)
            d = np.maximum(N - 1.0, 1.0)
            se[s:e] = np.sqrt(np.stack([v[:, 1], v[:, 3], v[:, 4], v_tot], axis=1) / d)
        if progress:
            progress(e, n_samples, time.perf_counter() - t0)
    if return_se:
        return inp, tal, se, time.perf_counter() - t0, n_samples * int(packets)
    return inp, tal, time.perf_counter() - t0, n_samples * int(packets)


class MLP:
    """Handle on one ismc_mlp (weights live in device memory)."""

    def __init__(self, n_in=5, hidden=256, depth=4, n_out=4, max_batch=4096, device=0):
        self.device = device
        self._L = _lib.load()
        L = self._L
        L.ismc_mlp_create.argtypes = [ctypes.c_int, ctypes.c_int32, ctypes.c_int32, ctypes.c_int32,
                                      ctypes.c_int32, ctypes.c_int32, ctypes.POINTER(ctypes.c_void_p)]
        L.ismc_mlp_create.restype = ctypes.c_int
        L.ismc_mlp_destroy.argtypes = [ctypes.c_void_p]
        L.ismc_mlp_destroy.restype = None
        L.ismc_mlp_last_error.argtypes = [ctypes.c_void_p]
        L.ismc_mlp_last_error.restype = ctypes.c_char_p
        L.ismc_mlp_n_params.argtypes = [ctypes.c_void_p]
        L.ismc_mlp_n_params.restype = ctypes.c_int64
        for name in ("ismc_mlp_set_params", "ismc_mlp_get_params"):
            getattr(L, name).argtypes = [ctypes.c_void_p, ctypes.c_void_p]
            getattr(L, name).restype = ctypes.c_int
        L.ismc_mlp_forward.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int32, ctypes.c_void_p]
        L.ismc_mlp_forward.restype = ctypes.c_int
        L.ismc_mlp_gradients.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int32,
                                         ctypes.c_void_p, ctypes.POINTER(ctypes.c_float)]
        L.ismc_mlp_gradients.restype = ctypes.c_int
        L.ismc_mlp_load_dataset.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int64]
        L.ismc_mlp_load_dataset.restype = ctypes.c_int
        L.ismc_mlp_train.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int32, ctypes.c_int32,
                                     ctypes.c_int32, ctypes.c_float, ctypes.c_float, ctypes.c_float,
                                     ctypes.c_float, ctypes.c_void_p, ctypes.POINTER(ctypes.c_float)]
        L.ismc_mlp_train.restype = ctypes.c_int
        L.ismc_mlp_train_step.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int32,
                                          ctypes.c_float, ctypes.c_float, ctypes.c_float, ctypes.c_float,
                                          ctypes.POINTER(ctypes.c_float)]
        L.ismc_mlp_train_step.restype = ctypes.c_int
        L.ismc_mlp_last_ms.argtypes = [ctypes.c_void_p, ctypes.POINTER(ctypes.c_float)]
        L.ismc_mlp_last_ms.restype = ctypes.c_int
        self._h = ctypes.c_void_p()
        rc = L.ismc_mlp_create(device, n_in, hidden, depth, n_out, max_batch, ctypes.byref(self._h))
        if rc != 0:
            raise RuntimeError("ismc_mlp_create failed with status %d (needs a GPU)" % rc)
        self.arch = (n_in, hidden, depth, n_out)
        self.max_batch = max_batch
        self.n_params = int(L.ismc_mlp_n_params(self._h))

    def close(self):
        if self._h:
            self._L.ismc_mlp_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, rc, what):
        if rc != 0:
            msg = self._L.ismc_mlp_last_error(self._h).decode(errors="replace")
            raise RuntimeError("%s: %s (status %d)" % (what, msg, rc))

    def set_params(self, flat):
        f = np.ascontiguousarray(flat, dtype=np.float32)
        assert f.size == self.n_params
        self._check(self._L.ismc_mlp_set_params(self._h, f.ctypes.data), "set_params")

    def get_params(self):
        f = np.zeros(self.n_params, dtype=np.float32)
        self._check(self._L.ismc_mlp_get_params(self._h, f.ctypes.data), "get_params")
        return f

    def forward(self, x):
        x = np.ascontiguousarray(x, dtype=np.float32)
        y = np.zeros((x.shape[0], self.arch[3]), dtype=np.float32)
        self._check(self._L.ismc_mlp_forward(self._h, x.ctypes.data, x.shape[0], y.ctypes.data), "forward")
        return y

    def last_ms(self):
        ms = ctypes.c_float()
        self._L.ismc_mlp_last_ms(self._h, ctypes.byref(ms))
        return ms.value

    def gradients(self, x, y):
        x = np.ascontiguousarray(x, dtype=np.float32)
        y = np.ascontiguousarray(y, dtype=np.float32)
        g = np.zeros(self.n_params, dtype=np.float32)
        loss = ctypes.c_float()
        self._check(self._L.ismc_mlp_gradients(self._h, x.ctypes.data, y.ctypes.data, x.shape[0],
                                               g.ctypes.data, ctypes.byref(loss)), "gradients")
        return float(loss.value), g

    def train_step(self, x, y, lr=1e-3, beta1=0.9, beta2=0.999, eps=1e-8):
        """One Adam step on a host batch (rows a multiple of 64); returns the
        batch's mean squared error before the step."""
        x = np.ascontiguousarray(x, dtype=np.float32)
        y = np.ascontiguousarray(y, dtype=np.float32)
        loss = ctypes.c_float()
        self._check(self._L.ismc_mlp_train_step(self._h, x.ctypes.data, y.ctypes.data, x.shape[0], lr, beta1,
                                                beta2, eps, ctypes.byref(loss)), "train_step")
        return float(loss.value)

    def load_dataset(self, x, y):
        self._x = np.ascontiguousarray(x, dtype=np.float32)
        self._y = np.ascontiguousarray(y, dtype=np.float32)
        self._check(self._L.ismc_mlp_load_dataset(self._h, self._x.ctypes.data, self._y.ctypes.data,
                                                  self._x.shape[0]), "load_dataset")

    def train(self, epochs, batch=4096, lr=1e-3, beta1=0.9, beta2=0.999, eps=1e-8, seed=0, n_rows=None):
        n = int(n_rows or self._x.shape[0])
        steps = n // batch
        rng = np.random.default_rng(seed)
        perm = np.concatenate([rng.permutation(n)[:steps * batch] for _ in range(epochs)]).astype(np.int32)
        losses = np.zeros(epochs, dtype=np.float32)
        ms = ctypes.c_float()
        self._check(self._L.ismc_mlp_train(self._h, perm.ctypes.data, epochs, steps, batch, lr, beta1,
                                           beta2, eps, losses.ctypes.data, ctypes.byref(ms)), "train")
        return losses, ms.value


@dataclass
class Surrogate:
    """Trained network plus the transform between tallies and outputs."""
    mlp: MLP
    train_info: dict

    def predict(self, mua, mus, g, n, t):
        x = features(mua, mus, g, n, t)
        z = self.mlp.forward(x.reshape(-1, 5))
        p = expit(z)
        return {k: p[:, i] for i, k in enumerate(OUTPUTS)}

    def predict_log10(self, log_mua, log_mus, g, n, t, outputs=("R_d", "T_d")):
        """predict() for points given as log10 mu_a, log10 mu_s and scalar
        g, n, t, with only `outputs` activated: the inputs are scaled
        straight into float32 (no 10**x followed by log10) for the large
        candidate grids of the inverse solver's seeds."""
        la = np.asarray(log_mua, dtype=np.float64).ravel()
        ls = np.asarray(log_mus, dtype=np.float64).ravel()
        lo, hi = RANGES[:, 0], RANGES[:, 1]
        x = np.empty((la.size, 5), dtype=np.float32)
        x[:, 0] = 2.0 * (la - lo[0]) / (hi[0] - lo[0]) - 1.0
        x[:, 1] = 2.0 * (ls - lo[1]) / (hi[1] - lo[1]) - 1.0
        for j, v in ((2, g), (3, n), (4, t)):
            x[:, j] = 2.0 * (float(v) - lo[j]) / (hi[j] - lo[j]) - 1.0
        z = self.mlp.forward(x)
        return {k: expit(z[:, OUTPUTS.index(k)]) for k in outputs}

    def save(self, path):
        np.savez(path, params=self.mlp.get_params(), arch=np.array(self.mlp.arch),
                 ranges=RANGES, info=np.array([repr(self.train_info)]))

    @classmethod
    def load(cls, path, device=0, max_batch=65536):
        d = np.load(path, allow_pickle=False)
        arch = tuple(int(v) for v in d["arch"])
        mlp = MLP(*arch, max_batch=max_batch, device=device)
        mlp.set_params(d["params"])
        return cls(mlp, {"loaded_from": os.path.basename(path)})


def init_params(arch, seed=0):
    """He-style initialisation (N(0, 2/fan_in) for SiLU layers, 1/fan_in at
    the output), flat logical layout."""
    rng = np.random.default_rng(seed)
    dims = [arch[0]] + [arch[1]] * arch[2] + [arch[3]]
    parts = []
    for i in range(len(dims) - 1):
        k, n = dims[i], dims[i + 1]
        scale = np.sqrt((2.0 if i < len(dims) - 2 else 1.0) / k)
        parts += [rng.normal(0, scale, k * n), np.zeros(n)]
    return np.concatenate(parts).astype(np.float32)


def packaged_weights():
    """Path of the trained weights shipped with the package (2^20 MC-labelled
    samples at 1e5 packets each)."""
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "data", "surrogate.npz")
    if not os.path.exists(path):
        raise FileNotFoundError("no packaged surrogate weights at %s; train them with "
                                "python -m ismc_amd.train_surrogate --out %s" % (path, path))
    return path


def train_surrogate(inputs, tallies, epochs=60, batch=4096, lr=2e-3, seed=0, holdout=0.05,
                    device=0, lr_decay_every=20):
    """Fit the MLP to MC tallies (MSE on logits); returns (Surrogate,
    held-out errors)."""
    n = tallies.shape[0]
    x = features(inputs["mu_a"], inputs["mu_s"], inputs["g"], inputs["n"], inputs["t"])
    y = logit(tallies).astype(np.float32)
    rng = np.random.default_rng(seed)
    order = rng.permutation(n)
    n_hold = max(batch, int(n * holdout))
    hold, tr = order[:n_hold], order[n_hold:]
    mlp = MLP(max_batch=batch, device=device)
    mlp.set_params(init_params(mlp.arch, seed))
    mlp.load_dataset(x[tr], y[tr])
    losses, ms = [], 0.0
    done = 0
    while done < epochs:
        e = min(lr_decay_every, epochs - done)
        l, m = mlp.train(e, batch=batch, lr=lr * (0.3 ** (done // lr_decay_every)), seed=seed + done)
        losses += list(map(float, l))
        ms += m
        done += e
    pred = mlp.forward(x[hold])
    err = {k: float(np.sqrt(np.mean((expit(pred[:, i]) - tallies[hold, i]) ** 2)))
           for i, k in enumerate(OUTPUTS)}
    rel = {k: float(np.median(np.abs(expit(pred[:, i]) - tallies[hold, i]) /
                              np.maximum(tallies[hold, i], 1e-3))) for i, k in enumerate(OUTPUTS)}
    info = {"epochs": epochs, "batch": batch, "lr": lr, "losses": losses, "train_ms": ms,
            "n_train": int(tr.size), "n_holdout": int(hold.size), "holdout_rmse": err,
            "holdout_median_rel": rel}
    return Surrogate(mlp, info), info


HOLDOUT_SEED = 20_000_003   # input draws and Philox seed of the fresh held-out set


def holdout_report(sur, n_samples=16384, packets=100_000, seed=HOLDOUT_SEED, engine: Optional[Engine] = None):
    """The network against fresh MC labels: n_samples inputs drawn over the
    training box with a seed the training never used, each labelled by
    `packets` packets of this engine.  Per output: RMSE of the prediction
    against the labels, the labels' median and RMS MC standard error, their
    ratio, and the network's own RMSE with the labels' noise removed,
    sqrt(max(RMSE^2 - mean SE^2, 0)).  Also the label generation's
    seconds and packets/s without the autotuner's one-time timing of this
    workload class (reported beside them), and its kernel time."""
    eng = engine or default_engine()
    st = {}
    inp, tal, se, secs, npk = generate_dataset(n_samples, packets, seed=seed, engine=eng, return_se=True,
                                               stats=st)
    p = sur.predict(inp["mu_a"], inp["mu_s"], inp["g"], inp["n"], inp["t"])
    out = {}
    for i, k in enumerate(OUTPUTS):
        d = p[k] - tal[:, i]
        rmse = float(np.sqrt(np.mean(d * d)))
        msq = float(np.mean(se[:, i] ** 2))
        med = float(np.median(se[:, i]))
        out[k] = {"rmse": rmse, "median_label_se": med, "rms_label_se": msq ** 0.5,
                  "rmse_over_median_se": rmse / med if med > 0 else None,
                  "rmse_over_rms_se": rmse / msq ** 0.5 if msq > 0 else None,
                  "network_rmse_noise_removed": max(rmse * rmse - msq, 0.0) ** 0.5}
    gen = secs - st.get("tuning_s", 0.0)
    return {"samples": n_samples, "packets_per_label": int(packets), "seed": seed,
            "label_packets": int(npk), "label_seconds": gen, "label_packets_per_s": npk / gen,
            "label_tuning_seconds": st.get("tuning_s", 0.0),
            "label_kernel_seconds": st["kernel_ms"] / 1e3,
            "label_kernel_packets_per_s": npk / (st["kernel_ms"] / 1e3), "label_events": st["events"],
            "outputs": out, "inputs": inp, "labels": tal, "label_se": se}


def time_training(x, y, epochs=10, batch=4096, device=0, seed=0):
    """Device ms per Adam step of the network's training on (x, y) (a fresh
    network; the packaged one is not touched): `epochs` passes of
    len(x) // batch steps in one ismc_mlp_train call."""
    mlp = MLP(max_batch=batch, device=device)
    try:
        mlp.set_params(init_params(mlp.arch, seed))
        mlp.load_dataset(x, y)
        steps = x.shape[0] // batch
        losses, ms = mlp.train(epochs, batch=batch, lr=1e-3, seed=seed)
        return {"steps": epochs * steps, "batch": batch, "ms_per_step": ms / (epochs * steps),
                "loss_first_last": [float(losses[0]), float(losses[-1])]}
    finally:
        mlp.close()
