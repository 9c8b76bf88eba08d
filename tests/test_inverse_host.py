"""Inverse pipeline host logic (CPU) with an analytic stand-in for the GPU
transport and for the surrogate."""
import numpy as np

from ismc_amd import inverse
from test_minimization_host import AnalyticEngine


class AnalyticSurrogate:
    def predict(self, mua, mus, g, n, t):
        mua = np.atleast_1d(mua)
        mus = np.atleast_1d(mus)
        s = mus * (1 - g)
        Rd = 0.5 * s / (s + 8 * mua + 5)
        Td = 0.6 * np.exp(-0.08 * mua - 0.01 * s)
        return {"R_d": Rd, "T_d": Td, "T_u": 0 * Rd, "R_tot": Rd}


def np_grid_argmin(pR, pT, R, T):
    """The global seed search in numpy (its GPU form, grid_argmin, is
    checked against this in test_gpu_inverse.py)."""
    R = np.asarray(R, dtype=np.float64)
    T = np.asarray(T, dtype=np.float64)
    c = (((np.asarray(pR, dtype=np.float32)[None, :] - R[:, None]) / R[:, None]) ** 2 +
         ((np.asarray(pT, dtype=np.float32)[None, :] - T[:, None]) / np.maximum(T, 1e-6)[:, None]) ** 2)
    return np.argmin(c, axis=1)


def test_seed_refine_check_on_analytic_model():
    eng = AnalyticEngine()
    truths = np.array([(0.05, 2.0), (0.5, 20.0), (3.0, 80.0), (0.2, 150.0)])
    probs = []
    for a, s in truths:
        t = eng.tallies(inverse.problem(0.5, 0.5).configs(a, s)[0][0])
        probs.append(inverse.problem(t["R_d"], t["T_d"]))
    R = [p.measured.R_d for p in probs]
    T = [p.measured.T_d for p in probs]
    seeds = inverse.surrogate_seeds(AnalyticSurrogate(), R, T, grid=64, polish=True, argmin=np_grid_argmin)
    assert np.allclose(seeds, truths, rtol=1e-3)
    sol, nev, st = inverse.refine_batch(probs, seeds * 1.05, packets=1000, engine=eng, tol=1e-6,
                                        step=0.1, max_evals=400)
    assert np.allclose(sol, truths, rtol=1e-4)
    miss, rR, rT, _ = inverse.check(probs, sol, 1000, engine=eng)
    assert not miss.any() and rR.max() < 1e-4
    ea, es = inverse.errors(sol, truths[:, 0], truths[:, 1])
    s = inverse.summary(ea, es)
    assert s["frac_both_within_1pct"] == 1.0


def test_defect_correction_converges_with_an_exact_model():
    eng = AnalyticEngine()
    truths = np.array([(0.05, 2.0), (0.5, 20.0), (3.0, 80.0)])
    probs = []
    for a, s in truths:
        t = eng.tallies(inverse.problem(0.5, 0.5).configs(a, s)[0][0])
        probs.append(inverse.problem(t["R_d"], t["T_d"]))
    sol, nev, st = inverse.refine_newton(probs, truths * [1.2, 0.85], AnalyticSurrogate(), iters=8,
                                         packets=1000, engine=eng, s_cut=1e-6)
    assert np.allclose(sol, truths, rtol=1e-5)
    assert st["launches"] == 8 and eng.calls == [3] * 8


def test_defect_correction_leaves_unresolved_directions_alone():
    # at mu_a = 0.05 the model's slope in ln mu_a is ~0.04 (below the cut):
    # only the mu_s' direction is corrected and the residual still falls
    eng = AnalyticEngine()
    truth = np.array([(0.05, 2.0)])
    t = eng.tallies(inverse.problem(0.5, 0.5).configs(0.05, 2.0)[0][0])
    probs = [inverse.problem(t["R_d"], t["T_d"])]
    start = truth * [1.5, 0.8]
    sol, _, _ = inverse.refine_newton(probs, start, AnalyticSurrogate(), iters=8, packets=1000, engine=eng)
    def resid(x):
        v = eng.tallies(probs[0].configs(*x)[0][0])
        return np.hypot(v["R_d"] / t["R_d"] - 1, v["T_d"] / t["T_d"] - 1)
    assert resid(sol[0]) < 0.2 * resid(start[0])
    assert abs(sol[0, 1] / 2.0 - 1) < abs(start[0, 1] / 2.0 - 1)


class WrongSlopeSurrogate(AnalyticSurrogate):
    def predict(self, mua, mus, g, n, t):
        p = super().predict(mua, mus, g, n, t)
        # slopes of the wrong sign: every Gauss-Newton step goes uphill
        return {k: 2 * np.median(v) - v for k, v in p.items()}


def test_defect_correction_rejects_uphill_steps():
    eng = AnalyticEngine()
    t = eng.tallies(inverse.problem(0.5, 0.5).configs(0.5, 20.0)[0][0])
    probs = [inverse.problem(t["R_d"], t["T_d"])]
    start = np.array([(0.6, 17.0)])
    sol, _, st = inverse.refine_newton(probs, start, WrongSlopeSurrogate(), iters=5, packets=1000,
                                       engine=eng, s_cut=1e-6)
    assert np.allclose(sol, start)        # the best point evaluated is the start
    assert st["rejected"] == 4


def test_mc_newton_fallback_recovers_truth_from_far_starts():
    """The fallback's Gauss-Newton on (noise-free) residuals with
    forward-difference Jacobians reaches the truth from starts 2-3x off,
    launching only the solves still pending."""
    eng = AnalyticEngine()
    truths = np.array([(0.05, 2.0), (0.5, 20.0), (3.0, 80.0), (0.2, 150.0)])
    probs = []
    for a, s in truths:
        t = eng.tallies(inverse.problem(0.5, 0.5).configs(a, s)[0][0])
        probs.append(inverse.problem(t["R_d"], t["T_d"]))
    start = truths * np.array([[2.5, 0.4], [0.4, 2.0], [1.8, 1.8], [0.5, 0.5]])
    sol, nev, st = inverse.refine_mc_newton(probs, start, iters=30, packets=1000, engine=eng)
    assert np.allclose(sol, truths, rtol=2e-3)
    assert st["launches"] <= 30 and np.all(nev % 3 == 0)
    # every launch held three points per pending solve, and fewer as solves finished
    assert all(n % 3 == 0 for n in eng.calls) and eng.calls[-1] <= eng.calls[0]


def test_predict_log10_equals_predict():
    # the seeds' fast path (inputs built from log10 coordinates, two outputs)
    # gives the network the same float32 inputs as predict()
    from ismc_amd import surrogate

    class Recording:
        device = 0

        def __init__(self):
            self.seen = []

        def forward(self, x):
            x = np.asarray(x, dtype=np.float32)
            self.seen.append(x.copy())
            return np.stack([x[:, 0] - 0.5 * x[:, 1], x[:, 1] * x[:, 4] - x[:, 0], x[:, 2] + x[:, 3], x[:, 3]],
                            1).astype(np.float32)

    sur = surrogate.Surrogate(Recording(), {})
    rng = np.random.default_rng(3)
    la, ls = rng.uniform(-2, 1, 5000), rng.uniform(0, 3, 5000)
    a = sur.predict(10 ** la, 10 ** ls, 0.8, 1.4, 0.1)
    b = sur.predict_log10(la, ls, 0.8, 1.4, 0.1)
    assert set(b) == {"R_d", "T_d"}
    assert np.max(np.abs(sur.mlp.seen[0] - sur.mlp.seen[1])) <= 2 ** -22
    for k in b:
        assert np.allclose(a[k], b[k], rtol=0, atol=1e-6)
