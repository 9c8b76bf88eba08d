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


class NoisyStreamEngine(AnalyticEngine):
    """The analytic model plus a smooth offset that depends on each point's
    Philox stream (as MC noise does): runs on different streams find
    different minima, so the plain solver's stream bookkeeping shows."""

    def simulate_batch_array(self, cfg, n, seed=0, streams=None, use_graph=False, **kw):
        s, q, npk = super().simulate_batch_array(cfg, n, seed, streams, use_graph, **kw)
        st = np.asarray(streams, dtype=np.float64)
        f = 1.0 + 0.02 * np.sin(0.7 * st)
        s = s.copy()
        s[:, 1] *= f
        s[:, 3] *= 2.0 - f
        return s, s.copy(), npk


def test_plain_lockstep_equals_sequential_with_runs_and_retries():
    """minimize.o's loop (runs averaged, a result outside the accept window
    retried on a new stream, at most max_retries times per run) gives the
    same solutions run in lock step as one evaluation at a time."""
    eng = NoisyStreamEngine()
    truths = np.array([(0.05, 2.0), (0.5, 20.0), (3.0, 80.0)])
    probs = []
    for a, s in truths:
        t = AnalyticEngine.tallies(inverse.problem(0.5, 0.5).configs(a, s)[0][0])
        probs.append(inverse.problem(t["R_d"], t["T_d"]))
    # a window that part of the first problem's runs miss (its mu_a is poorly
    # determined under the stream-dependent offset)
    window = ((0.1, 20.0), (0.1, 400.0))
    sol, nev, retries = inverse.plain_batch(probs, runs=3, packets=1000, engine=eng, window=window,
                                            max_retries=2)
    # the loop written out: each run retried on its next stream until inside
    # the window or out of retries, the accepted runs averaged
    for p in range(len(probs)):
        acc, n_retry = [], 0
        for r in range(3):
            for att in range(3):
                x, _, _, _ = inverse.solve_plain(probs[p], packets=1000, engine=eng,
                                                 stream=inverse._plain_stream(120_000_000, p, 3, r, 3, att))
                if (0.1 <= x[0] <= 20.0 and 0.1 <= x[1] <= 400.0) or att == 2:
                    break
                n_retry += 1
            acc.append(x)
        assert np.array_equal(np.mean(acc, axis=0), sol[p]) and retries[p] == n_retry
        x, n, _ = inverse.plain_sequential(probs[p], p, len(probs), runs=3, packets=1000, engine=eng,
                                           window=window, max_retries=2)
        assert np.array_equal(x, sol[p]) and n == nev[p]
    assert retries[0] > 0
    # runs on different streams differ, and their mean is the solution
    one, _, _ = inverse.plain_batch(probs, runs=1, packets=1000, engine=eng, window=window, max_retries=2)
    assert not np.array_equal(one, sol)


def test_wilson_and_ratio_intervals():
    lo, hi = inverse.wilson(3, 96)
    assert 0.0106 < lo < 0.0108 and 0.0878 < hi < 0.0880       # 3/96: [1.07 %, 8.79 %]
    assert inverse.wilson(0, 50)[0] == 0.0 and inverse.wilson(50, 50)[1] == 1.0
    lo, hi = inverse.ratio_ci(339, 4096, 35, 512)
    r = (339 / 4096) / (35 / 512)
    # Katz: exp(+-1.96 sqrt(1/339 - 1/4096 + 1/35 - 1/512)) = 0.7149, 1.3988
    assert abs(lo / r - 0.7149) < 1e-3 and abs(hi / r - 1.3988) < 1e-3


def _analytic_problems(eng, truths):
    probs = []
    for a, s in truths:
        t = eng.tallies(inverse.problem(0.5, 0.5).configs(a, s)[0][0])
        probs.append(inverse.problem(t["R_d"], t["T_d"]))
    return probs


def test_check_screens_clear_fits_and_judges_the_rest():
    # exact solutions clear the screen (one launch of every solution at the
    # refinement's packets); solutions 20 % off in mu_s' do not, go to the
    # full check (a second launch of just those) and miss; the one-stage
    # check gives the same verdicts in one launch of every solution
    eng = AnalyticEngine()
    truths = np.array([(0.05, 2.0), (0.5, 20.0), (3.0, 80.0), (0.2, 150.0), (1.0, 10.0)])
    probs = _analytic_problems(eng, truths)
    sol = truths * np.array([[1, 1], [1, 1.2], [1, 1], [1, 0.8], [1, 1]])
    miss, rR, rT, used = inverse.check(probs, sol, 1_000_000, engine=eng, refine_packets=100_000)
    assert eng.calls == [5, 2]
    assert miss.tolist() == [False, True, False, True, False]
    assert used == 5 * 100_000 + 2 * 1_000_000
    assert rR[[0, 2, 4]].max() < 1e-12 and rR[1] > 0.01
    eng.calls.clear()
    miss1, _, _, used1 = inverse.check(probs, sol, 1_000_000, engine=eng, refine_packets=100_000, screen_z=None)
    assert eng.calls == [5] and miss1.tolist() == miss.tolist() and used1 == 5 * 1_000_000


def test_defect_correction_stops_below_the_step_tolerance():
    # with an exact model the steps shrink quadratically: once a solve's
    # next step is below step_tol it runs in no later launch, and it stops
    # within about that step of the truth
    eng = AnalyticEngine()
    truths = np.array([(0.05, 2.0), (0.5, 20.0), (3.0, 80.0)])
    probs = _analytic_problems(eng, truths)
    sol, nev, st = inverse.refine_newton(probs, truths * [1.2, 0.85], AnalyticSurrogate(), iters=8,
                                         packets=1000, engine=eng, s_cut=1e-6, step_tol=1e-3)
    assert st["launches"] < 8 and nev.max() < 8 and nev.min() >= 2
    assert eng.calls == sorted(eng.calls, reverse=True) and eng.calls[0] == 3
    assert np.allclose(sol, truths, rtol=3e-3)
