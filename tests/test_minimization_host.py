"""Minimization driver host logic (CPU), with a stand-in engine whose
tallies are an analytic function of (mu_a, mu_s') instead of a GPU launch."""
import math
import os

import numpy as np
import pytest

from ismc_amd import ini as ini_mod
from ismc_amd import minimization as M
from ismc_amd.config import Sphere
from ismc_amd.engine import Result

HERE = os.path.dirname(os.path.abspath(__file__))


class AnalyticEngine:
    """R_d, T_d, R_s, T_u as smooth functions of the first layer's mu_a, mu_s'."""

    def __init__(self):
        self.calls = []

    @staticmethod
    def tallies(cfg):
        L = cfg.layers[0]
        a, s = L.mu_a, L.mu_s * (1 - L.g)
        Rd = 0.5 * s / (s + 8 * a + 5)
        Td = 0.6 * math.exp(-0.08 * a - 0.01 * s)
        return {"A": 1 - Rd - Td, "R_d": Rd if cfg.R.present or not cfg.T.present else Rd,
                "R_s": 0.03, "T_d": Td, "T_u": 0.001 * math.exp(-a)}

    def simulate_batch_array(self, cfg, n, seed=0, streams=None, use_graph=False, **kw):
        self.calls.append(len(cfg))
        a, mus, g = cfg["mu_a"][:, 0], cfg["mu_s"][:, 0], cfg["g"][:, 0]
        s = mus * (1 - g)
        Rd = 0.5 * s / (s + 8 * a + 5)
        Td = 0.6 * np.exp(-0.08 * a - 0.01 * s)
        tal = np.stack([1 - Rd - Td, Rd, np.full_like(Rd, 0.03), Td, 0.001 * np.exp(-a)], 1)
        npk = np.full(len(cfg), int(n), dtype=np.uint64)
        return tal * n, tal * n, npk

    def last_launch(self):
        return {"kernel_ms": 0.25, "grid": 1, "block": 256, "chunks": 1}

    def simulate_batch(self, cfgs, n, seed=0, streams=None, use_graph=False, **kw):
        self.calls.append(len(cfgs))
        out = []
        for c in cfgs:
            t = self.tallies(c)
            out.append(Result(n_packets=int(n), mean=t, se={k: 0 for k in t}, sum=t,
                              sumsq=t, events={}, kernel_ms=0.0))
        return out


def problem(measured, separate=False):
    S = Sphere(True, 0.98, 0.0039, 0.0039, 0.001)
    return M.Problem(n=1.4, g=0.8, t=0.1, R=S, T=S, measured=measured, separate=separate)


def test_batch_cost_counts_kernel_time_on_the_array_path():
    eng = AnalyticEngine()
    probs = [problem(M.Measured(R_d=0.3, T_d=0.4)) for _ in range(3)]
    cost = M.BatchCost(eng, probs, 1000)
    cost([(1.0, 10.0), (2.0, 5.0), (0.5, 20.0)], [0, 1, 2])
    cost([(1.0, 11.0)], [1])
    assert cost.launches == 2 and cost.kernel_ms == pytest.approx(0.5)


def test_cost_formula():
    m = M.Measured(R_d=0.3, T_d=0.4, T_u=0.001)
    v = {"R_d": 0.33, "R_s": 0.5, "T_d": 0.38, "T_u": 0.002}
    exp = math.sqrt((0.03 / 0.3) ** 2 + (0.02 / 0.4) ** 2 + 0.001 ** 2)
    assert M.cost_from(v, m) == pytest.approx(exp)
    assert M.cost_from(v, M.Measured(R_s=0.4)) == pytest.approx(0.1)


def test_separate_configs():
    p = problem(M.Measured(R_d=0.3, T_d=0.4, T_u=0.001), separate=True)
    cfgs, src = p.configs(1.0, 10.0)
    assert len(cfgs) == 3
    assert cfgs[0].R.present and not cfgs[0].T.present and src["R_d"] == (0, "R_d")
    assert not cfgs[1].R.present and cfgs[1].T.present and src["T_d"] == (1, "T_d")
    assert not cfgs[2].R.present and not cfgs[2].T.present and src["T_u"] == (2, "T_u")
    assert cfgs[0].layers[0].mu_s == pytest.approx(50.0)
    cfgs, _ = problem(M.Measured(R_d=0.3)).configs(1.0, 10.0)
    assert len(cfgs) == 1 and cfgs[0].R.present and cfgs[0].T.present


def test_infeasible_points_are_penalised():
    eng = AnalyticEngine()
    bc = M.BatchCost(eng, [problem(M.Measured(R_d=0.3))], 1000)
    v = bc([np.array([-1.0, 5.0]), np.array([1.0, 0.0]), np.array([1.0, 10.0])], [0, 0, 0])
    assert v[0] > M.PENALTY and v[1] >= M.PENALTY and v[2] < 1
    assert eng.calls == [1]


def test_recovers_truth_and_batch_equals_sequential():
    eng = AnalyticEngine()
    truths = [(0.5, 8.0), (2.0, 20.0), (1.0, 40.0)]
    probs = []
    for a, s in truths:
        t = eng.tallies(problem(M.Measured()).configs(a, s)[0][0])
        probs.append(problem(M.Measured(R_d=t["R_d"], T_d=t["T_d"])))
    res, stats = M.solve_batch(probs, [(1.0, 15.0)] * 3, n_packets=1000, step=5.0, tol=1e-6,
                               runs=1, engine=eng)
    for (a, s), r in zip(truths, res):
        assert r.mu_a == pytest.approx(a, rel=1e-4) and r.mus_reduced == pytest.approx(s, rel=1e-4)
    # every round is one launch holding all pending evaluations
    assert stats["launches"] == len(eng.calls) and max(eng.calls) == 3
    for p, r in zip(probs, res):
        x, fx, nf, _ = M.solve_sequential(p, (1.0, 15.0), n_packets=1000, step=5.0, tol=1e-6,
                                          engine=AnalyticEngine())
        assert np.allclose(x, [r.mu_a, r.mus_reduced], rtol=1e-12)


def test_accept_window_retries_are_bounded():
    eng = AnalyticEngine()
    t = eng.tallies(problem(M.Measured()).configs(2.0, 20.0)[0][0])
    p = problem(M.Measured(R_d=t["R_d"], T_d=t["T_d"]))
    res, _ = M.solve_batch([p], [(1.0, 15.0)], n_packets=1000, tol=1e-5, runs=2,
                           window=((5.0, 6.0), (0.0, 100.0)), max_retries=3, engine=eng)
    assert res[0].retries == 3 and len(res[0].runs_mu_a) == 2
    out = res[0].stdout().split()
    assert len(out) == 4 and float(out[0]) == pytest.approx(res[0].mu_a, abs=1e-6)


def test_problems_from_ini():
    f = ini_mod.load(os.path.join(HERE, "golden", "sample_lookup.ini"))
    probs = M.problems_from_ini(f, samples=5, rng_seed=1)
    assert len(probs) == 5
    assert all(p.measured.R_d is not None and p.measured.T_d is not None and p.measured.R_s is None
               for p in probs)
    assert np.std([p.n for p in probs]) > 0


def test_config_array_matches_config_objects():
    """The template-patched C configurations are byte for byte those of
    Problem.configs(...).to_c(), for several templates and a subset."""
    from ismc_amd.engine import CONFIG_DTYPE
    rng = np.random.default_rng(3)
    S1 = Sphere(True, 0.98, 0.0039, 0.0039, 0.001)
    S2 = Sphere(True, 0.95, 0.01, 0.002, 0.003)
    probs = [M.Problem(n=n, g=g, t=t, R=R, T=T, measured=M.Measured(R_d=0.2, T_d=0.3))
             for n, g, t, R, T in [(1.4, 0.8, 0.1, S1, S1), (1.33, 0.9, 0.2, S2, S1), (1.5, 1.0, 0.05, S1, S2),
                                   (1.4, 0.8, 0.1, S1, S1), (1.37, 0.0, 0.3, S2, S2)]]
    a = 10 ** rng.uniform(-2, 1, len(probs))
    s = 10 ** rng.uniform(0, 2, len(probs))
    arr = M.config_array(probs, a, s)
    for i, p in enumerate(probs):
        ref = np.frombuffer(bytes(p.configs(float(a[i]), float(s[i]))[0][0].to_c()), dtype=CONFIG_DTYPE)[0]
        assert arr[i].tobytes() == ref.tobytes()
    ids = np.array([4, 1, 1])
    sub = M.config_array(probs, a[ids], s[ids], ids=ids)
    for j, i in enumerate(ids):
        assert sub[j].tobytes() == arr[i].tobytes()
