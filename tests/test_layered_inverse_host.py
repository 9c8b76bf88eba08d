"""Multilayer inverse host logic (CPU) with an analytic stand-in for the
transport whose readings depend on the order of the layers."""
import math

import numpy as np
import pytest

from ismc_amd import layered_inverse as LI
from ismc_amd.config import Layer, SampleConfig
from ismc_amd.engine import Result
from ismc_amd.workloads import standard_sphere


class AnalyticStackEngine:
    def __init__(self):
        self.launches = []

    @staticmethod
    def tallies(cfg):
        w = [0.6, 0.3, 0.1]
        Rd = Td_exp = Tu_exp = 0.0
        for k, L in enumerate(cfg.layers):
            s = L.mu_s * (1 - L.g)
            Rd += w[k] * s / (s + 3 * L.mu_a + 2)
            Td_exp += (0.3 * L.mu_a + 0.02 * s) * L.t / 0.1
            Tu_exp += (L.mu_a + L.mu_s) * L.t
        sph = cfg.R.present or cfg.T.present
        return {"A": 0.0, "R_d": Rd if sph else 0.0, "R_s": 0.04, "T_d": 0.6 * math.exp(-Td_exp) if sph else 0.0,
                "T_u": math.exp(-Tu_exp)}

    def simulate_batch(self, cfgs, n, seed=0, streams=None, use_graph=False, **kw):
        self.launches.append(len(cfgs))
        out = []
        for c in cfgs:
            t = self.tallies(c)
            out.append(Result(n_packets=int(n), mean=t, se={k: 0 for k in t}, sum=t, sumsq=t, events={},
                              kernel_ms=0.0))
        return out


def stack(a0=0.5, s0=8.0, a1=2.0, s1=15.0):
    g = 0.8
    return SampleConfig(layers=[Layer(n=1.4, g=g, t=0.05, mu_a=a0, mu_s=s0 / (1 - g)),
                                Layer(n=1.4, g=g, t=0.1, mu_a=a1, mu_s=s1 / (1 - g))],
                        R=standard_sphere(), T=standard_sphere())


UNK = [LI.Unknown(0, "mu_a"), LI.Unknown(1, "mu_a"), LI.Unknown(1, "mus_reduced")]


def test_configs_and_measurement_setups():
    p = LI.LayeredProblem(stack(), UNK, {"R_d": 0.3, "T_d": 0.2, "R_d_back": 0.25, "T_u": 1e-4})
    cfgs, src = p.configs([0.5, 2.0, 15.0])
    assert len(cfgs) == 3
    assert src["R_d_back"] == (1, "R_d") and src["T_u"] == (2, "T_u")
    assert cfgs[1].layers[0].t == 0.1 and cfgs[1].layers[1].t == 0.05      # reversed
    assert not cfgs[2].R.present and not cfgs[2].T.present
    assert cfgs[0].layers[1].mu_s == pytest.approx(15.0 / 0.2)
    assert np.allclose(p.truth_of(stack().layers), [0.5, 2.0, 15.0])
    with pytest.raises(ValueError):
        LI.LayeredProblem(stack(), UNK, {"R_total": 0.3})
    with pytest.raises(ValueError):
        LI.Unknown(0, "g")


def test_three_unknowns_recovered_in_one_batch():
    eng = AnalyticStackEngine()
    truths = [stack(), stack(0.2, 8.0, 4.0, 30.0), stack(1.5, 8.0, 0.7, 9.0)]
    probs = []
    for tr in truths:
        p = LI.LayeredProblem(tr, UNK, {"R_d": 0, "T_d": 0, "R_d_back": 0, "T_u": 0})
        p.measured = LI.measure(p, tr.layers, engine=eng)
        probs.append(p)
    guesses = [p.truth_of(tr.layers) * [1.4, 0.7, 1.3] for p, tr in zip(probs, truths)]
    eng.launches = []
    sol, cost, nev, st = LI.solve(probs, guesses, packets=1000, step=0.3, tol=1e-6, engine=eng)
    for p, tr, x in zip(probs, truths, sol):
        assert np.allclose(x, p.truth_of(tr.layers), rtol=1e-3)
    # every round of pending evaluations is one launch
    assert st["launches"] == len(eng.launches) < nev.sum()
    assert max(eng.launches) <= 3 * len(probs)


def test_restarts_continue_only_while_they_make_progress():
    # Loose runs (tol 1e-2) stop short on the analytic readings; restarts
    # from their own result refine them.  With restart_rtol a problem is
    # restarted only while its last run lowered its cost by > that fraction,
    # so the problems leave the batch at different restarts: the cost must
    # still see each point under its own problem's id.
    eng = AnalyticStackEngine()
    truths = [stack(), stack(0.2, 8.0, 4.0, 30.0), stack(1.5, 8.0, 0.7, 9.0)]
    probs = []
    for tr in truths:
        p = LI.LayeredProblem(tr, UNK, {"R_d": 0, "T_d": 0, "R_d_back": 0, "T_u": 0})
        p.measured = LI.measure(p, tr.layers, engine=eng)
        probs.append(p)
    starts = [p.truth_of(tr.layers) * f for p, tr, f in zip(probs, truths, ([1.4, 0.7, 1.3], [1.02, 1.0, 1.0],
                                                                             [0.6, 1.5, 1.8]))]
    one = LI.solve(probs, starts, packets=1000, step=0.3, tol=1e-2, engine=eng)
    many = LI.solve(probs, starts, packets=1000, step=0.3, tol=1e-2, restarts=10, restart_rtol=0.1, engine=eng)
    fixed = LI.solve(probs, starts, packets=1000, step=0.3, tol=1e-2, restarts=10, engine=eng)
    assert np.all(many[1] <= one[1])
    assert np.all(many[2] > one[2]) and many[2].sum() < fixed[2].sum()
    # problems leave at different restarts (evaluation counts differ)
    assert len(set(many[2].tolist())) > 1
    # each reported cost is the cost of its own solution
    cost = LI.LayeredCost(eng, probs, 1000, streams=np.arange(len(probs)))
    again = cost([np.log(x) for x in many[0]], list(range(len(probs))))
    assert np.allclose(again, many[1])
