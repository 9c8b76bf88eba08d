"""Inverse problem for a layer stack with several unknown optical properties.

The reference's solvers fix a single layer (CUDA/ISMLMC-minimization.cu:233,
CUDA/ISMLMC-lookup.cu:735) although its forward model (ISMLMC, CUDA/
CUDAISMLMC.h:656) takes a stack.  Here any mu_a or mu_s' of any turbid
layer can be unknown; PRAXIS minimizes the Monte Carlo cost in as many
dimensions (log coordinates), and every pending evaluation of every solve
goes into one GPU launch.

A stack of two or more unknown layers is not determined by one (R_d, T_d)
pair, so a problem can use up to three measurement set-ups:
  front   both spheres, light entering the first layer: R_d, T_d
  back    both spheres, the stack reversed (sample turned over): R_d_back, T_d_back
  bare    no spheres, light entering the first layer: R_s, T_u (T_u fixes the
          total attenuation sum(mu_t t))
The cost extends the reference's (CUDA/ISMLMC-minimization.cu:119-129):
relative errors of the diffuse readings, absolute errors of R_s and T_u,
over the quantities measured.
"""
import math
from dataclasses import dataclass, replace
from typing import Dict, List, Optional, Sequence

import numpy as np

from . import praxis as _praxis
from .config import Layer, SampleConfig, Sphere, mus_from_reduced
from .engine import Engine, default_engine
from .minimization import PENALTY, REL_FLOOR

RELATIVE = ("R_d", "T_d", "R_d_back", "T_d_back")
ABSOLUTE = ("R_s", "T_u")
QUANTITIES = RELATIVE + ABSOLUTE


@dataclass(frozen=True)
class Unknown:
    layer: int
    quantity: str        # "mu_a" or "mus_reduced"

    def __post_init__(self):
        if self.quantity not in ("mu_a", "mus_reduced"):
            raise ValueError("unknown quantity %r (mu_a or mus_reduced)" % self.quantity)


class LayeredProblem:
    """A known stack `template` whose `unknowns` are to be recovered from
    `measured` (a dict over QUANTITIES)."""

    def __init__(self, template: SampleConfig, unknowns: Sequence[Unknown], measured: Dict[str, float],
                 box=(1e-4, 1e4)):
        self.template = template
        self.unknowns = list(unknowns)
        self.measured = {k: float(v) for k, v in measured.items() if v is not None}
        bad = set(self.measured) - set(QUANTITIES)
        if bad:
            raise ValueError("unknown measured quantities %s" % sorted(bad))
        if len(self.unknowns) < 1 or len(self.measured) < 1:
            raise ValueError("need at least one unknown and one measurement")
        for u in self.unknowns:
            if not (0 <= u.layer < len(template.layers)) or template.layers[u.layer].mu_s == 0.0:
                raise ValueError("layer %d is not a turbid layer of the stack" % u.layer)
        self.box = (math.log(box[0]), math.log(box[1]))

    def stack(self, x) -> List[Layer]:
        """The layers with the unknowns set to x (linear units)."""
        layers = [replace(L) for L in self.template.layers]
        for u, v in zip(self.unknowns, x):
            L = layers[u.layer]
            if u.quantity == "mu_a":
                layers[u.layer] = replace(L, mu_a=float(v), mu_t=None)
            else:
                layers[u.layer] = replace(L, mu_s=mus_from_reduced(float(v), L.g), mu_t=None)
        return layers

    def truth_of(self, layers: Sequence[Layer]):
        """The unknowns' values in a given stack (for synthetic tests)."""
        out = []
        for u in self.unknowns:
            L = layers[u.layer]
            out.append(L.mu_a if u.quantity == "mu_a" else L.mu_s * (1 - L.g))
        return np.array(out)

    def configs(self, x):
        """The simulations one evaluation needs, and the (config index, tally)
        feeding each measured quantity."""
        t = self.template
        layers = self.stack(x)
        mk = lambda L, R, T: replace(t, layers=list(L), R=R, T=T)
        no = lambda S: Sphere(**{**vars(S), "present": False})
        cfgs, src = [], {}
        m = self.measured
        if "R_d" in m or "T_d" in m:
            src["R_d"], src["T_d"] = (len(cfgs), "R_d"), (len(cfgs), "T_d")
            cfgs.append(mk(layers, t.R, t.T))
        if "R_d_back" in m or "T_d_back" in m:
            src["R_d_back"], src["T_d_back"] = (len(cfgs), "R_d"), (len(cfgs), "T_d")
            cfgs.append(mk(layers[::-1], t.R, t.T))
        if "R_s" in m or "T_u" in m:
            src["R_s"], src["T_u"] = (len(cfgs), "R_s"), (len(cfgs), "T_u")
            cfgs.append(mk(layers, no(t.R), no(t.T)))
        return cfgs, src

    def cost_from(self, values):
        c = 0.0
        for q, M in self.measured.items():
            d = values[q] - M
            c += (d / max(M, REL_FLOOR)) ** 2 if q in RELATIVE else d * d
        return math.sqrt(c)


class LayeredCost:
    """MC cost of many (problem, point) pairs, points in log coordinates,
    in one launch; points outside a problem's box cost PENALTY plus their
    distance to it.  Each solve keeps its own Philox streams."""

    def __init__(self, engine: Engine, problems, n_packets, seed=1234, streams=None, use_graph=True):
        self.engine = engine
        self.problems = list(problems)
        self.n_packets = int(n_packets)
        self.seed = seed
        self.streams = list(streams) if streams is not None else list(range(len(self.problems)))
        self.use_graph = use_graph
        self.launches = 0
        self.packets = 0

    def __call__(self, points, ids):
        cfgs, meta, st = [], [], []
        out = np.zeros(len(points))
        for slot, (y, i) in enumerate(zip(points, ids)):
            y = np.asarray(y, dtype=np.float64)
            lo, hi = self.problems[i].box
            over = np.maximum(lo - y, 0) + np.maximum(y - hi, 0)
            if not np.all(np.isfinite(y)) or np.any(over > 0):
                out[slot] = PENALTY + float(np.sum(np.nan_to_num(over, nan=1e6)))
                continue
            c, src = self.problems[i].configs(np.exp(y))
            meta.append((slot, i, len(cfgs), src))
            cfgs += c
            st += [self.streams[i] * 4 + k for k in range(len(c))]
        if cfgs:
            res = self.engine.simulate_batch(cfgs, self.n_packets, seed=self.seed, streams=st,
                                             use_graph=self.use_graph)
            self.launches += 1
            self.packets += self.n_packets * len(cfgs)
            for slot, i, base, src in meta:
                vals = {q: res[base + k].mean[tally] for q, (k, tally) in src.items()}
                out[slot] = self.problems[i].cost_from(vals)
        return out


def solve(problems: Sequence[LayeredProblem], guesses, packets=1_000_000, step=0.5, tol=1e-3,
          max_evals=2000, restarts=0, restart_rtol=None, seed=1234, stream_base=0,
          engine: Optional[Engine] = None, log=None):
    """PRAXIS per problem in log coordinates from `guesses` (linear units),
    all problems in lock step.  With `restarts`, PRAXIS starts again from
    its own result (fresh directions, step `step`, the same cost surface)
    up to that many times: in three or more dimensions a long narrow valley
    of the MC cost can stop it early.  With `restart_rtol`, a problem is
    started again only while its last run lowered its cost by more than
    that fraction, so restarts go on while they make progress and stop
    where they do not.  Returns (solutions [P, n] linear, costs,
    evaluations per solve, stats)."""
    eng = engine or default_engine()
    if len({len(p.unknowns) for p in problems}) != 1:
        raise ValueError("all problems of one batch need the same number of unknowns")
    cost = LayeredCost(eng, problems, packets, seed=seed,
                       streams=np.arange(len(problems)) + stream_base)
    x = [np.log(np.asarray(g, dtype=np.float64)) for g in guesses]
    f = np.full(len(problems), np.inf)
    on_round = None
    if log:
        on_round = lambda r, pending: (log("    round %d: %d solves pending" % (r, pending))
                                       if r % 20 == 0 else None)
    nev = np.zeros(len(problems), dtype=np.int64)
    active = list(range(len(problems)))
    for _ in range(restarts + 1):
        if not active:
            break
        # PRAXIS numbers the runs it is given 0..k-1; the cost needs the problems'
        run = list(active)
        res = _praxis.minimize_batch([x[i] for i in run], lambda pts, ids: cost(pts, [run[k] for k in ids]),
                                     tol, step, max_evals=max_evals, on_round=on_round)
        active = []
        for i, (xi, fi, ni) in zip(run, res):
            nev[i] += ni
            if restart_rtol is None or fi < f[i] * (1.0 - restart_rtol):
                active.append(i)
            x[i], f[i] = xi, fi
    sol = np.array([np.exp(xi) for xi in x])
    return sol, f.copy(), nev, {"launches": cost.launches, "packets": cost.packets}


def measure(problem: LayeredProblem, layers: Sequence[Layer], packets=10_000_000, seed=99, stream=0,
            engine: Optional[Engine] = None):
    """Synthetic measurement of a known stack: the readings `problem`
    measures, simulated with `packets` photons (for tests and studies)."""
    eng = engine or default_engine()
    truth = problem.truth_of(layers)
    cfgs, src = problem.configs(truth)
    # the stack as given (the unknowns' values are the stack's own)
    res = eng.simulate_batch(cfgs, packets, seed=seed, streams=[stream * 4 + k for k in range(len(cfgs))])
    return {q: res[k].mean[t] for q, (k, t) in src.items() if q in problem.measured}
