"""Inverse solver: recover (mu_a, mu_s') of a single-layer sample from
measured sphere readings by minimizing a Monte Carlo cost with PRAXIS.

cost()            the reference cost (CUDA/ISMLMC-minimization.cu:89-131,
                  Minimization/minimization.py:201-238): the Euclidean norm
                  of relative R_d / T_d errors and absolute R_s / T_u errors,
                  over the quantities that were measured.  With single-sphere
                  (separate) measurements each evaluation runs up to three
                  configurations (R sphere only, T sphere only, no sphere).
solve()           minimize.o's run loop (ISMLMC-minimization.cu:457-516):
                  `runs` PRAXIS minimizations from the same guess, results
                  outside the accept window are retried (a bounded number of
                  times, the reference loops for ever), mean and standard
                  error of the accepted results.
solve_batch()     many independent problems at once: one PRAXIS generator
                  per problem and run, all advancing in lock step so that
                  every round of pending evaluations is one batched GPU
                  launch.  Each solver keeps its own Philox stream (common
                  random numbers), so its cost surface is a fixed function.
run_minimization() minimization.py's outer loop over input-uncertainty
                  samples (:132-283), every sample and run in the same batch.
"""
import math
from dataclasses import dataclass, field
from typing import List, Optional, Sequence

import numpy as np

from . import praxis as _praxis
from .config import Layer, SampleConfig, Sphere, mus_from_reduced
from .engine import CONFIG_DTYPE, Engine, default_engine

PENALTY = 1.0e3


@dataclass
class Measured:
    """Measured sphere readings; None marks a quantity that was not measured."""
    R_d: Optional[float] = None
    R_s: Optional[float] = None
    T_d: Optional[float] = None
    T_u: Optional[float] = None


@dataclass
class Problem:
    """A single-layer sample with known n, g, t between known spheres."""
    n: float
    g: float
    t: float
    R: Sphere
    T: Sphere
    measured: Measured
    specular_included: bool = False
    separate: bool = False
    roulette_threshold: float = 1e-4

    def configs(self, mu_a, mus_reduced):
        """The simulations one cost evaluation needs, and which tally of
        which of them feeds each measured quantity."""
        mu_s = mus_from_reduced(mus_reduced, self.g)
        L = [Layer(n=self.n, g=self.g, t=self.t, mu_a=mu_a, mu_s=mu_s)]
        mk = lambda R, T: SampleConfig(layers=L, R=R, T=T, specular_included=self.specular_included,
                                       roulette_threshold=self.roulette_threshold)
        m = self.measured
        if not self.separate:
            return [mk(self.R, self.T)], {"R_d": (0, "R_d"), "R_s": (0, "R_s"),
                                           "T_d": (0, "T_d"), "T_u": (0, "T_u")}
        cfgs, src = [], {}
        if m.R_d is not None:
            src["R_d"] = (len(cfgs), "R_d")
            cfgs.append(mk(Sphere(**{**vars(self.R), "present": True}),
                           Sphere(**{**vars(self.T), "present": False})))
        if m.T_d is not None:
            src["T_d"] = (len(cfgs), "T_d")
            cfgs.append(mk(Sphere(**{**vars(self.R), "present": False}),
                           Sphere(**{**vars(self.T), "present": True})))
        if m.R_s is not None or m.T_u is not None:
            src["R_s"] = (len(cfgs), "R_s")
            src["T_u"] = (len(cfgs), "T_u")
            cfgs.append(mk(Sphere(**{**vars(self.R), "present": False}),
                           Sphere(**{**vars(self.T), "present": False})))
        return cfgs, src

    def template_key(self):
        """Everything of a (non-separate) evaluation's configuration except
        mu_a and mu_s."""
        sph = lambda S: (S.present, S.pw, S.fs, S.fp, S.fd, S.f, S.angle)
        return (self.n, self.g, self.t, sph(self.R), sph(self.T), self.specular_included,
                self.roulette_threshold)


_TEMPLATES = {}


def _template(problem: Problem):
    """The C configuration of a problem's evaluations with mu_a = mu_s = 0
    (numpy record of CONFIG_DTYPE), cached per template key."""
    key = problem.template_key()
    rec = _TEMPLATES.get(key)
    if rec is None:
        cfgs, _ = problem.configs(0.0, 0.0)
        rec = np.frombuffer(bytes(cfgs[0].to_c()), dtype=CONFIG_DTYPE)[0].copy()
        if len(_TEMPLATES) > 64:
            _TEMPLATES.clear()
        _TEMPLATES[key] = rec
    return rec


def config_array(problems: Sequence["Problem"], mu_a, mus_reduced, ids=None, _cache=None):
    """The C configurations (CONFIG_DTYPE array) of non-separate problems'
    evaluations at (mu_a, mu_s'), one per problem (or per entry of `ids`,
    indices into `problems`, with mu_a and mus_reduced given per entry),
    patched into the cached templates: byte for byte what
    problem.configs(mu_a, mu_s')[0][0].to_c() gives, without building the
    objects.  `_cache` (a dict) keeps the problems' template index between
    calls on the same problem list."""
    if _cache is not None and _cache.get("problems") is problems:
        tpl, index, g = _cache["tpl"], _cache["index"], _cache["g"]
    else:
        keys, recs, idx = {}, [], []
        for p in problems:
            if p.separate:
                raise ValueError("config_array takes non-separate problems only")
            k = p.template_key()
            if k not in keys:
                keys[k] = len(recs)
                recs.append(_template(p))
            idx.append(keys[k])
        tpl = np.array(recs, dtype=CONFIG_DTYPE)
        index = np.array(idx, dtype=np.int64)
        g = np.array([p.g for p in problems], dtype=np.float64)
        if _cache is not None:
            _cache.update(problems=problems, tpl=tpl, index=index, g=g)
    if ids is not None:
        ids = np.asarray(ids, dtype=np.int64)
        index, g = index[ids], g[ids]
    mu_a = np.asarray(mu_a, dtype=np.float64)
    musr = np.asarray(mus_reduced, dtype=np.float64)
    cfg = tpl[index]
    # mus_from_reduced, elementwise
    mu_s = np.where(g == 1, musr, musr / np.where(g == 1, 1.0, 1.0 - g))
    cfg["mu_a"][:, 0] = mu_a
    cfg["mu_s"][:, 0] = mu_s
    cfg["mu_t"][:, 0] = mu_a + mu_s
    return cfg


REL_FLOOR = 1e-6


def cost_from(values, measured: Measured):
    """sqrt(((R_d - M_R)/M_R)^2 + (R_s - M_S)^2 + ((T_d - M_T)/M_T)^2 + (T_u - M_U)^2)
    over the measured terms (ISMLMC-minimization.cu:119-129).  The relative
    terms divide by max(M, 1e-6): a measured 0 (a thick absorbing sample)
    would divide by zero in the reference."""
    c = 0.0
    if measured.R_d is not None:
        c += ((values["R_d"] - measured.R_d) / max(measured.R_d, REL_FLOOR)) ** 2
    if measured.R_s is not None:
        c += (values["R_s"] - measured.R_s) ** 2
    if measured.T_d is not None:
        c += ((values["T_d"] - measured.T_d) / max(measured.T_d, REL_FLOOR)) ** 2
    if measured.T_u is not None:
        c += (values["T_u"] - measured.T_u) ** 2
    return math.sqrt(c)


def feasible(x):
    return np.all(np.isfinite(x)) and x[0] >= 0.0 and x[1] > 0.0


class BatchCost:
    """Evaluates the MC cost of many (problem, point) pairs in one launch.
    Infeasible points (mu_a < 0 or mu_s' <= 0, where the transport is
    undefined) get PENALTY plus their distance to the feasible set."""

    def __init__(self, engine: Engine, problems: Sequence[Problem], n_packets, seed=1234,
                 streams=None, use_graph=True):
        self.engine = engine
        self.problems = list(problems)
        self.n_packets = int(n_packets)
        self.seed = seed
        self.streams = list(streams) if streams is not None else list(range(len(self.problems)))
        self.use_graph = use_graph
        self.launch = {}  # launch-shape knobs (idle_lanes, bound_lanes) for this cost's launches
        self.evaluations = 0
        self.launches = 0
        self.packets = 0
        self.kernel_ms = 0.0

    def _measured_arrays(self):
        if getattr(self, "_meas", None) is None:
            nan = float("nan")
            get = lambda q: np.array([getattr(p.measured, q) if getattr(p.measured, q) is not None else nan
                                      for p in self.problems], dtype=np.float64)
            self._meas = {q: get(q) for q in ("R_d", "R_s", "T_d", "T_u")}
            self._plain = all(not p.separate for p in self.problems)
            self._cfg_cache = {}
        return self._meas

    def _eval_feasible(self, X, ids):
        """Costs of feasible points X [k, 2] (mu_a, mu_s') of problems ids:
        one launch, configurations patched into cached C templates, the
        reference cost formula on arrays (same operations in the same order
        as cost_from, so the same values)."""
        ids = np.asarray(ids, dtype=np.int64)
        k = len(ids)
        cfg = config_array(self.problems, X[:, 0], X[:, 1], ids, self._cfg_cache)
        st = np.asarray(self.streams, dtype=np.int64)[ids] * 4
        sums, _, npk = self.engine.simulate_batch_array(cfg, self.n_packets, seed=self.seed, streams=st,
                                                        use_graph=self.use_graph, **self.launch)
        self.launches += 1
        self.packets += self.n_packets * k
        self.kernel_ms += self.engine.last_launch()["kernel_ms"]
        mean = sums / npk[:, None].astype(np.float64)
        m = self._meas
        c = np.zeros(k)
        for q, col, rel in (("R_d", 1, True), ("R_s", 2, False), ("T_d", 3, True), ("T_u", 4, False)):
            M = m[q][ids]
            has = ~np.isnan(M)
            d = mean[:, col] - M
            term = d / np.maximum(M, REL_FLOOR) if rel else d
            c = c + np.where(has, term * term, 0.0)
        return np.sqrt(c)

    def __call__(self, points, ids):
        self._measured_arrays()
        if self._plain and len(points):
            X = np.asarray(points, dtype=np.float64).reshape(len(points), 2)
            ok = np.all(np.isfinite(X), axis=1) & (X[:, 0] >= 0.0) & (X[:, 1] > 0.0)
            out = np.empty(len(points))
            Xc = np.where(np.isfinite(X), X, -np.inf)
            out[~ok] = PENALTY + np.sqrt(np.sum(np.minimum(Xc[~ok], 0.0) ** 2, axis=1))
            self.evaluations += len(points)
            if ok.any():
                out[ok] = self._eval_feasible(X[ok], np.asarray(ids)[ok])
            return out
        cfgs, meta, st = [], [], []
        out = np.zeros(len(points))
        for slot, (x, i) in enumerate(zip(points, ids)):
            x = np.asarray(x, dtype=np.float64)
            if not feasible(x):
                out[slot] = PENALTY + float(np.sum(np.minimum(x, 0.0) ** 2)) ** 0.5
                continue
            c, src = self.problems[i].configs(float(x[0]), float(x[1]))
            meta.append((slot, i, len(cfgs), src))
            cfgs += c
            st += [self.streams[i] * 4 + k for k in range(len(c))]
        self.evaluations += len(points)
        if cfgs:
            res = self.engine.simulate_batch(cfgs, self.n_packets, seed=self.seed, streams=st,
                                             use_graph=self.use_graph)
            self.launches += 1
            self.packets += self.n_packets * len(cfgs)
            self.kernel_ms += res[0].kernel_ms
            for slot, i, base, src in meta:
                vals = {q: res[base + k].mean[t] for q, (k, t) in src.items()}
                out[slot] = cost_from(vals, self.problems[i].measured)
        return out


@dataclass
class SolveResult:
    mu_a: float
    mu_a_stderr: float
    mus_reduced: float
    mus_reduced_stderr: float
    runs_mu_a: List[float] = field(default_factory=list)
    runs_mus_reduced: List[float] = field(default_factory=list)
    evaluations: int = 0
    retries: int = 0

    def stdout(self):
        """minimize.o's four-line output (ISMLMC-minimization.cu:516)."""
        return "%f\n%f\n%f\n%f\n" % (self.mu_a, self.mu_a_stderr, self.mus_reduced,
                                     self.mus_reduced_stderr)


def _stats(vals):
    v = np.asarray(vals, dtype=np.float64)
    m = float(v.mean())
    se = math.sqrt(float(((v - m) ** 2).sum()) / v.size) / math.sqrt(v.size)
    return m, se


def solve_batch(problems: Sequence[Problem], guesses, n_packets=10000, step=5.0, tol=0.1,
                runs=1, window=((0.0, 1e4), (0.0, 1e4)), max_retries=5, seed=1234,
                engine: Optional[Engine] = None, max_evals=2000, stream_base=0):
    """PRAXIS for every (problem, run) at once, one GPU launch per round.
    guesses: one (mu_a, mu_s') per problem.  Returns (results, stats)."""
    eng = engine or default_engine()
    P = len(problems)
    guesses = np.asarray(guesses, dtype=np.float64).reshape(P, 2)
    accepted = [[] for _ in range(P)]
    retries = [0] * P
    evals = [0] * P
    todo = [(p, r) for p in range(P) for r in range(runs)]
    attempt = 0
    stats = {"launches": 0, "packets": 0, "kernel_ms": 0.0, "evaluations": 0, "rounds": 0}
    while todo:
        # one stream per (problem, run, attempt): common random numbers within a solve
        inst = [problems[p] for p, _ in todo]
        streams = [stream_base + (p * max(runs, 1) + r) * (max_retries + 1) + attempt for p, r in todo]
        cost = BatchCost(eng, inst, n_packets, seed=seed, streams=streams)
        res = _praxis.minimize_batch([guesses[p] for p, _ in todo], cost, tol, step,
                                     max_evals=max_evals)
        stats["launches"] += cost.launches
        stats["packets"] += cost.packets
        stats["kernel_ms"] += cost.kernel_ms
        stats["evaluations"] += cost.evaluations
        nxt = []
        for (p, r), (x, fx, nf) in zip(todo, res):
            evals[p] += nf
            (a0, a1), (s0, s1) = window
            if a0 <= x[0] <= a1 and s0 <= x[1] <= s1:
                accepted[p].append(x)
            elif retries[p] < max_retries:
                retries[p] += 1
                nxt.append((p, r))
            else:
                accepted[p].append(x)   # bounded retries: keep the last result
        todo = nxt
        attempt += 1
    out = []
    for p in range(P):
        xs = np.array(accepted[p])
        ma, sa = _stats(xs[:, 0])
        ms, ss = _stats(xs[:, 1])
        out.append(SolveResult(ma, sa, ms, ss, list(xs[:, 0]), list(xs[:, 1]), evals[p], retries[p]))
    return out, stats


def solve(problem: Problem, guess, **kw):
    """minimize.o for one problem (runs, accept window, mean/stderr)."""
    res, stats = solve_batch([problem], [guess], **kw)
    return res[0]


def solve_sequential(problem: Problem, guess, n_packets=10000, step=5.0, tol=0.1, seed=1234,
                     engine: Optional[Engine] = None, max_evals=2000, stream=0):
    """The reference's execution model on the GPU: one PRAXIS, one MC
    launch per cost evaluation (CUDA/ISMLMC-minimization.cu:476 -> cost() ->
    ISMLMC())."""
    eng = engine or default_engine()
    cost = BatchCost(eng, [problem], n_packets, seed=seed, streams=[stream], use_graph=False)
    x, fx, nf = _praxis.minimize(lambda pt: cost([pt], [0])[0], guess, tol, step,
                                 max_evals=max_evals)
    return x, fx, nf, cost


def problems_from_ini(ini, samples=None, rng_seed=None):
    """minimization.py's input-uncertainty samples (:132-152): n, sphere
    parameters, angles and the measurements drawn from normal
    distributions; one Problem per sample."""
    S = int(samples if samples is not None else ini.minimization.get("input_distributions_samples", 1))
    rng = np.random.default_rng(rng_seed)
    nrm = lambda mu, sd: rng.normal(mu, sd, S)
    n_d = nrm(ini.n, ini.n_unc)
    Rp = [nrm(getattr(ini.R, k), getattr(ini.R, k + "_unc")) for k in ("pw", "fs", "fp", "fd")]
    Tp = [nrm(getattr(ini.T, k), getattr(ini.T, k + "_unc")) for k in ("pw", "fs", "fp", "fd")]
    Ra = nrm(ini.R.angle_deg, ini.R.angle_unc_deg)
    Ta = nrm(ini.T.angle_deg, ini.T.angle_unc_deg)
    meas = {}
    for key, m in (("R_d", ini.M_R), ("R_s", ini.M_S), ("T_d", ini.M_T), ("T_u", ini.M_U)):
        meas[key] = nrm(m.first, m.first_unc) if m is not None else None
    probs = []
    for s in range(S):
        R = ini.R.sphere(Rp[0][s], Rp[1][s], Rp[2][s], Rp[3][s], Ra[s])
        T = ini.T.sphere(Tp[0][s], Tp[1][s], Tp[2][s], Tp[3][s], Ta[s])
        m = Measured(**{k: (float(v[s]) if v is not None else None) for k, v in meas.items()})
        probs.append(Problem(n=float(n_d[s]), g=ini.g, t=ini.t, R=R, T=T, measured=m,
                             specular_included=ini.specular_included, separate=ini.separate))
    return probs


def run_minimization(ini, samples=None, engine: Optional[Engine] = None, seed=1234, rng_seed=None):
    """The whole minimization.py run: every input-uncertainty sample and
    run of PRAXIS in one lock-step batch.  Returns mean and std over samples
    of the per-sample mean (mu_a, mu_s') (minimization.py:273-283)."""
    M = ini.minimization
    probs = problems_from_ini(ini, samples, rng_seed)
    guess = (M["mua_guess"], M["mus_guess"])
    res, stats = solve_batch(probs, [guess] * len(probs), n_packets=int(M["N"]), step=M["step"],
                             tol=M["tolerance"], runs=int(M.get("runs", 1)),
                             window=((M["mua_min"], M["mua_max"]), (M["mus_min"], M["mus_max"])),
                             seed=seed, engine=engine)
    mua = np.array([r.mu_a for r in res])
    mus = np.array([r.mus_reduced for r in res])
    return {"mu_a": float(mua.mean()), "mu_a_std": float(mua.std()),
            "mus_reduced": float(mus.mean()), "mus_reduced_std": float(mus.std()),
            "per_sample": res, "stats": stats}
