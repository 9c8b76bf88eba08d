"""Batched inverse solving with a learned seed.

Pipeline for many measurements (R_d, T_d of a single-layer sample with
known g, n, t between the standard spheres):

1. seed      the surrogate is evaluated once on a dense (mu_a, mu_s) grid at
             the sample's (g, n, t), then on a fine local grid around each
             measurement's best point; each measurement takes the grid point
             of least relative misfit;
2. refine    defect correction (`refine_newton`): Gauss-Newton steps on the
             Monte Carlo residuals with the surrogate's Jacobian, every
             solve's MC in one transport launch per iteration, each solve on
             its own Philox stream (common random numbers); or PRAXIS on the
             MC cost from the seed (`refine_batch`);
3. judge     one independent MC run per solution at `check_packets`; a
             solution whose R_d or T_d misses the measurement by more than
             the combined MC error is re-solved from where it stands by
             Gauss-Newton on the MC residuals alone, the Jacobian from MC
             differences (`refine_mc_newton`, the fallback).

The plain method (`solve_plain`) is the reference's: sequential PRAXIS in
(mu_a, mu_s') with one MC launch per cost evaluation
(CUDA/ISMLMC-minimization.cu:467-488).
"""
import math
import time
from dataclasses import dataclass, field
from typing import Optional

import numpy as np

from . import minimization as M
from . import praxis as _praxis
from .config import mus_from_reduced
from .engine import TALLIES, Engine, default_engine
from .workloads import standard_sphere

G, N_INDEX, THICK = 0.8, 1.40, 0.1


def problem(R_d, T_d, g=G, n=N_INDEX, t=THICK):
    return M.Problem(n=n, g=g, t=t, R=standard_sphere(), T=standard_sphere(),
                     measured=M.Measured(R_d=float(R_d), T_d=float(T_d)))


_RT = [TALLIES.index("R_d"), TALLIES.index("T_d")]

# Launch shape for this module's sample family (one bare turbid layer):
# packets hit an interface about twice per history, so lanes waiting at one
# are served after 8 of them gather rather than 16.  Measured on MI355X,
# 4096 points x 1e5 packets: 86.2 ms against 90.6 ms with the library's
# defaults of the time (8 idle, 16 bound lanes) and 87.0 with 4 bound lanes; 4096 x
# 1e6: 773 against 800 and 779 ms.  Every solver here uses it, the plain
# one included.  Tallies do not depend on it.
LAUNCH = {"idle_lanes": 8, "bound_lanes": 8}


def _last_ms(eng):
    last = getattr(eng, "last_launch", None)
    return float(last()["kernel_ms"]) if last is not None else 0.0


def mc_rt(eng, problems, mua, musr, packets, seed, streams, ids=None, cache=None, use_graph=True):
    """Mean and standard error ([k, 2] each, columns R_d, T_d) of each
    problem's set-up at (mu_a, mu_s'), or of problems[ids]: one launch, the
    configurations patched into cached C templates (the same bytes as
    problem.configs()), the standard error as Result.se computes it."""
    cfg = M.config_array(problems, mua, musr, ids=ids, _cache=cache)
    sums, sq, npk = eng.simulate_batch_array(cfg, packets, seed=seed, streams=streams, use_graph=use_graph,
                                             **LAUNCH)
    n = npk.astype(np.float64)[:, None]
    mean = sums / n
    var = np.maximum(sq / n - mean ** 2, 0.0)
    se = np.sqrt(var / np.maximum(n - 1, 1))
    return mean[:, _RT], se[:, _RT]


def make_measurements(count, packets=1_000_000, noise=0.01, seed=0, engine: Optional[Engine] = None,
                      g=G, n=N_INDEX, t=THICK):
    """Ground truths log-uniform in mu_a 0.01..10 /cm, mu_s 1..1000 /cm;
    sphere readings simulated with `packets` photons, then multiplied by
    (1 + noise * N(0, 1))."""
    eng = engine or default_engine()
    rng = np.random.default_rng(seed)
    mua = 10 ** rng.uniform(-2, 1, count)
    mus = 10 ** rng.uniform(0, 3, count)
    musr = mus * (1 - g)
    base = problem(0.5, 0.5, g, n, t)
    cfgs = [base.configs(a, s)[0][0] for a, s in zip(mua, musr)]
    res = eng.simulate_batch(cfgs, packets, seed=seed + 101, streams=np.arange(count, dtype=np.uint32) + 50_000_000)
    Rd = np.array([r.mean["R_d"] for r in res])
    Td = np.array([r.mean["T_d"] for r in res])
    Rn = Rd * (1 + noise * rng.standard_normal(count))
    Tn = Td * (1 + noise * rng.standard_normal(count))
    return {"mu_a": mua, "mus_reduced": musr, "R_d_true": Rd, "T_d_true": Td, "R_d": Rn, "T_d": Tn,
            "kernel_ms": res[0].kernel_ms}


def grid_argmin(pR, pT, R, T, device=0):
    """For each measurement (R[m], T[m]) the index of the grid cell with the
    least relative misfit ((pR - R)/R)^2 + ((pT - T)/max(T, 1e-6))^2, on the
    GPU (ismc_grid_argmin): numpy's argmin of that cost over the grid, bit
    for bit (FP64, first index on ties)."""
    from . import _lib
    L = _lib.load()
    pR = np.ascontiguousarray(pR, dtype=np.float32)
    pT = np.ascontiguousarray(pT, dtype=np.float32)
    R = np.ascontiguousarray(R, dtype=np.float64)
    T = np.ascontiguousarray(T, dtype=np.float64)
    if pR.shape != pT.shape or R.shape != T.shape or pR.ndim != 1 or R.ndim != 1:
        raise ValueError("pR, pT and R, T must be 1-d arrays of equal length")
    out = np.zeros(R.size, dtype=np.int32)
    rc = L.ismc_grid_argmin(int(device), pR.ctypes.data, pT.ctypes.data, int(pR.size), R.ctypes.data,
                            T.ctypes.data, int(R.size), out.ctypes.data)
    if rc != 0:
        raise RuntimeError("ismc_grid_argmin: %s" % L.ismc_table_last_error().decode())
    return out


def surrogate_seeds(sur, R_d, T_d, g=G, n=N_INDEX, t=THICK, grid=128, fine=16, polish=False, argmin=None):
    """Initial (mu_a, mu_s') for each measurement from the surrogate: the
    best of a grid x grid log grid over mu_a 0.01..10 /cm, mu_s 1..1000 /cm
    (every measurement against every cell: a global search, so a seed lands
    on the branch of (R_d, T_d) nearest the measurement), then the best of a
    fine x fine grid spanning the neighbouring cells around it; optionally
    polished by PRAXIS on the surrogate.  The global search runs on the GPU
    (`grid_argmin`); `argmin`, a function of the same signature, replaces
    it in host-only tests of this logic."""
    R_d = np.asarray(R_d, dtype=np.float64)
    T_d = np.asarray(T_d, dtype=np.float64)
    Tf = np.maximum(T_d, 1e-6)

    def predict(la_, ls_):
        # log10 coordinates in; predict_log10 where the model has it
        if hasattr(sur, "predict_log10"):
            return sur.predict_log10(la_, ls_, g, n, t)
        return sur.predict(10 ** la_, 10 ** ls_, g, n, t)

    def best_of(cand_a, cand_s, rows):
        # cand_*: [m, c] log10 coordinates of c candidates for each of m rows
        pr = predict(cand_a.ravel(), cand_s.ravel())
        pR = pr["R_d"].reshape(cand_a.shape)
        pT = pr["T_d"].reshape(cand_a.shape)
        c = ((pR - R_d[rows, None]) / R_d[rows, None]) ** 2 + ((pT - T_d[rows, None]) / Tf[rows, None]) ** 2
        k = np.argmin(c, axis=1)
        r = np.arange(len(rows))
        return cand_a[r, k], cand_s[r, k]

    la = np.linspace(-2, 1, grid)
    ls = np.linspace(0, 3, grid)
    A, Sg = np.meshgrid(la, ls, indexing="ij")
    pred = predict(A.ravel(), Sg.ravel())
    # every measurement against every cell, on the GPU
    search = argmin or (lambda pR, pT, R, T: grid_argmin(pR, pT, R, T, device=sur.mlp.device))
    k = search(pred["R_d"], pred["T_d"], R_d, T_d)
    ba, bs = A.ravel()[k], Sg.ravel()[k]
    if fine > 1:
        ha, hs = la[1] - la[0], ls[1] - ls[0]
        off = np.linspace(-1, 1, fine)
        Oa, Os = np.meshgrid(off, off, indexing="ij")
        for s in range(0, R_d.size, 1024):
            rows = np.arange(s, min(s + 1024, R_d.size))
            ca = np.clip(ba[rows, None] + ha * Oa.ravel()[None, :], -2, 1)
            cs = np.clip(bs[rows, None] + hs * Os.ravel()[None, :], 0, 3)
            ba[rows], bs[rows] = best_of(ca, cs, rows)
    x0 = np.stack([ba, bs], 1) * math.log(10)   # ln mu_a, ln mu_s
    if polish:
        def ev(points, ids):
            P = np.array(points)
            pr = sur.predict(np.exp(P[:, 0]), np.exp(P[:, 1]), g, n, t)
            i = np.array(ids)
            return (((pr["R_d"] - R_d[i]) / R_d[i]) ** 2 +
                    ((pr["T_d"] - T_d[i]) / np.maximum(T_d[i], 1e-6)) ** 2)
        res = _praxis.minimize_batch(list(x0), ev, 1e-4, 0.05, max_evals=200)
        x0 = np.array([r[0] for r in res])
    mua = np.exp(x0[:, 0])
    musr = np.exp(x0[:, 1]) * (1 - g)
    return np.stack([mua, musr], 1)


# search box in (ln mu_a, ln mu_s'); outside it the cost is a penalty
LOG_BOX = np.array([[math.log(1e-4), math.log(1e3)], [math.log(1e-3), math.log(1e4)]])


class LogCost(M.BatchCost):
    """MC cost of points given as (ln mu_a, ln mu_s'); points outside
    LOG_BOX get PENALTY plus their distance to it.  Launches with LAUNCH."""

    def __init__(self, *args, **kw):
        super().__init__(*args, **kw)
        self.launch = dict(LAUNCH)

    def __call__(self, points, ids):
        Y = np.asarray(points, dtype=np.float64).reshape(len(points), 2)
        over = np.maximum(LOG_BOX[:, 0] - Y, 0) + np.maximum(Y - LOG_BOX[:, 1], 0)
        inside = np.all(np.isfinite(Y), axis=1) & ~np.any(over > 0, axis=1)
        out = np.empty(len(points))
        out[~inside] = M.PENALTY + np.sum(np.nan_to_num(over[~inside], nan=1e6), axis=1)
        if inside.any():
            out[inside] = super().__call__(np.exp(Y[inside]), np.asarray(ids)[inside])
        return out


def refine_batch(problems, seeds, packets=100_000, step=0.03, tol=3e-3, seed=7,
                 engine: Optional[Engine] = None, max_evals=80, stream_base=0, log=None):
    """Batched PRAXIS on the MC cost in log coordinates from `seeds`
    ((mu_a, mu_s') per problem); with one problem this is the sequential
    one-launch-per-evaluation solver.  Returns (solutions [P, 2],
    evaluations, stats)."""
    eng = engine or default_engine()
    x0 = [np.log(np.asarray(s, dtype=np.float64)) for s in seeds]
    cost = LogCost(eng, problems, packets, seed=seed,
                   streams=np.arange(len(problems)) + stream_base)
    t0 = time.perf_counter()

    def on_round(r, pending):
        if log and r % 10 == 0:
            log("    round %d: %d solves pending, %.1f s" % (r, pending, time.perf_counter() - t0))
    res = _praxis.minimize_batch(x0, cost, tol, step, max_evals=max_evals, on_round=on_round)
    sol = np.array([np.exp(r[0]) for r in res])
    nev = np.array([r[2] for r in res])
    return sol, nev, {"launches": cost.launches, "packets": cost.packets, "kernel_ms": cost.kernel_ms}


def refine_newton(problems, seeds, sur, iters=6, packets=100_000, seed=7, engine: Optional[Engine] = None,
                  stream_base=0, g=G, n=N_INDEX, t=THICK, radius=0.5, s_cut=0.05, log=None, step_tol=0.0):
    """Defect correction with the surrogate's Jacobian, in y = (ln mu_a, ln mu_s):
    r = (MC(y) - M) / M are the relative residuals of R_d, T_d (every solve's
    MC in one launch per iteration, common random numbers per solve) and J
    the surrogate's relative Jacobian by central differences.  The step is
    the Gauss-Newton step -J^+ r with singular values of J below `s_cut`
    dropped (a direction the network cannot resolve, e.g. mu_a when
    absorption hardly changes R_d and T_d, is left alone), limited to a
    trust radius per solve.  A trial whose residual norm is not smaller is
    rejected and the radius shrinks fourfold; an accepted one doubles it.
    A solve whose next step is shorter than step_tol (in ln mu) stops
    there and runs in no later launch (0: every solve runs every
    iteration).  Returns (best solutions, evaluations per solve, stats)."""
    eng = engine or default_engine()
    P = len(problems)
    Mv = np.array([[p.measured.R_d, p.measured.T_d] for p in problems])
    Mv = np.maximum(Mv, 1e-6)
    y_acc = np.stack([np.log(np.asarray(seeds)[:, 0]), np.log(np.asarray(seeds)[:, 1] / (1 - g))], 1)
    y_try = y_acc.copy()
    r_acc = np.full((P, 2), np.inf)
    rad = np.full(P, float(radius))
    streams = np.arange(P) + stream_base
    stats = {"launches": 0, "packets": 0, "kernel_ms": 0.0, "rejected": 0}
    # difference step well above the BF16 network's output resolution (~4e-3)
    h = 0.05
    lo, hi = np.array([[-9.2, -6.9]]), np.array([[6.9, 9.2]])
    cache = {}
    active = np.ones(P, dtype=bool)
    nev = np.zeros(P, dtype=np.int64)
    for it in range(iters):
        ids = np.where(active)[0]
        if len(ids) == 0:
            break
        mua, mus = np.exp(y_try[ids, 0]), np.exp(y_try[ids, 1])
        Fi, _ = mc_rt(eng, problems, mua, mus * (1 - g), packets, seed, streams[ids],
                      ids=None if len(ids) == P else ids, cache=cache)
        F = np.full((P, 2), np.inf)
        F[ids] = Fi
        nev[ids] += 1
        stats["launches"] += 1
        stats["packets"] += packets * len(ids)
        stats["kernel_ms"] += _last_ms(eng)
        r = (F - Mv) / Mv
        better = active & (np.linalg.norm(r, axis=1) < np.linalg.norm(r_acc, axis=1))
        if it > 0:
            stats["rejected"] += int((active & ~better).sum())
        y_acc[better] = y_try[better]
        r_acc[better] = r[better]
        rad = np.where(better, np.minimum(rad * 2.0, 1.0), np.where(active, rad * 0.25, rad))
        if it == iters - 1:
            break
        # surrogate Jacobian d(R, T)/dy at the accepted points, all solves at once
        pts = np.concatenate([y_acc + [h, 0], y_acc - [h, 0], y_acc + [0, h], y_acc - [0, h]])
        pr = sur.predict(np.exp(pts[:, 0]), np.exp(pts[:, 1]), g, n, t)
        V = np.stack([pr["R_d"], pr["T_d"]], 1).reshape(4, P, 2)
        J = np.stack([(V[0] - V[1]) / (2 * h), (V[2] - V[3]) / (2 * h)], 2) / Mv[:, :, None]
        U, S, Wt = np.linalg.svd(J)
        Sinv = np.where(S > s_cut, 1.0 / np.maximum(S, 1e-300), 0.0)
        dy = -np.einsum("pji,pj,pjk,pk->pi", Wt, Sinv, U, r_acc)
        nrm = np.linalg.norm(dy, axis=1)
        dy *= np.minimum(1.0, rad / np.maximum(nrm, 1e-300))[:, None]
        if step_tol > 0:
            active &= np.linalg.norm(dy, axis=1) >= step_tol
        y_try = np.where(active[:, None], np.clip(y_acc + dy, lo, hi), y_acc)
        if log:
            log("    defect correction %d: median |r| %.2e, median |dy| %.2e, %d rejected, %d running" %
                (it, float(np.median(np.abs(r_acc).max(1))), float(np.median(nrm)), int((active & ~better).sum()),
                 int(active.sum())))
    sol = np.stack([np.exp(y_acc[:, 0]), np.exp(y_acc[:, 1]) * (1 - g)], 1)
    return sol, nev, stats


def refine_mc_newton(problems, start, iters=8, packets=100_000, seed=7, engine: Optional[Engine] = None,
                     stream_base=0, h=0.05, radius=0.5, s_cut=1e-3, step_tol=1e-3, log=None):
    """Gauss-Newton on the MC residuals alone, for solves the surrogate
    cannot guide (the fallback after the MC check).  y = (ln mu_a, ln mu_s'),
    r = (MC(y) - M) / M the relative residuals of R_d and T_d.  Each
    iteration is one launch that runs, for every pending solve, its trial
    point y and the points y + h e_1, y + h e_2 on the solve's own Philox
    stream: r and the forward-difference Jacobian J come from common random
    numbers (runs that share their histories differ smoothly in y).  A
    trial whose residual norm is below the accepted one's is accepted with
    its J and the trust radius doubles; otherwise the radius shrinks
    fourfold and the next step starts again from the accepted point and its
    J.  The step is -J^+ r with singular values of J below `s_cut` dropped
    (far below the surrogate path's cut: an MC difference has no network
    resolution limit, a weakly resolved direction such as mu_a of a
    scattering-dominated sample is still worth following, and the
    acceptance test rejects a step that noise spoils), cut to the trust
    radius and kept inside LOG_BOX.  A solve stops when its
    step is shorter than `step_tol` or its radius falls below it.  Returns
    (best solutions [P, 2] as (mu_a, mu_s'), MC evaluations per solve,
    stats)."""
    eng = engine or default_engine()
    P = len(problems)
    Mv = np.maximum(np.array([[p.measured.R_d, p.measured.T_d] for p in problems], dtype=np.float64), 1e-6)
    y_acc = np.log(np.asarray(start, dtype=np.float64).reshape(P, 2))
    y_try = y_acc.copy()
    r_acc = np.full((P, 2), np.inf)
    J_acc = np.zeros((P, 2, 2))
    rad = np.full(P, float(radius))
    pending = np.ones(P, dtype=bool)
    nev = np.zeros(P, dtype=np.int64)
    streams = (np.arange(P) + stream_base).astype(np.uint32)
    lo, hi = LOG_BOX[:, 0], LOG_BOX[:, 1] - h
    cache = {}
    stats = {"launches": 0, "packets": 0, "kernel_ms": 0.0, "rejected": 0}
    for it in range(iters):
        idx = np.where(pending)[0]
        if len(idx) == 0:
            break
        k = len(idx)
        yt = y_try[idx]
        pts = np.concatenate([yt, yt + [h, 0.0], yt + [0.0, h]])
        F, _ = mc_rt(eng, problems, np.exp(pts[:, 0]), np.exp(pts[:, 1]), packets, seed,
                     np.tile(streams[idx], 3), ids=np.tile(idx, 3), cache=cache)
        stats["launches"] += 1
        stats["packets"] += 3 * k * packets
        stats["kernel_ms"] += _last_ms(eng)
        nev[idx] += 3
        F = F.reshape(3, k, 2)
        Mk = Mv[idx]
        r = (F[0] - Mk) / Mk
        J = np.stack([(F[1] - F[0]) / h, (F[2] - F[0]) / h], 2) / Mk[:, :, None]
        better = np.linalg.norm(r, axis=1) < np.linalg.norm(r_acc[idx], axis=1)
        stats["rejected"] += int((~better).sum()) if it > 0 else 0
        acc = idx[better]
        y_acc[acc] = yt[better]
        r_acc[acc] = r[better]
        J_acc[acc] = J[better]
        rad[idx] = np.where(better, np.minimum(rad[idx] * 2.0, 1.0), rad[idx] * 0.25)
        U, S, Wt = np.linalg.svd(J_acc[idx])
        Sinv = np.where(S > s_cut, 1.0 / np.maximum(S, 1e-300), 0.0)
        dy = -np.einsum("pji,pj,pjk,pk->pi", Wt, Sinv, U, r_acc[idx])
        nrm = np.linalg.norm(dy, axis=1)
        dy *= np.minimum(1.0, rad[idx] / np.maximum(nrm, 1e-300))[:, None]
        y_try[idx] = np.clip(y_acc[idx] + dy, lo, hi)
        done = (np.linalg.norm(dy, axis=1) < step_tol) | (rad[idx] < step_tol)
        pending[idx[done]] = False
        if log:
            log("    MC Gauss-Newton %d: %d solves, median |r| %.2e, %d rejected, %d done" %
                (it, k, float(np.median(np.abs(r_acc[idx]).max(1))), int((~better).sum()), int(done.sum())))
    return np.exp(y_acc), nev, stats


def check(problems, sol, packets=1_000_000, seed=9, engine: Optional[Engine] = None,
          refine_packets=100_000, z=4.0, floor=0.002, screen_z=1.5):
    """Independent MC at each solution.  A solution misses when R_d or T_d
    differs from the measurement by more than z combined MC standard errors
    (this check's and the refinement's, which scales as sqrt(packets /
    refine_packets)) plus a relative floor.

    Two stages: every solution is first re-simulated at refine_packets on
    streams of its own, and passes when both readings lie within screen_z
    combined standard errors (that run's and the refinement's) plus the
    floor; only the others are judged at `packets` as above (on the same
    streams and seed as a one-stage check, so their verdicts are the ones
    it gives).  A solution that the full check would fail is off by more
    than z (1 + refine_packets / packets)^(1/2) refinement standard errors
    (4.2 for z = 4 and packets = 10 refine_packets) beyond the floor; it
    passes the screen only where the screen's reading falls short of that
    by more than 4.2 - screen_z 2^(1/2) = 2.08 of its standard errors,
    odds of 1.9 % at the threshold and less beyond it.  A solution that
    fits (residual of about one refinement standard error) passes with
    odds of about 0.87 per reading, so about three in four skip the full
    check.  screen_z=None checks every solution at `packets`.  Returns
    (miss, relative misfits R, T, packets), the misfits from the last
    stage each solution ran."""
    eng = engine or default_engine()
    sol = np.asarray(sol, dtype=np.float64)
    P = len(problems)
    Mv = np.array([[p.measured.R_d, p.measured.T_d] for p in problems], dtype=np.float64)
    rel = np.zeros((P, 2))
    todo = np.arange(P)
    used = 0
    if screen_z is not None and refine_packets < packets and P:
        F1, SE1 = mc_rt(eng, problems, sol[:, 0], sol[:, 1], refine_packets, seed + 1,
                        np.arange(P, dtype=np.uint32) + 95_000_000, use_graph=False)
        used += P * refine_packets
        d1 = np.abs(F1 - Mv)
        rel = d1 / np.maximum(Mv, 1e-12)
        clear = np.all(d1 <= screen_z * math.sqrt(2.0) * SE1 + floor * Mv, axis=1)
        todo = np.where(~clear)[0]
    miss = np.zeros(P, dtype=bool)
    if len(todo):
        F, SE = mc_rt(eng, problems, sol[todo, 0], sol[todo, 1], packets, seed,
                      todo.astype(np.uint32) + 90_000_000, ids=todo, use_graph=False)
        used += len(todo) * packets
        k = math.sqrt(1.0 + packets / float(refine_packets))
        d = np.abs(F - Mv[todo])
        rel[todo] = d / np.maximum(Mv[todo], 1e-12)
        miss[todo] = np.any(d > z * k * SE + floor * Mv[todo], axis=1)
    return miss, rel[:, 0], rel[:, 1], used


GENERIC_GUESS = (0.3, 5.0)


def solve_plain(problem_, guess=GENERIC_GUESS, packets=1_000_000, step=1.0, tol=3e-3, seed=3,
                engine: Optional[Engine] = None, stream=0, max_evals=400):
    """Plain MC minimization, the reference's execution model: one PRAXIS,
    one transport launch per cost evaluation, from a generic guess (here in
    log coordinates with a tolerance that reaches ~0.3 %, so that its
    accuracy is not limited by the optimizer)."""
    t0 = time.perf_counter()
    sol, nev, st = refine_batch([problem_], [guess], packets=packets, step=step, tol=tol, seed=seed,
                                engine=engine, max_evals=max_evals, stream_base=stream)
    return sol[0], int(nev[0]), time.perf_counter() - t0, st


# minimize.o's accept window for the synthetic measurements' box (truth mu_a
# 0.01..10 /cm, mu_s' 0.2..200 /cm), a factor of two wider on every side
PLAIN_WINDOW = ((0.005, 20.0), (0.1, 400.0))


def _plain_stream(stream_base, p, count, r, runs, attempt):
    # run r's attempt of problem p; run 0's first attempt is stream_base + p,
    # the stream the single-run plain solver has always used
    return stream_base + p + count * (r + runs * attempt)


def plain_batch(problems, runs=1, packets=1_000_000, tol=3e-3, engine: Optional[Engine] = None,
                stream_base=120_000_000, window=PLAIN_WINDOW, max_retries=5, guess=GENERIC_GUESS,
                log=None):
    """The plain solver on every problem at once, in lock step: `runs`
    PRAXIS minimizations per problem from the generic guess, each on its
    own Philox stream, a result outside `window` retried on a new stream
    (at most `max_retries` times, then kept), and the mean of the accepted
    runs as the solution (minimize.o's loop, CUDA/ISMLMC-minimization.cu:
    467-512; with runs=1 and no retry, solve_plain).  Each PRAXIS sees the
    same cost values as when run alone (common random numbers per stream,
    tallies independent of the launch), so every solution equals the
    sequential solver's bit for bit; only the time differs.  Returns
    (solutions [P, 2], evaluations per problem, retries per problem)."""
    eng = engine or default_engine()
    P = len(problems)
    accepted = [[None] * runs for _ in range(P)]
    nev = np.zeros(P, dtype=np.int64)
    retries = np.zeros((P, runs), dtype=np.int64)   # per run, as minimize.o retries each run
    todo = [(p, r) for p in range(P) for r in range(runs)]
    attempt = 0
    (a0, a1), (s0, s1) = window
    while todo:
        # one lock-step batch per attempt; streams per (problem, run, attempt)
        cost_probs = [problems[p] for p, _ in todo]
        streams = np.array([_plain_stream(stream_base, p, P, r, runs, attempt) for p, r in todo])
        x0 = [np.log(np.asarray(guess, dtype=np.float64))] * len(todo)
        cost = LogCost(eng, cost_probs, packets, seed=3, streams=streams)
        res = _praxis.minimize_batch(x0, cost, tol, 1.0, max_evals=400)
        nxt = []
        for k, ((p, r), (x, fx, nf)) in enumerate(zip(todo, res)):
            xs = np.exp(x)
            nev[p] += nf
            if (a0 <= xs[0] <= a1 and s0 <= xs[1] <= s1) or retries[p, r] >= max_retries:
                accepted[p][r] = xs
            else:
                retries[p, r] += 1
                nxt.append((p, r))
        if log:
            log("    plain (runs %d), attempt %d: %d solves, %d to retry" % (runs, attempt, len(todo), len(nxt)))
        todo = nxt
        attempt += 1
    sol = np.array([np.mean(a, axis=0) for a in accepted])
    return sol, nev, retries.sum(axis=1)


def plain_sequential(problem_, p, count, runs=1, packets=1_000_000, tol=3e-3, engine=None,
                     stream_base=120_000_000, window=PLAIN_WINDOW, max_retries=5):
    """One problem (index p of `count`) through plain_batch's loop one
    PRAXIS and one launch per evaluation at a time, as minimize.o runs it.
    Returns (solution, evaluations, seconds)."""
    t0 = time.perf_counter()
    (a0, a1), (s0, s1) = window
    acc, nev = [], 0
    for r in range(runs):
        attempt = 0
        while True:
            x, nf, _, _ = solve_plain(problem_, packets=packets, tol=tol, engine=engine,
                                      stream=_plain_stream(stream_base, p, count, r, runs, attempt))
            nev += nf
            if (a0 <= x[0] <= a1 and s0 <= x[1] <= s1) or attempt >= max_retries:
                break
            attempt += 1
        acc.append(x)
    return np.mean(acc, axis=0), nev, time.perf_counter() - t0


def wilson(k, n, z=1.96):
    """Wilson score interval of a binomial proportion k / n."""
    if n <= 0:
        return [0.0, 1.0]
    p = k / n
    d = 1 + z * z / n
    c = (p + z * z / (2 * n)) / d
    h = z * math.sqrt(p * (1 - p) / n + z * z / (4 * n * n)) / d
    return [max(0.0, c - h), min(1.0, c + h)]


def ratio_ci(k1, n1, k2, n2, z=1.96):
    """95 % interval of (k1/n1) / (k2/n2) for independent binomial counts
    (Katz's log method; 0.5 added to a zero count)."""
    a, b = (k1 or 0.5), (k2 or 0.5)
    r = (a / n1) / (b / n2)
    se = math.sqrt(max(0.0, 1 / a - 1 / n1 + 1 / b - 1 / n2))
    return [r * math.exp(-z * se), r * math.exp(z * se)]


def errors(sol, truth_mua, truth_musr):
    ea = np.abs(sol[:, 0] - truth_mua) / truth_mua
    es = np.abs(sol[:, 1] - truth_musr) / truth_musr
    return ea, es


def summary(ea, es):
    return {"median_mu_a": float(np.median(ea)), "median_mus_reduced": float(np.median(es)),
            "p90_mu_a": float(np.percentile(ea, 90)), "p90_mus_reduced": float(np.percentile(es, 90)),
            "frac_both_within_1pct": float(np.mean((ea <= 0.01) & (es <= 0.01))),
            "frac_both_within_2pct": float(np.mean((ea <= 0.02) & (es <= 0.02)))}


# Defect correction stops a solve once its next step is shorter than this
# in ln mu (0.1 %): a tenth of the accuracy the solves are judged at, and
# below what the refinement's MC evaluations resolve.  Measured on 4096
# solves against refining every solve six times: 4.1 evaluations per solve
# instead of 6 at 1 % noise (solves within 1 % of both parameters 339
# against 344, median errors and fit unchanged) and 2.7 at 0.1 % noise
# (1647 against 1635).
REFINE_STEP_TOL = 1e-3


def _solve_seeded(probs, sur, seeds, method, refine_packets, check_packets, eng, log, stream_base,
                  fallback="newton"):
    t0 = time.perf_counter()
    if method == "newton":
        sol, nev, st = refine_newton(probs, seeds, sur, packets=refine_packets, engine=eng,
                                     stream_base=stream_base, log=log, step_tol=REFINE_STEP_TOL)
    else:
        sol, nev, st = refine_batch(probs, seeds, packets=refine_packets, engine=eng, log=log,
                                    stream_base=stream_base)
    t_ref = time.perf_counter() - t0
    miss, rR, rT, chk_packets = check(probs, sol, check_packets, engine=eng, refine_packets=refine_packets)
    t_chk = time.perf_counter() - t0 - t_ref
    n_miss = int(miss.sum())
    fb_time = 0.0
    n_repaired = 0
    if n_miss:
        # MC minimization from the refined point: the misses are mostly
        # measurements whose best fit is poorly conditioned (or not exact
        # within the noise), not wrong basins
        idx = np.where(miss)[0]
        sub = [probs[i] for i in idx]
        log("  fallback (%s): MC minimization from the refined point for %d misses" % (fallback, n_miss))
        tf = time.perf_counter()
        if fallback == "newton":
            fb, _, _ = refine_mc_newton(sub, sol[idx], packets=refine_packets, engine=eng,
                                        stream_base=stream_base + 10_000_000, log=log)
        else:
            fb, _, _ = refine_batch(sub, sol[idx], packets=refine_packets, step=0.3, tol=3e-3,
                                    max_evals=150, engine=eng, stream_base=stream_base + 10_000_000, log=log)
        sol[idx] = fb
        fb_time = time.perf_counter() - tf
        # judge the fallback's results too (reported, not timed)
        miss2, _, _, _ = check([probs[i] for i in idx], fb, check_packets, seed=11, engine=eng,
                               refine_packets=refine_packets)
        n_repaired = int((~miss2).sum())
    return sol, {"refine_s": t_ref, "check_s": t_chk, "fallback_s": fb_time, "misses": n_miss,
                 "misses_repaired_by_fallback": n_repaired,
                 "mean_evals": float(np.mean(nev)), "launches": st["launches"],
                 "refine_packets": st["packets"], "check_packets": chk_packets,
                 "misfit_R_median": float(np.median(rR)), "misfit_T_median": float(np.median(rT))}


FIT_TOL = 0.003


def fit_cost(problems, sol, packets, eng, seed=13):
    """Cost (relative R_d, T_d misfit) of each solution against its measurement,
    MC at `packets` on streams independent of every solver's."""
    sol = np.asarray(sol, dtype=np.float64)
    F, _ = mc_rt(eng, problems, sol[:, 0], sol[:, 1], packets, seed,
                 np.arange(len(problems), dtype=np.uint32) + 200_000_000, use_graph=False)
    return np.array([M.cost_from({"R_d": f[0], "T_d": f[1]}, p.measured) for f, p in zip(F, problems)])


def run(sur, count=4096, meas_packets=1_000_000, refine_packets=100_000, check_packets=1_000_000,
        plain_count=8, plain_packets=1_000_000, seed=0, engine: Optional[Engine] = None,
        methods=("newton", "praxis"), with_ablation=True, log=print, noise=0.01, plain_tol=3e-3,
        fallback="newton", plain=None):
    """Surrogate-seeded batched solves vs plain MC minimization on the same
    GPU, over `count` synthetic measurements (`noise` relative Gaussian
    noise on R_d and T_d).  `plain` lists (runs, accuracy sample, timed
    sample) triples of the plain solver (default: one run, `plain_count`
    measurements for both); entry "plain" is the one-run solver, the
    others "plain_runsN".  Every within-1 % rate carries a Wilson 95 %
    interval."""
    eng = engine or default_engine()
    if plain is None:
        plain = ((1, plain_count, plain_count),)
    meas = make_measurements(count, meas_packets, noise=noise, seed=seed, engine=eng)
    probs = [problem(r, t) for r, t in zip(meas["R_d"], meas["T_d"])]
    out = {"count": count, "noise": noise, "meas_packets": meas_packets, "refine_packets": refine_packets,
           "check_packets": check_packets, "plain_packets": plain_packets, "plain_tol": plain_tol}
    log("measurements ready (%d at %d packets)" % (count, meas_packets))
    # the truth's own misfit to the noisy measurement (MC at check_packets)
    cost_truth = fit_cost(probs, np.stack([meas["mu_a"], meas["mus_reduced"]], 1), check_packets, eng)
    for k, method in enumerate(methods):
        t0 = time.perf_counter()
        seeds = surrogate_seeds(sur, meas["R_d"], meas["T_d"])
        t_seed = time.perf_counter() - t0
        log("seeded refinement: %s (seeds %.1f s)" % (method, t_seed))
        sol, info = _solve_seeded(probs, sur, seeds, method, refine_packets, check_packets, eng, log,
                                  stream_base=k * 40_000_000, fallback=fallback)
        info["fallback"] = fallback
        total = t_seed + info["refine_s"] + info["check_s"] + info["fallback_s"]
        info["seed_s"] = t_seed
        info["seconds"] = total
        info["solves_per_s"] = count / total
        info["seed_errors"] = summary(*errors(seeds, meas["mu_a"], meas["mus_reduced"]))
        ea, es = errors(sol, meas["mu_a"], meas["mus_reduced"])
        info["errors"] = summary(ea, es)
        k_ok = int(np.sum((ea <= 0.01) & (es <= 0.01)))
        info["within_1pct"] = {"count": k_ok, "of": count, "frac": k_ok / count, "ci95": wilson(k_ok, count)}
        c = fit_cost(probs, sol, check_packets, eng)
        info["fits_as_well_as_truth"] = float(np.mean(c <= cost_truth + FIT_TOL))
        out["seeded_" + method] = info
        log("  %s: %.1f solves/s, %d misses, fit %.3f, errors %s" %
            (method, count / total, info["misses"], info["fits_as_well_as_truth"], info["errors"]))
    if with_ablation:
        log("ablation: batched PRAXIS from the generic guess (no surrogate)")
        t1 = time.perf_counter()
        sol2, nev2, st2 = refine_batch(probs, [GENERIC_GUESS] * count, packets=refine_packets, step=1.0,
                                       max_evals=400, engine=eng, stream_base=100_000_000, log=log)
        t2 = time.perf_counter() - t1
        c2 = fit_cost(probs, sol2, check_packets, eng)
        out["batched_unseeded"] = {"solves_per_s": count / t2, "seconds": t2,
                                   "mean_evals": float(nev2.mean()),
                                   "fits_as_well_as_truth": float(np.mean(c2 <= cost_truth + FIT_TOL)),
                                   "errors": summary(*errors(sol2, meas["mu_a"], meas["mus_reduced"]))}
        log("  unseeded: %.1f solves/s, %.1f evals/solve" % (count / t2, nev2.mean()))
    for runs, count_p, timed in plain:
        key = "plain" if runs == 1 else "plain_runs%d" % runs
        out[key] = _plain_stats(probs, meas, cost_truth, runs, min(count_p, count), min(timed, count_p),
                                plain_packets, plain_tol, check_packets, eng, log)
    for method in methods:
        out["speedup_%s_vs_plain" % method] = out["seeded_" + method]["solves_per_s"] / out["plain"]["solves_per_s"]
    return out


def rate_ci(secs, n_boot=4000, seed=0):
    """95 % bootstrap interval of a rate len(secs) / sum(secs) measured
    over independent solves (resampling the solves; the per-solve time is
    heavy-tailed, so the interval is asymmetric)."""
    secs = np.asarray(secs, dtype=np.float64)
    rng = np.random.default_rng(seed)
    idx = rng.integers(0, len(secs), size=(n_boot, len(secs)))
    rates = len(secs) / secs[idx].sum(axis=1)
    lo, hi = np.percentile(rates, [2.5, 97.5])
    return [float(lo), float(hi)]


def _plain_stats(probs, meas, cost_truth, runs, count, timed, packets, tol, check_packets, eng, log):
    """Accuracy of the plain solver over the first `count` measurements (run
    in lock step, plain_batch) and its solves/s from the first `timed` of
    them solved one evaluation at a time (plain_sequential), whose
    solutions must equal the lock-step ones."""
    t0 = time.perf_counter()
    sol, nev, retries = plain_batch(probs[:count], runs=runs, packets=packets, tol=tol, engine=eng, log=log)
    t_batch = time.perf_counter() - t0
    seq, secs, sevals = [], [], []
    for k in range(timed):
        x, nf, dt = plain_sequential(probs[k], k, count, runs=runs, packets=packets, tol=tol, engine=eng)
        seq.append(x)
        secs.append(dt)
        sevals.append(nf)
        log("  plain (runs %d) solve %d: %.2f s, %d evaluations" % (runs, k, dt, nf))
    same = bool(np.array_equal(np.array(seq), sol[:timed])) if timed else None
    c = fit_cost(probs[:count], sol, check_packets, eng)
    ea, es = errors(sol, meas["mu_a"][:count], meas["mus_reduced"][:count])
    k_ok = int(np.sum((ea <= 0.01) & (es <= 0.01)))
    return {"runs": runs, "count": count, "timed": timed,
            "timed_as": "sequential: one PRAXIS per run, one %d-packet launch per cost evaluation, from "
                        "the generic guess%s" % (packets, "; mean of %d accepted runs" % runs if runs > 1 else ""),
            "solves_per_s": timed / sum(secs) if timed else None, "seconds_timed": sum(secs),
            # a solve's cost is heavy-tailed (a path through long-history
            # parameters can take seconds), so the rate is total solves over
            # total time and the spread is reported beside it
            "seconds_per_solve": ({"median": float(np.median(secs)), "p90": float(np.percentile(secs, 90)),
                                   "max": float(np.max(secs))} if timed else None),
            "solves_per_s_ci95": rate_ci(secs) if timed else None,
            "mean_evals_timed": float(np.mean(sevals)) if timed else None,
            "mean_evals": float(np.mean(nev)), "retries": int(retries.sum()),
            "sequential_equals_lockstep": same, "lockstep_s": t_batch, "packets_per_eval": packets,
            "fits_as_well_as_truth": float(np.mean(c <= cost_truth[:count] + FIT_TOL)),
            "within_1pct": {"count": k_ok, "of": count, "frac": k_ok / count, "ci95": wilson(k_ok, count)},
            "errors": summary(ea, es)}
