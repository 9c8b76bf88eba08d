"""Lookup-table sweeps over (mu_a, mu_s').

sample_space()   one grid of single-layer samples x runs (and, for
                 single-sphere experiments, an R-sphere and a T-sphere
                 configuration per point), all in one batched launch:
                 singleLayerSampleSpace, CUDA/ISMLMC-lookup.cu:153-421,
                 and its statistics array_avg / array_stderr /
                 array_poisson (:23-110).
write_csv()      the reference's two CSV tables (write_data, :113-150).
run_lookup()     the input-uncertainty outer loop of Lookup/lookup.py:131-222
                 (n and sphere parameters drawn from normal distributions),
                 with every sample's grid folded into the same launch.
adaptive_sweep() uneven packet counts per point chosen from a pilot run so
                 that the relative standard error of total reflectance
                 reaches a target, in one launch after the pilot.
"""
import math
from dataclasses import dataclass, field
from typing import List, Optional

import numpy as np

from .config import Layer, SampleConfig, Sphere, mus_from_reduced
from .engine import CONFIG_DTYPE, TALLIES, Engine, default_engine


@dataclass
class SweepResult:
    mua: np.ndarray
    mus_reduced: np.ndarray
    R_d: np.ndarray          # [N_mua, N_mus] mean over runs
    R_d_unc: np.ndarray      # standard error (or Poisson) as the reference reports it
    T_d: np.ndarray
    T_d_unc: np.ndarray
    runs_R_d: np.ndarray     # [runs, N_mua, N_mus]
    runs_T_d: np.ndarray
    n_packets: int           # per simulation
    kernel_ms: float
    total_packets: int
    extra: dict = field(default_factory=dict)


def _single(n, g, t, mu_a, mu_s, R, T, specular_included, roulette_threshold):
    return SampleConfig(layers=[Layer(n=n, g=g, t=t, mu_a=mu_a, mu_s=mu_s)], R=R, T=T,
                        specular_included=specular_included,
                        roulette_threshold=roulette_threshold)


def _grid_configs(n, g, t, mua, mus_reduced, R, T, specular_included, separate, roulette_threshold):
    """Configs for one grid in (i, j) order; with `separate` each point has
    an R-sphere-only and a T-sphere-only configuration (ISMLMC-lookup.cu:317,338)."""
    cfgs = []
    for a in mua:
        for sr in mus_reduced:
            mu_s = mus_from_reduced(sr, g)
            if separate:
                Rs = Sphere(**{**vars(R), "present": True})
                Ts = Sphere(**{**vars(T), "present": True})
                cfgs.append(_single(n, g, t, a, mu_s, Rs, Sphere(**{**vars(T), "present": False}),
                                    specular_included, roulette_threshold))
                cfgs.append(_single(n, g, t, a, mu_s, Sphere(**{**vars(R), "present": False}), Ts,
                                    specular_included, roulette_threshold))
            else:
                cfgs.append(_single(n, g, t, a, mu_s, R, T, specular_included, roulette_threshold))
    return cfgs


def _reduce_grid(results, rows, cols, runs, separate, spheres_present):
    """Per-run R_d/T_d tables from the batch results of one grid."""
    per = 2 if separate else 1
    Rd = np.zeros((runs, rows, cols))
    Td = np.zeros((runs, rows, cols))
    k = 0
    for r in range(runs):
        for i in range(rows):
            for j in range(cols):
                if separate:
                    Rd[r, i, j] = results[k].mean["R_d"]
                    Td[r, i, j] = results[k + 1].mean["T_d"]
                elif spheres_present:
                    Rd[r, i, j] = results[k].mean["R_d"]
                    Td[r, i, j] = results[k].mean["T_d"]
                else:
                    Rd[r, i, j] = results[k].mean["R_d"] + results[k].mean["R_s"]
                    Td[r, i, j] = results[k].mean["T_d"] + results[k].mean["T_u"]
                k += per
    return Rd, Td


def array_stderr(meta):
    """Population standard deviation over runs / sqrt(runs) (ISMLMC-lookup.cu:46-80)."""
    runs = meta.shape[0]
    return np.sqrt(((meta - meta.mean(axis=0)) ** 2).sum(axis=0) / runs) / math.sqrt(runs)


def array_poisson(meta, N):
    """sqrt(x / N) of the last run (ISMLMC-lookup.cu:83-110)."""
    return np.sqrt(np.maximum(meta[-1], 0.0) / N)


def sample_space(n, g, t, mua, mus_reduced, R: Sphere, T: Sphere, N, runs, separate=False,
                 specular_included=False, seed=1234, engine: Optional[Engine] = None,
                 roulette_threshold=1e-4, fp64=False, chunk_packets=0, blocks_per_cu=0):
    eng = engine or default_engine()
    mua = np.asarray(mua, dtype=np.float64)
    mus_reduced = np.asarray(mus_reduced, dtype=np.float64)
    poisson = runs == 0
    runs_eff = 1 if poisson else int(runs)
    grid = _grid_configs(n, g, t, mua, mus_reduced, R, T, specular_included, separate,
                         roulette_threshold)
    cfgs = grid * runs_eff
    streams = np.arange(len(cfgs), dtype=np.uint32)
    res = eng.simulate_batch(cfgs, int(N), seed=seed, streams=streams, fp64=fp64,
                             chunk_packets=chunk_packets, blocks_per_cu=blocks_per_cu)
    Rd, Td = _reduce_grid(res, mua.size, mus_reduced.size, runs_eff, separate,
                          bool(R.present or T.present))
    if poisson:
        Ru, Tu = array_poisson(Rd, N), array_poisson(Td, N)
    else:
        Ru, Tu = array_stderr(Rd), array_stderr(Td)
    return SweepResult(mua=mua, mus_reduced=mus_reduced, R_d=Rd.mean(axis=0), R_d_unc=Ru,
                       T_d=Td.mean(axis=0), T_d_unc=Tu, runs_R_d=Rd, runs_T_d=Td, n_packets=int(N),
                       kernel_ms=res[0].kernel_ms if res else 0.0,
                       total_packets=int(N) * len(cfgs))


def _fmt(x):
    """C++ ostream default formatting of a float (6 significant digits)."""
    return "%g" % float(np.float32(x))


def write_csv(path, header, mua, mus_reduced, data, first=True):
    """One table in the layout of write_data (ISMLMC-lookup.cu:113-150);
    first=False appends (the standard-error table follows the means)."""
    with open(path, "w" if first else "a") as f:
        f.write('"%s","mua"' % header + ',""' * len(mus_reduced) + "\n")
        f.write('"mus\'",""' + "".join("," + _fmt(v) for v in mus_reduced) + "\n")
        for i, a in enumerate(mua):
            f.write('"",' + _fmt(a) + "".join("," + _fmt(v) for v in data[i]) + "\n")


def write_tables(result: SweepResult, directory="."):
    """Diffuse_Reflectance.csv and Diffuse_Transmittance.csv (ISMLMC-lookup.cu:410-418)."""
    import os
    pr = os.path.join(directory, "Diffuse_Reflectance.csv")
    pt = os.path.join(directory, "Diffuse_Transmittance.csv")
    write_csv(pr, "Average Diffuse Reflectance", result.mua, result.mus_reduced, result.R_d, True)
    write_csv(pt, "Average Diffuse Transmittance", result.mua, result.mus_reduced, result.T_d, True)
    write_csv(pr, "Standard Error in Diffuse Reflectance", result.mua, result.mus_reduced,
              result.R_d_unc, False)
    write_csv(pt, "Standard Error in Diffuse Transmittance", result.mua, result.mus_reduced,
              result.T_d_unc, False)
    return pr, pt


def read_csv_table(path, rows):
    """Parse a table written by write_csv the way Lookup/lookup.py:189-213 does:
    returns (means[rows, cols], stderr[rows, cols])."""
    lines = [l.rstrip("\n").split(",") for l in open(path)]
    def block(start):
        return np.array([[float(x) for x in l[2:]] for l in lines[start + 2:start + 2 + rows]])
    return block(0), block(rows + 2)


def run_lookup(ini, samples=None, seed=1234, engine: Optional[Engine] = None, rng_seed=None):
    """lookup.py's CUDA path (Lookup/lookup.py:131-222): draw `samples` sets of
    (n, sphere parameters, angles) from their normal distributions and sweep
    the grid for each; all samples share one batched launch.  Returns the
    mean tables, their std over samples, and the per-sample tables."""
    eng = engine or default_engine()
    L = ini.lookup
    S = int(samples if samples is not None else L.get("input_distributions_samples", 1))
    N, runs = ini.lookup_photons()
    mua, mus_r = ini.lookup_grid()
    rng = np.random.default_rng(rng_seed)
    draws = {
        "n": rng.normal(ini.n, ini.n_unc, S),
        "R_pw": rng.normal(ini.R.pw, ini.R.pw_unc, S), "R_fs": rng.normal(ini.R.fs, ini.R.fs_unc, S),
        "R_fp": rng.normal(ini.R.fp, ini.R.fp_unc, S), "R_fd": rng.normal(ini.R.fd, ini.R.fd_unc, S),
        "R_angle": rng.normal(ini.R.angle_deg, ini.R.angle_unc_deg, S),
        "T_pw": rng.normal(ini.T.pw, ini.T.pw_unc, S), "T_fs": rng.normal(ini.T.fs, ini.T.fs_unc, S),
        "T_fp": rng.normal(ini.T.fp, ini.T.fp_unc, S), "T_fd": rng.normal(ini.T.fd, ini.T.fd_unc, S),
        "T_angle": rng.normal(ini.T.angle_deg, ini.T.angle_unc_deg, S),
    }
    poisson = runs == 0
    runs_eff = 1 if poisson else runs
    cfgs, per_sample = [], None
    for s in range(S):
        R = ini.R.sphere(draws["R_pw"][s], draws["R_fs"][s], draws["R_fp"][s], draws["R_fd"][s],
                         draws["R_angle"][s])
        T = ini.T.sphere(draws["T_pw"][s], draws["T_fs"][s], draws["T_fp"][s], draws["T_fd"][s],
                         draws["T_angle"][s])
        grid = _grid_configs(float(draws["n"][s]), ini.g, ini.t, mua, mus_r, R, T,
                             ini.specular_included, ini.separate, 1e-4) * runs_eff
        per_sample = len(grid)
        cfgs += grid
    res = eng.simulate_batch(cfgs, N, seed=seed, streams=np.arange(len(cfgs), dtype=np.uint32))
    rows, cols = len(mua), len(mus_r)
    Rd = np.zeros((S, rows, cols))
    Td = np.zeros((S, rows, cols))
    for s in range(S):
        part = res[s * per_sample:(s + 1) * per_sample]
        r_runs, t_runs = _reduce_grid(part, rows, cols, runs_eff, ini.separate,
                                      bool(ini.R.present or ini.T.present))
        Rd[s] = r_runs.mean(axis=0)
        Td[s] = t_runs.mean(axis=0)
    return {"mua": np.array(mua), "mus_reduced": np.array(mus_r), "R_d": Rd.mean(axis=0),
            "R_d_std": Rd.std(axis=0), "T_d": Td.mean(axis=0), "T_d_std": Td.std(axis=0),
            "samples_R_d": Rd, "samples_T_d": Td, "draws": draws, "N": N, "runs": runs,
            "kernel_ms": res[0].kernel_ms, "total_packets": N * len(cfgs)}


def adaptive_sweep(cfgs: List[SampleConfig], rse=0.005, pilot=10_000, n_min=10_000,
                   n_max=1_000_000, seed=1234, engine: Optional[Engine] = None,
                   chunk_packets=0, blocks_per_cu=0, autotune=False, topups=3, margin=1.05):
    """Uneven packet counts per configuration: a pilot of `pilot` packets
    estimates total reflectance R = R_d + R_s and its per-packet variance
    (a packet leaves its weight in exactly one exit tally, so
    Var(R_d + R_s) = E[R_d^2] + E[R_s^2] - R^2); each point then gets
    N = clamp(ceil(Var / (R rse)^2), n_min, n_max) packets in total, the
    remaining N - pilot in one launch on packet ids after the pilot's.
    The pilot's estimate is noisy, so a point whose relative standard error
    of R is still above `rse` after that launch, and that has not reached
    n_max, is topped up: up to `topups` more launches, each running every
    such point's shortfall, ceil(margin Var / (R rse)^2) - N (at most
    n_max - N), on the packet ids after its last ones, re-estimated from all
    its packets so far.  The configurations go to the library once, as an
    array, and the tallies come back as arrays (no per-point objects until
    the result list)."""
    eng = engine or default_engine()
    npts = len(cfgs)
    streams = np.arange(npts, dtype=np.uint32)
    arr = np.frombuffer(b"".join(bytes(c.to_c()) for c in cfgs), dtype=CONFIG_DTYPE).copy()

    events = {}

    def launch(n_packets, first=None):
        kw = {"chunk_packets": chunk_packets, "blocks_per_cu": blocks_per_cu}
        if autotune:
            kw = eng.tuner.params(cfgs, n_packets, seed=seed)
        out = eng.simulate_batch_array(arr, n_packets, seed=seed, streams=streams, first_packet_ids=first, **kw)
        if hasattr(eng, "last_events"):
            for k, v in eng.last_events().items():
                events[k] = events.get(k, 0) + v
        return out, eng.last_launch()["kernel_ms"]

    iRd, iRs = TALLIES.index("R_d"), TALLIES.index("R_s")

    def rse_of(S, Q, tot):
        Rm = S[:, iRd] / tot + S[:, iRs] / tot
        varR = (Q[:, iRd] + Q[:, iRs]) / tot - Rm * Rm
        with np.errstate(divide="ignore", invalid="ignore"):
            r = np.where(Rm > 0, np.sqrt(np.maximum(varR, 0) / tot) / Rm, float("inf"))
        return r, Rm, varR

    (s1, q1, n1), pilot_ms = launch(pilot)
    n1 = n1.astype(np.float64)
    # the plan, in the scalar form's order of operations (same packet counts)
    R = s1[:, iRd] / n1 + s1[:, iRs] / n1
    var = (q1[:, iRd] + q1[:, iRs]) / n1 - R * R
    with np.errstate(divide="ignore", invalid="ignore"):
        n = np.where(R <= 0, float(n_max), np.ceil(np.maximum(var, 0.0) / (R * rse) ** 2))
    need = np.clip(n, n_min, n_max).astype(np.int64)
    extra = np.maximum(need - pilot, 0)
    if extra.sum() > 0:
        (s2, q2, n2), main_ms = launch(extra.astype(np.uint64), first=np.full(npts, pilot, dtype=np.uint64))
    else:
        s2, q2, n2, main_ms = np.zeros_like(s1), np.zeros_like(q1), np.zeros(npts, dtype=np.uint64), 0.0
    S, Q = s1 + s2, q1 + q2
    done = n1.astype(np.int64) + n2.astype(np.int64)
    topup_ms, topup_packets, topup_launches = 0.0, 0, 0
    for _ in range(int(topups)):
        r, Rm, varR = rse_of(S, Q, done.astype(np.float64))
        short = (r > rse * (1 + 1e-9)) & (done < n_max) & (Rm > 0)
        if not short.any():
            break
        with np.errstate(divide="ignore", invalid="ignore"):
            want = np.ceil(margin * np.maximum(varR, 0.0) / (Rm * rse) ** 2)
        add = np.where(short, np.clip(want - done, 1, n_max - done), 0).astype(np.int64)
        (s3, q3, n3), ms = launch(add.astype(np.uint64), first=done.astype(np.uint64))
        S, Q = S + s3, Q + q3
        done = done + n3.astype(np.int64)
        topup_ms += ms
        topup_packets += int(add.sum())
        topup_launches += 1
    tot = done.astype(np.float64)[:, None]
    mean = S / tot
    se = np.sqrt(np.maximum(Q / tot - mean ** 2, 0.0) / np.maximum(tot - 1, 1))
    rse_R, _, _ = rse_of(S, Q, tot[:, 0])
    mean_l, se_l, tot_l, rse_l = mean.tolist(), se.tolist(), done.tolist(), rse_R.tolist()
    out = [{"n_packets": tot_l[k], "mean": dict(zip(TALLIES, mean_l[k])), "se": dict(zip(TALLIES, se_l[k])),
            "rse_R": rse_l[k]} for k in range(npts)]
    capped = done >= n_max
    info = {"pilot_kernel_ms": pilot_ms, "main_kernel_ms": main_ms, "topup_kernel_ms": topup_ms,
            "topup_launches": topup_launches, "topup_packets": topup_packets,
            "total_packets": int(done.sum()), "planned": need, "capped_points": int(capped.sum()),
            "max_rse_R_uncapped": float(rse_R[~capped].max()) if (~capped).any() else 0.0,
            "events": events}
    return out, info
