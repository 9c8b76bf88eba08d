"""Lookup tables (R_d, T_d) -> (mu_a, mu_s') from a simulated grid, their
uncertainty maps, and lookups of measurements (Lookup/lookup.py:296-560).

    sweep = lookup.run_lookup(ini)                       # simulated grid (GPU)
    tabs = make_tables(sweep["R_d"], sweep["T_d"], sweep["R_d_std"], sweep["T_d_std"],
                       sweep["mua"], sweep["mus_reduced"], P=1000)
    mua_unc, mus_unc = table_uncertainty(tabs, M=200)    # GPU
    res = measure(tabs, mua_unc, mus_unc, R_avg, R_unc, T_avg, T_unc, U=1000)

Tables hold R and T in percent on the grid linspace(0, 100, P) and are
indexed [i (R)][j (T)] internally; `reference_orientation` gives the
transposed [T][R] arrays that makelookup returns (lookup.py:486).
"""
import math
from dataclasses import dataclass
from typing import Optional

import numpy as np

from . import _lib


@dataclass
class Tables:
    mua: np.ndarray        # [P][P] mu_a at (R_i, T_j); NaN outside the simulated hull
    mus: np.ndarray        # [P][P] mu_s'
    rstd: np.ndarray       # [P][P] interpolated std of R_d (percent)
    tstd: np.ndarray       # [P][P] interpolated std of T_d (percent)
    Rvals: np.ndarray      # [P] grid, percent
    Tvals: np.ndarray
    mua_points: np.ndarray  # [n, 2] simulated (R, T) points, percent, sorted by (R, T)
    mus_points: np.ndarray

    def reference_orientation(self):
        """(mua, mus) as makelookup returns them: [T][R]."""
        return self.mua.T, self.mus.T


def make_tables(refl, trans, refl_std, trans_std, mu_a, mu_s_, P=1000, method="cubic",
                skip_first=False) -> Tables:
    """Scattered simulation points (100 R, 100 T) -> value, sorted by R then
    T, interpolated onto a P x P grid over [0, 100]^2 with scipy's griddata
    (lookup.py:300-345).  refl[i][j] is the simulation at mu_a[i],
    mu_s_[j].  makelookup skips row 0 and column 0 of the simulated grid
    (lookup.py:306-311), dropping the smallest mu_a and mu_s' points;
    skip_first=True reproduces that, the default keeps every point."""
    from scipy import interpolate
    refl, trans = np.asarray(refl, dtype=np.float64), np.asarray(trans, dtype=np.float64)
    refl_std, trans_std = np.asarray(refl_std, dtype=np.float64), np.asarray(trans_std, dtype=np.float64)
    mu_a, mu_s_ = np.asarray(mu_a, dtype=np.float64), np.asarray(mu_s_, dtype=np.float64)
    r0 = 1 if skip_first else 0
    R = (refl[r0:, r0:] * 100).ravel()
    T = (trans[r0:, r0:] * 100).ravel()
    A = np.broadcast_to(mu_a[r0:, None], refl[r0:, r0:].shape).ravel()
    S = np.broadcast_to(mu_s_[None, r0:], refl[r0:, r0:].shape).ravel()
    RS = (refl_std[r0:, r0:] * 100).ravel()
    TS = (trans_std[r0:, r0:] * 100).ravel()
    order = np.lexsort((T, R))          # stable: by R, then T
    pts = np.stack([R[order], T[order]], 1)
    gx, gy = np.mgrid[0:100:complex(P), 0:100:complex(P)]
    interp = lambda v: interpolate.griddata(pts, v[order], (gx, gy), method=method)
    vals = np.linspace(0, 100, P)
    return Tables(mua=interp(A), mus=interp(S), rstd=interp(RS), tstd=interp(TS),
                  Rvals=vals, Tvals=vals.copy(), mua_points=pts, mus_points=pts.copy())


def table_uncertainty(tabs: Tables, M=200, seed=1234, device=0):
    """Percent uncertainty of mu_a and mu_s' per cell (lookup.py:347-376),
    on the GPU (ismc_table_uncertainty).  Returns ([P][P], [P][P]) indexed
    [i (R)][j (T)]; NaN where the reference leaves NaN."""
    L = _lib.load()
    P = tabs.Rvals.size
    arrs = [np.ascontiguousarray(x, dtype=np.float64) for x in
            (tabs.mua, tabs.mus, tabs.rstd, tabs.tstd, tabs.Rvals, tabs.Tvals)]
    ua = np.empty((P, P))
    us = np.empty((P, P))
    rc = L.ismc_table_uncertainty(int(device), *[a.ctypes.data for a in arrs], int(P), int(M),
                                  int(seed), ua.ctypes.data, us.ctypes.data)
    if rc != _lib.OK:
        raise RuntimeError("ismc_table_uncertainty: %s" % L.ismc_table_last_error().decode())
    return ua, us


def nearest_index(grid, v):
    """argmin_k |grid[k] - v| for a uniform increasing grid (first index on
    ties, like find_nearest, lookup.py:296-299), vectorized and O(1) per
    value: round on the grid, then compare the neighbours' actual
    distances."""
    grid = np.asarray(grid, dtype=np.float64)
    v = np.asarray(v, dtype=np.float64)
    P = grid.size
    step = (grid[-1] - grid[0]) / (P - 1)
    k0 = np.clip((v - grid[0]) / step + 0.5, 0, P - 1).astype(np.int64)
    best = np.maximum(k0 - 1, 0)
    bd = np.abs(grid[best] - v)
    for k in (k0, np.minimum(k0 + 1, P - 1)):
        d = np.abs(grid[k] - v)
        better = d < bd
        best = np.where(better, k, best)
        bd = np.where(better, d, bd)
    return best


def lookup(table_RT, Rvals, Tvals, R, T):
    """Value of an [R][T] table at the grid point nearest (R, T) (percent),
    with the grid values used (lookup.py:433-457)."""
    i = nearest_index(Rvals, R)
    j = nearest_index(Tvals, T)
    return table_RT[i, j], Rvals[i], Tvals[j]


def measure(tabs: Tables, mua_unc, mus_unc, R_avg, R_unc, T_avg, T_unc, U=1000, rng=None):
    """One measurement (percent) through the tables (lookup.py:520-560):
    U draws R ~ N(R_avg, R_unc), T ~ N(T_avg, T_unc) looked up in the mu_a
    and mu_s' tables; the result is their mean, with a percent uncertainty
    combining the table's own (at the grid point nearest the measurement)
    and the spread of the draws.  The reference reads the table uncertainty
    at the last draw instead of the measurement itself; here it is read at
    the measurement."""
    rng = rng if rng is not None else np.random.default_rng()
    R = rng.normal(R_avg, R_unc, U)
    T = rng.normal(T_avg, T_unc, U)
    a, _, _ = lookup(tabs.mua, tabs.Rvals, tabs.Tvals, R, T)
    s, _, _ = lookup(tabs.mus, tabs.Rvals, tabs.Tvals, R, T)
    ua1, _, _ = lookup(mua_unc, tabs.Rvals, tabs.Tvals, R_avg, T_avg)
    us1, _, _ = lookup(mus_unc, tabs.Rvals, tabs.Tvals, R_avg, T_avg)
    mua, mus = float(np.mean(a)), float(np.mean(s))
    ua2 = float(np.std(a)) * 100 / mua
    us2 = float(np.std(s)) * 100 / mus
    ua = math.sqrt(float(ua1) ** 2 + ua2 ** 2)
    us = math.sqrt(float(us1) ** 2 + us2 ** 2)
    return {"mu_a": mua, "mu_a_unc": mua * ua / 100, "mus_reduced": mus, "mus_reduced_unc": mus * us / 100,
            "lines": ["mua = %f +/- %f for R_d = %f +/- %f and T_d = %f +/- %f"
                      % (mua, mua * ua / 100, R_avg, R_unc, T_avg, T_unc),
                      "mus' = %f +/- %f for R_d = %f and +/- %f T_d = %f +/- %f"
                      % (mus, mus * us / 100, R_avg, R_unc, T_avg, T_unc)]}


def save(path, tabs: Tables, mua_unc, mus_unc, extra: Optional[dict] = None):
    """The arrays lookup.py writes to <prefix>_lookup_tables.h5 (lookup.py:511),
    same keys and [T][R] orientation, as an .npz (h5py is not a dependency)."""
    d = {"mua_lookup": tabs.mua.T, "mus_lookup": tabs.mus.T, "mua_lookup_uncertainty": mua_unc.T,
         "mus_lookup_uncertainty": mus_unc.T, "Rvals": tabs.Rvals, "Tvals": tabs.Tvals,
         "mua_points": tabs.mua_points, "mus_points": tabs.mus_points}
    d.update(extra or {})
    np.savez(path, **d)


def run(ini, seed=1234, engine=None, rng_seed=None, out_prefix=None):
    """lookup.py's whole run (lookup.py:491-560) from an .ini: the simulated
    grid over every input-uncertainty sample (one batched launch), the P x P
    tables (P = [lookup] dimensions), their uncertainty maps (M =
    simulation_distributions_samples draws per cell), and every
    [experiment] measurement looked up with U =
    measurement_distributions_samples draws.  Measurements are fractions in
    the .ini and percent in the tables, as in the reference."""
    from . import lookup as sweep_mod
    L = ini.lookup
    P = int(L.get("dimensions", 1000))
    M = int(L.get("simulation_distributions_samples", 200))
    U = int(L.get("measurement_distributions_samples", 1000))
    sweep = sweep_mod.run_lookup(ini, seed=seed, engine=engine, rng_seed=rng_seed)
    tabs = make_tables(sweep["R_d"], sweep["T_d"], sweep["R_d_std"], sweep["T_d_std"], sweep["mua"],
                       sweep["mus_reduced"], P=P)
    ua, us = table_uncertainty(tabs, M=M, seed=seed, device=engine.device if engine is not None else 0)
    rng = np.random.default_rng(rng_seed)
    results = []
    if ini.M_R is not None and ini.M_T is not None:
        n = max(len(ini.M_R.values), len(ini.M_T.values))
        # a single value (or uncertainty) applies to every measurement; a
        # missing uncertainty is 0
        pick = lambda xs, k: xs[k] if len(xs) > 1 else (xs[0] if xs else 0.0)
        for k in range(n):
            results.append(measure(tabs, ua, us, 100 * pick(ini.M_R.values, k), 100 * pick(ini.M_R.uncs, k),
                                   100 * pick(ini.M_T.values, k), 100 * pick(ini.M_T.uncs, k), U=U, rng=rng))
    if out_prefix:
        save(out_prefix + "_lookup_tables.npz", tabs, ua, us,
             extra={"muas": sweep["mua"], "muss": sweep["mus_reduced"], "n": ini.n, "n_unc": ini.n_unc,
                    "g": ini.g, "t": ini.t})
    return {"tables": tabs, "mua_unc": ua, "mus_unc": us, "measurements": results, "sweep": sweep}
