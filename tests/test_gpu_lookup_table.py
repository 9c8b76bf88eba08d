"""Lookup-table uncertainty map on the GPU (ismc_table_uncertainty) against
the oracle's literal restatement of lookup.py:347-376."""
import numpy as np
import pytest

from ismc_amd import lookup, lookup_table as LT
from ismc_amd.workloads import standard_sphere
from oracle import table_ref

pytestmark = pytest.mark.gpu


def _tables(P, engine):
    # a real simulated grid (single layer, dual spheres), 12 x 12 points
    res = lookup.sample_space(1.4, 0.8, 0.1, 0.1 * 1.5 ** np.arange(12), 2.0 * 1.5 ** np.arange(12),
                              standard_sphere(), standard_sphere(), N=20000, runs=3, seed=3, engine=engine)
    std_R = res.runs_R_d.std(axis=0)
    std_T = res.runs_T_d.std(axis=0)
    return LT.make_tables(res.R_d, res.T_d, std_R, std_T, res.mua, res.mus_reduced, P=P)


def test_uncertainty_matches_oracle(engine):
    tabs = _tables(80, engine)
    ua, us = LT.table_uncertainty(tabs, M=50, seed=21)
    ra, rs = table_ref.uncertainty(tabs.mua, tabs.mus, tabs.rstd, tabs.tstd, tabs.Rvals, tabs.Tvals, 50, 21)
    assert np.array_equal(np.isnan(ua), np.isnan(ra)) and np.array_equal(np.isnan(us), np.isnan(rs))
    ok = np.isfinite(ra)
    assert ok.sum() > 500
    # identical draws and lookups; the one-pass shifted variance differs from
    # numpy's two-pass std only in rounding (numpy's std of identical values
    # is ~1e-14 %, the kernel's exactly 0)
    da = np.abs(ua[ok] - ra[ok])
    ds = np.abs(us[ok] - rs[ok])
    print("cells %d, max abs diff %.2e / %.2e (percent)" % (ok.sum(), da.max(), ds.max()))
    assert np.mean(da <= 1e-9 * np.abs(ra[ok]) + 1e-10) > 0.999
    assert np.mean(ds <= 1e-9 * np.abs(rs[ok]) + 1e-10) > 0.999


def test_full_size_table(engine):
    tabs = _tables(1000, engine)
    ua, us = LT.table_uncertainty(tabs, M=200, seed=5)
    valid = np.isfinite(tabs.rstd) & np.isfinite(tabs.tstd) & (tabs.rstd >= 0) & (tabs.tstd >= 0)
    assert np.all(np.isnan(ua[~valid]))
    # 100 std / centre: the sign of the centre value (cubic interpolation can
    # undershoot below zero near the hull, there the reference's ratio is negative too)
    fin = np.isfinite(ua)
    assert fin.sum() > 1000
    assert np.all((ua[fin] >= 0) == (tabs.mua[fin] > 0) | (ua[fin] == 0))
    rng = np.random.default_rng(0)
    idx = np.argwhere(valid & np.isfinite(ua))
    cells = [tuple(c) for c in idx[rng.choice(len(idx), 60, replace=False)]]
    ra, rs = table_ref.uncertainty(tabs.mua, tabs.mus, tabs.rstd, tabs.tstd, tabs.Rvals, tabs.Tvals,
                                   200, 5, cells=cells)
    for c in cells:
        assert ua[c] == pytest.approx(ra[c], rel=1e-9, abs=1e-10)
        assert us[c] == pytest.approx(rs[c], rel=1e-9, abs=1e-10)


def test_run_from_ini(engine, tmp_path):
    import os
    from ismc_amd import ini as ini_mod
    here = os.path.dirname(os.path.abspath(__file__))
    f = ini_mod.load(os.path.join(here, "golden", "sample_lookup.ini"))
    out = LT.run(f, seed=3, engine=engine, rng_seed=1, out_prefix=str(tmp_path / "TEST"))
    tabs = out["tables"]
    assert tabs.mua.shape == (100, 100) and out["mua_unc"].shape == (100, 100)
    assert len(out["measurements"]) == 2
    assert all(len(m["lines"]) == 2 for m in out["measurements"])
    d = np.load(str(tmp_path / "TEST_lookup_tables.npz"), allow_pickle=False)
    for key in ("mua_lookup", "mus_lookup", "mua_lookup_uncertainty", "mus_lookup_uncertainty", "Rvals",
                "Tvals", "mua_points", "muas"):
        assert key in d
    assert np.array_equal(d["mua_lookup"], tabs.mua.T, equal_nan=True)
    # a simulated point read back through its own table cell
    sw = out["sweep"]
    r, t = 100 * sw["R_d"][1, 2], 100 * sw["T_d"][1, 2]
    a, _, _ = LT.lookup(tabs.mua, tabs.Rvals, tabs.Tvals, r, t)
    if np.isfinite(a):
        assert abs(a / sw["mua"][1] - 1) < 0.15
