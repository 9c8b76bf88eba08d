"""Lookup-table construction and lookups (CPU): the oracle's Philox and
loop, the O(1) nearest-index rule, griddata tables from a known forward
model, and the measurement formula."""
import numpy as np
import pytest

from ismc_amd import lookup_table as LT
from oracle import table_ref
from test_oracle_kat import KAT


@pytest.mark.parametrize("ctr,key,expected", KAT)
def test_numpy_philox_kat(ctr, key, expected):
    out = table_ref.philox4x32(*[np.array([c]) for c in ctr], *key)
    assert tuple(int(x[0]) for x in out) == expected


def test_normals_are_standard_normal():
    z = np.concatenate([np.concatenate(table_ref.normals(c, 200, 99)) for c in range(200)])
    assert abs(z.mean()) < 0.01 and abs(z.std() - 1) < 0.01
    # odd M: the last pair's second sample is dropped
    a, b = table_ref.normals(5, 7, 3)
    assert a.size == 7 and b.size == 7


@pytest.mark.parametrize("P", [2, 3, 7, 101, 1000])
def test_nearest_index_is_argmin(P):
    g = np.linspace(0, 100, P)
    rng = np.random.default_rng(P)
    h = 100 / (P - 1)
    v = np.concatenate([rng.uniform(-20, 120, 5000), g, (g[:-1] + g[1:]) / 2,
                        g + 1e-13, g - 1e-13, [-1e9, 1e9, 0.0, 100.0, 50.0 + h / 2]])
    fast = LT.nearest_index(g, v)
    slow = np.array([table_ref.find_nearest(g, x) for x in v])
    assert np.array_equal(fast, slow)


def _forward(a, s):
    """A smooth, monotone stand-in for the transport: R, T in (0, 1)."""
    R = 0.5 * s / (s + 3 * a + 2)
    T = 0.6 * np.exp(-0.3 * a - 0.02 * s)
    return R, T


def test_tables_invert_a_known_forward_model():
    mua = 0.2 * 1.25 ** np.arange(16)
    mus = 2.0 * 1.25 ** np.arange(16)
    A, S = np.meshgrid(mua, mus, indexing="ij")
    R, T = _forward(A, S)
    tabs = LT.make_tables(R, T, 0.01 * R, 0.01 * T, mua, mus, P=400)
    # held-out points inside the simulated range come back close to their inputs
    rng = np.random.default_rng(0)
    a = np.exp(rng.uniform(np.log(mua[2]), np.log(mua[-3]), 200))
    s = np.exp(rng.uniform(np.log(mus[2]), np.log(mus[-3]), 200))
    r, t = _forward(a, s)
    va, _, _ = LT.lookup(tabs.mua, tabs.Rvals, tabs.Tvals, 100 * r, 100 * t)
    vs, _, _ = LT.lookup(tabs.mus, tabs.Rvals, tabs.Tvals, 100 * r, 100 * t)
    ok = np.isfinite(va) & np.isfinite(vs)
    assert ok.mean() > 0.9
    assert np.median(np.abs(va[ok] / a[ok] - 1)) < 0.05
    assert np.median(np.abs(vs[ok] / s[ok] - 1)) < 0.05
    # outside the convex hull of the simulated points: NaN (griddata)
    assert np.isnan(tabs.mua[-1, -1])
    # skip_first drops the first row and column of simulated points
    t2 = LT.make_tables(R, T, 0.01 * R, 0.01 * T, mua, mus, P=50, skip_first=True)
    assert t2.mua_points.shape[0] == 15 * 15 and tabs.mua_points.shape[0] == 16 * 16
    assert tabs.reference_orientation()[0].shape == (400, 400)


def test_oracle_loop_closed_forms():
    P = 41
    g = np.linspace(0, 100, P)
    const = np.full((P, P), 3.0)
    rstd = np.full((P, P), 4.0)
    tstd = np.full((P, P), 2.0)
    rstd[0, 0] = np.nan
    tstd[1, 1] = -1.0
    ua, us = table_ref.uncertainty(const, 2 * const, rstd, tstd, g, g, 64, 5,
                                   cells=[(0, 0), (1, 1), (20, 20)])
    assert np.isnan(ua[0, 0]) and np.isnan(ua[1, 1]) and ua[20, 20] == 0.0 and us[20, 20] == 0.0
    # a table linear in R: std of the looked-up values ~ slope * R_std (grid step 2.5
    # rounds the draws; with std 10 the rounding adds ~2.5/sqrt(12) in quadrature)
    lin = np.tile(10.0 + 0.5 * g[:, None], (1, P))
    r10 = np.full((P, P), 10.0)
    ua, _ = table_ref.uncertainty(lin, lin, r10, r10, g, g, 4000, 7, cells=[(20, 20)])
    centre = 10.0 + 0.5 * g[20]
    expect = 100 * 0.5 * np.sqrt(10.0 ** 2 + 2.5 ** 2 / 12) / centre
    assert ua[20, 20] == pytest.approx(expect, rel=0.05)


def test_measure_combines_uncertainties():
    P = 101
    g = np.linspace(0, 100, P)
    tabs = LT.Tables(mua=np.full((P, P), 2.0), mus=np.full((P, P), 20.0), rstd=np.zeros((P, P)),
                     tstd=np.zeros((P, P)), Rvals=g, Tvals=g.copy(), mua_points=np.zeros((1, 2)),
                     mus_points=np.zeros((1, 2)))
    ua = np.full((P, P), 3.0)
    us = np.full((P, P), 4.0)
    r = LT.measure(tabs, ua, us, 30.0, 1.0, 20.0, 1.0, U=500, rng=np.random.default_rng(1))
    # constant tables: no spread from the draws, only the table's own uncertainty
    assert r["mu_a"] == 2.0 and r["mu_a_unc"] == pytest.approx(0.06)
    assert r["mus_reduced"] == 20.0 and r["mus_reduced_unc"] == pytest.approx(0.8)
    assert r["lines"][0].startswith("mua = 2.000000 +/- 0.060000 for R_d = 30.000000")


def test_c_entry_point_rejects_bad_arguments():
    # argument checks run before any HIP call, so this needs no GPU
    from ismc_amd import _lib
    L = _lib.load()
    P = 8
    z = np.zeros((P, P))
    g = np.linspace(0, 100, P)
    out = np.zeros((P, P))
    args = lambda grid, p, m: (0, z.ctypes.data, z.ctypes.data, z.ctypes.data, z.ctypes.data,
                               grid.ctypes.data, g.ctypes.data, p, m, 1, out.ctypes.data, out.ctypes.data)
    assert L.ismc_table_uncertainty(*args(g, 1, 10)) == _lib.EINVAL
    assert L.ismc_table_uncertainty(*args(g, P, 0)) == _lib.EINVAL
    other = np.linspace(0, 50, P)
    assert L.ismc_table_uncertainty(*args(other, P, 10)) == _lib.EINVAL
    assert b"same increasing uniform grid" in L.ismc_table_last_error()
