"""Batched launches on the GPU: the lookup sweep driver, the adaptive
(pilot-sized) sweep, and hipGraph replay of a repeated batch."""
import os

import numpy as np
import pytest

from ismc_amd import ini as ini_mod
from ismc_amd import lookup
from ismc_amd.config import Sphere, mus_from_reduced
from ismc_amd.workloads import lookup_grid, single_layer, standard_sphere

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))


def test_sample_space_matches_single_runs(engine):
    R, T = standard_sphere(), standard_sphere()
    mua, mus = [0.1, 1.0], [10.0, 30.0, 90.0]
    res = lookup.sample_space(1.4, 0.8, 0.1, mua, mus, R, T, N=20000, runs=2, seed=5, engine=engine)
    assert res.R_d.shape == (2, 3) and res.runs_R_d.shape == (2, 2, 3)
    # point (i, j) of run r is batch entry (r * 6 + i * 3 + j) with that stream id
    for r in range(2):
        for i, a in enumerate(mua):
            for j, s in enumerate(mus):
                k = r * 6 + i * 3 + j
                one = engine.simulate(lookup._single(1.4, 0.8, 0.1, a, mus_from_reduced(s, 0.8), R, T, False, 1e-4),
                                      20000, seed=5, stream=k)
                assert res.runs_R_d[r, i, j] == one.mean["R_d"]
                assert res.runs_T_d[r, i, j] == one.mean["T_d"]
    assert np.all(res.R_d_unc >= 0) and np.all(res.R_d > 0)


def test_separate_spheres_and_no_spheres(engine):
    R, T = standard_sphere(), standard_sphere()
    sep = lookup.sample_space(1.4, 0.8, 0.1, [1.0], [10.0], R, T, N=50000, runs=0, separate=True,
                              seed=3, engine=engine)
    # runs == 0: Poisson uncertainty sqrt(x / N)
    assert np.allclose(sep.R_d_unc, np.sqrt(sep.R_d / 50000))
    none = lookup.sample_space(1.4, 0.8, 0.1, [1.0], [10.0], Sphere(False, angle=0.0),
                               Sphere(False, angle=0.0), N=50000, runs=1, seed=3, engine=engine)
    # without spheres the table holds total R and T (ISMLMC-lookup.cu:299-306)
    assert none.R_d[0, 0] > sep.R_d[0, 0] * 0.5
    assert 0 < none.R_d[0, 0] + none.T_d[0, 0] <= 1


def test_graph_replay_is_bitwise_identical(engine):
    cfgs = [single_layer(a, s) for a in (0.1, 1.0, 5.0) for s in (10.0, 100.0)]
    eager = engine.simulate_batch(cfgs, 30000, seed=17)
    before = engine.graph_stats()
    for rep in range(4):
        g = engine.simulate_batch(cfgs, 30000, seed=17, use_graph=True)
        assert [x.sum for x in g] == [x.sum for x in eager]
        if rep == 0:
            # a shape's first call runs eagerly; it is captured when it comes back
            assert engine.graph_stats() == before
        # unrelated eager work between replays
        engine.simulate(single_layer(2.0, 50.0), 10000, seed=rep)
    # a replay with new configurations of the same shape reads them from staging
    cfgs2 = [single_layer(a * 2, s) for a in (0.1, 1.0, 5.0) for s in (10.0, 100.0)]
    g2 = engine.simulate_batch(cfgs2, 30000, seed=17, use_graph=True)
    e2 = engine.simulate_batch(cfgs2, 30000, seed=17)
    assert [x.sum for x in g2] == [x.sum for x in e2]
    after = engine.graph_stats()
    assert after["captures"] - before["captures"] == 1
    assert after["launches"] - before["launches"] == 4


def test_adaptive_sweep_reaches_target_rse(engine):
    mua, mus, cfgs = lookup_grid(4, 4)
    out, info = lookup.adaptive_sweep(cfgs, rse=0.005, pilot=10000, n_min=10000, n_max=1_000_000,
                                      seed=11, engine=engine)
    for o in out:
        assert o["rse_R"] <= 0.005 or o["n_packets"] == 1_000_000
        assert 10000 <= o["n_packets"] <= 1_000_000
    assert info["total_packets"] == sum(o["n_packets"] for o in out)
    assert info["max_rse_R_uncapped"] <= 0.005


def test_adaptive_sweep_matches_per_point_plan(engine):
    """The array form of the sweep plans and reports exactly what the
    per-point form does: a pilot launch, N = clamp(ceil(Var / (R rse)^2)),
    the rest on packet ids after the pilot's; and the autotuned launch
    shape changes nothing."""
    import math
    mua, mus, cfgs = lookup_grid(3, 4)
    pilot, n_min, n_max, rse = 5000, 8000, 400_000, 0.004
    out, info = lookup.adaptive_sweep(cfgs, rse=rse, pilot=pilot, n_min=n_min, n_max=n_max, seed=21, engine=engine,
                                      topups=0)
    tuned, _ = lookup.adaptive_sweep(cfgs, rse=rse, pilot=pilot, n_min=n_min, n_max=n_max, seed=21, engine=engine,
                                     autotune=True, topups=0)
    streams = np.arange(len(cfgs), dtype=np.uint32)
    p = engine.simulate_batch(cfgs, pilot, seed=21, streams=streams)
    need = []
    for r in p:
        R = r.mean["R_d"] + r.mean["R_s"]
        var = (r.sumsq["R_d"] + r.sumsq["R_s"]) / r.n_packets - R * R
        n = n_max if R <= 0 else int(math.ceil(max(var, 0.0) / (R * rse) ** 2))
        need.append(min(max(n, n_min), n_max))
    assert list(info["planned"]) == need
    extra = np.maximum(np.array(need) - pilot, 0).astype(np.uint64)
    m = engine.simulate_batch(cfgs, extra, seed=21, streams=streams,
                              first_packet_ids=np.full(len(cfgs), pilot, dtype=np.uint64))
    for k, o in enumerate(out):
        tot = p[k].n_packets + m[k].n_packets
        assert o["n_packets"] == tot == need[k]
        for t in o["mean"]:
            assert o["mean"][t] == (p[k].sum[t] + m[k].sum[t]) / tot
        assert tuned[k] == o
    assert info["total_packets"] == sum(need)


def test_adaptive_sweep_top_up_runs_the_next_packet_ids(engine):
    """Top-up launches run each short point's next packet ids: the sweep's
    tallies equal one run of the point's final packet count on its stream,
    bit for bit, and every point below the cap meets the target."""
    mua, mus, cfgs = lookup_grid(3, 4)
    out, info = lookup.adaptive_sweep(cfgs, rse=0.004, pilot=2000, n_min=2000, n_max=400_000, seed=23,
                                      engine=engine)
    assert info["topup_launches"] >= 1 and info["topup_packets"] > 0
    assert info["max_rse_R_uncapped"] <= 0.004
    for k, o in enumerate(out):
        assert o["rse_R"] <= 0.004 or o["n_packets"] == 400_000
        one = engine.simulate(cfgs[k], o["n_packets"], seed=23, stream=k)
        for t in one.mean:
            assert one.sum[t] / one.n_packets == o["mean"][t], (k, t)


def test_full_lookup_grid_sweep(engine):
    # The 64 x 64 lookup grid at its full size (mu_a 0.01-10, mu_s 1-1000,
    # adaptive 1e4-1e6 packets per point, 0.5 % RSE target), with
    # size-independent checks: every point reaches the target or the cap,
    # the packet total adds up, and sampled points equal single-point runs
    # of the same packet ids on the same stream, bit for bit.
    mua, mus, cfgs = lookup_grid(64, 64)
    out, info = lookup.adaptive_sweep(cfgs, rse=0.005, pilot=10000, n_min=10000, n_max=1_000_000,
                                      seed=4321, engine=engine, autotune=True)
    n = np.array([o["n_packets"] for o in out])
    rse = np.array([o["rse_R"] for o in out])
    assert np.all((rse <= 0.005) | (n == 1_000_000)) and np.all((n >= 10000) & (n <= 1_000_000))
    assert info["max_rse_R_uncapped"] <= 0.005 and info["capped_points"] == int((n == 1_000_000).sum())
    assert info["total_packets"] == n.sum() > 4096 * 10000
    for k in (0, 63, 2080, 4032, 4095):
        one = engine.simulate(cfgs[k], int(n[k]), seed=4321, stream=k)
        for t in one.mean:
            assert one.sum[t] / one.n_packets == out[k]["mean"][t], (k, t)


def test_run_lookup_from_ini(engine, tmp_path):
    f = ini_mod.load(os.path.join(HERE, "golden", "sample_lookup.ini"))
    out = lookup.run_lookup(f, seed=2, engine=engine, rng_seed=4)
    assert out["R_d"].shape == (3, 4) and out["samples_R_d"].shape == (3, 3, 4)
    assert np.all(out["R_d_std"] > 0)
    res = lookup.sample_space(f.n, f.g, f.t, *f.lookup_grid(), f.R.sphere(), f.T.sphere(),
                              N=20000, runs=2, engine=engine)
    pr, pt = lookup.write_tables(res, str(tmp_path))
    m, s = lookup.read_csv_table(pr, 3)
    assert np.allclose(m, res.R_d, rtol=1e-5) and np.allclose(s, res.R_d_unc, rtol=1e-5, atol=1e-12)


def test_autotuned_launch_is_bitwise_identical(engine):
    from ismc_amd.autotune import AutoTuner
    cfgs = [single_layer(a, s) for a in (0.05, 1.0, 8.0) for s in (5.0, 50.0, 400.0)]
    npk = np.array([20000, 3000, 90000, 1000, 50000, 7000, 40000, 2000, 60000], dtype=np.uint64)
    base = engine.simulate_batch(cfgs, npk, seed=77)
    tuner = AutoTuner(engine, budget_packets=200_000, chunks=(512, 4096, 65536), blocks=(0, 1),
                      idle=(4, 8), bound=(8, 16), repeats=1)
    kw = tuner.params(cfgs, npk)
    assert kw["chunk_packets"] in (512, 4096, 65536) and kw["blocks_per_cu"] in (0, 1)
    assert kw["idle_lanes"] in (4, 8)
    assert len(tuner.report(cfgs, npk)) == 6
    for b, c, i, bl in ((0, 512, 1, 1), (1, 65536, 64, 64), (1, 512, 16, 3), (0, 65536, 5, 40)):
        r = engine.simulate_batch(cfgs, npk, seed=77, blocks_per_cu=b, chunk_packets=c, idle_lanes=i,
                                  bound_lanes=bl)
        for x, y in zip(r, base):
            assert x.sum == y.sum and x.sumsq == y.sumsq and x.events == y.events
    r = engine.simulate_batch(cfgs, npk, seed=77, autotune=True)
    assert all(x.sum == y.sum for x, y in zip(r, base))


def test_many_chunks_per_wave_are_bitwise_identical(engine):
    """Shapes where every wave runs many small chunks of several points, so
    its two table slots change hands often: tallies and event counts equal
    the default launch's bit for bit."""
    cfgs = [single_layer(a, s) for a in (0.05, 1.0, 8.0) for s in (5.0, 50.0, 400.0)]
    npk = np.array([20000, 3000, 90000, 1000, 50000, 7000, 40000, 2000, 60000], dtype=np.uint64)
    base = engine.simulate_batch(cfgs, npk, seed=77)
    for b, c, i, bl in ((1, 0, 4, 8), (2, 0, 8, 16), (1, 64, 16, 8), (2, 128, 4, 16), (3, 64, 32, 8),
                        (4, 256, 8, 8)):
        r = engine.simulate_batch(cfgs, npk, seed=77, blocks_per_cu=b, chunk_packets=c, idle_lanes=i,
                                  bound_lanes=bl)
        bad = [k for k, (x, y) in enumerate(zip(r, base)) if x.sum != y.sum or x.events != y.events]
        assert not bad, "blocks/CU %d, chunk %d, idle %d, bound %d: points %s differ" % (b, c, i, bl, bad)


def test_points_without_packets_are_passed_over(engine):
    """A batch with points of 0 packets at the start, in the middle and at the
    end: those points tally nothing, and every other point equals a batch
    of the non-empty points alone (same streams), for single-chunk and
    many-chunk shapes."""
    cfgs = [single_layer(a, s) for a in (0.05, 1.0, 8.0) for s in (5.0, 50.0)]
    npk = np.array([0, 20000, 0, 0, 7000, 3000, 0], dtype=np.uint64)
    cfg7 = [cfgs[k % len(cfgs)] for k in range(len(npk))]
    streams = np.arange(len(npk), dtype=np.uint32) + 40
    nz = np.where(npk > 0)[0]
    ref = engine.simulate_batch([cfg7[k] for k in nz], npk[nz], seed=5, streams=streams[nz])
    for chunk in (0, 64, 100000):
        r = engine.simulate_batch(cfg7, npk, seed=5, streams=streams, chunk_packets=chunk)
        for k in range(len(npk)):
            if npk[k] == 0:
                assert r[k].n_packets == 0 and all(v == 0 for v in r[k].sum.values())
        for j, k in enumerate(nz):
            assert r[k].sum == ref[j].sum and r[k].events == ref[j].events, (chunk, k)
