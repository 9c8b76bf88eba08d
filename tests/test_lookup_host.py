"""Lookup driver host logic (CPU): .ini loading, grid, statistics and the
CSV table layout of write_data (CUDA/ISMLMC-lookup.cu:113-150)."""
import math
import os

import numpy as np
import pytest

from ismc_amd import ini as ini_mod
from ismc_amd import lookup

HERE = os.path.dirname(os.path.abspath(__file__))
INI = os.path.join(HERE, "golden", "sample_lookup.ini")


def test_ini_loader_keys_and_units():
    f = ini_mod.load(INI)
    assert (f.n, f.g, f.t) == (1.40, 0.8, 0.1)
    assert f.R.present and f.T.present and not f.separate
    assert f.M_R.values == [0.30, 0.25] and f.M_R.first_unc == 0.003
    assert f.M_U.first == 0.0            # a measured 0.0 stays measured
    assert f.M_S is None
    mua, mus = f.lookup_grid()
    assert mua == [0.5, 1.0, 2.0] and mus == [5.0, 15.0, 45.0, 135.0]   # mus_base used for mu_s'
    assert f.lookup_photons() == (20000, 2)
    T = f.T.sphere()
    assert T.angle == pytest.approx(math.radians(0.5))   # degrees converted
    cfg = f.sample(1.0, 15.0)
    assert cfg.layers[0].mu_s == pytest.approx(75.0)      # mu_s' / (1 - g)


def test_lookup_photons_from_uncertainty():
    txt = open(INI).read().replace("N           = 20000\nruns        = 2\n",
                                   "uncertainty = 0.001\nestimate = 0.5\n")
    f = ini_mod.load(txt)
    assert f.lookup_photons() == (500000, 0)


def test_statistics_match_reference_formulas():
    meta = np.array([[[1.0, 2.0]], [[3.0, 6.0]]])
    # array_stderr: sqrt(sum (x - mean)^2 / runs) / sqrt(runs)
    # means (2, 4); squared deviations summed (2, 8)
    assert np.allclose(lookup.array_stderr(meta), [[math.sqrt(2 / 2) / math.sqrt(2), math.sqrt(8 / 2) / math.sqrt(2)]])
    assert np.allclose(lookup.array_poisson(meta, 100), np.sqrt(meta[-1] / 100))


def test_csv_layout(tmp_path):
    mua = [0.5, 1.0]
    mus = [10.0, 12.0, 14.4]
    d = np.array([[0.1, 0.2, 0.3], [0.4, 0.5, 1.25e-5]])
    p = tmp_path / "t.csv"
    lookup.write_csv(str(p), "Average Diffuse Reflectance", mua, mus, d, True)
    lookup.write_csv(str(p), "Standard Error in Diffuse Reflectance", mua, mus, d / 10, False)
    lines = open(p).read().splitlines()
    assert lines[0] == '"Average Diffuse Reflectance","mua","","",""'
    assert lines[1] == '"mus\'","",10,12,14.4'
    assert lines[2] == '"",0.5,0.1,0.2,0.3'
    assert lines[3] == '"",1,0.4,0.5,1.25e-05'
    assert lines[4].startswith('"Standard Error in Diffuse Reflectance"')
    m, s = lookup.read_csv_table(str(p), 2)
    assert np.allclose(m, d, rtol=1e-6) and np.allclose(s, d / 10, rtol=1e-6)


class BernoulliEngine:
    """Stand-in for the batch launch: every packet leaves weight 1 in R_d
    with probability p_R and in R_s with probability 0.02 (so sum == sum of
    squares), p_R depending on the point's mu_a; records each launch."""

    def __init__(self, p_R):
        self.p_R = np.asarray(p_R, dtype=np.float64)
        self.launches = []

    def simulate_batch_array(self, arr, n_packets, seed=0, streams=None, first_packet_ids=None, **kw):
        npts = arr.shape[0]
        n = np.broadcast_to(np.asarray(n_packets, dtype=np.uint64), (npts,)).copy()
        first = None if first_packet_ids is None else np.asarray(first_packet_ids).copy()
        self.launches.append({"n": n, "first": first, "kw": kw, "mu_a": arr["mu_a"][:, 0].copy()})
        s = np.zeros((npts, 5))
        s[:, 1] = self.p_R * n
        s[:, 2] = 0.02 * n * (self.p_R > 0)
        return s, s.copy(), n

    def last_launch(self):
        return {"kernel_ms": 1.5}


class BiasedPilotEngine(BernoulliEngine):
    """The same stand-in, except that the first launch (the pilot) reports
    R_d 1.6 times too high: the plan then asks too few packets and the
    statistics of pilot + main miss the target."""

    def simulate_batch_array(self, arr, n_packets, seed=0, streams=None, first_packet_ids=None, **kw):
        s, q, n = super().simulate_batch_array(arr, n_packets, seed, streams, first_packet_ids, **kw)
        if len(self.launches) == 1:
            s[:, 1] *= 1.6
            q[:, 1] *= 1.6
        return s, q, n


def test_adaptive_sweep_plan_and_statistics():
    """The pilot estimates R = R_d + R_s and its per-packet variance; each
    point gets N = clamp(ceil(Var / (R rse)^2), n_min, n_max) packets, the
    rest of them in a second launch on packet ids after the pilot's."""
    from ismc_amd.workloads import lookup_grid
    _, _, cfgs = lookup_grid(2, 3)
    p_R = np.array([0.3, 0.0, 0.05, 0.9, 0.6, 0.001])
    eng = BernoulliEngine(p_R)
    pilot, n_min, n_max, rse = 1000, 2000, 500_000, 0.01
    out, info = lookup.adaptive_sweep(cfgs, rse=rse, pilot=pilot, n_min=n_min, n_max=n_max, seed=3,
                                      engine=eng, chunk_packets=512)
    need = []
    for p in p_R:
        R = p + 0.02 * (p > 0)
        var = R - R * R                                   # weights 0/1 in one tally or none
        n = n_max if R <= 0 else math.ceil(max(var, 0.0) / (R * rse) ** 2)
        need.append(min(max(n, n_min), n_max))
    assert list(info["planned"]) == need
    assert len(eng.launches) == 2
    assert np.all(eng.launches[0]["n"] == pilot) and eng.launches[0]["first"] is None
    assert list(eng.launches[1]["n"]) == [max(k - pilot, 0) for k in need]
    assert np.all(eng.launches[1]["first"] == pilot)
    assert eng.launches[0]["kw"] == {"chunk_packets": 512, "blocks_per_cu": 0}
    assert info["total_packets"] == sum(need) and info["pilot_kernel_ms"] == 1.5
    for k, o in enumerate(out):
        R = p_R[k] + 0.02 * (p_R[k] > 0)
        assert o["n_packets"] == need[k]
        assert o["mean"]["R_d"] == pytest.approx(p_R[k], rel=1e-12)
        if R > 0:
            assert o["rse_R"] == pytest.approx(math.sqrt((R - R * R) / need[k]) / R, rel=1e-9)
            assert o["rse_R"] <= rse * 1.000001 or need[k] == n_max
        else:
            assert o["rse_R"] == float("inf") and need[k] == n_max


def test_adaptive_sweep_tops_up_points_short_of_the_target():
    """After the main launch a point whose measured RSE of R is above the
    target (and below the cap) gets its shortfall, ceil(margin Var /
    (R rse)^2) - N, in a further launch on the packet ids after its last
    ones; every uncapped point then meets the target."""
    from ismc_amd.workloads import lookup_grid
    _, _, cfgs = lookup_grid(2, 3)
    p_R = np.array([0.3, 0.05, 0.5, 0.9, 0.6, 0.01])
    eng = BiasedPilotEngine(p_R)
    rse, n_max = 0.01, 200_000
    out, info = lookup.adaptive_sweep(cfgs, rse=rse, pilot=1000, n_min=1000, n_max=n_max, seed=3, engine=eng)
    assert info["topup_launches"] >= 1 and len(eng.launches) == 2 + info["topup_launches"]
    main = eng.launches[1]
    top = eng.launches[2]
    done = 1000 + main["n"].astype(np.int64)
    assert np.all(top["first"] == done)
    assert info["topup_packets"] == sum(int(l["n"].sum()) for l in eng.launches[2:])
    for k, o in enumerate(out):
        assert o["n_packets"] == 1000 + sum(int(l["n"][k]) for l in eng.launches[1:])
        assert o["rse_R"] <= rse * (1 + 1e-9) or o["n_packets"] == n_max
    assert info["max_rse_R_uncapped"] <= rse * (1 + 1e-9)
    assert info["total_packets"] == sum(o["n_packets"] for o in out)
