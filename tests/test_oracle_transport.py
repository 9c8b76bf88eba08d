"""Oracle transport against closed forms, a literature known answer and
invariants (CPU).  These pin the composite propagator (MLMC.py:76-202,
ISMC.py:87-160, ISMLMC.py:101-347) where the mathematics fixes the answer."""
import math

import numpy as np
import pytest

from oracle import oracle as O

NO_SPHERE = {"present": 0, "angle": 0.0}


def slab(n, mu_a, mu_s, g, t, **kw):
    d = {"layers": [{"n": n, "g": g, "t": t, "mu_a": mu_a, "mu_s": mu_s}],
         "R": dict(NO_SPHERE), "T": dict(NO_SPHERE), "specular_included": 0}
    d.update(kw)
    return d


@pytest.mark.parametrize("flags", [0, O.MIRROR])
def test_absorbing_slab_closed_form(flags):
    # mu_s -> 0: every packet is absorbed whole at its first interaction, so
    # only the ballistic multiple-reflection series leaves the slab.
    n, mu_a, t = 1.5, 5.0, 0.1
    r = ((n - 1) / (n + 1)) ** 2
    E = math.exp(-mu_a * t)
    Tu = (1 - r) ** 2 * E / (1 - r * r * E * E)
    Rs = r + (1 - r) ** 2 * r * E * E / (1 - r * r * E * E)
    N = 200000
    res = O.simulate(slab(n, mu_a, 1e-9, 0.0, t), N, seed=11, flags=flags, threads=4)
    for key, exp in (("T_u", Tu), ("R_s", Rs), ("A", 1 - Tu - Rs)):
        se = math.sqrt(exp * (1 - exp) / N)
        assert abs(res["mean"][key] - exp) < 3 * se + 1e-6, (key, res["mean"][key], exp)
    assert res["mean"]["R_d"] < 1e-6 and res["mean"]["T_d"] < 1e-6


def test_matched_boundary_known_answer():
    # van de Hulst's tabulated case (n=1, albedo 0.9, optical thickness 2,
    # g=0.75), the classic MCML check, here with the reference's sample
    # definition of MC.py:22-26: mu_a=10, mu_s'=22.5 -> mu_s=90, t=0.02 cm.
    N = 1000000
    res = O.simulate(slab(1.0, 10.0, 90.0, 0.75, 0.02), N, seed=5, threads=8)
    m = res["mean"]
    # unscattered light is exactly the ballistic fraction exp(-mu_t t)
    Tb = math.exp(-2.0)
    assert abs(m["T_u"] - Tb) < 3 * math.sqrt(Tb * (1 - Tb) / N)
    Rt = m["R_d"] + m["R_s"]
    Tt = m["T_d"] + m["T_u"]
    # published: R_d = 0.09739, T_t = 0.66096 (van de Hulst 1980 / MCML)
    assert Rt == pytest.approx(0.09739, abs=0.0008)
    assert Tt == pytest.approx(0.66096, abs=0.0012)


def test_energy_conservation_per_packet():
    spec = {"layers": [{"n": 1.52, "g": 0, "t": .1, "mu_a": 0, "mu_s": 0},
                       {"n": 1.40, "g": .9, "t": .1, "mu_a": 1, "mu_s": 100},
                       {"n": 1.52, "g": 0, "t": .1, "mu_a": 0, "mu_s": 0}],
            "R": {"present": 1, "pw": .98, "fs": .0039, "fp": .0039, "fd": .001},
            "T": {"present": 1, "pw": .98, "fs": .0039, "fp": .0039, "fd": .001}}
    res = O.simulate(spec, 20000, seed=3, per_packet=True)
    tot = res["per_packet"][:, :5].sum(axis=1)
    # roulette keeps the expectation: mean 1 within its standard error
    assert abs(tot.mean() - 1) < 4 * tot.std() / math.sqrt(tot.size) + 1e-12
    # packets that were never rouletted carry exactly weight 1
    assert np.mean(np.abs(tot - 1) < 1e-9) > 0.5
    # CUDA-compat roulette (CUDAISMLMC.h:275, integer 1/m) only ever loses weight
    rc = O.simulate(spec, 20000, seed=3, per_packet=True, flags=O.CUDA_ROULETTE)
    tc = rc["per_packet"][:, :5].sum(axis=1)
    assert np.all(tc <= 1 + 1e-12) and np.all(tc >= 1 - 2e-4)


def test_mirror_equals_substep_per_packet():
    # ISMC.py's mirror boundary and MLMC.py's substep boundary are the same
    # geometry for one slab: with the same draws the histories coincide.
    spec = slab(1.4, 1.0, 100.0, 0.9, 0.1, roulette_threshold=1e-4,
                R={"present": 1, "pw": .98, "fs": .0039, "fp": .0039, "fd": .001},
                T={"present": 1, "pw": .98, "fs": .0039, "fp": .0039, "fd": .001})
    a = O.simulate(spec, 5000, seed=8, per_packet=True)["per_packet"]
    b = O.simulate(spec, 5000, seed=8, per_packet=True, flags=O.MIRROR)["per_packet"]
    same = np.all(np.abs(a - b) < 1e-9, axis=1)
    assert same.mean() > 0.999


def test_fp32_oracle_tracks_fp64():
    spec = slab(1.4, 1.0, 100.0, 0.9, 0.1)
    a = O.simulate(spec, 5000, seed=8, per_packet=True)["per_packet"]
    b = O.simulate(spec, 5000, seed=8, per_packet=True, flags=O.FP32)["per_packet"]
    same = np.all(np.abs(a[:, :5] - b[:, :5]) < 1e-4, axis=1)
    assert same.mean() > 0.97


def test_grazing_glass_entry_cannot_trap_a_packet():
    # Packet 44473 of seed 2024 in FP32 reaches the lower slide at grazing
    # incidence from the sample; without the glass-entry guard it bounces
    # between total internal reflection at the air face and a rounding-made
    # TIR at the sample face until the event cap stops it.
    spec = {"layers": [{"n": 1.52, "g": 0, "t": .1, "mu_a": 0, "mu_s": 0},
                       {"n": 1.40, "g": .9, "t": .1, "mu_a": 1, "mu_s": 100},
                       {"n": 1.52, "g": 0, "t": .1, "mu_a": 0, "mu_s": 0}],
            "R": {"present": 1, "pw": .98, "fs": 0.003921629, "fp": 0.003921629, "fd": 0.000977518},
            "T": {"present": 1, "pw": .98, "fs": 0.003921629, "fp": 0.003921629, "fd": 0.000977518}}
    r = O.simulate(spec, 1, seed=2024, first_pid=44473, flags=O.FP32, per_packet=True)
    assert r["n_aborted"] == 0 and r["per_packet"][0, 5] < 100000
    assert abs(r["per_packet"][0, :5].sum() - 1) < 1e-6 or r["per_packet"][0, :5].sum() < 1e-3
    big = O.simulate(spec, 2_000_000, seed=2024, flags=O.FP32, threads=8)
    assert big["n_aborted"] == 0


SPHERE = {"present": 1, "pw": .98, "fs": 0.003921629, "fp": 0.003921629, "fd": 0.000977518}
GLASS = {"n": 1.52, "g": 0, "t": .1, "mu_a": 0, "mu_s": 0}
SAMPLE = {"n": 1.40, "g": .9, "t": .1, "mu_a": 1, "mu_s": 100}


def _stack(*layers):
    return {"layers": list(layers), "R": dict(SPHERE), "T": dict(SPHERE)}


@pytest.mark.parametrize("name,spec", [
    ("slide|sample|slide", _stack(GLASS, SAMPLE, GLASS)),
    ("sample|slide|sample", _stack(dict(SAMPLE, t=.05), GLASS,
                                   {"n": 1.33, "g": .8, "t": .05, "mu_a": .5, "mu_s": 50})),
    ("dense sample|slide", _stack({"n": 1.60, "g": .7, "t": .1, "mu_a": .3, "mu_s": 30},
                                  dict(GLASS, n=1.45, t=.2))),
])
def test_slide_closed_form_matches_bounces(name, spec):
    # ORACLE_ETALON takes a slide's bounces as the geometric series that the
    # face-by-face simulation (the reference's, restated by default) sums
    # one bounce at a time: the tallies agree within 4 combined SE (15
    # comparisons) on independent streams, the interactions per packet too,
    # and the boundary events fall.
    N = 1_000_000
    a = O.simulate(spec, N, seed=31, threads=8)
    b = O.simulate(spec, N, seed=32, threads=8, flags=O.ETALON)
    for k in a["mean"]:
        se = math.hypot(a["se"][k], b["se"][k])
        assert abs(a["mean"][k] - b["mean"][k]) < 4 * se + 1e-12, (name, k, a["mean"][k], b["mean"][k], se)
    ia, ib = a["events"]["interactions"] / N, b["events"]["interactions"] / N
    assert abs(ia - ib) < 0.01 * ia, (ia, ib)
    assert b["events"]["boundary_hits"] < 0.8 * a["events"]["boundary_hits"]


def test_slide_closed_form_changes_nothing_without_slides():
    # no slide (no glass, or glass next to glass): per packet identical
    for spec in (_stack(SAMPLE), _stack(GLASS, dict(GLASS, n=1.45), SAMPLE),
                 _stack(dict(SAMPLE, t=.03), {"n": 1.44, "g": .8, "t": .05, "mu_a": 2, "mu_s": 80})):
        a = O.simulate(spec, 3000, seed=4, per_packet=True)["per_packet"]
        b = O.simulate(spec, 3000, seed=4, per_packet=True, flags=O.ETALON)["per_packet"]
        assert np.array_equal(a, b)


def test_slide_closed_form_single_packet_paths():
    # slide|sample|slide with the closed form: the boundary events are the
    # sample's crossings into a slide only (the glass bounces are gone), and
    # each packet's weight is conserved up to roulette (which moves at most
    # 1e-4 of it, times m - 1 = 9 on survival)
    spec = _stack(GLASS, SAMPLE, GLASS)
    r = O.simulate(spec, 20000, seed=6, per_packet=True, flags=O.ETALON)
    pp = r["per_packet"]
    assert np.all(np.abs(pp[:, :5].sum(1) - 1) < 1e-3)
    assert r["events"]["boundary_hits"] / 20000 < 2.6
