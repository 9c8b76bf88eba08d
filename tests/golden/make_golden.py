"""Write golden fixtures by running the reference's own pure-Python propagators.

Run in the build container only (the reference is not present where the GPU
tests run):

    python tests/golden/make_golden.py [--ref /root/reference] [--jobs 8] [case ...]

It imports `ISMC.py` (single slab, ISMC.py:164-406) and `ISMLMC.py` +
`MLMC.py` (multilayer with glass slides, ISMLMC.py:101-347, MLMC.py:76-202)
from the reference checkout, runs each case as `batches` independent batches
of `n` packets (Python's `random` seeded per batch so a fixture can be
regenerated), and writes one JSON file per case holding the inputs (in the
plain-dict spec form used by the oracle and the tests) and the per-batch
tallies (A, R_d, R_s, T_d, T_u).  Only these inputs/outputs are committed;
nothing of the reference's source travels with them.

These fixtures pin the composite statistical behaviour of the reference path
(glass-entry zero-length interaction, double-counted slide Fresnel,
uniform-cosine re-incidence, roulette thresholds, mirror vs substep
boundaries), which no closed form covers.
"""
import argparse
import json
import math
import os
import sys
import time
from multiprocessing import Pool

HERE = os.path.dirname(os.path.abspath(__file__))


def cap_fraction(r_port, r_sphere):
    """Spherical-cap port fraction (1 - sqrt(1 - (r/R)^2)) / 2."""
    return (1.0 - math.sqrt(1.0 - (r_port / r_sphere) ** 2)) / 2.0


FS = cap_fraction(0.5, 4.0)
FD = cap_fraction(0.25, 4.0)
SPH0 = {"present": 1, "pw": 0.98, "fs": FS, "fp": FS, "fd": FD, "angle": 0.0}
GLASS = {"n": 1.52, "g": 0.0, "t": 0.1, "mu_a": 0.0, "mu_s": 0.0}


def _layer(n, g, t, mu_a, mu_s):
    return {"n": n, "g": g, "t": t, "mu_a": mu_a, "mu_s": mu_s}


# case name -> (propagator, spec, n per batch, batches)
#   "ismlmc" = ISMLMC.ISMLMC (multilayer, MLMC.MC, roulette 1e-4)
#   "ismc"   = ISMC.ISMC (single slab mirror, roulette 1e-3)
CASES = {
    # slab between two glass slides, dual spheres
    "cfg0_slides_dual": ("ismlmc", {
        "layers": [GLASS, _layer(1.40, 0.9, 0.1, 1.0, 100.0), GLASS],
        "R": SPH0, "T": SPH0, "specular_included": 0}, 100000, 10),
    # three turbid layers, dual spheres
    "cfg1_three_layer_dual": ("ismlmc", {
        "layers": [_layer(1.37, 0.8, 0.01, 0.5, 150.0), _layer(1.40, 0.9, 0.1, 2.0, 80.0),
                   _layer(1.44, 0.85, 0.2, 0.2, 120.0)],
        "R": SPH0, "T": SPH0, "specular_included": 0}, 50000, 10),
    # ISMC.py's own example (ISMC.py:19-50): R sphere, no T sphere (angle in radians)
    "ismc_example": ("ismc", {
        "layers": [_layer(1.4, 0.5, 0.135, 1.78, 31.2 / 0.5)],
        "R": {"present": 1, "pw": 0.99, "fs": 0.02351, "fp": 0.054638, "fd": 0.02351, "angle": math.pi / 10},
        "T": {"present": 0, "pw": 0.99, "fs": 0.054638, "fp": 0.02351, "fd": 0.02351, "angle": 0.22 * math.pi / 180},
        "specular_included": 0, "roulette_threshold": 1e-3}, 50000, 10),
    # the slab-between-slides sample without slides, through ISMC.py, dual spheres
    "cfg0_sample_ismc_dual": ("ismc", {
        "layers": [_layer(1.40, 0.9, 0.1, 1.0, 100.0)],
        "R": SPH0, "T": SPH0, "specular_included": 0, "roulette_threshold": 1e-3}, 50000, 10),
    # slab between slides with the 8.1 degree beam (specular_included)
    "cfg0_slides_specular": ("ismlmc", {
        "layers": [GLASS, _layer(1.40, 0.9, 0.1, 1.0, 100.0), GLASS],
        "R": SPH0, "T": SPH0, "specular_included": 1}, 20000, 10),
    # no spheres: angle thresholds (radians), slides
    "slides_no_spheres": ("ismlmc", {
        "layers": [GLASS, _layer(1.40, 0.9, 0.1, 1.0, 100.0), GLASS],
        "R": {"present": 0, "angle": 0.1}, "T": {"present": 0, "angle": 0.05},
        "specular_included": 0}, 20000, 10),
    # separate-sphere configurations used by the minimization cost (minimization.py:209-222)
    "example_R_only": ("ismc", {
        "layers": [_layer(1.4, 0.5, 0.135, 2.0, 60.0)],
        "R": {"present": 1, "pw": 0.99, "fs": 0.02351, "fp": 0.054638, "fd": 0.02351, "angle": 0.0},
        "T": {"present": 0, "pw": 0.99, "fs": 0.054638, "fp": 0.02351, "fd": 0.02351, "angle": 0.0},
        "specular_included": 0, "roulette_threshold": 1e-3}, 20000, 10),
    "example_T_only": ("ismc", {
        "layers": [_layer(1.4, 0.5, 0.135, 2.0, 60.0)],
        "R": {"present": 0, "pw": 0.99, "fs": 0.02351, "fp": 0.054638, "fd": 0.02351, "angle": 0.0},
        "T": {"present": 1, "pw": 0.99, "fs": 0.054638, "fp": 0.02351, "fd": 0.02351, "angle": 0.0},
        "specular_included": 0, "roulette_threshold": 1e-3}, 20000, 10),
    # 64x64 lookup-grid corners/centre (single layer, g=0.8, n=1.4, t=0.1, dual spheres), ISMC.py
    "grid_mua0.01_mus1": ("ismc", {
        "layers": [_layer(1.40, 0.8, 0.1, 0.01, 1.0)],
        "R": SPH0, "T": SPH0, "specular_included": 0, "roulette_threshold": 1e-3}, 20000, 10),
    "grid_mua0.1_mus31.6": ("ismc", {
        "layers": [_layer(1.40, 0.8, 0.1, 0.1, 31.6227766)],
        "R": SPH0, "T": SPH0, "specular_included": 0, "roulette_threshold": 1e-3}, 20000, 10),
    "grid_mua10_mus1": ("ismc", {
        "layers": [_layer(1.40, 0.8, 0.1, 10.0, 1.0)],
        "R": SPH0, "T": SPH0, "specular_included": 0, "roulette_threshold": 1e-3}, 20000, 10),
    "grid_mua1_mus300": ("ismc", {
        "layers": [_layer(1.40, 0.8, 0.1, 1.0, 300.0)],
        "R": SPH0, "T": SPH0, "specular_included": 0, "roulette_threshold": 1e-3}, 5000, 10),
    # multilayer path on a single turbid layer (same physics as the lookup kernel)
    "grid_mua1_mus100_ismlmc": ("ismlmc", {
        "layers": [_layer(1.40, 0.8, 0.1, 1.0, 100.0)],
        "R": SPH0, "T": SPH0, "specular_included": 0}, 20000, 10),
}


def _run_batch(args):
    ref, prop, spec, n, seed = args
    sys.path.insert(0, ref)
    import random
    random.seed(seed)
    layers = spec["layers"]
    Rs, Ts = spec["R"], spec["T"]

    def sph(d):
        fs, fp, fd = d.get("fs", 0.0), d.get("fp", 0.0), d.get("fd", 0.0)
        return (bool(d.get("present", 0)), d.get("pw", 0.0), fs, fp, fd, fs + fp + fd,
                d.get("angle", 0.0))

    Rp, Tp = sph(Rs), sph(Ts)
    t0 = time.time()
    if prop == "ismc":
        import ISMC
        L = layers[0]
        out = ISMC.ISMC(n, L["n"], L["g"], L["t"], L["mu_a"], L["mu_s"], L["mu_a"] + L["mu_s"],
                        Rp[0], Rp[1], Rp[2], Rp[3], Rp[4], Rp[5], Rp[6],
                        bool(spec.get("specular_included", 0)),
                        Tp[0], Tp[1], Tp[2], Tp[3], Tp[4], Tp[5], Tp[6])
    else:
        import ISMLMC
        col = lambda k: [L[k] for L in layers]
        mu_t = [L["mu_a"] + L["mu_s"] for L in layers]
        out = ISMLMC.ISMLMC(n, col("n"), col("g"), col("t"), col("mu_a"), col("mu_s"), mu_t,
                            Rp[0], Rp[1], Rp[2], Rp[3], Rp[4], Rp[5], Rp[6],
                            bool(spec.get("specular_included", 0)),
                            Tp[0], Tp[1], Tp[2], Tp[3], Tp[4], Tp[5], Tp[6])
    return [float(x) for x in out], time.time() - t0


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--ref", default="/root/reference")
    ap.add_argument("--jobs", type=int, default=8)
    ap.add_argument("--scale", type=float, default=1.0, help="multiply packets per batch")
    ap.add_argument("--out", default=HERE)
    ap.add_argument("cases", nargs="*")
    a = ap.parse_args()
    names = a.cases or list(CASES)
    jobs = []
    for name in names:
        prop, spec, n, batches = CASES[name]
        n = max(1, int(n * a.scale))
        for b in range(batches):
            jobs.append((name, (a.ref, prop, spec, n, 1000 + b)))
    with Pool(a.jobs) as pool:
        res = pool.map(_run_batch, [j[1] for j in jobs], chunksize=1)
    by = {}
    for (name, args), (tallies, dt) in zip(jobs, res):
        by.setdefault(name, []).append((args[4], tallies, dt))
    for name, rows in by.items():
        prop, spec, n, batches = CASES[name]
        n = max(1, int(n * a.scale))
        doc = {
            "case": name,
            "propagator": {"ismc": "ISMC.py:164 ISMC()", "ismlmc": "ISMLMC.py:101 ISMLMC()"}[prop],
            "spec": spec,
            "n_per_batch": n,
            "python_seeds": [r[0] for r in rows],
            "tallies_order": ["A", "R_d", "R_s", "T_d", "T_u"],
            "batches": [r[1] for r in rows],
            "seconds_per_batch": [round(r[2], 2) for r in rows],
        }
        path = os.path.join(a.out, name + ".json")
        with open(path, "w") as f:
            json.dump(doc, f, indent=1)
        print("wrote", path, "mean", [round(sum(r[1][i] for r in rows) / len(rows), 6) for i in range(5)])


if __name__ == "__main__":
    main()
