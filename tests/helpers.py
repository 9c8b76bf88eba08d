"""Shared helpers for the tests (fixtures, spec conversion, z-scores)."""
import glob
import json
import math
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
GOLDEN = os.path.join(HERE, "golden")
TALLIES = ("A", "R_d", "R_s", "T_d", "T_u")


def golden_cases():
    return sorted(os.path.splitext(os.path.basename(p))[0]
                  for p in glob.glob(os.path.join(GOLDEN, "*.json")))


def load_golden(name):
    with open(os.path.join(GOLDEN, name + ".json")) as f:
        d = json.load(f)
    b = np.array(d["batches"], dtype=np.float64)
    d["mean"] = b.mean(axis=0)
    d["se"] = b.std(axis=0, ddof=1) / math.sqrt(b.shape[0])
    d["mirror"] = d["propagator"].startswith("ISMC.py")
    return d


def spec_of(cfg):
    """Plain-dict spec (oracle form) of an ismc_amd SampleConfig."""
    def sph(s):
        return {"present": int(s.present), "pw": s.pw, "fs": s.fs, "fp": s.fp, "fd": s.fd,
                "f": s.f, "angle": s.angle}
    return {
        "layers": [{"n": L.n, "g": L.g, "t": L.t, "mu_a": L.mu_a, "mu_s": L.mu_s,
                    **({"mu_t": L.mu_t} if L.mu_t is not None else {})} for L in cfg.layers],
        "R": sph(cfg.R), "T": sph(cfg.T), "specular_included": int(cfg.specular_included),
        "n_ambient": cfg.n_ambient, "roulette_threshold": cfg.roulette_threshold,
        "roulette_m": cfg.roulette_m,
    }


def zscores(m1, se1, m2, se2):
    m1, se1, m2, se2 = map(lambda a: np.asarray(a, dtype=np.float64), (m1, se1, m2, se2))
    den = np.sqrt(se1 ** 2 + se2 ** 2)
    return np.where(den > 0, (m1 - m2) / np.where(den > 0, den, 1), 0.0)


def vec(d):
    return np.array([d[k] for k in TALLIES], dtype=np.float64)
