"""The pure-Python restatement timed as the CPU baseline reproduces the
reference fixture (within 3 combined SE)."""
import numpy as np

from helpers import load_golden, zscores
from oracle import ismc_py


def test_python_restatement_matches_fixture():
    d = load_golden("cfg0_slides_dual")
    runs = np.array([ismc_py.simulate(d["spec"], 3000, seed=s) for s in range(6)])
    m = runs.mean(axis=0)
    se = runs.std(axis=0, ddof=1) / np.sqrt(runs.shape[0])
    z = zscores(m, se, d["mean"], d["se"])
    assert np.all(np.abs(z) < 3.5), np.round(z, 2)
    assert abs(m.sum() - 1) < 0.01
