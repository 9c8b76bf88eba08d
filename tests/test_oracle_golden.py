"""The oracle against golden fixtures written by running the reference's own
Python propagators (tests/golden/make_golden.py): every tally within 3
combined Monte Carlo standard errors (batch-of-10 SE for the reference)."""
import os

import numpy as np
import pytest

from helpers import golden_cases, load_golden, vec, zscores
from oracle import oracle as O

THREADS = max(1, min(8, os.cpu_count() or 1))


@pytest.mark.parametrize("case", golden_cases())
def test_oracle_matches_reference_fixture(case):
    d = load_golden(case)
    n = 10 * d["n_per_batch"]
    res = O.simulate(d["spec"], n, seed=99, flags=O.MIRROR if d["mirror"] else 0, threads=THREADS)
    z = zscores(vec(res["mean"]), vec(res["se"]), d["mean"], d["se"])
    assert np.all(np.abs(z) < 3.0), dict(zip(("A", "R_d", "R_s", "T_d", "T_u"), np.round(z, 2)))
