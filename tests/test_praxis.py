"""PRAXIS (ismc_amd.praxis) against the reference's own PRAXIS
(CUDA/praxis.cpp, compiled from the reference sources into oracle/_ref by
oracle/build_ref.py) on deterministic objectives (CPU).

Pinned: the whole evaluation sequence (every point, in order, to the last
bit), the minimum and the evaluation count are identical to the reference's
on every objective and setting below, including its singular value
decompositions (Golub-Reinsch, same signs) and its random steps."""
import ctypes
import os

import numpy as np
import pytest

from ismc_amd import praxis as P

LIB = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle", "_ref",
                   "libpraxis_ref.so")
OBJ = ctypes.CFUNCTYPE(ctypes.c_double, ctypes.POINTER(ctypes.c_double), ctypes.c_int)

FUNCS = {
    "quadratic": (lambda x: (x[0] - 1) ** 2 + 10 * (x[1] + 2) ** 2 + 0.5 * x[0] * x[1], [0.0, 0.0]),
    "rosenbrock": (lambda x: 100 * (x[1] - x[0] ** 2) ** 2 + (1 - x[0]) ** 2, [-1.2, 1.0]),
    "beale": (lambda x: (1.5 - x[0] + x[0] * x[1]) ** 2 + (2.25 - x[0] + x[0] * x[1] ** 2) ** 2
              + (2.625 - x[0] + x[0] * x[1] ** 3) ** 2, [1.0, 1.0]),
    "quadratic4": (lambda x: sum((i + 1) * (x[i] - i) ** 2 for i in range(4)) + x[0] * x[3],
                   [3.0, -1.0, 2.0, 0.5]),
    # a narrow curved valley: ill-conditioned, so PRAXIS takes random steps
    "valley": (lambda x: (x[0] - 0.3) ** 2 + 1e4 * (x[1] - x[0] ** 3) ** 2, [2.0, -1.0]),
    "powell4": (lambda x: (x[0] + 10 * x[1]) ** 2 + 5 * (x[2] - x[3]) ** 2 + (x[1] - 2 * x[2]) ** 4
                + 10 * (x[0] - x[3]) ** 4, [3.0, -1.0, 0.0, 1.0]),
}
SETTINGS = [(1e-6, 1.0), (0.1, 5.0), (1e-3, 0.03)]


def ref_run(f, x0, t0, h0, cap=20000):
    L = ctypes.CDLL(LIB)
    L.praxis_ref_run.restype = ctypes.c_double
    L.praxis_ref_run.argtypes = [ctypes.c_double, ctypes.c_double, ctypes.c_int,
                                 ctypes.POINTER(ctypes.c_double), OBJ, ctypes.POINTER(ctypes.c_long),
                                 ctypes.POINTER(ctypes.c_double), ctypes.c_long]
    n = len(x0)
    x = (ctypes.c_double * n)(*x0)
    nf = ctypes.c_long()
    tr = (ctypes.c_double * (cap * (n + 1)))()
    cb = OBJ(lambda p, nn: f(np.array([p[i] for i in range(nn)])))
    fx = L.praxis_ref_run(t0, h0, n, x, cb, ctypes.byref(nf), tr, cap)
    k = min(nf.value, cap)
    return np.array(x[:]), fx, nf.value, np.array(tr[:k * (n + 1)]).reshape(k, n + 1)


def own_run(f, x0, t0, h0):
    pts = []

    def g(p):
        v = f(p)
        pts.append(np.r_[p, v])
        return v
    x, fx, nf = P.minimize(g, x0, t0, h0)
    return x, fx, nf, np.array(pts)


@pytest.mark.skipif(not os.path.exists(LIB), reason="reference PRAXIS not built (needs the reference checkout)")
@pytest.mark.parametrize("t0,h0", SETTINGS)
@pytest.mark.parametrize("name", sorted(FUNCS))
def test_praxis_matches_reference(name, t0, h0):
    f, x0 = FUNCS[name]
    xr, fr, nr, Tr = ref_run(f, x0, t0, h0)
    xm, fm, nm, Tm = own_run(f, x0, t0, h0)
    same = 0
    while same < min(len(Tr), len(Tm)) and np.array_equal(Tr[same], Tm[same]):
        same += 1
    assert same == len(Tr) == len(Tm), "evaluation %d differs" % same
    assert nm == nr and np.array_equal(xm, xr) and fm == fr


def test_batch_equals_single():
    f, _ = FUNCS["rosenbrock"]
    starts = [[-1.2, 1.0], [0.5, 0.5], [2.0, -1.0], [0.0, 3.0]]
    single = [P.minimize(f, s, 1e-5, 1.0) for s in starts]
    calls = []

    def ev(points, ids):
        calls.append(len(points))
        return [f(p) for p in points]
    batch = P.minimize_batch(starts, ev, 1e-5, 1.0)
    for (xs, fs, ns), (xb, fb, nb) in zip(single, batch):
        assert np.array_equal(xs, xb) and fs == fb and ns == nb
    assert max(calls) == 4 and len(calls) == max(n for _, _, n in single)


def test_park_miller_sequence():
    # Park & Miller (1988): from seed 1 the 10,000th value is 1043618065
    r = P._ParkMiller(1)
    for _ in range(10000):
        r()
    assert r.seed == 1043618065
