"""Oracle physics pieces against closed forms (CPU).

Fresnel: fresnel_snell (MLMC.py:45-54) against normal-incidence values,
total internal reflection, Brewster's angle, reciprocity and grazing
incidence; the glass-slide entrance formula (ISMLMC.py:79-93); the
Henyey-Greenstein sampler (MLMC.py:183-184) against its first two
Legendre moments; the sphere return probability (ISMLMC.py:96-97)."""
import math

import numpy as np
import pytest

from oracle import oracle as O


def r0(n1, n2):
    return ((n2 - n1) / (n2 + n1)) ** 2


@pytest.mark.parametrize("n1,n2,expected", [(1.0, 1.40, 0.0277778), (1.0, 1.52, 0.0425800),
                                            (1.52, 1.40, 0.00168887), (1.0, 1.37, 0.0243729)])
def test_fresnel_normal_incidence(n1, n2, expected):
    assert O.fresnel(n1, n2, 1.0) == pytest.approx(expected, abs=5e-7)
    # just below the 0.99999 shortcut the full formula must agree with it
    assert O.fresnel(n1, n2, 0.999989) == pytest.approx(expected, rel=1e-4)


def test_fresnel_total_internal_reflection():
    n1, n2 = 1.4, 1.0
    crit = math.sqrt(1 - (n2 / n1) ** 2)   # cos of the critical angle
    assert O.fresnel(n1, n2, crit * 0.999) == 1.0
    assert O.fresnel(n1, n2, crit * 1.001) < 1.0


def test_fresnel_brewster_angle():
    n1, n2 = 1.0, 1.5
    tb = math.atan(n2 / n1)
    tt = math.asin(n1 * math.sin(tb) / n2)
    rs = (math.sin(tb - tt) / math.sin(tb + tt)) ** 2
    # at Brewster's angle only the s-polarised half is reflected
    assert O.fresnel(n1, n2, math.cos(tb)) == pytest.approx(0.5 * rs, rel=1e-9)
    # textbook value for the s term: ((n1 cos - n2 cos_t)/(n1 cos + n2 cos_t))^2
    ci, ct = math.cos(tb), math.cos(tt)
    assert rs == pytest.approx(((n1 * ci - n2 * ct) / (n1 * ci + n2 * ct)) ** 2, rel=1e-12)


@pytest.mark.parametrize("mu", [0.2, 0.5, 0.8, 0.95])
def test_fresnel_reciprocity(mu):
    n1, n2 = 1.0, 1.4
    ti = math.acos(mu)
    tt = math.asin(n1 * math.sin(ti) / n2)
    assert O.fresnel(n1, n2, mu) == pytest.approx(O.fresnel(n2, n1, math.cos(tt)), rel=1e-10)


def test_fresnel_grazing_and_fp32():
    assert O.fresnel(1.0, 1.4, 1e-9) == pytest.approx(1.0, abs=1e-6)
    for mu in (0.1, 0.3, 0.7, 0.99):
        assert O.fresnel(1.0, 1.4, mu, fp32=True) == pytest.approx(O.fresnel(1.0, 1.4, mu), abs=2e-6)


def test_glass_entrance_closed_form():
    spec = {"layers": [{"n": 1.52, "g": 0, "t": .1, "mu_a": 0, "mu_s": 0},
                       {"n": 1.40, "g": .9, "t": .1, "mu_a": 1, "mu_s": 100},
                       {"n": 1.52, "g": 0, "t": .1, "mu_a": 0, "mu_s": 0}]}
    r1, r2 = r0(1.0, 1.52), r0(1.52, 1.40)
    expected = r1 + (1 - r1) ** 2 * r2 / (1 - r1 * r2)
    assert expected == pytest.approx(0.0441282, abs=5e-7)
    assert O.entrance_reflectance(spec, 1.0) == pytest.approx(expected, rel=1e-12)
    assert O.entrance_reflectance(spec, -1.0) == pytest.approx(expected, rel=1e-12)
    spec1 = {"layers": [{"n": 1.40, "g": .9, "t": .1, "mu_a": 1, "mu_s": 100}]}
    assert O.entrance_reflectance(spec1, 1.0) == pytest.approx(r0(1, 1.4), rel=1e-12)


def test_glass_guard_exit_index():
    # The glass-entry guard keeps n^2 (1 - mu_z^2) inside a glass layer below
    # the square of its exit index: on each side the least index from the
    # neighbour out to the medium that bounds the layer's run of glass
    # layers (that medium included), the larger of the two sides.  A packet
    # below it crosses every face on that side.
    def T(n):
        return {"n": n, "g": .5, "t": .01, "mu_a": .1, "mu_s": 10}

    def G(n):
        return {"n": n, "g": 0, "t": .01, "mu_a": 0, "mu_s": 0}

    assert O.glass_guard_index({"layers": [G(1.52), T(1.40)]}, 0) == 1.40          # air | slide | sample
    assert O.glass_guard_index({"layers": [T(1.33), G(1.52), T(1.40)]}, 1) == 1.40  # between two samples
    s = {"layers": [G(1.4675), G(1.4693), T(1.433)]}                               # two slides on a sample
    assert [O.glass_guard_index(s, k) for k in (0, 1)] == [1.433, 1.433]
    # a run with a low-index layer inside: from 1.6318 and 1.4731 the way
    # out to the left passes 1.1919, so the bound is the right side's 1.1968
    s = {"layers": [T(1.3755), G(1.444), G(1.1919), G(1.6318), G(1.4731), T(1.1968)]}
    assert [O.glass_guard_index(s, k) for k in (1, 2, 3, 4)] == [1.3755, 1.3755, 1.1968, 1.1968]
    s = {"n_ambient": 1.33, "layers": [G(1.2967), G(1.1711), T(1.2029)]}             # denser ambient
    assert [O.glass_guard_index(s, k) for k in (0, 1)] == [1.33, 1.2967]


def test_glass_entrance_under_a_denser_ambient():
    # An ambient medium denser than the first slide (n 1.33 over glass
    # 1.2967, glass 1.1711 behind it): the series holds at normal incidence,
    # and where the ambient face reflects totally (sin^2 > (1.2967/1.33)^2 at
    # the ambient-side cosine, both terms taken there) the entrance reflects
    # totally, although the far face then reflects totally too and the
    # series reads 0/0
    spec = {"n_ambient": 1.33,
            "layers": [{"n": 1.2967, "g": 0, "t": .005, "mu_a": 0, "mu_s": 0},
                       {"n": 1.1711, "g": 0, "t": .01, "mu_a": 0, "mu_s": 0},
                       {"n": 1.2029, "g": -.46, "t": .017, "mu_a": .02, "mu_s": .018}]}
    r1, r2 = r0(1.33, 1.2967), r0(1.2967, 1.1711)
    assert O.entrance_reflectance(spec, 1.0) == pytest.approx(r1 + (1 - r1) ** 2 * r2 / (1 - r1 * r2), rel=1e-12)
    mu_c = math.sqrt(1 - (1.2967 / 1.33) ** 2)
    for mu in (0.01, 0.1, 0.5 * mu_c, 0.99 * mu_c):
        assert O.entrance_reflectance(spec, mu) == 1.0
    # past mu_c the far face alone reflects totally up to mu = 0.43 (the
    # series is then r1 + (1 - r1) = 1); at 0.5 both faces transmit
    assert O.entrance_reflectance(spec, 1.01 * mu_c) == pytest.approx(1.0, abs=1e-15)
    r1, r2 = O.fresnel(1.33, 1.2967, 0.5), O.fresnel(1.2967, 1.1711, 0.5)
    assert r1 < 1 and r2 < 1
    assert O.entrance_reflectance(spec, 0.5) == pytest.approx(r1 + (1 - r1) ** 2 * r2 / (1 - r1 * r2), rel=1e-12)


@pytest.mark.parametrize("g", [-0.5, 0.0, 0.3, 0.8, 0.9, 0.95])
def test_hg_moments(g):
    # E[cos] = g and E[P2(cos)] = g^2 for Henyey-Greenstein; the sampler is the
    # inverse CDF, so the expectations are integrals over u in (0, 1).
    u = (np.arange(200000) + 0.5) / 200000
    c = np.array([O.hg_cos(g, x) for x in u[::50]])
    uu = u[::50]
    del uu
    c_full = np.vectorize(O.hg_cos)(g, u)
    assert c_full.mean() == pytest.approx(g, abs=2e-4)
    assert (1.5 * c_full ** 2 - 0.5).mean() == pytest.approx(g * g, abs=5e-4)
    assert np.all(np.abs(c) <= 1 + 1e-12)


def test_hg_cdf_ks():
    # KS distance between sampled cos(theta) and the analytic HG CDF
    g = 0.8
    rng = np.random.default_rng(3)
    u = rng.random(20000)
    c = np.sort(np.vectorize(O.hg_cos)(g, u))
    cdf = (1 - g * g) / (2 * g) * (1 / np.sqrt(1 + g * g - 2 * g * c) - 1 / (1 + g))
    ks = np.max(np.abs(cdf - (np.arange(1, c.size + 1) / c.size)))
    assert ks < 1.63 / math.sqrt(c.size)   # 1% critical value


def test_sphere_return_closed_form():
    fs = (1 - math.sqrt(1 - (0.5 / 4) ** 2)) / 2
    fd = (1 - math.sqrt(1 - (0.25 / 4) ** 2)) / 2
    assert fs == pytest.approx(0.003921629, abs=1e-9)
    assert fd == pytest.approx(0.000977518, abs=1e-9)
    f = 2 * fs + fd
    p = 0.98 * fs / (1 - 0.98 * (1 - f))
    assert p == pytest.approx(0.134169, abs=2e-6)
