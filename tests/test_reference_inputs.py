"""The reference's own input files, read by the .ini loader: both
Lookup/example.ini and Minimization/example.ini (identical files with
multi-line measurement lists, ';' comments, trailing blanks and a
whitespace-only line inside [minimization]) give every key its value, and
the drivers' derived quantities follow.  The expected values are the ones
those files hold; the test skips where the reference checkout is absent."""
import math
import os

import pytest

from ismc_amd import ini as ini_mod
from ismc_amd import minimization as M

REF = "/root/reference"
FILES = [os.path.join(REF, "Minimization", "example.ini"), os.path.join(REF, "Lookup", "example.ini")]


@pytest.fixture(params=FILES, ids=["Minimization", "Lookup"])
def ini(request):
    if not os.path.exists(request.param):
        pytest.skip("reference checkout not present")
    return ini_mod.load(request.param)


def test_sample_and_spheres(ini):
    assert (ini.n, ini.n_unc, ini.g, ini.t) == (1.4, 0.1, 0.5, 0.135)
    R, T = ini.R, ini.T
    assert R.present and T.present and not ini.specular_included and ini.separate
    assert (R.pw, R.pw_unc) == (0.99, 0.005)
    assert (R.fs, R.fs_unc, R.fp, R.fp_unc, R.fd, R.fd_unc) == (0.023510, 0.0011755, 0.054638, 0.0027319,
                                                                 0.023510, 0.0011755)
    assert (R.angle_deg, R.angle_unc_deg) == (0.0, 0.0)
    assert (T.pw, T.pw_unc) == (0.99, 0.005)
    assert (T.fs, T.fs_unc, T.fp, T.fp_unc, T.fd, T.fd_unc) == (0.054638, 0.0027319, 0.023510, 0.0011755,
                                                                 0.023510, 0.0011755)
    assert (T.angle_deg, T.angle_unc_deg) == (0.219523, 0.000643)
    # degrees become radians; f is the sum of the three port fractions
    s = T.sphere()
    assert s.angle == pytest.approx(0.219523 * math.pi / 180, rel=1e-15)
    assert s.f_total == pytest.approx(0.054638 + 0.023510 + 0.023510, rel=1e-15)


def test_experiment_lists(ini):
    assert ini.M_R.values == [0.2716, 0.3000] and ini.M_R.uncs == [0.0002, 0.0010]
    assert ini.M_T.values == [0.0853, 0.1000] and ini.M_T.uncs == [0.0001, 0.0020]
    assert ini.M_U.values == [0.000022, 0.0020] and ini.M_U.uncs == [0.000003, 0.0001]
    assert ini.M_S is None  # commented out, key included
    assert (ini.M_R.first, ini.M_R.first_unc) == (0.2716, 0.0002)


def test_lookup_section(ini):
    L = ini.lookup
    assert L["filename_prefix"] == "EXAMPLE" and L["dimensions"] == 1000
    assert (L["R_d_max"], L["R_d_min"], L["T_d_max"], L["T_d_min"]) == (55, 5, 30, 0)
    assert (L["perc_unc_max"], L["perc_unc_min"]) == (15, 5)
    assert (L["N"], L["runs"]) == (100000, 1) and "uncertainty" not in L and "estimate" not in L
    assert (L["input_distributions_samples"], L["simulation_distributions_samples"],
            L["measurement_distributions_samples"]) == (100, 200, 1000)
    assert (L["N_mua"], L["N_mus"], L["mua_start"], L["mua_base"], L["mus_start"], L["mus_base"]) == \
        (13, 11, 0.5, 1.2, 10, 1.2)
    assert ini.lookup_photons() == (100000, 1)
    mua, mus = ini.lookup_grid()
    assert len(mua) == 13 and len(mus) == 11
    assert mua[0] == 0.5 and mua[12] == pytest.approx(0.5 * 1.2 ** 12) and mus[10] == pytest.approx(10 * 1.2 ** 10)


def test_minimization_section(ini):
    m = ini.minimization
    assert (m["N"], m["runs"], m["input_distributions_samples"]) == (10000, 5, 100)
    assert (m["mua_guess"], m["mus_guess"]) == (5, 45)
    assert (m["mua_min"], m["mua_max"], m["mus_min"], m["mus_max"]) == (1, 3, 20, 40)
    assert (m["step"], m["tolerance"]) == (5, 0.1)


def test_problems_from_reference_ini(ini):
    """minimization.py's input-uncertainty samples from the reference file:
    separate (single-sphere) problems whose cost needs the R-sphere, the
    T-sphere and the no-sphere simulations (M_R, M_T and M_U measured)."""
    probs = M.problems_from_ini(ini, samples=4, rng_seed=1)
    assert len(probs) == 4 and all(p.separate and p.g == 0.5 and p.t == 0.135 for p in probs)
    cfgs, src = probs[0].configs(2.0, 30.0)
    assert len(cfgs) == 3 and set(src) == {"R_d", "T_d", "R_s", "T_u"}
    assert cfgs[0].R.present and not cfgs[0].T.present and cfgs[1].T.present and not cfgs[1].R.present
    assert not cfgs[2].R.present and not cfgs[2].T.present
    assert cfgs[0].layers[0].mu_s == pytest.approx(60.0)
