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
