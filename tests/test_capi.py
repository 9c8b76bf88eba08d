"""The C-ABI library (no GPU calls): it loads, exports every function
declared in include/ismc.h, validates configurations on the host, and the
Python-side config maps onto its struct."""
import ctypes
import math
import os
import re

import pytest

from ismc_amd import _lib
from ismc_amd.config import Layer, SampleConfig, Sphere, cap_port_fraction, mus_from_reduced
from ismc_amd.workloads import slab_between_slides, three_layer_sample

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions(header):
    txt = open(os.path.join(ROOT, "include", header)).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"^[A-Za-z_][\w \*]*?\b(\w+)\s*\(", txt, flags=re.M)))


def test_every_header_has_its_library():
    assert sorted(h for h in os.listdir(os.path.join(ROOT, "include")) if h.endswith(".h")) == \
        ["ismc.h", "ismc_diag.h"]


def test_library_exports_every_declared_function():
    L = _lib.load()
    names = declared_functions("ismc.h")
    assert "ismc_simulate" in names and "ismc_simulate_batch" in names
    for n in names:
        assert hasattr(L, n), "missing export " + n
    assert set(_lib.EXPORTS) <= set(names)


def test_product_library_exports_no_probe():
    L = _lib.load()
    for n in declared_functions("ismc_diag.h"):
        assert not hasattr(L, n), "probe %s is exported by the product library" % n


def test_diag_library_exports_every_declared_function():
    from ismc_amd import profiling
    D = profiling.diag()
    names = declared_functions("ismc_diag.h")
    assert "ismc_copy_probe" in names and "ismc_issue_rate" in names
    for n in names:
        assert hasattr(D, n), "missing export " + n
    assert D.ismc_issue_rate_ops() > 0


def test_struct_sizes_match_header_layout():
    # 2 ints + 3 doubles + 6 arrays of 16 doubles + 2 spheres (8 + 6 doubles)
    assert ctypes.sizeof(_lib.Sphere) == 8 + 6 * 8
    assert ctypes.sizeof(_lib.Config) == 8 + 3 * 8 + 6 * 16 * 8 + 2 * ctypes.sizeof(_lib.Sphere)
    assert ctypes.sizeof(_lib.Tally) == 10 * 8 + 5 * 8 + 16


def test_validate_accepts_workloads():
    for cfg in (slab_between_slides(), three_layer_sample()):
        rc, msg = _lib.validate(cfg.to_c())
        assert rc == 0, msg


@pytest.mark.parametrize("mutate,code", [
    (lambda c: setattr(c, "n_layers", 0), _lib.EINVAL),
    (lambda c: c.mu_a.__setitem__(1, -1.0), _lib.EINVAL),
    (lambda c: c.g.__setitem__(1, 1.0), _lib.EINVAL),
    (lambda c: c.n.__setitem__(0, float("nan")), _lib.EINVAL),
    (lambda c: c.t.__setitem__(2, 0.0), _lib.EINVAL),
    (lambda c: setattr(c.R, "pw", 2.0), _lib.EINVAL),
    (lambda c: setattr(c, "roulette_threshold", 0.5), _lib.ERANGE),
])
def test_validate_rejects_bad_configs(mutate, code):
    c = slab_between_slides().to_c()
    mutate(c)
    rc, msg = _lib.validate(c)
    assert rc == code and msg


def test_cap_fractions_reproduce_ini_values():
    # Lookup/example.ini:21,25 for 1.27 cm / 1.905 cm ports on an 8.382 cm sphere
    assert cap_port_fraction(1.27, 8.382 / 2) == pytest.approx(0.023510, abs=5e-7)
    assert cap_port_fraction(1.905, 8.382 / 2) == pytest.approx(0.054638, abs=5e-7)
    assert cap_port_fraction(0.5, 4.0) == pytest.approx(0.003921629, abs=1e-9)


def test_config_to_c_roundtrip():
    cfg = SampleConfig(layers=[Layer.glass(), Layer.from_reduced(1.4, 0.9, 0.1, 1.0, 10.0), Layer.glass()],
                       R=Sphere(True, 0.98, 0.01, 0.02, 0.003), T=Sphere(False, angle=0.1))
    c = cfg.to_c()
    assert c.n_layers == 3 and c.mu_s[1] == pytest.approx(100.0) and c.mu_t[1] == pytest.approx(101.0)
    assert c.R.present == 1 and c.R.f == pytest.approx(0.033) and c.T.angle == 0.1
    assert mus_from_reduced(5.0, 1.0) == 5.0
    assert Sphere(True, 0.98, 0.0039, 0.0039, 0.001).return_probability() == pytest.approx(
        0.98 * 0.0039 / (1 - 0.98 * (1 - 0.0088)))


def test_engine_fails_loudly_without_gpu():
    L = _lib.load()
    ctx = ctypes.c_void_p()
    rc = L.ismc_create(0, ctypes.byref(ctx))
    if rc == 0:
        L.ismc_destroy(ctx)
        pytest.skip("a GPU is present")
    from ismc_amd import Engine
    with pytest.raises(RuntimeError):
        Engine(0)
