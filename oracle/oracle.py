"""ORACLE — Python binding of the C restatement in ismc_oracle.c.

Test infrastructure only: tests/, __graft_entry__.smoke() and bench.py's
cpu_baseline leg may import it, and only as the checker.  The product
package (ismc_amd) never imports anything from oracle/.

Configs are plain dicts so that the oracle shares no code with the product:

    spec = {
        "layers": [{"n": 1.52, "g": 0.0, "t": 0.1, "mu_a": 0.0, "mu_s": 0.0}, ...],
        "n_ambient": 1.0,
        "R": {"present": 1, "pw": .98, "fs": .0039, "fp": .0039, "fd": .00098,
              "f": None, "angle": 0.0},          # f defaults to fs+fp+fd
        "T": {...},
        "specular_included": 0,
        "roulette_threshold": 1e-4, "roulette_m": 10,
    }

Layer "mu_t" defaults to mu_a + mu_s (MLMC.py:30-33); mu_s == 0 marks glass.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "libismc_oracle.so")

MAX_LAYERS = 16
FP32 = 1
MIRROR = 2
CUDA_ROULETTE = 4
ETALON = 8
TALLIES = ("A", "R_d", "R_s", "T_d", "T_u")
EVENTS = ("interactions", "boundary_hits", "entrances", "sphere_decisions", "reinjections")


class _Sphere(ctypes.Structure):
    _fields_ = [("present", ctypes.c_int32), ("_pad", ctypes.c_int32),
                ("pw", ctypes.c_double), ("fs", ctypes.c_double), ("fp", ctypes.c_double),
                ("fd", ctypes.c_double), ("f", ctypes.c_double), ("angle", ctypes.c_double)]


class _Config(ctypes.Structure):
    _fields_ = [("n_layers", ctypes.c_int32), ("specular_included", ctypes.c_int32),
                ("n_ambient", ctypes.c_double), ("roulette_threshold", ctypes.c_double),
                ("roulette_m", ctypes.c_double),
                ("n", ctypes.c_double * MAX_LAYERS), ("g", ctypes.c_double * MAX_LAYERS),
                ("t", ctypes.c_double * MAX_LAYERS), ("mu_a", ctypes.c_double * MAX_LAYERS),
                ("mu_s", ctypes.c_double * MAX_LAYERS), ("mu_t", ctypes.c_double * MAX_LAYERS),
                ("R", _Sphere), ("T", _Sphere)]


class _Tally(ctypes.Structure):
    _fields_ = [("sum", ctypes.c_double * 5), ("sumsq", ctypes.c_double * 5),
                ("events", ctypes.c_uint64 * 5), ("n_packets", ctypes.c_uint64),
                ("n_aborted", ctypes.c_uint64)]


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        _lib = ctypes.CDLL(_LIB_PATH)
        _lib.oracle_simulate.restype = ctypes.c_int
        _lib.oracle_simulate.argtypes = [ctypes.POINTER(_Config), ctypes.c_uint64, ctypes.c_uint64,
                                         ctypes.c_uint64, ctypes.c_uint32, ctypes.c_uint32,
                                         ctypes.POINTER(_Tally), ctypes.c_void_p]
        _lib.oracle_simulate_omp.restype = ctypes.c_int
        _lib.oracle_simulate_omp.argtypes = [ctypes.POINTER(_Config), ctypes.c_uint64,
                                             ctypes.c_uint64, ctypes.c_uint64, ctypes.c_uint32,
                                             ctypes.c_uint32, ctypes.POINTER(_Tally), ctypes.c_int]
        _lib.oracle_philox4x32_10.restype = None
        _lib.oracle_philox4x32_10.argtypes = [ctypes.POINTER(ctypes.c_uint32),
                                              ctypes.POINTER(ctypes.c_uint32),
                                              ctypes.POINTER(ctypes.c_uint32)]
        _lib.oracle_fresnel.restype = ctypes.c_double
        _lib.oracle_fresnel.argtypes = [ctypes.c_double] * 3
        _lib.oracle_fresnel_f.restype = ctypes.c_float
        _lib.oracle_fresnel_f.argtypes = [ctypes.c_float] * 3
        _lib.oracle_entrance_reflectance.restype = ctypes.c_double
        _lib.oracle_entrance_reflectance.argtypes = [ctypes.POINTER(_Config), ctypes.c_double]
        _lib.oracle_glass_guard_index.restype = ctypes.c_double
        _lib.oracle_glass_guard_index.argtypes = [ctypes.POINTER(_Config), ctypes.c_int]
        _lib.oracle_hg_cos.restype = ctypes.c_double
        _lib.oracle_hg_cos.argtypes = [ctypes.c_double, ctypes.c_double]
    return _lib


ROCRAND_XCHECK = os.path.join(_HERE, "libphilox_rocrand.so")


def build():
    """Compile ismc_oracle.c with gcc (OpenMP) into oracle/libismc_oracle.so,
    and the host-only rocRAND Philox harness with hipcc when it is present."""
    import shutil
    import subprocess
    src = os.path.join(_HERE, "ismc_oracle.c")
    tmp = _LIB_PATH + ".tmp"
    subprocess.check_call(["gcc", "-O2", "-fopenmp", "-fPIC", "-shared", "-o", tmp, src, "-lm"])
    os.replace(tmp, _LIB_PATH)
    hipcc = shutil.which("hipcc")
    if hipcc:
        tmp = ROCRAND_XCHECK + ".tmp"
        subprocess.check_call([hipcc, "-O2", "-fPIC", "-shared", "-o", tmp,
                               os.path.join(_HERE, "philox_rocrand_xcheck.cpp")])
        os.replace(tmp, ROCRAND_XCHECK)


def rocrand_philox(ctr, key):
    """rocRAND's Philox4x32-10 rounds on (ctr, key) (host harness)."""
    L = ctypes.CDLL(ROCRAND_XCHECK)
    o = (ctypes.c_uint32 * 4)()
    L.rocrand_philox4x32_10_rounds((ctypes.c_uint32 * 4)(*ctr), (ctypes.c_uint32 * 2)(*key), o)
    return tuple(int(x) for x in o)


def _sphere(d):
    s = _Sphere()
    d = d or {}
    s.present = int(bool(d.get("present", 0)))
    for k in ("pw", "fs", "fp", "fd", "angle"):
        setattr(s, k, float(d.get(k, 0.0) or 0.0))
    f = d.get("f")
    s.f = float(f) if f is not None else s.fs + s.fp + s.fd
    return s


def make_config(spec):
    c = _Config()
    layers = spec["layers"]
    if not 1 <= len(layers) <= MAX_LAYERS:
        raise ValueError("1..%d layers" % MAX_LAYERS)
    c.n_layers = len(layers)
    c.specular_included = int(bool(spec.get("specular_included", 0)))
    c.n_ambient = float(spec.get("n_ambient", 1.0))
    c.roulette_threshold = float(spec.get("roulette_threshold", 1e-4))
    c.roulette_m = float(spec.get("roulette_m", 10))
    for i, L in enumerate(layers):
        c.n[i] = L["n"]
        c.g[i] = L["g"]
        c.t[i] = L["t"]
        c.mu_a[i] = L["mu_a"]
        c.mu_s[i] = L["mu_s"]
        c.mu_t[i] = L.get("mu_t", L["mu_a"] + L["mu_s"])
    c.R = _sphere(spec.get("R"))
    c.T = _sphere(spec.get("T"))
    return c


def philox(ctr, key):
    c = (ctypes.c_uint32 * 4)(*ctr)
    k = (ctypes.c_uint32 * 2)(*key)
    o = (ctypes.c_uint32 * 4)()
    lib().oracle_philox4x32_10(c, k, o)
    return tuple(int(x) for x in o)


def fresnel(ni, nt, mu_z, fp32=False):
    if fp32:
        return lib().oracle_fresnel_f(ni, nt, mu_z)
    return lib().oracle_fresnel(ni, nt, mu_z)


def entrance_reflectance(spec, mu_z):
    c = make_config(spec)
    return lib().oracle_entrance_reflectance(ctypes.byref(c), mu_z)


def glass_guard_index(spec, layer):
    c = make_config(spec)
    return lib().oracle_glass_guard_index(ctypes.byref(c), int(layer))


def hg_cos(g, u):
    return lib().oracle_hg_cos(g, u)


def _result(t):
    n = t.n_packets
    s = np.array(t.sum[:])
    q = np.array(t.sumsq[:])
    mean = s / n if n else s * 0
    var = (q / n - mean ** 2) if n else q * 0
    se = np.sqrt(np.maximum(var, 0.0) / max(n - 1, 1))
    return {
        "n_packets": int(n),
        "n_aborted": int(t.n_aborted),
        "mean": dict(zip(TALLIES, mean.tolist())),
        "se": dict(zip(TALLIES, se.tolist())),
        "sum": dict(zip(TALLIES, s.tolist())),
        "sumsq": dict(zip(TALLIES, q.tolist())),
        "events": dict(zip(EVENTS, [int(x) for x in t.events])),
    }


def simulate(spec, n_packets, seed=1234, first_pid=0, stream=0, flags=0, per_packet=False,
             threads=1):
    """Run the oracle; returns a dict with mean/se per tally, event counts
    and (optionally) an (n, 6) per-packet array {A, R_d, R_s, T_d, T_u, events}."""
    c = make_config(spec)
    t = _Tally()
    pp = None
    if per_packet:
        if threads != 1:
            raise ValueError("per_packet output is single-threaded")
        pp = np.zeros((n_packets, 6), dtype=np.float64)
        rc = lib().oracle_simulate(ctypes.byref(c), n_packets, seed, first_pid, stream, flags,
                                   ctypes.byref(t), pp.ctypes.data)
    elif threads == 1:
        rc = lib().oracle_simulate(ctypes.byref(c), n_packets, seed, first_pid, stream, flags,
                                   ctypes.byref(t), None)
    else:
        rc = lib().oracle_simulate_omp(ctypes.byref(c), n_packets, seed, first_pid, stream, flags,
                                       ctypes.byref(t), threads)
    if rc != 0:
        raise ValueError("oracle_simulate rejected the config (rc=%d)" % rc)
    out = _result(t)
    if per_packet:
        out["per_packet"] = pp
    return out
