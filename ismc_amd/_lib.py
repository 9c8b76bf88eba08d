"""ctypes binding of libismc.so (declarations in include/ismc.h).

The library is built in-tree by ismc_amd.build (hipcc, gfx950).  Loading
fails loudly when it is missing or when no GPU is visible: there is no CPU
fallback anywhere in the package.
"""
import ctypes
import os
import threading

MAX_LAYERS = 16
FLAG_FP64 = 0x1
FLAG_CUDA_ROULETTE = 0x4
FLAG_SLIDE_BOUNCES = 0x8

OK, EINVAL, EDEVICE, ENOMEM, ERANGE = 0, -1, -2, -3, -4
_STATUS = {EINVAL: "invalid argument", EDEVICE: "HIP error", ENOMEM: "out of memory",
           ERANGE: "size out of range"}

LIB_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "libismc.so")

EXPORTS = ("ismc_create", "ismc_destroy", "ismc_last_error", "ismc_device_info_get",
           "ismc_simulate", "ismc_simulate_batch", "ismc_simulate_batch_static",
           "ismc_simulate_packets", "ismc_last_launch",
           "ismc_eval_physics", "ismc_validate_config", "ismc_graph_stats",
           "ismc_table_uncertainty", "ismc_table_last_error", "ismc_grid_argmin")


class Sphere(ctypes.Structure):
    _fields_ = [("present", ctypes.c_int32), ("_pad", ctypes.c_int32),
                ("pw", ctypes.c_double), ("fs", ctypes.c_double), ("fp", ctypes.c_double),
                ("fd", ctypes.c_double), ("f", ctypes.c_double), ("angle", ctypes.c_double)]


class Config(ctypes.Structure):
    _fields_ = [("n_layers", ctypes.c_int32), ("specular_included", ctypes.c_int32),
                ("n_ambient", ctypes.c_double), ("roulette_threshold", ctypes.c_double),
                ("roulette_m", ctypes.c_double),
                ("n", ctypes.c_double * MAX_LAYERS), ("g", ctypes.c_double * MAX_LAYERS),
                ("t", ctypes.c_double * MAX_LAYERS), ("mu_a", ctypes.c_double * MAX_LAYERS),
                ("mu_s", ctypes.c_double * MAX_LAYERS), ("mu_t", ctypes.c_double * MAX_LAYERS),
                ("R", Sphere), ("T", Sphere)]


class Tally(ctypes.Structure):
    _fields_ = [("sum", ctypes.c_double * 5), ("sumsq", ctypes.c_double * 5),
                ("events", ctypes.c_uint64 * 5), ("n_packets", ctypes.c_uint64),
                ("n_aborted", ctypes.c_uint64)]


class Tune(ctypes.Structure):
    _fields_ = [("blocks_per_cu", ctypes.c_int32), ("chunk_packets", ctypes.c_int32),
                ("use_graph", ctypes.c_int32), ("idle_lanes", ctypes.c_int32),
                ("bound_lanes", ctypes.c_int32), ("_reserved", ctypes.c_int32)]


class DeviceInfo(ctypes.Structure):
    _fields_ = [("name", ctypes.c_char * 256), ("arch", ctypes.c_char * 64),
                ("n_cu", ctypes.c_int32), ("clock_khz", ctypes.c_int32),
                ("total_mem", ctypes.c_int64), ("warp_size", ctypes.c_int32),
                ("_pad", ctypes.c_int32)]


_lib = None
_lock = threading.Lock()


def load():
    """Load libismc.so and declare its signatures (no GPU needed to load)."""
    global _lib
    with _lock:
        if _lib is not None:
            return _lib
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(
                "libismc.so is not built (%s); run `python -c 'import __graft_entry__ as g; "
                "g.build()'` or `python -m ismc_amd.build`" % LIB_PATH)
        L = ctypes.CDLL(LIB_PATH)
        P = ctypes.POINTER
        vp = ctypes.c_void_p
        L.ismc_create.argtypes = [ctypes.c_int, P(vp)]
        L.ismc_create.restype = ctypes.c_int
        L.ismc_destroy.argtypes = [vp]
        L.ismc_destroy.restype = None
        L.ismc_last_error.argtypes = [vp]
        L.ismc_last_error.restype = ctypes.c_char_p
        L.ismc_device_info_get.argtypes = [vp, P(DeviceInfo)]
        L.ismc_device_info_get.restype = ctypes.c_int
        L.ismc_simulate.argtypes = [vp, P(Config), ctypes.c_uint64, ctypes.c_uint64,
                                    ctypes.c_uint64, ctypes.c_uint32, ctypes.c_uint32, P(Tally)]
        L.ismc_simulate.restype = ctypes.c_int
        L.ismc_simulate_batch.argtypes = [vp, P(Config), P(ctypes.c_uint64), P(ctypes.c_uint64),
                                          P(ctypes.c_uint32), ctypes.c_int32, ctypes.c_uint64,
                                          ctypes.c_uint32, P(Tune), P(Tally)]
        L.ismc_simulate_batch.restype = ctypes.c_int
        L.ismc_simulate_batch_static.argtypes = L.ismc_simulate_batch.argtypes
        L.ismc_simulate_batch_static.restype = ctypes.c_int
        L.ismc_simulate_packets.argtypes = [vp, P(Config), ctypes.c_uint64, ctypes.c_uint64,
                                            ctypes.c_uint64, ctypes.c_uint32, ctypes.c_uint32,
                                            vp, P(Tally)]
        L.ismc_simulate_packets.restype = ctypes.c_int
        L.ismc_last_launch.argtypes = [vp, P(ctypes.c_float), P(ctypes.c_int32),
                                       P(ctypes.c_int32), P(ctypes.c_int64)]
        L.ismc_last_launch.restype = ctypes.c_int
        L.ismc_eval_physics.argtypes = [vp, ctypes.c_int32, P(Config), vp, ctypes.c_int32,
                                        ctypes.c_uint32, vp]
        L.ismc_eval_physics.restype = ctypes.c_int
        L.ismc_validate_config.argtypes = [P(Config), ctypes.c_char_p, ctypes.c_int32]
        L.ismc_validate_config.restype = ctypes.c_int
        L.ismc_graph_stats.argtypes = [vp, P(ctypes.c_int64), P(ctypes.c_int64)]
        L.ismc_graph_stats.restype = ctypes.c_int
        L.ismc_table_uncertainty.argtypes = [ctypes.c_int] + [vp] * 6 + [ctypes.c_int32, ctypes.c_int32,
                                                                         ctypes.c_uint64, vp, vp]
        L.ismc_table_uncertainty.restype = ctypes.c_int
        L.ismc_table_last_error.argtypes = []
        L.ismc_table_last_error.restype = ctypes.c_char_p
        L.ismc_grid_argmin.argtypes = [ctypes.c_int, vp, vp, ctypes.c_int32, vp, vp, ctypes.c_int32, vp]
        L.ismc_grid_argmin.restype = ctypes.c_int
        _lib = L
        return L


def validate(cfg_c):
    """(status, message) of ismc_validate_config; needs no GPU."""
    L = load()
    buf = ctypes.create_string_buffer(512)
    rc = L.ismc_validate_config(ctypes.byref(cfg_c), buf, 512)
    return rc, buf.value.decode(errors="replace")


def check(ctx_ptr, rc, what):
    if rc == OK:
        return
    L = load()
    msg = L.ismc_last_error(ctx_ptr) if ctx_ptr else b""
    msg = msg.decode(errors="replace") if msg else ""
    kind = _STATUS.get(rc, "status %d" % rc)
    if rc in (EINVAL, ERANGE):
        raise ValueError("%s: %s (%s)" % (what, msg or kind, kind))
    if rc == ENOMEM:
        raise MemoryError("%s: %s" % (what, msg or kind))
    raise RuntimeError("%s: %s (%s)" % (what, msg or kind, kind))
