"""Host-side handle on the MI355X transport (one libismc context per Engine)."""
import ctypes
import os
import secrets
import threading
from dataclasses import dataclass, field
from typing import Dict, Optional, Sequence

import numpy as np

from . import _lib
from .config import SampleConfig

TALLIES = ("A", "R_d", "R_s", "T_d", "T_u")
# numpy views of the C structs (include/ismc.h)
CONFIG_DTYPE = np.dtype(_lib.Config)
TALLY_DTYPE = np.dtype(_lib.Tally)
EVENTS = ("interactions", "boundary_hits", "entrances", "sphere_decisions", "reinjections")


@dataclass
class Result:
    """Tallies of one configuration, as fractions of the launched packets
    (CUDAISMLMC.h:754), with their Monte Carlo standard errors."""
    n_packets: int
    mean: Dict[str, float]
    se: Dict[str, float]
    sum: Dict[str, float]
    sumsq: Dict[str, float]
    events: Dict[str, int]
    n_aborted: int = 0
    kernel_ms: float = 0.0
    extra: dict = field(default_factory=dict)

    @property
    def tuple5(self):
        """(A, R_diffuse, R_specular, T_diffuse, T_direct), ISMC.py:406."""
        return tuple(self.mean[k] for k in TALLIES)

    def poisson_unc(self):
        """sqrt(x / N) per tally, the reference's reported uncertainty
        (CUDA/ISMLMC.cu:353-357, ISMC.py:419-428)."""
        n = max(self.n_packets, 1)
        return {k: float(np.sqrt(max(v, 0.0) / n)) for k, v in self.mean.items()}


def _result(t: _lib.Tally, kernel_ms=0.0) -> Result:
    n = int(t.n_packets)
    s = np.array(t.sum[:], dtype=np.float64)
    q = np.array(t.sumsq[:], dtype=np.float64)
    mean = s / n if n else np.zeros(5)
    var = np.maximum(q / n - mean ** 2, 0.0) if n else np.zeros(5)
    se = np.sqrt(var / max(n - 1, 1))
    return Result(n_packets=n, mean=dict(zip(TALLIES, mean.tolist())),
                  se=dict(zip(TALLIES, se.tolist())), sum=dict(zip(TALLIES, s.tolist())),
                  sumsq=dict(zip(TALLIES, q.tolist())),
                  events=dict(zip(EVENTS, [int(x) for x in t.events])),
                  n_aborted=int(t.n_aborted), kernel_ms=kernel_ms)


def _results(outs: np.ndarray, kernel_ms=0.0):
    """_result for every row of a TALLY_DTYPE array, the arithmetic done on
    the whole array at once."""
    n = outs["n_packets"].astype(np.int64)
    s = outs["sum"].astype(np.float64)
    q = outs["sumsq"].astype(np.float64)
    nf = np.maximum(n, 1).astype(np.float64)[:, None]
    mean = np.where(n[:, None] > 0, s / nf, 0.0)
    var = np.where(n[:, None] > 0, np.maximum(q / nf - mean ** 2, 0.0), 0.0)
    se = np.sqrt(var / np.maximum(nf - 1.0, 1.0))
    rows = zip(n.tolist(), mean.tolist(), se.tolist(), s.tolist(), q.tolist(),
               outs["events"].tolist(), outs["n_aborted"].tolist())
    return [Result(n_packets=ni, mean=dict(zip(TALLIES, m)), se=dict(zip(TALLIES, e)),
                   sum=dict(zip(TALLIES, si)), sumsq=dict(zip(TALLIES, qi)),
                   events=dict(zip(EVENTS, ev)), n_aborted=ab, kernel_ms=kernel_ms)
            for ni, m, e, si, qi, ev, ab in rows]


def config_array(cfgs: Sequence[SampleConfig]) -> np.ndarray:
    """The configurations as one C-contiguous CONFIG_DTYPE array (the C
    struct of include/ismc.h, one record per configuration)."""
    return np.frombuffer(b"".join(bytes(c.to_c()) for c in cfgs), dtype=CONFIG_DTYPE).copy()


class Engine:
    """Owns one libismc context on one GPU.  Not thread-safe: use one
    Engine per host thread."""

    def __init__(self, device: int = 0):
        self._L = _lib.load()
        self._ctx = ctypes.c_void_p()
        rc = self._L.ismc_create(int(device), ctypes.byref(self._ctx))
        if rc != _lib.OK:
            raise RuntimeError("ismc_create(device=%d) failed with status %d: no usable HIP "
                               "device (this package has no CPU path)" % (device, rc))
        self.device = device

    def close(self):
        if self._ctx:
            self._L.ismc_destroy(self._ctx)
            self._ctx = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # ------------------------------------------------------------------
    def device_info(self):
        d = _lib.DeviceInfo()
        _lib.check(self._ctx, self._L.ismc_device_info_get(self._ctx, ctypes.byref(d)), "device_info")
        return {"name": d.name.decode(), "arch": d.arch.decode(), "n_cu": d.n_cu,
                "clock_khz": d.clock_khz, "total_mem": d.total_mem, "warp_size": d.warp_size}

    def last_launch(self):
        ms = ctypes.c_float()
        grid = ctypes.c_int32()
        blk = ctypes.c_int32()
        nch = ctypes.c_int64()
        self._L.ismc_last_launch(self._ctx, ctypes.byref(ms), ctypes.byref(grid),
                                 ctypes.byref(blk), ctypes.byref(nch))
        return {"kernel_ms": ms.value, "grid": grid.value, "block": blk.value,
                "chunks": nch.value}

    @staticmethod
    def _flags(fp64, cuda_roulette, slide_bounces=False):
        return ((_lib.FLAG_FP64 if fp64 else 0) | (_lib.FLAG_CUDA_ROULETTE if cuda_roulette else 0) |
                (_lib.FLAG_SLIDE_BOUNCES if slide_bounces else 0))

    def simulate(self, cfg: SampleConfig, n_packets: int, seed: int = 1234,
                 first_packet_id: int = 0, stream: int = 0, fp64: bool = False,
                 cuda_roulette: bool = False, slide_bounces: bool = False) -> Result:
        """One configuration, n_packets packets.  slide_bounces simulates a
        glass slide's internal bounces face by face as the reference does
        (by default they are taken in closed form: same tallies in
        distribution, fewer boundary events)."""
        c = cfg.to_c()
        t = _lib.Tally()
        rc = self._L.ismc_simulate(self._ctx, ctypes.byref(c), int(n_packets), int(seed),
                                   int(first_packet_id), int(stream),
                                   self._flags(fp64, cuda_roulette, slide_bounces), ctypes.byref(t))
        _lib.check(self._ctx, rc, "ismc_simulate")
        return _result(t, self.last_launch()["kernel_ms"])

    def simulate_packets(self, cfg: SampleConfig, n_packets: int, seed: int = 1234,
                         first_packet_id: int = 0, stream: int = 0, fp64: bool = False,
                         cuda_roulette: bool = False, slide_bounces: bool = False):
        """Result plus an (n, 6) array of per-packet {A, R_d, R_s, T_d, T_u, events}."""
        c = cfg.to_c()
        t = _lib.Tally()
        out = np.zeros((int(n_packets), 6), dtype=np.float64)
        rc = self._L.ismc_simulate_packets(self._ctx, ctypes.byref(c), int(n_packets), int(seed),
                                           int(first_packet_id), int(stream),
                                           self._flags(fp64, cuda_roulette, slide_bounces),
                                           out.ctypes.data_as(ctypes.c_void_p), ctypes.byref(t))
        _lib.check(self._ctx, rc, "ismc_simulate_packets")
        return _result(t, self.last_launch()["kernel_ms"]), out

    def simulate_batch(self, cfgs: Sequence[SampleConfig], n_packets, seed: int = 1234,
                       first_packet_ids=None, streams=None, fp64: bool = False,
                       cuda_roulette: bool = False, slide_bounces: bool = False, blocks_per_cu: int = 0,
                       chunk_packets: int = 0, use_graph: bool = False, autotune: bool = False,
                       idle_lanes: int = 0, bound_lanes: int = 0):
        """Many configurations, uneven packet counts, one persistent launch.
        use_graph captures the launch sequence as a hipGraph the second time
        a call repeats an earlier one's shape and seed, and replays it from
        then on (the uploads inside the graph read the new configurations
        from pinned staging); a shape seen once runs without a graph.
        autotune takes the launch shape from this engine's AutoTuner
        (timed on first use per workload class; tallies are identical for
        every launch shape)."""
        npts = len(cfgs)
        if npts == 0:
            return []
        if autotune:
            kw = self.tuner.params(cfgs, n_packets, fp64=fp64, seed=seed)
            blocks_per_cu, chunk_packets = kw["blocks_per_cu"], kw["chunk_packets"]
            idle_lanes = kw.get("idle_lanes", idle_lanes)
            bound_lanes = kw.get("bound_lanes", bound_lanes)
        arr = config_array(cfgs)
        npk = np.ascontiguousarray(np.broadcast_to(np.asarray(n_packets, dtype=np.uint64), (npts,)))
        first = None
        if first_packet_ids is not None:
            first = np.ascontiguousarray(np.broadcast_to(np.asarray(first_packet_ids, dtype=np.uint64), (npts,)))
        st = None
        if streams is not None:
            st = np.ascontiguousarray(np.broadcast_to(np.asarray(streams, dtype=np.uint32), (npts,)))
        tune = _lib.Tune(int(blocks_per_cu), int(chunk_packets), int(bool(use_graph)), int(idle_lanes),
                         int(bound_lanes), 0)
        outs = np.zeros(npts, dtype=TALLY_DTYPE)
        P64 = ctypes.POINTER(ctypes.c_uint64)
        P32 = ctypes.POINTER(ctypes.c_uint32)
        rc = self._L.ismc_simulate_batch(
            self._ctx, arr.ctypes.data_as(ctypes.POINTER(_lib.Config)), npk.ctypes.data_as(P64),
            first.ctypes.data_as(P64) if first is not None else None,
            st.ctypes.data_as(P32) if st is not None else None,
            npts, int(seed), self._flags(fp64, cuda_roulette, slide_bounces), ctypes.byref(tune),
            outs.ctypes.data_as(ctypes.POINTER(_lib.Tally)))
        _lib.check(self._ctx, rc, "ismc_simulate_batch")
        return _results(outs, self.last_launch()["kernel_ms"])

    def simulate_batch_array(self, configs: np.ndarray, n_packets, seed: int = 1234, streams=None,
                             fp64: bool = False, use_graph: bool = False, idle_lanes: int = 0,
                             bound_lanes: int = 0, first_packet_ids=None, chunk_packets: int = 0,
                             blocks_per_cu: int = 0):
        """simulate_batch on configurations already laid out as a numpy array
        of the C struct (dtype CONFIG_DTYPE), for callers that patch a few
        fields of many copies of a template.  Returns (sums [P, 5],
        sums of squares [P, 5], packets [P]) as arrays, in TALLIES order."""
        npts = int(configs.shape[0])
        if npts == 0:
            z = np.zeros((0, 5))
            return z, z, np.zeros(0, dtype=np.uint64)
        if configs.dtype != CONFIG_DTYPE or not configs.flags.c_contiguous:
            raise ValueError("configs must be a C-contiguous array of CONFIG_DTYPE")
        npk = np.ascontiguousarray(np.broadcast_to(np.asarray(n_packets, dtype=np.uint64), (npts,)))
        st = None
        if streams is not None:
            st = np.ascontiguousarray(np.broadcast_to(np.asarray(streams, dtype=np.uint32), (npts,)))
        first = None
        if first_packet_ids is not None:
            first = np.ascontiguousarray(np.broadcast_to(np.asarray(first_packet_ids, dtype=np.uint64), (npts,)))
        tune = _lib.Tune(int(blocks_per_cu), int(chunk_packets), int(bool(use_graph)), int(idle_lanes),
                         int(bound_lanes), 0)
        outs = np.zeros(npts, dtype=TALLY_DTYPE)
        P64 = ctypes.POINTER(ctypes.c_uint64)
        rc = self._L.ismc_simulate_batch(
            self._ctx, configs.ctypes.data_as(ctypes.POINTER(_lib.Config)), npk.ctypes.data_as(P64),
            first.ctypes.data_as(P64) if first is not None else None,
            st.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32)) if st is not None else None, npts, int(seed),
            self._flags(fp64, False), ctypes.byref(tune), outs.ctypes.data_as(ctypes.POINTER(_lib.Tally)))
        _lib.check(self._ctx, rc, "ismc_simulate_batch")
        # event counts of the launch (interactions, boundary hits, entrances,
        # sphere decisions, re-injections), summed over its points
        self._last_events = dict(zip(EVENTS, (int(x) for x in outs["events"].sum(axis=0))))
        return outs["sum"], outs["sumsq"], outs["n_packets"]

    def last_events(self):
        """Event counts of the last simulate_batch_array launch, summed over its points."""
        return dict(getattr(self, "_last_events", dict.fromkeys(EVENTS, 0)))

    @property
    def tuner(self):
        if getattr(self, "_tuner", None) is None:
            from .autotune import AutoTuner
            self._tuner = AutoTuner(self)
        return self._tuner

    def graph_stats(self):
        cap = ctypes.c_int64()
        lau = ctypes.c_int64()
        self._L.ismc_graph_stats(self._ctx, ctypes.byref(cap), ctypes.byref(lau))
        return {"captures": cap.value, "launches": lau.value}

    def eval_physics(self, op: int, inputs, cfg: Optional[SampleConfig] = None, fp64=False):
        x = np.ascontiguousarray(np.asarray(inputs, dtype=np.float64).reshape(-1, 4))
        out = np.zeros((x.shape[0], 2), dtype=np.float64)
        c = cfg.to_c() if cfg is not None else None
        rc = self._L.ismc_eval_physics(self._ctx, int(op), ctypes.byref(c) if c is not None else None,
                                       x.ctypes.data_as(ctypes.c_void_p), x.shape[0],
                                       self._flags(fp64, False), out.ctypes.data_as(ctypes.c_void_p))
        _lib.check(self._ctx, rc, "ismc_eval_physics")
        return out


_default = None
_default_lock = threading.Lock()


def default_engine() -> Engine:
    global _default
    with _default_lock:
        if _default is None:
            _default = Engine(int(os.environ.get("ISMC_DEVICE", "0")))
        return _default


def fresh_seed() -> int:
    """A random 64-bit seed: the reference propagators are unseeded."""
    return secrets.randbits(64)
