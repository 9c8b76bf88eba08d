"""Autotuner host logic with a stand-in engine whose kernel time is a known
function of the launch shape."""
import numpy as np

from ismc_amd import autotune, workloads
from ismc_amd.engine import CONFIG_DTYPE


class TimedEngine:
    """Times a launch as a known function of its shape; the tuner reads the
    time back through last_launch, as from the real engine."""

    def __init__(self, best_chunk=4096, best_blocks=2, best_idle=8):
        self.best = (best_blocks, best_chunk, best_idle)
        self.calls = []
        self.ms = 0.0

    def kernel_ms(self, blocks_per_cu, chunk_packets, idle_lanes, bound_lanes):
        return 1.0 + abs(np.log2(max(chunk_packets, 1)) - np.log2(self.best[1])) + \
            (0.0 if blocks_per_cu == self.best[0] else 0.5) + \
            0.1 * abs(np.log2(max(idle_lanes, 1)) - np.log2(self.best[2])) + \
            0.05 * abs(np.log2(max(bound_lanes, 1)) - 3)

    def simulate_batch_array(self, arr, n_packets, seed=0, streams=None, fp64=False, blocks_per_cu=0,
                             chunk_packets=0, idle_lanes=0, bound_lanes=0, **kw):
        assert arr.dtype == CONFIG_DTYPE and arr.flags.c_contiguous
        npk = np.broadcast_to(np.asarray(n_packets), (arr.shape[0],))
        self.calls.append((blocks_per_cu, chunk_packets, int(npk.sum()), idle_lanes))
        self.ms = self.kernel_ms(blocks_per_cu, chunk_packets, idle_lanes, bound_lanes)
        z = np.zeros((arr.shape[0], 5))
        return z, z, npk.astype(np.uint64)

    def last_launch(self):
        return {"kernel_ms": self.ms}


def test_picks_fastest_shape_and_caches():
    eng = TimedEngine()
    t = autotune.AutoTuner(eng, budget_packets=1 << 20)
    cfgs = [workloads.single_layer(0.1 * (k + 1), 10.0) for k in range(8)]
    kw = t.params(cfgs, 1 << 22)
    assert kw == {"blocks_per_cu": 2, "chunk_packets": 4096, "idle_lanes": 8, "bound_lanes": 8}
    n = len(eng.calls)
    # coordinate search: each knob's values once (the shape shared by two
    # stages is timed once)
    shapes = len(autotune.CHUNKS) + len(autotune.IDLE) + len(autotune.BOUND) + len(autotune.BLOCKS) - 3
    assert n == 1 + shapes * t.repeats
    assert len(t.report(cfgs, 1 << 22)) == shapes
    # every timing launch stays within the budget
    assert all(c[2] <= 1 << 20 for c in eng.calls)
    # same workload class: cached, no new launches
    assert t.params(cfgs, 3 << 21) == kw and len(eng.calls) == n
    # another class (16x the packets per point) is tuned afresh
    t.params(cfgs, 1 << 26)
    assert len(eng.calls) == 2 * n


def test_sample_counts_keep_proportions():
    c = autotune.sample_counts([10, 1000, 100000], 3, budget=1000)
    assert c.tolist() == [1, 9, 990]
    assert autotune.sample_counts(500, 4, budget=1 << 20).tolist() == [500] * 4


def test_small_workloads_skip_chunks_larger_than_any_point():
    eng = TimedEngine(best_chunk=512)
    t = autotune.AutoTuner(eng)
    cfgs = [workloads.single_layer(1.0, 10.0)] * 4
    kw = t.params(cfgs, 3000)
    assert kw["chunk_packets"] == 512
    assert max(c[1] for c in eng.calls) <= 2048


def test_signature_separates_glass_and_precision():
    a = [workloads.single_layer(1.0, 10.0)]
    b = [workloads.slab_between_slides()]
    assert autotune.signature(a, 1e6) != autotune.signature(b, 1e6)
    assert autotune.signature(a, 1e6) != autotune.signature(a, 1e6, fp64=True)
    assert autotune.signature(a, 1.1e6) == autotune.signature(a, 1.5e6)


def test_knobs_stay_unless_a_shape_is_clearly_faster():
    # a candidate measured only 0.5 % faster than the default does not move
    # the knob; one 5 % faster does
    class Close(TimedEngine):
        def kernel_ms(self, blocks_per_cu, chunk_packets, idle_lanes, bound_lanes):
            return 1.0 - (0.005 if idle_lanes == 16 else 0.0) - (0.05 if bound_lanes == 16 else 0.0)
    cfgs = [workloads.single_layer(0.1 * (k + 1), 10.0) for k in range(8)]
    kw = autotune.AutoTuner(Close(), budget_packets=1 << 20).params(cfgs, 1 << 22)
    assert kw["idle_lanes"] == 10 and kw["bound_lanes"] == 16
    kw0 = autotune.AutoTuner(Close(), budget_packets=1 << 20, min_gain=0.0).params(cfgs, 1 << 22)
    assert kw0["idle_lanes"] == 16
