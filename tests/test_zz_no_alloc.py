"""The no-allocation entry point, ismc_simulate_batch_static, through the C
ABI: on two calls in a row (the second smaller, so the shared scratch array
still holds the first call's larger contents) its tallies equal those of
ismc_simulate_batch bit for bit, in FP32 and in FP64, and a batch beyond
its limits is refused before any device work, leaving the context usable."""
import ctypes

import pytest

from ismc_amd import _lib
from ismc_amd.workloads import single_layer, slab_between_slides, three_layer_sample

pytestmark = pytest.mark.gpu


def call(fn, ctx, cfgs, npk, first, streams, seed, flags=0):
    n = len(cfgs)
    arr = (_lib.Config * n)(*[c.to_c() for c in cfgs])
    out = (_lib.Tally * n)()
    tune = _lib.Tune(0, 0, 0, 0, 0, 0)
    rc = fn(ctx, arr, (ctypes.c_uint64 * n)(*npk), (ctypes.c_uint64 * n)(*first),
            (ctypes.c_uint32 * n)(*streams), n, seed, flags, ctypes.byref(tune), out)
    return rc, out


def rows(out):
    return [(tuple(t.sum), tuple(t.sumsq), tuple(t.events), t.n_packets, t.n_aborted) for t in out]


def both(engine, cfgs, npk, first, streams, seed, flags=0):
    L = _lib.load()
    print("no-alloc case: %d points, %s packets, flags %d" % (len(cfgs), list(npk), flags), flush=True)
    rs, out_s = call(L.ismc_simulate_batch_static, engine._ctx, cfgs, npk, first, streams, seed, flags)
    assert rs == 0, L.ismc_last_error(engine._ctx)
    rb, out_b = call(L.ismc_simulate_batch, engine._ctx, cfgs, npk, first, streams, seed, flags)
    assert rb == 0, L.ismc_last_error(engine._ctx)
    assert rows(out_s) == rows(out_b)
    return out_s


def test_static_batch_equals_batch_on_two_calls(engine):
    big = [slab_between_slides(), three_layer_sample(), single_layer(1.0, 10.0), single_layer(0.1, 300.0)]
    out = both(engine, big, [300_000, 200_000, 0, 123_457], [0, 1000, 5, 7], [3, 4, 5, 6], seed=21)
    assert [t.n_packets for t in out] == [300_000, 200_000, 0, 123_457]
    assert 0.99 < sum(out[0].sum) / 300_000 < 1.01
    out = both(engine, [single_layer(2.0, 50.0)], [50_000], [11], [9], seed=22)
    assert out[0].n_packets == 50_000 and out[0].sum[1] > 0


def test_static_batch_fp64(engine):
    both(engine, [slab_between_slides(), single_layer(1.0, 10.0)], [20_000, 30_000], [0, 0], [0, 1],
         seed=5, flags=_lib.FLAG_FP64)
    both(engine, [three_layer_sample()], [10_000], [3], [2], seed=6, flags=_lib.FLAG_FP64)


def test_static_batch_limits(engine):
    L = _lib.load()
    limit = 1024
    cfgs = [single_layer(1.0, 10.0)] * (limit + 1)
    rc, _ = call(L.ismc_simulate_batch_static, engine._ctx, cfgs, [10] * (limit + 1), [0] * (limit + 1),
                 list(range(limit + 1)), 1)
    assert rc == _lib.ERANGE and b"1024" in L.ismc_last_error(engine._ctx)
    rc, _ = call(L.ismc_simulate_batch_static, engine._ctx, cfgs[:2], [10, (1 << 31) + 1], [0, 0], [0, 1], 1)
    assert rc == _lib.ERANGE and b"2^31" in L.ismc_last_error(engine._ctx)
    # the largest batch it takes, then the context still runs both paths
    both(engine, cfgs[:limit], [100] * limit, [0] * limit, list(range(limit)), seed=7)
