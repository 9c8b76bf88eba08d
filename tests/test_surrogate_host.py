"""Surrogate label generation's host arithmetic with a stand-in engine: the
labels and standard errors formed on a batch's tally arrays equal those of
the per-point Results that simulate_batch returns for the same tallies."""
import numpy as np

from ismc_amd import surrogate as S
from ismc_amd.engine import CONFIG_DTYPE, EVENTS, TALLY_DTYPE, _results


class TallyEngine:
    """Returns seeded tallies shaped like the kernel's (sums within [0, N],
    sums of squares between sum^2 / N and sum) and records the launches."""

    def __init__(self):
        self.rng = np.random.default_rng(7)
        self.launches = []
        self.outs = []

        class Tuner:
            def params(_, cfgs, n_packets, fp64=False, seed=0):
                return {"blocks_per_cu": 0, "chunk_packets": 512, "idle_lanes": 10, "bound_lanes": 12}
        self.tuner = Tuner()

    def simulate_batch_array(self, arr, n_packets, seed=0, streams=None, **kw):
        assert arr.dtype == CONFIG_DTYPE and arr.flags.c_contiguous
        P = arr.shape[0]
        npk = np.broadcast_to(np.asarray(n_packets, dtype=np.uint64), (P,))
        self.launches.append((P, int(seed), streams.tolist(), kw))
        o = np.zeros(P, dtype=TALLY_DTYPE)
        o["n_packets"] = npk
        frac = self.rng.dirichlet(np.ones(5), P)
        o["sum"] = frac * npk[:, None]
        o["sumsq"] = o["sum"] * self.rng.uniform(frac, 1.0)
        o["events"] = self.rng.integers(0, 1000, (P, 5))
        self.outs.append(o)
        self.ms = 2.5
        return o["sum"].copy(), o["sumsq"].copy(), o["n_packets"].copy()

    def last_launch(self):
        return {"kernel_ms": self.ms}

    def last_events(self):
        return dict(zip(EVENTS, (int(x) for x in self.outs[-1]["events"].sum(axis=0))))


def test_labels_and_errors_match_per_point_results():
    eng = TallyEngine()
    st = {}
    inp, tal, se, secs, npk = S.generate_dataset(700, 3000, seed=4, engine=eng, return_se=True,
                                                 points_per_launch=256, stats=st)
    assert npk == 700 * 3000 and [l[0] for l in eng.launches] == [256, 256, 188]
    assert all(l[1] == 4 + 7 for l in eng.launches)
    assert sum((l[2] for l in eng.launches), []) == list(range(700))
    assert all(l[3]["chunk_packets"] == 512 for l in eng.launches)
    res = [r for o in eng.outs for r in _results(o)]
    for i, r in enumerate(res):
        m = r.mean
        assert tal[i].tolist() == [m["R_d"], m["T_d"], m["T_u"], m["R_d"] + m["R_s"]]
        N = r.n_packets
        rt = m["R_d"] + m["R_s"]
        v_tot = max((r.sumsq["R_d"] + r.sumsq["R_s"]) / N - rt * rt, 0.0)
        np.testing.assert_allclose(se[i], [r.se["R_d"], r.se["T_d"], r.se["T_u"], (v_tot / (N - 1)) ** 0.5],
                                   rtol=1e-12)
    assert st["kernel_ms"] == 3 * 2.5
    assert st["events"]["interactions"] == int(sum(o["events"][:, 0].sum() for o in eng.outs))
    assert st["tuning_s"] >= 0.0


def test_without_autotune_chunks_follow_the_packet_count():
    eng = TallyEngine()
    S.generate_dataset(10, 3000, seed=1, engine=eng, autotune=False)
    assert eng.launches[0][3] == {"chunk_packets": 3000}
