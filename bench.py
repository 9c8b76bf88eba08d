"""Photon-transport throughput on one MI355X.

A step is one persistent launch of the transport kernel over N packets of
the slab-between-glass-slides workload (mu_a=1/cm, mu_s=100/cm, g=0.9,
n=1.40, 0.1 cm sample between 0.1 cm n=1.52 slides, dual spheres of radius
4 cm with 0.5 cm sample/source ports, 0.25 cm detector port, wall 0.98;
FP32 transport, Philox seed 1234 + step).  Prints one JSON line, which
also carries the three-layer and FP64 rates, the lookup sweep and tables,
and the metric's second half: surrogate-seeded inverse solves/s against
the plain MC minimizer (`--inverse N` measurements, 4096 by default;
`--no-extra` keeps only the transport step).

    python bench.py [--gpus 1] [--steps 5] [--warmup 1] [--packets 1073741824] [--inverse 4096]
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

# VALU lane-slots per event of a packet history (simple op = 1, quarter-rate
# transcendental = 4, one Philox4x32-10 uniform = 37):
#   interaction 174, interface hit 87, entrance 100, sphere decision 40,
#   re-injection 130, packet set-up 10.
# Peak: a SIMD issues a full-rate FP32 operation for 32 lanes per cycle (a
# wave64 instruction in 2 cycles, measured with ismc_issue_rate), so
# 4 SIMDs x 32 lanes per CU per cycle: 256 CUs x 128 x 2.4 GHz = 78.6 Tslot/s,
# MI355X's published FP32 vector peak of 157.3 TFLOP/s with an FMA as 2.
SLOTS = {"interactions": 174, "boundary_hits": 87, "entrances": 100,
         "sphere_decisions": 40, "reinjections": 130}
SLOTS_PER_PACKET = 10
LANES_PER_CU_CYCLE = 4 * 32


def slots_per_packet(events, n):
    return SLOTS_PER_PACKET + sum(SLOTS[k] * events[k] for k in SLOTS) / float(n)


def cpu_baseline(seconds):
    """Time the pure-Python restatement of the reference CPU propagator
    (single-threaded, like ISMLMC.py) on a bounded sample of the same
    workload, and the C oracle on all host cores."""
    from ismc_amd.workloads import slab_between_slides
    from oracle import ismc_py, oracle
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    from helpers import spec_of
    spec = spec_of(slab_between_slides())
    n, t0 = 0, time.perf_counter()
    chunk = 2000
    while time.perf_counter() - t0 < seconds:
        ismc_py.simulate(spec, chunk, seed=n)
        n += chunk
    py_rate = n / (time.perf_counter() - t0)
    cores = len(os.sched_getaffinity(0))
    threads = min(cores, 16)
    m = 2_000_000
    rates = {}
    for name, flags in (("fp64", 0), ("fp32", oracle.FP32)):
        t0 = time.perf_counter()
        oracle.simulate(spec, m, seed=1, threads=threads, flags=flags)
        rates[name] = m / (time.perf_counter() - t0)
    return {"value": py_rate, "unit": "packets/s", "cores": 1, "kind": "port",
            "sample": "%d packets of the pure-Python restatement (ISMLMC.py/MLMC.py structure, "
                      "math+random) of the same slab-between-slides workload" % n,
            "c_oracle": {"value": rates["fp64"], "unit": "packets/s", "cores": threads,
                         "sample": "%d packets, C oracle FP64, OpenMP" % m},
            "c_oracle_fp32": {"value": rates["fp32"], "unit": "packets/s", "cores": threads,
                              "sample": "%d packets, C oracle FP32, OpenMP" % m}}


def cpu_inverse_baseline(packets_per_eval=2000, target_packets=1_000_000):
    """The reference's CPU inverse path: scipy's Powell (minimization.py:264-266)
    minimizing the relative R_d/T_d cost of the pure-Python restatement, on
    1 core.  Bounded sample: one solve for (mu_a 1 /cm, mu_s' 10 /cm; g 0.8,
    n 1.40, t 0.1 cm, dual spheres) at `packets_per_eval` packets per cost
    evaluation (fixed seed per evaluation), from the plain solver's generic
    guess in log coordinates; solves/s is extrapolated to `target_packets`
    packets per evaluation (the plain GPU solver's N) from the measured
    Python packets/s and the solve's evaluation count."""
    import math
    import numpy as np
    from scipy.optimize import minimize
    from ismc_amd import inverse
    from ismc_amd.minimization import cost_from
    from oracle import ismc_py, oracle
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    from helpers import spec_of
    probe = inverse.problem(0.3, 0.5)
    truth = oracle.simulate(spec_of(probe.configs(1.0, 10.0)[0][0]), 1_000_000, seed=11,
                            threads=min(len(os.sched_getaffinity(0)), 16))["mean"]
    prob = inverse.problem(truth["R_d"], truth["T_d"])
    stats = {"evals": 0, "packets": 0, "seconds": 0.0}

    def cost(y):
        a, s = math.exp(y[0]), math.exp(y[1])
        t0 = time.perf_counter()
        A, Rd, Rs, Td, Tu = ismc_py.simulate(spec_of(prob.configs(a, s)[0][0]), packets_per_eval, seed=5)
        stats["seconds"] += time.perf_counter() - t0
        stats["evals"] += 1
        stats["packets"] += packets_per_eval
        return cost_from({"R_d": Rd, "T_d": Td}, prob.measured)

    t0 = time.perf_counter()
    res = minimize(cost, np.log(inverse.GENERIC_GUESS), method="Powell", tol=3e-3)
    wall = time.perf_counter() - t0
    rate = stats["packets"] / stats["seconds"]
    secs_per_solve = stats["evals"] * target_packets / rate
    return {"value": 1.0 / secs_per_solve, "unit": "solves/s", "cores": 1, "kind": "port",
            "sample": "one scipy Powell solve on the pure-Python restatement at %d packets per "
                      "evaluation: %d evaluations, %.1f s, %.3g packets/s; extrapolated to %d "
                      "packets per evaluation" % (packets_per_eval, stats["evals"], wall, rate,
                                                   target_packets),
            "solution": [math.exp(res.x[0]), math.exp(res.x[1])]}


def lookup_sweep(eng):
    """64 x 64 single-layer lookup grid (mu_a 0.01..10 /cm, mu_s 1..1000 /cm,
    g 0.8, n 1.40, t 0.1 cm, dual spheres), adaptive packet counts: a 1e4
    pilot, then N_ij = clamp(Var/(R 0.005)^2, 1e4, 1e6) in one launch.  The
    first sweep tunes the launch shape; the second is timed."""
    from ismc_amd import lookup
    from ismc_amd.workloads import lookup_grid
    _, _, cfgs = lookup_grid(64, 64)
    t0 = time.perf_counter()
    lookup.adaptive_sweep(cfgs, engine=eng, autotune=True)
    t_tune = time.perf_counter() - t0
    t0 = time.perf_counter()
    res, info = lookup.adaptive_sweep(cfgs, engine=eng, autotune=True, seed=4321)
    wall = time.perf_counter() - t0
    kern = (info["pilot_kernel_ms"] + info["main_kernel_ms"] + info["topup_kernel_ms"]) / 1e3
    rse = [r["rse_R"] for r in res]
    n = info["total_packets"]
    W = slots_per_packet(info["events"], n)
    out = {"points": len(cfgs), "total_packets": n, "wall_s": wall,
           "kernel_s": kern, "packets_per_s": n / wall,
           "kernel_packets_per_s": n / kern,
           "max_rse_R": max(rse), "max_rse_R_uncapped": info["max_rse_R_uncapped"],
           "capped_points": info["capped_points"], "topup_launches": info["topup_launches"],
           "topup_packets": info["topup_packets"], "first_sweep_with_tuning_s": t_tune,
           "roofline": roofline_of(W, n, kern, eng, info["events"]),
           "tuned": {str(k): v for k, v in eng.tuner.cache.items()}}
    return out, res


def roofline_of(W, packets, kernel_s, eng, events):
    """Slot-roofline figures of transport launches: W slots per packet,
    `packets` packets in `kernel_s` seconds of kernel time."""
    info = eng.device_info()
    peak = info["n_cu"] * LANES_PER_CU_CYCLE * info["clock_khz"] * 1e3 / 1e12
    achieved = W * packets / kernel_s / 1e12
    return {"bound": "valu", "achieved": achieved, "peak": peak, "unit": "Tslot/s", "frac": achieved / peak,
            "slots_per_packet": W, "events_per_packet": {k: v / packets for k, v in events.items()}}


def lookup_tables(sweep_res, P=1000, M=200, cpu_cells=12):
    """Lookup tables from the 64 x 64 sweep (lookup.py:300-376): griddata onto
    a P x P (R_d, T_d) grid on the host, then the per-cell uncertainty map
    (M draws per cell) on the GPU.  CPU baseline: the reference's per-cell
    loop (oracle restatement, argmin over the grid per draw) on `cpu_cells`
    cells, extrapolated to every cell the map covers."""
    import numpy as np
    from ismc_amd import lookup_table as LT
    from ismc_amd.workloads import lookup_grid
    mua, mus, _ = lookup_grid(64, 64)
    mus_r = mus * (1 - 0.8)
    grid = lambda key, f: np.array([f(r, key) for r in sweep_res]).reshape(64, 64)
    Rd, Td = grid("R_d", lambda r, k: r["mean"][k]), grid("T_d", lambda r, k: r["mean"][k])
    Rs, Ts = grid("R_d", lambda r, k: r["se"][k]), grid("T_d", lambda r, k: r["se"][k])
    t0 = time.perf_counter()
    tabs = LT.make_tables(Rd, Td, Rs, Ts, mua, mus_r, P=P)
    t_grid = time.perf_counter() - t0
    LT.table_uncertainty(tabs, M=8, seed=1)       # first call: module load
    t0 = time.perf_counter()
    ua, us = LT.table_uncertainty(tabs, M=M, seed=2)
    t_unc = time.perf_counter() - t0
    cells = int(np.isfinite(ua).sum())
    from oracle import table_ref
    idx = np.argwhere(np.isfinite(ua))[:: max(1, cells // cpu_cells)][:cpu_cells]
    t0 = time.perf_counter()
    table_ref.uncertainty(tabs.mua, tabs.mus, tabs.rstd, tabs.tstd, tabs.Rvals, tabs.Tvals, M, 2,
                          cells=[tuple(c) for c in idx])
    t_cpu = (time.perf_counter() - t0) / len(idx)
    return {"P": P, "M": M, "cells_with_uncertainty": cells, "griddata_s": t_grid,
            "uncertainty_map_s": t_unc, "draws_per_s": cells * M / t_unc,
            "cpu_baseline": {"value": 1.0 / (t_cpu * cells), "unit": "uncertainty maps/s", "cores": 1,
                             "kind": "port", "seconds_per_map": t_cpu * cells,
                             "sample": "%d cells of the reference's per-cell loop (oracle "
                                       "restatement, numpy argmin per draw), extrapolated to %d "
                                       "cells" % (len(idx), cells)}}


def surrogate_bench(eng, n_samples=16384, packets=100_000):
    """The packaged network against fresh MC labels (the surrogate-training
    workload's box: mu_a 0.01-10 /cm, mu_s 1-1000 /cm, g 0.6-0.95, n
    1.33-1.50, t 0.05-0.5 cm, log-uniform / uniform; 1e5 packets per label,
    a seed the training never used), the label generation's packets/s, and
    the device time of one training step (batch 4096) of a fresh network on
    those labels; the label launches' kernel time and slot roofline (their
    histories are long: mu_s t reaches 500)."""
    from ismc_amd import surrogate as S
    sur = S.Surrogate.load(S.packaged_weights())
    rep = S.holdout_report(sur, n_samples=n_samples, packets=packets, engine=eng)
    ev = rep.pop("label_events")
    npk = rep["label_packets"]
    rep["label_roofline"] = roofline_of(slots_per_packet(ev, npk), npk, rep["label_kernel_seconds"], eng, ev)
    inp = rep.pop("inputs")
    lab = rep.pop("labels")
    rep.pop("label_se")
    x = S.features(inp["mu_a"], inp["mu_s"], inp["g"], inp["n"], inp["t"])
    rep["training_step"] = S.time_training(x, S.logit(lab).astype("float32"), device=eng.device)
    return rep


def inverse_bench(eng, count):
    """Surrogate-seeded batched inverse solves vs the plain solver (sequential
    PRAXIS, one MC launch per cost evaluation) on this GPU, in two regimes:
    measurements with 1 % noise (simulated at 1e6 packets; refinement at
    1e5, judge and plain solver at 1e6 packets per evaluation), and with
    0.1 % noise (simulated at 1e7; refinement at 1e6, judge and plain solver
    at 1e7), where a 1 % recovered error is within reach.  The plain
    solver's accuracy comes from a lock-step batch of its solves (bit for bit
    the sequential solutions), its solves/s from a sequentially timed
    sample whose solutions are checked against the batch; at 1 % noise also
    minimize.o's form, the mean of 5 window-accepted runs."""
    from ismc_amd import inverse
    from ismc_amd.surrogate import Surrogate, packaged_weights
    sur = Surrogate.load(packaged_weights())
    log = lambda *x: print(*x, file=sys.stderr, flush=True)
    out = {}
    for key, methods, kw in (
            ("noise_1pct", ("newton", "praxis"),
             dict(count=count, noise=0.01, plain=((1, 512, 96), (5, 128, 8)))),
            ("noise_0.1pct", ("newton",),
             dict(count=count, noise=0.001, plain=((1, 128, 32),), meas_packets=10_000_000,
                  refine_packets=1_000_000, check_packets=10_000_000, plain_packets=10_000_000,
                  plain_tol=1e-3))):
        r = inverse.run(sur, engine=eng, methods=methods, with_ablation=False, log=log, **kw)
        p = r["plain"]
        plain_ok = p["solves_per_s"] * p["within_1pct"]["frac"]
        entry = {"measurements": r["count"], "noise": r["noise"],
                 "packets": {"measurement": r["meas_packets"], "refinement": r["refine_packets"],
                             "judge": r["check_packets"], "plain_per_evaluation": r["plain_packets"]},
                 "speedups_against": "plain (runs=1, the faster plain form)"}
        for k in r:
            if k.startswith("plain"):
                entry[k] = r[k]
        for m in methods:
            n = r["seeded_" + m]
            w, pw = n["within_1pct"], p["within_1pct"]
            sps_ratio = n["solves_per_s"] / p["solves_per_s"]
            lo, hi = inverse.ratio_ci(w["count"], w["of"], pw["count"], pw["of"])
            plo, phi = p["solves_per_s_ci95"]
            entry[{"newton": "defect_correction", "praxis": "praxis"}[m]] = {
                "solves_per_s": n["solves_per_s"], "speedup_vs_plain": sps_ratio,
                # from the plain rate's bootstrap interval (its per-solve time is heavy-tailed)
                "speedup_vs_plain_ci95": [n["solves_per_s"] / phi, n["solves_per_s"] / plo],
                "within_1pct": w,
                # solves that recover both mu_a and mu_s' within 1 %, per second, and that
                # rate against the plain solver's, with a 95 % interval from both binomial counts
                "within_1pct_solves_per_s": n["solves_per_s"] * w["frac"],
                "within_1pct_speedup_vs_plain": (n["solves_per_s"] * w["frac"] / plain_ok) if plain_ok > 0 else None,
                # both binomial counts and the plain rate's bootstrap interval (product of
                # the two intervals' ends: conservative)
                "within_1pct_speedup_ci95": [n["solves_per_s"] / phi * lo, n["solves_per_s"] / plo * hi],
                "fits_as_well_as_truth": n["fits_as_well_as_truth"], "errors": n["errors"],
                "misses": n["misses"], "misses_repaired_by_fallback": n["misses_repaired_by_fallback"],
                "fallback": n["fallback"], "mc_evaluations_per_solve": n["mean_evals"],
                "seconds": {k: n[k + "_s"] for k in ("seed", "refine", "check", "fallback")}}
        out[key] = entry
    print("bench: timing the CPU inverse baseline", file=sys.stderr, flush=True)
    out["cpu_baseline"] = cpu_inverse_baseline()
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--packets", type=int, default=1 << 30)
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-extra", action="store_true")
    ap.add_argument("--inverse", type=int, default=4096,
                    help="surrogate-seeded inverse solves run against the plain solver (the metric's "
                         "second half; 0 skips them)")
    ap.add_argument("--copy-probe", type=int, default=0,
                    help="also run a plain HBM copy of this many bytes (byte-counter calibration)")
    a = ap.parse_args()
    if a.gpus != 1:
        print("bench.py measures one GPU only: the transport runs on a single MI355X here "
              "(packets shard by packet id across GPUs, with one sum of the tallies at the end)", file=sys.stderr)
        sys.exit(2)

    from ismc_amd import Engine
    from ismc_amd.workloads import slab_between_slides, three_layer_sample

    eng = Engine(0)
    info = eng.device_info()
    cfg = slab_between_slides()
    N = int(a.packets)
    print("bench: %s (%s, %d CUs @ %d MHz), %d packets/step" %
          (info["name"], info["arch"], info["n_cu"], info["clock_khz"] // 1000, N),
          file=sys.stderr, flush=True)
    for i in range(a.warmup):
        eng.simulate(cfg, N, seed=1000 + i)
    kernel_ms = []
    events = None
    t0 = time.perf_counter()
    for i in range(a.steps):
        r = eng.simulate(cfg, N, seed=1234 + i)
        kernel_ms.append(r.kernel_ms)
        if events is None:
            events = r.events
    elapsed = time.perf_counter() - t0
    value = N * a.steps / elapsed
    launch = eng.last_launch()

    traffic = None
    prof = os.path.join(ROOT, "profiles", "transport_slab_between_slides_counters.json")
    if os.path.exists(prof):
        with open(prof) as f:
            traffic = json.load(f)["hbm_bytes_per_packet"] * N
    W = slots_per_packet(events, N)
    avg_s = sum(kernel_ms) / len(kernel_ms) / 1e3
    achieved = W * N / avg_s / 1e12
    peak = info["n_cu"] * LANES_PER_CU_CYCLE * info["clock_khz"] * 1e3 / 1e12
    out = {
        # value is the first half (packets/s); the second half is under "inverse"
        "metric": "photon packets/s on 1 GPU; inverse solves/s at \u22641% \u03bca/\u03bcs' error vs plain MC",
        "value": value,
        "unit": "packets/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": elapsed / a.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32",
        "data": "synthetic",
        "config": {"workload": "single-layer slab between glass slides, dual integrating spheres",
                   "packets_per_step": N, "layers": "glass 1.52/0.1 | mu_a 1, mu_s 100, g 0.9, "
                   "n 1.40, 0.1 cm | glass 1.52/0.1", "spheres": "R=4 cm, ports 0.5/0.5/0.25 cm, pw 0.98",
                   "roulette": "1e-4, 1 in 10", "seed": 1234,
                   "slides": "bounces inside each slide in closed form (one boundary event per crossing; "
                             "ISMC_FLAG_SLIDE_BOUNCES gives the reference's face-by-face events, same tallies)",
                   "grid_blocks": launch["grid"],
                   "block_threads": launch["block"]},
        "roofline": {"bound": "valu", "achieved": achieved, "peak": peak, "unit": "Tslot/s",
                     "frac": achieved / peak, "traffic": traffic,
                     "traffic_note": "HBM bytes per launch: rocprofv3 FETCH_SIZE x2.00 + WRITE_SIZE x1.00 "
                                     "(copy-calibrated) at 2^28 packets, per packet x N",
                     "peak_note": "n_CU x 4 SIMDs x 32 lanes x clock: a SIMD issues a full-rate FP32 "
                                  "wave64 instruction in 2 cycles (ismc_issue_rate); at 64 lanes per CU "
                                  "per cycle this fraction would be %.2f" % (2 * achieved / peak),
                     "slots_per_packet": W, "kernel_ms_avg": avg_s * 1e3,
                     "kernel_packets_per_s": N / avg_s,
                     "events_per_packet": {k: v / N for k, v in events.items()}},
    }
    if not a.no_extra:
        n3 = 100_000_000
        r3 = eng.simulate(three_layer_sample(), n3, seed=77)
        out["three_layer_packets_per_s"] = n3 / (r3.kernel_ms / 1e3)
        out["three_layer_roofline"] = roofline_of(slots_per_packet(r3.events, n3), n3, r3.kernel_ms / 1e3,
                                                  eng, r3.events)
        n64 = 1 << 26
        r64 = eng.simulate(cfg, n64, seed=78, fp64=True)
        out["fp64_kernel_packets_per_s"] = n64 / (r64.kernel_ms / 1e3)
    if not a.no_extra:
        out["lookup_sweep"], sweep_res = lookup_sweep(eng)
        out["lookup_tables"] = lookup_tables(sweep_res)
    if not a.no_extra:
        print("bench: surrogate against fresh labels", file=sys.stderr, flush=True)
        out["surrogate"] = surrogate_bench(eng)
    if a.inverse and not a.no_extra:
        out["inverse"] = inverse_bench(eng, a.inverse)
    if a.copy_probe:
        from ismc_amd import profiling
        ms = profiling.copy_probe(eng.device, a.copy_probe)
        out["copy_probe"] = {"bytes": a.copy_probe, "kernel_ms": ms,
                             "GB_per_s": 2 * a.copy_probe / ms / 1e6}
    if not a.no_cpu:
        print("bench: timing the CPU baseline (%.0f s)" % a.cpu_seconds, file=sys.stderr, flush=True)
        out["cpu_baseline"] = cpu_baseline(a.cpu_seconds)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
