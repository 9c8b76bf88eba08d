"""Summaries of rocprofv3 output for the transport kernel.

Reads the CSV directories written by

    rocprofv3 --kernel-trace --stats --output-format csv -d KT -o kt -- python3 bench.py ...
    rocprofv3 --pmc FETCH_SIZE --output-format csv -d F -o f -- python3 bench.py ... --copy-probe BYTES
    rocprofv3 --pmc WRITE_SIZE --output-format csv -d W -o w -- python3 bench.py ... --copy-probe BYTES
    rocprofv3 --pmc <SQ/GRBM activity counters> --output-format csv -d A -o a -- python3 bench.py ...

and prints one JSON object: the kernel's average duration from the trace,
HBM bytes per launch from FETCH_SIZE / WRITE_SIZE (each scaled by what the
same counter reads for the plain copy kernel of known size), and the
activity counters per packet.

    python -m ismc_amd.profiling --packets 268435456 --trace-packets 1073741824 \\
        --copy-bytes 1073741824 --trace KT --fetch F --write W --pmc A --pmc B > counters.json

`--mlp-steps N` times N training steps of the surrogate network (batch
4096) on synthetic rows.  `--issue-rates` instead measures the device's VALU issue cost per wave64
instruction for single instruction kinds (ismc_issue_rate) and prints them
as JSON: the per-cycle figures behind the transport kernel's roofline.
The probes (ismc_issue_rate, ismc_copy_probe) live in libismc_diag.so
(include/ismc_diag.h), beside the product library and not in it.

    python -m ismc_amd.profiling --issue-rates
"""
import argparse
import csv
import glob
import json
import os
import sys

KERNEL = "transport_kernel<float, false, true>"  # the slab between slides has slides
COPY = "copy_probe_kernel"


def _csv(directory, suffix):
    hits = sorted(glob.glob(os.path.join(directory, "*" + suffix)))
    if not hits:
        raise FileNotFoundError("no *%s under %s" % (suffix, directory))
    with open(hits[0], newline="") as f:
        return list(csv.DictReader(f))


def counters(directory, kernel):
    """Average value per dispatch of every counter collected for `kernel`."""
    per = {}
    for row in _csv(directory, "counter_collection.csv"):
        if kernel not in row["Kernel_Name"]:
            continue
        d = per.setdefault(row["Counter_Name"], {})
        d[row["Dispatch_Id"]] = d.get(row["Dispatch_Id"], 0.0) + float(row["Counter_Value"])
    return {k: sum(v.values()) / len(v) for k, v in per.items() if v}


def dispatch_info(directory, kernel):
    for row in _csv(directory, "counter_collection.csv"):
        if kernel in row["Kernel_Name"]:
            return {"grid": int(row["Grid_Size"]), "workgroup": int(row["Workgroup_Size"]),
                    "vgpr": int(row["VGPR_Count"]), "sgpr": int(row["SGPR_Count"]),
                    "lds_bytes": int(row["LDS_Block_Size"])}
    return {}


def kernel_stats(directory):
    return [{"name": r["Name"], "calls": int(r["Calls"]), "avg_ns": float(r["AverageNs"]),
             "min_ns": float(r["MinNs"]), "max_ns": float(r["MaxNs"]), "percent": float(r["Percentage"])}
            for r in _csv(directory, "kernel_stats.csv")]


def summarize(packets, copy_bytes, trace=None, fetch=None, write=None, pmc=(), kernel=KERNEL,
              trace_packets=None):
    """`packets`: packets per launch of the counter runs (the per-packet
    figures); `trace_packets`: per launch of the kernel-trace run, if that
    run used another size (the packets/s figure)."""
    out = {"kernel": kernel, "packets": packets}
    if trace:
        st = kernel_stats(trace)
        out["trace"] = st
        k = [s for s in st if kernel in s["name"]]
        if k:
            tp = trace_packets or packets
            out["avg_ns"] = k[0]["avg_ns"]
            out["trace_packets"] = tp
            out["packets_per_s"] = tp / (k[0]["avg_ns"] * 1e-9)
    if fetch and write:
        f_copy = counters(fetch, COPY)["FETCH_SIZE"]
        w_copy = counters(write, COPY)["WRITE_SIZE"]
        fs = copy_bytes / 1024.0 / f_copy
        ws = copy_bytes / 1024.0 / w_copy
        fk = counters(fetch, kernel)["FETCH_SIZE"]
        wk = counters(write, kernel)["WRITE_SIZE"]
        out.update({"fetch_scale": fs, "write_scale": ws,
                    "hbm_fetch_bytes": fk * 1024.0 * fs, "hbm_write_bytes": wk * 1024.0 * ws})
        out["hbm_bytes_per_packet"] = (out["hbm_fetch_bytes"] + out["hbm_write_bytes"]) / packets
        out.setdefault("raw", {}).update({"FETCH_SIZE": fk, "WRITE_SIZE": wk})
    raw = out.setdefault("raw", {})
    for d in pmc:
        raw.update(counters(d, kernel))
        out.setdefault("dispatch", dispatch_info(d, kernel))
    per = lambda k: raw[k] / packets if k in raw else None
    out["per_packet"] = {
        "valu_wave_insts": per("SQ_INSTS_VALU"),
        "trans_f32_wave_insts": per("SQ_INSTS_VALU_TRANS_F32"),
        "int64_wave_insts": per("SQ_INSTS_VALU_INT64"),
        "salu_insts": per("SQ_INSTS_SALU"),
        "lds_insts": per("SQ_INSTS_LDS"),
    }
    if "SQ_THREAD_CYCLES_VALU" in raw and "SQ_ACTIVE_INST_VALU" in raw:
        lanes = raw["SQ_THREAD_CYCLES_VALU"] / raw["SQ_ACTIVE_INST_VALU"]
        out["active_lanes_per_valu_cycle"] = lanes
        out["simd_lane_efficiency"] = lanes / 64.0
    if "SQ_WAVES" in raw:
        out["waves"] = raw["SQ_WAVES"]
    return out


DIAG_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "libismc_diag.so")
_diag_lib = None


def diag():
    """Load libismc_diag.so (the measurement probes) and declare its
    signatures; no GPU is needed to load it."""
    global _diag_lib
    if _diag_lib is None:
        import ctypes
        if not os.path.exists(DIAG_PATH):
            raise RuntimeError("libismc_diag.so is not built (%s); run `python -m ismc_amd.build`" % DIAG_PATH)
        L = ctypes.CDLL(DIAG_PATH)
        P = ctypes.POINTER
        L.ismc_issue_rate_ops.argtypes = []
        L.ismc_issue_rate_ops.restype = ctypes.c_int32
        L.ismc_issue_rate_name.argtypes = [ctypes.c_int32]
        L.ismc_issue_rate_name.restype = ctypes.c_char_p
        L.ismc_issue_rate.argtypes = [ctypes.c_int, ctypes.c_int32, P(ctypes.c_double), P(ctypes.c_double)]
        L.ismc_issue_rate.restype = ctypes.c_int
        L.ismc_copy_probe.argtypes = [ctypes.c_int, ctypes.c_int64, P(ctypes.c_float)]
        L.ismc_copy_probe.restype = ctypes.c_int
        L.ismc_diag_last_error.argtypes = []
        L.ismc_diag_last_error.restype = ctypes.c_char_p
        _diag_lib = L
    return _diag_lib


def copy_probe(device, nbytes):
    """Device ms of a plain device-to-device copy of nbytes (the byte
    counters' calibration: copy_probe_kernel in a profile run)."""
    import ctypes
    L = diag()
    ms = ctypes.c_float()
    if L.ismc_copy_probe(int(device), int(nbytes), ctypes.byref(ms)) != 0:
        raise RuntimeError("ismc_copy_probe: %s" % L.ismc_diag_last_error().decode())
    return ms.value


def issue_rates(device=0):
    """Wave64 instructions per second, shader clock and cycles per
    instruction for each instruction kind of ismc_issue_rate."""
    import ctypes

    from .engine import Engine
    L = diag()
    n_cu = Engine(device).device_info()["n_cu"]
    out = []
    for k in range(L.ismc_issue_rate_ops()):
        rate, ghz = ctypes.c_double(), ctypes.c_double()
        rc = L.ismc_issue_rate(int(device), k, ctypes.byref(rate), ctypes.byref(ghz))
        if rc != 0:
            raise RuntimeError("ismc_issue_rate(%d): %s" % (k, L.ismc_diag_last_error().decode()))
        simds = 4 * n_cu
        out.append({"op": L.ismc_issue_rate_name(k).decode(), "wave_insts_per_s": rate.value,
                    "shader_ghz": ghz.value,
                    "cycles_per_wave_inst": simds * ghz.value * 1e9 / rate.value,
                    "lanes_per_simd_cycle": 64.0 * rate.value / (simds * ghz.value * 1e9)})
    return {"n_cu": n_cu, "simds": 4 * n_cu, "ops": out}


def mlp_steps(steps=200, batch=4096, rows=65536, device=0, seed=0):
    """Training steps of a fresh surrogate network (5 -> 256 x 4 -> 4) on
    synthetic rows: `steps` Adam steps of `batch` rows in one ismc_mlp_train
    call after a one-epoch warm-up; returns device ms per step and the
    step's dense FLOP against the BF16 MFMA peak."""
    import numpy as np

    from . import surrogate as S
    rng = np.random.default_rng(seed)
    x = rng.uniform(-1, 1, (rows, 5)).astype(np.float32)
    y = np.stack([np.sin(2 * x[:, 0]) + x[:, 1] * x[:, 2], np.tanh(x[:, 3] - x[:, 4]),
                  x[:, 0] ** 2 - x[:, 1], np.cos(x[:, 2] + x[:, 4])], 1).astype(np.float32)
    m = S.MLP(max_batch=batch, device=device)
    try:
        m.set_params(S.init_params(m.arch, seed))
        m.load_dataset(x, y)
        per_epoch = rows // batch
        m.train(1, batch=batch, lr=1e-3, seed=seed)
        epochs = max(1, steps // per_epoch)
        losses, ms = m.train(epochs, batch=batch, lr=1e-3, seed=seed + 1)
    finally:
        m.close()
    n = epochs * per_epoch
    # logical network: 5 -> 256 x 4 -> 4; forward 2 flops per weight per row,
    # backward twice that (dZ W^T and A^T dZ)
    weights = 5 * 256 + 3 * 256 * 256 + 256 * 4
    flop = 3 * 2 * weights * batch
    ms_step = ms / n
    return {"steps": n, "batch": batch, "ms_per_step": ms_step, "flop_per_step": flop,
            "tflops": flop / (ms_step * 1e-3) / 1e12, "mfma_peak_tflops": 2500.0,
            "loss_first_last": [float(losses[0]), float(losses[-1])]}


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--issue-rates", action="store_true",
                    help="measure VALU issue costs per instruction kind instead")
    ap.add_argument("--mlp-steps", type=int, default=0,
                    help="time this many surrogate training steps (batch 4096) instead")
    ap.add_argument("--device", type=int, default=0)
    ap.add_argument("--packets", type=int, help="packets per launch of the counter runs")
    ap.add_argument("--trace-packets", type=int, help="packets per launch of the trace run (default --packets)")
    ap.add_argument("--copy-bytes", type=int, default=1 << 30)
    ap.add_argument("--trace")
    ap.add_argument("--fetch")
    ap.add_argument("--write")
    ap.add_argument("--pmc", action="append", default=[])
    ap.add_argument("--kernel", default=KERNEL)
    a = ap.parse_args(argv)
    if a.issue_rates:
        json.dump(issue_rates(a.device), sys.stdout, indent=1)
        print()
        return
    if a.mlp_steps:
        json.dump(mlp_steps(a.mlp_steps, device=a.device), sys.stdout, indent=1)
        print()
        return
    if a.packets is None:
        ap.error("--packets is required")
    json.dump(summarize(a.packets, a.copy_bytes, a.trace, a.fetch, a.write, a.pmc, a.kernel,
                        a.trace_packets), sys.stdout, indent=1)
    print()


if __name__ == "__main__":
    main()
