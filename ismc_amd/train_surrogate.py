"""Generate MC-labelled data, train the surrogate, and run the inverse
benchmark on one GPU.

    python -m ismc_amd.train_surrogate --samples 262144 --packets 20000 \\
        --epochs 60 --out weights.npz [--inverse 4096] [--json result.json]
"""
import argparse
import json
import sys
import time

import numpy as np

from . import inverse, surrogate
from .engine import Engine


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--samples", type=int, default=1 << 18)
    ap.add_argument("--packets", type=int, default=20000)
    ap.add_argument("--epochs", type=int, default=60)
    ap.add_argument("--batch", type=int, default=4096)
    ap.add_argument("--lr", type=float, default=2e-3)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--out", default="")
    ap.add_argument("--load", default="", help="skip training, load weights")
    ap.add_argument("--data", default="", help="train on labels saved by --save-data instead of generating them")
    ap.add_argument("--save-data", default="", help="save the generated inputs and labels (float32 .npz)")
    ap.add_argument("--inverse", type=int, default=0, help="measurements for the inverse benchmark")
    ap.add_argument("--plain", type=int, default=8)
    ap.add_argument("--refine-packets", type=int, default=100000)
    ap.add_argument("--json", default="")
    ap.add_argument("--no-ablation", action="store_true", help="skip the unseeded batched solves")
    ap.add_argument("--methods", default="newton,praxis")
    ap.add_argument("--noise", type=float, default=0.01, help="relative noise on the synthetic measurements")
    ap.add_argument("--meas-packets", type=int, default=1_000_000)
    ap.add_argument("--check-packets", type=int, default=1_000_000)
    ap.add_argument("--plain-packets", type=int, default=1_000_000)
    ap.add_argument("--plain-tol", type=float, default=3e-3)
    a = ap.parse_args(argv)
    log = lambda *x: print(*x, file=sys.stderr, flush=True)
    eng = Engine(0)
    result = {}
    if a.load:
        sur = surrogate.Surrogate.load(a.load)
    else:
        if a.data:
            d = np.load(a.data, allow_pickle=False)
            inp = {k: d[k].astype(np.float64) for k in ("mu_a", "mu_s", "g", "n", "t")}
            tal = d["labels"].astype(np.float64)
            a.packets = int(d["packets"])
            result["dataset"] = {"samples": int(tal.shape[0]), "packets_per_sample": a.packets,
                                 "loaded_from": a.data}
        else:
            log("generating %d samples x %d packets" % (a.samples, a.packets))
            inp, tal, secs, packets = surrogate.generate_dataset(
                a.samples, a.packets, seed=a.seed, engine=eng,
                progress=lambda d, n, t: log("  labels %d/%d  %.1f s" % (d, n, t)))
            result["dataset"] = {"samples": a.samples, "packets_per_sample": a.packets,
                                 "seconds": secs, "packets_per_s": packets / secs}
            log("dataset: %.1f s (%.3g packets/s)" % (secs, packets / secs))
            if a.save_data:
                np.savez(a.save_data, labels=tal.astype(np.float32), packets=np.int64(a.packets),
                         **{k: v.astype(np.float32) for k, v in inp.items()})
        t0 = time.perf_counter()
        sur, info = surrogate.train_surrogate(inp, tal, epochs=a.epochs, batch=a.batch, lr=a.lr,
                                              seed=a.seed, lr_decay_every=max(1, a.epochs // 3))
        info["wall_s"] = time.perf_counter() - t0
        # hold-out error in units of the labels' own MC standard error
        info["label_se_R_d_median"] = float(np.median(np.sqrt(tal[:, 0] * (1 - tal[:, 0]) / a.packets)))
        result["training"] = {k: v for k, v in info.items() if k != "losses"}
        result["training"]["loss_first_last"] = [info["losses"][0], info["losses"][-1]]
        log("training: %s" % json.dumps(result["training"]))
        if a.out:
            sur.train_info = result["training"]
            sur.save(a.out)
    if a.inverse:
        result["inverse"] = inverse.run(sur, count=a.inverse, plain_count=a.plain,
                                        refine_packets=a.refine_packets, engine=eng, log=log,
                                        methods=tuple(a.methods.split(",")),
                                        with_ablation=not a.no_ablation, noise=a.noise,
                                        meas_packets=a.meas_packets, check_packets=a.check_packets,
                                        plain_packets=a.plain_packets, plain_tol=a.plain_tol)
    out = json.dumps(result)
    print(out, flush=True)
    if a.json:
        with open(a.json, "w") as f:
            f.write(out)


if __name__ == "__main__":
    main()
