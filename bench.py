#!/usr/bin/env python3
"""Benchmark of the DSGD++ epoch on one MI355X.

Workload (default): 480,189 x 17,770 matrix, 100M training + 1.4M test
nonzeros, Zipf(1.0) row/column degrees (capped), integer values 1..5 from a
rank-20 ground truth, rank 50, FP32 factors, Nzsl_L2(0.05) / None /
Nzsl_L2(0.05), BoldDriver(0.01), WOR points, COWOR strata, 128 x 256 blocks.
One step = one DSGD++ epoch over all training points (schedule + visit orders
+ 256 stratum launches).  Inputs are resident in HBM before timing starts.

Prints ONE JSON line: metric/value (nnz updates/s), roofline of the stratum
kernel (algorithmic bytes per update = (8 + s_x) + 4 r s_f, HIP-event time),
and the CPU baseline (the oracle, an FP64 C++ restatement of the reference's
mfdsgdpp run with d = 128 on the host cores, one epoch on the same data).
"""
import argparse
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", default="zipf100m", choices=["zipf100m", "quickstart1m"])
    ap.add_argument("--shrink", type=int, default=1)
    ap.add_argument("--precision", default=None)
    ap.add_argument("--kernel", type=int, default=0)
    ap.add_argument("--tasks", type=int, default=None)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-threads", type=int, default=None)
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="only run a plain 1 GiB device copy (HBM counter calibration)")
    args = ap.parse_args()
    if args.calibrate_copy:
        calibrate_copy()
        return

    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; --gpus above 1 is not supported "
                                   "(multi-GPU sharding is a design note only)"}))
        sys.exit(2)

    import numpy as np
    import torch
    import dsgdpp_amd as D
    from dsgdpp_amd import data

    if not torch.cuda.is_available():
        log("no GPU visible")
        sys.exit(3)

    t0 = time.time()
    name = args.workload
    if name == "zipf100m":
        p = data.zipf_100m(device="cuda", shrink=args.shrink)
        update, decay0, tasks = "Nzsl_L2(0.05)", 0.01, args.tasks or 128
        precision = args.precision or "FP32"
    else:
        p = data.quickstart_1m(device="cuda", shrink=args.shrink)
        update, decay0, tasks = "Nzsl_L2(1)", 0.001, args.tasks or 4
        precision = args.precision or "FP64_STRICT"
    torch.cuda.synchronize()
    m, n, r = p["m"], p["n"], p["rank"]
    nnz, nnz_test = len(p["train"][2]), len(p["test"][2])
    log(f"[bench] generated {name}: {m}x{n} nnz={nnz} test={nnz_test} r={r} in {time.time() - t0:.1f}s")

    t0 = time.time()
    g = D.DsgdPp(update=update, loss=update, rank=r, tasks_per_rank=tasks, sgd_order="WOR",
                 stratum_order="COWOR", seed=12345, precision=precision, kernel=args.kernel)
    g.load_coo(D.TRAIN, m, n, *p["train"])
    g.load_coo(D.TEST, m, n, *p["test"])
    g.set_factors(p["W0"], p["H0"])
    log(f"[bench] loaded + blocked d={tasks} ({tasks}x{2 * tasks}) in {time.time() - t0:.1f}s")

    bold = D.BoldDriver(decay0)
    cur = g.loss(D.TRAIN)
    prev = None
    eps_hist = []

    def one_epoch():
        nonlocal prev, cur
        eps = bold(prev, cur)
        st = g.epoch(eps)
        prev, cur = cur, g.loss(D.TRAIN)
        eps_hist.append(eps)
        return st

    for k in range(args.warmup):
        st = one_epoch()
        log(f"[bench] warmup {k + 1}/{args.warmup}: epoch {st['seconds'] * 1e3:.1f} ms, loss {cur:.6g}")

    torch.cuda.synchronize()
    stats = []
    wall = 0.0
    for k in range(args.steps):
        torch.cuda.synchronize()
        a = time.perf_counter()
        eps = bold(prev, cur)
        st = g.epoch(eps)
        torch.cuda.synchronize()
        wall += time.perf_counter() - a
        stats.append(st)
        prev, cur = cur, g.loss(D.TRAIN)
        log(f"[bench] step {k + 1}/{args.steps}: epoch {st['seconds'] * 1e3:.1f} ms "
            f"(perm {st['perm_seconds'] * 1e3:.1f}, strata {st['sgd_seconds'] * 1e3:.1f}), loss {cur:.6g}")
    test = g.loss(D.TEST)
    rmse = math.sqrt(test / nnz_test)

    ms_per_step = wall / args.steps * 1e3
    value = nnz * args.steps / wall
    s_f = {"FP64_STRICT": 8, "FP64": 8, "FP32": 4, "BF16": 2}[precision]
    s_x = 8 if precision.startswith("FP64") else 4
    b_upd = (8 + s_x) + 4 * r * s_f
    sgd_s = sum(s["sgd_seconds"] for s in stats)
    launches = sum(s["launches"] for s in stats)
    per_launch_bytes = nnz * b_upd * args.steps / launches
    avg_launch_s = sgd_s / launches
    achieved = per_launch_bytes / avg_launch_s / 1e9

    cpu = None
    if not args.no_cpu_baseline:
        cpu = cpu_baseline(p, update, tasks, args.cpu_threads)

    out = {
        "metric": "nnz updates/s (DSGD++ epoch, 100M-nonzero Zipf workload)" if name == "zipf100m"
                  else "nnz updates/s (DSGD++ epoch, quick-start workload)",
        "value": value,
        "unit": "nnz updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": {"FP64_STRICT": "f64", "FP64": "f64", "FP32": "fp32", "BF16": "bf16-storage/fp32"}[precision],
        "data": f"synthetic ({p['rule']}, seed {p['seed']})",
        "config": {
            "workload": f"{name}: {m}x{n}, {nnz} train / {nnz_test} test nonzeros, rank {r}, "
                        f"{update}, BoldDriver({decay0}), WOR/COWOR, {tasks}x{2 * tasks} blocks, {precision}",
            "m": m, "n": n, "nnz": nnz, "rank": r, "blocks": [tasks, 2 * tasks], "precision": precision,
            "kernel": args.kernel,
        },
        "roofline": {
            "bound": "hbm",
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": measured_traffic(name, precision, r, stats[-1]["kernel"]),
            "kernel": KERNEL_NAMES.get(stats[-1]["kernel"], "?"),
            "bytes_per_update": b_upd,
            "avg_launch_us": avg_launch_s * 1e6,
        },
        "epoch_breakdown_ms": {
            "perm": sum(s["perm_seconds"] for s in stats) / args.steps * 1e3,
            "strata": sgd_s / args.steps * 1e3,
            "host_and_launch": ms_per_step - sum(s["device_seconds"] for s in stats) / args.steps * 1e3,
        },
        "train_loss": cur,
        "test_rmse": rmse,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out))


KERNEL_NAMES = {1: "k_stratum_seq", 2: "k_epoch_dataflow_seq", 3: "k_epoch_grouped"}


def calibrate_copy():
    """A plain device copy of known size: rocprofv3 FETCH_SIZE / WRITE_SIZE of
    this kernel give the counter-to-bytes factor (1 GiB read + 1 GiB written)."""
    import torch
    a = torch.empty(1 << 28, dtype=torch.float32, device="cuda").fill_(1.0)
    b = torch.empty_like(a)
    for _ in range(3):
        b.copy_(a)
    torch.cuda.synchronize()
    print(json.dumps({"copy_bytes": a.numel() * 4, "launches": 3}))


def measured_traffic(name, precision, rank, kernel):
    """HBM bytes per epoch launch from the committed rocprofv3 --pmc summary
    for this workload (FETCH_SIZE + WRITE_SIZE, scaled by the copy
    calibration), or None when no summary matches."""
    path = os.path.join(ROOT, "profiles", f"traffic_{KERNEL_NAMES.get(kernel, '?')}_{name}_{precision}_r{rank}.txt")
    if not os.path.exists(path):
        return None
    for line in open(path):
        if line.startswith("bytes_per_launch:"):
            return float(line.split(":")[1])
    return None


def cpu_baseline(p, update, tasks, threads):
    """The oracle (FP64 C++ restatement of mfdsgdpp) on the host cores: one
    epoch of the same workload with the same d x 2d blocking."""
    try:
        from oracle import pyoracle as po
    except Exception as e:   # noqa: BLE001
        return {"error": f"oracle unavailable: {e}"}
    cores = threads or min(16, len(os.sched_getaffinity(0)))
    t0 = time.time()
    o = po.Oracle(update=update, loss=update, rank=p["rank"], tasks_per_rank=tasks, seed=777,
                  threads=cores)
    o.load_coo(0, p["m"], p["n"], *p["train"])
    o.set_factors(p["W0"], p["H0"])
    log(f"[bench] cpu baseline: oracle loaded in {time.time() - t0:.1f}s, {cores} threads")
    secs = o.epoch(0.01)
    nnz = len(p["train"][2])
    log(f"[bench] cpu baseline: one epoch {secs:.2f}s")
    o.close()
    return {"value": nnz / secs, "unit": "nnz updates/s", "cores": cores, "kind": "port",
            "sample": f"one full FP64 epoch of the same workload ({nnz} updates, d={tasks}), "
                      f"{cores} threads, epoch time {secs:.3f}s"}


if __name__ == "__main__":
    main()
