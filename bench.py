#!/usr/bin/env python3
"""Benchmark of the DSGD++ epoch on one MI355X.

Workload (default): 480,189 x 17,770 matrix, 100M training + 1.4M test
nonzeros, Zipf(1.0) row/column degrees (capped), integer values 1..5 from a
rank-20 ground truth, rank 50, FP32 factors, Nzsl_L2(0.05) / None /
Nzsl_L2(0.05), BoldDriver(0.01), WOR points, COWOR strata, d = 256: 256 row
blocks x 512 column blocks, one task per CU.  One step = one DSGD++ epoch
over all training points (one persistent launch; the next epoch's schedule,
visit orders and groups are prepared beside it).  Inputs are resident in HBM
before timing starts.

Prints ONE JSON line: metric/value (nnz updates/s), roofline of the epoch
kernel (algorithmic bytes per update = (8 + s_x) + 4 r s_f, HIP-event time on
its stream), the metric's second half (time_to_target: wall time to the plain
FP64 GPU run's final test RMSE + 0.5 %, FP32 and BF16 with the learned step
controller against it, every epoch evaluating training and test loss as the
reference runner does), and the CPU baseline (the oracle, an FP64 C++
restatement of the reference's mfdsgdpp run, with d = --tasks-per-rank = the
host cores, whole epochs of the same data).
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
# The FP32 grouped kernel's cost per link of a column's update chain: clocks per
# four-slot group on the hottest task (DSGDPP_KPROF clock log of the 100M-nonzero
# workload), at the 2.4 GHz shader clock.
CHAIN_CLOCKS_PER_LINK = 300
CLOCK_GHZ = 2.4


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", default="zipf100m", choices=["zipf100m", "quickstart1m", "batch64"])
    ap.add_argument("--shrink", type=int, default=1)
    ap.add_argument("--precision", default=None)
    ap.add_argument("--kernel", type=int, default=None,
                    help="epoch kernel (default 0, automatic; batch64: 4, the grouped kernel that streams H)")
    ap.add_argument("--tasks", type=int, default=None)
    ap.add_argument("--no-graph", action="store_true", help="batch64: plain launches instead of graph replay")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-time-to-target", action="store_true",
                    help="skip the time-to-target runs of the default line")
    ap.add_argument("--cpu-threads", type=int, default=None)
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="only run a plain 1 GiB device copy (HBM counter calibration)")
    ap.add_argument("--report", choices=["precision", "time-to-rmse", "controller"], default=None,
                    help="instead of the throughput line: FP32/BF16 vs FP64 on identical sequences, or wall "
                         "time to the plain FP64 run's final test RMSE")
    ap.add_argument("--epochs", type=int, default=None)
    args = ap.parse_args()
    if args.calibrate_copy:
        calibrate_copy()
        return
    if args.report == "precision":
        report_precision(args)
        return
    if args.report == "time-to-rmse":
        report_time_to_rmse(args)
        return
    if args.report == "controller":
        report_controller(args)
        return
    if args.workload == "batch64":
        bench_batch(args)
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
        update, decay0, tasks = "Nzsl_L2(0.05)", 0.01, args.tasks or 256
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
                 stratum_order="COWOR", seed=12345, precision=precision, kernel=args.kernel or 0)
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
    g.close()

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

    # the hottest column's updates run one after another: the chain's length
    # times the cost per link bounds the FP32 grouped kernel from below
    chain = None
    if stats[-1]["kernel"] == 3 and precision == "FP32":
        import numpy as np
        hot = int(np.bincount(np.asarray(p["train"][1], dtype=np.int64)).max())
        chain_ms = hot * CHAIN_CLOCKS_PER_LINK / (CLOCK_GHZ * 1e6)
        chain = {"hot_column_updates": hot, "clocks_per_link": CHAIN_CLOCKS_PER_LINK, "clock_ghz": CLOCK_GHZ,
                 "ms": chain_ms, "frac_ceiling": per_launch_bytes / (chain_ms * 1e-3) / 1e9 / HBM_PEAK_GBS}

    ttt = None
    if name == "zipf100m" and not args.no_time_to_target:
        ttt = time_to_target(p, update, decay0, tasks, epochs=args.epochs or 30)

    cpu = None
    if not args.no_cpu_baseline:
        cpu = cpu_baseline(p, update, args.cpu_threads)

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
            "kernel": args.kernel or 0,
        },
        "roofline": {
            "bound": "hbm",
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": measured_traffic(name, precision, r, stats[-1]["kernel"], tasks),
            "kernel": KERNEL_NAMES.get(stats[-1]["kernel"], "?"),
            "bytes_per_update": b_upd,
            "avg_launch_us": avg_launch_s * 1e6,
        },
        "chain_bound": chain,
        "epoch_breakdown_ms": {
            "perm": sum(s["perm_seconds"] for s in stats) / args.steps * 1e3,
            "strata": sgd_s / args.steps * 1e3,
            "host_and_launch": ms_per_step - sum(s["device_seconds"] for s in stats) / args.steps * 1e3,
        },
        "train_loss": cur,
        "test_rmse": rmse,
        "time_to_target": ttt,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out))


KERNEL_NAMES = {1: "k_stratum_seq", 2: "k_epoch_dataflow_seq", 3: "k_epoch_grouped",
                4: "k_epoch_wide (FP64: k_epoch_wide64)"}


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


def measured_traffic(name, precision, rank, kernel, tasks):
    """HBM bytes per epoch launch from the committed rocprofv3 --pmc summary
    for this workload and blocking (FETCH_SIZE + WRITE_SIZE, scaled by the
    copy calibration), or None when no summary matches."""
    path = os.path.join(ROOT, "profiles", f"traffic_{KERNEL_NAMES.get(kernel, '?')}_{name}_{precision}_r{rank}.txt")
    if not os.path.exists(path):
        return None
    lines = open(path).read().splitlines()
    if not lines or f"d = {tasks})" not in lines[0]:
        return None
    for line in lines:
        if line.startswith("bytes_per_launch:"):
            return float(line.split(":")[1])
    return None


def _cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown CPU"


def _cpu_cores(threads):
    return threads or min(16, len(os.sched_getaffinity(0)))


def cpu_baseline(p, update, threads, min_seconds=10.0, max_epochs=40):
    """The oracle (FP64 C++ restatement of mfdsgdpp) on the host cores, run as
    the reference would be on them: d = --tasks-per-rank = the core count
    (d x 2d blocks), one worker thread per task.  Whole epochs of the same
    workload, as many as fit in about `min_seconds` (at least one, at most
    `max_epochs`)."""
    try:
        from oracle import pyoracle as po
    except Exception as e:   # noqa: BLE001
        return {"error": f"oracle unavailable: {e}"}
    cores = _cpu_cores(threads)
    tasks = cores
    t0 = time.time()
    o = po.Oracle(update=update, loss=update, rank=p["rank"], tasks_per_rank=tasks, seed=777,
                  threads=cores)
    o.load_coo(0, p["m"], p["n"], *p["train"])
    o.set_factors(p["W0"], p["H0"])
    log(f"[bench] cpu baseline: oracle loaded in {time.time() - t0:.1f}s, {cores} threads")
    nnz = len(p["train"][2])
    secs, epochs = 0.0, 0
    while epochs < max_epochs and (epochs == 0 or secs < min_seconds):
        secs += o.epoch(0.01)
        epochs += 1
    log(f"[bench] cpu baseline: {epochs} epochs in {secs:.2f}s")
    o.close()
    return {"value": epochs * nnz / secs, "unit": "nnz updates/s", "cores": cores, "kind": "port",
            "sample": f"{epochs} full FP64 epochs of the same workload ({nnz} updates each, eps 0.01), "
                      f"d = --tasks-per-rank = {tasks} ({tasks}x{2 * tasks} blocks), {cores} threads of "
                      f"nproc {os.cpu_count()} ({_cpu_model()}), {secs:.2f}s of epoch time"}


def cpu_time_to_rmse(p, update, decay0, target, threads, max_epochs):
    """The CPU side of the time-to-RMSE comparison: the oracle (FP64, d = the
    core count) with BoldDriver, training and test loss after every epoch as
    the reference runner evaluates them, until the test RMSE reaches
    `target` or `max_epochs` epochs have run."""
    from oracle import pyoracle as po
    import dsgdpp_amd as D
    cores = _cpu_cores(threads)
    o = po.Oracle(update=update, loss=update, rank=p["rank"], tasks_per_rank=cores, seed=12345,
                  threads=cores)
    o.load_coo(0, p["m"], p["n"], *p["train"])
    o.load_coo(1, p["m"], p["n"], *p["test"])
    o.set_factors(p["W0"], p["H0"])
    nt = len(p["test"][2])
    rule = D.BoldDriver(decay0)
    t0 = time.perf_counter()
    prev, cur = None, o.loss(0)
    hit, rmse, e = None, None, 0
    for e in range(1, max_epochs + 1):
        o.epoch(rule(prev, cur))
        prev, cur = cur, o.loss(0)
        rmse = math.sqrt(o.loss(1) / nt)
        log(f"[time-to-rmse] cpu epoch {e}: test rmse {rmse:.5f} at {time.perf_counter() - t0:.1f}s")
        if rmse <= target:
            hit = time.perf_counter() - t0
            break
    o.close()
    return {"epochs": e, "seconds": hit, "final_test_rmse": rmse, "reached": hit is not None,
            "cores": cores, "tasks_per_rank": cores, "cpu": _cpu_model()}


def _ctx(D, p, update, tasks, precision, seed, kernel=0):
    g = D.DsgdPp(update=update, loss=update, rank=p["rank"], tasks_per_rank=tasks, sgd_order="WOR",
                 stratum_order="COWOR", seed=seed, precision=precision, kernel=kernel)
    g.load_coo(D.TRAIN, p["m"], p["n"], *p["train"])
    if len(p["test"][2]):
        g.load_coo(D.TEST, p["m"], p["n"], *p["test"])
    g.set_factors(p["W0"], p["H0"])
    return g


def report_precision(args):
    """FP32 and BF16 factors against FP64 on identical stratum and point
    sequences (same seed) and identical step sizes (the FP64 run's BoldDriver
    sequence): per-epoch training loss and final test RMSE, relative error."""
    import torch
    import dsgdpp_amd as D
    from dsgdpp_amd import data
    epochs = args.epochs or 20
    p = data.quickstart_1m_r128(device="cuda", shrink=args.shrink)
    torch.cuda.synchronize()
    update, decay0, tasks = "Nzsl_L2(1)", 0.001, args.tasks or 4
    runs = {}
    eps_seq = None
    for prec in ("FP64_STRICT", "FP32", "BF16"):
        t0 = time.time()
        g = _ctx(D, p, update, tasks, prec, seed=12345)
        losses = [g.loss(D.TRAIN)]
        if eps_seq is None:
            bold, prev, eps_seq = D.BoldDriver(decay0), None, []
            for _ in range(epochs):
                eps_seq.append(bold(prev, losses[-1]))
                g.epoch(eps_seq[-1])
                prev = losses[-1]
                losses.append(g.loss(D.TRAIN))
        else:
            for e in eps_seq:
                g.epoch(e)
                losses.append(g.loss(D.TRAIN))
        rmse = math.sqrt(g.loss(D.TEST) / len(p["test"][2]))
        runs[prec] = dict(loss=losses, test_rmse=rmse)
        g.close()
        log(f"[precision] {prec}: final loss {losses[-1]:.8g}, test rmse {rmse:.6f} ({time.time() - t0:.1f}s)")
    ref = runs["FP64_STRICT"]
    out = {"report": "precision", "workload": f"quickstart_1m_r128: {p['m']}x{p['n']}, {len(p['train'][2])} train / "
                                              f"{len(p['test'][2])} test nonzeros, rank 128, {update}, "
                                              f"BoldDriver({decay0}) step sequence of the FP64 run, {tasks}x{2 * tasks} "
                                              f"blocks, {epochs} epochs, same seed",
           "fp64": {"final_train_loss": ref["loss"][-1], "test_rmse": ref["test_rmse"]}}
    for prec in ("FP32", "BF16"):
        r = runs[prec]
        rel = [abs(a - b) / abs(b) for a, b in zip(r["loss"], ref["loss"])]
        out[prec.lower()] = {"final_train_loss": r["loss"][-1], "test_rmse": r["test_rmse"],
                             "max_rel_err_train_loss": max(rel), "rel_err_train_loss_per_epoch": rel,
                             "rel_err_test_rmse": abs(r["test_rmse"] - ref["test_rmse"]) / ref["test_rmse"]}
    print(json.dumps(out))


def _run_to_target(D, p, update, tasks, prec, rule, target, max_epochs, seed=12345):
    """One arm: a fresh context (load not timed), then epochs with the training
    and test loss after each (sgd_impl.h:115-143) until the test RMSE reaches
    `target` (None: all `max_epochs`).  Returns [(seconds since start, rmse)]."""
    import torch
    g = _ctx(D, p, update, tasks, prec, seed=seed)
    nt = len(p["test"][2])
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    prev, cur = None, g.loss(D.TRAIN)
    hist = []
    for _ in range(max_epochs):
        g.epoch(rule(prev, cur))
        prev, cur = cur, g.loss(D.TRAIN)
        hist.append((time.perf_counter() - t0, math.sqrt(g.loss(D.TEST) / nt)))
        if target is not None and hist[-1][1] <= target:
            break
    g.close()
    if getattr(rule, "mlp", None) is not None:
        rule.mlp.close()
    return hist


def time_to_target(p, update, decay0, tasks, epochs=30, arms=(("FP32", "learned"), ("BF16", "learned"),
                                                              ("FP32", "bolddriver"))):
    """The metric's second half: plain DSGD++ in FP64 (the FP64 grouped
    kernel, per-epoch loss within 1e-9 of the reference arithmetic) with
    BoldDriver for `epochs` epochs sets the target, its final test RMSE +
    0.5 %; each arm (precision, step rule) runs until it reaches the target.
    Reported: seconds and epochs to the target, the RMSE reached, and the
    plain run's time to the target over the arm's."""
    import dsgdpp_amd as D
    from dsgdpp_amd import controller
    rules = {"bolddriver": D.BoldDriver, "learned": controller.LearnedDecay}
    t_all = time.time()
    h64 = _run_to_target(D, p, update, tasks, "FP64", D.BoldDriver(decay0), None, epochs)
    target = h64[-1][1] * 1.005
    t64, e64 = next((t, k + 1) for k, (t, r) in enumerate(h64) if r <= target)
    log(f"[time-to-target] FP64 plain: {len(h64)} epochs in {h64[-1][0]:.2f}s, final test rmse {h64[-1][1]:.5f}; "
        f"target {target:.5f} reached at epoch {e64}, {t64:.2f}s")
    out = {"target_test_rmse": target,
           "rule": "target = plain FP64 run's final test RMSE x 1.005; training and test loss every epoch; "
                   "context load not timed",
           "fp64_plain": {"precision": "FP64", "step_rule": f"BoldDriver({decay0})", "epochs": len(h64),
                          "seconds_total": h64[-1][0], "final_test_rmse": h64[-1][1],
                          "epochs_to_target": e64, "seconds_to_target": t64}}
    for prec, rname in arms:
        h = _run_to_target(D, p, update, tasks, prec, rules[rname](decay0), target, epochs)
        hit = [(t, k + 1) for k, (t, r) in enumerate(h) if r <= target]
        key = f"{prec.lower()}_{rname}"
        out[key] = {"precision": prec, "step_rule": rname, "reached": bool(hit),
                    "epochs_to_target": hit[0][1] if hit else None, "seconds_to_target": hit[0][0] if hit else None,
                    "test_rmse_reached": h[-1][1], "speedup_vs_fp64_plain": (t64 / hit[0][0]) if hit else None}
        log(f"[time-to-target] {key}: {out[key]}")
    chosen = out.get("fp32_learned") or {}
    out["chosen"] = "fp32_learned"
    out["speedup"] = chosen.get("speedup_vs_fp64_plain")
    out["bench_seconds"] = time.time() - t_all
    return out


def report_time_to_rmse(args):
    """The time_to_target object of the default line on its own (d = 128 by
    default, as earlier records), plus the CPU oracle's time to the same
    target."""
    import torch
    from dsgdpp_amd import data
    epochs = args.epochs or 30
    p = data.zipf_100m(device="cuda", shrink=args.shrink)
    torch.cuda.synchronize()
    update, decay0, tasks = "Nzsl_L2(0.05)", 0.01, args.tasks or 128
    res = time_to_target(p, update, decay0, tasks, epochs,
                         arms=(("FP32", "bolddriver"), ("FP32", "learned"), ("BF16", "bolddriver"),
                               ("BF16", "learned")))
    if not args.no_cpu_baseline:
        res["cpu_fp64_oracle"] = cpu_time_to_rmse(p, update, decay0, res["target_test_rmse"], args.cpu_threads,
                                                  epochs)
        log(f"[time-to-rmse] cpu_fp64_oracle: {res['cpu_fp64_oracle']}")
    print(json.dumps({"report": "time-to-rmse", "workload": f"zipf100m: {p['m']}x{p['n']}, "
                      f"{len(p['train'][2])} train / {len(p['test'][2])} test nonzeros, rank 50, {update}, start "
                      f"eps {decay0}, {tasks}x{2 * tasks} blocks", **res}))


def report_controller(args):
    """The step-size controller at large scale (dsgdpp_amd.data.controller_250m:
    1M x 625k, 250M training nonzeros, Zipf(1.1), rank 100, Nzsl_L2(0.05)):
    the plain FP64 run with BoldDriver(0.001) for `epochs` epochs sets the
    target (its final test RMSE + 0.5 %); FP32 with the learned controller and
    FP32 with BoldDriver, started from the same step, run until they reach it.
    Every epoch evaluates training and test loss.  d = 256 by default: the
    column blocks are too wide for an LDS tile (H rows stream from HBM,
    kernel 4) and the largest blocks keep their prep state in global scratch."""
    import torch
    import dsgdpp_amd as D
    from dsgdpp_amd import controller, data
    epochs = args.epochs or 20
    t0 = time.time()
    p = data.controller_250m(device="cuda", shrink=args.shrink)
    torch.cuda.synchronize()
    log(f"[controller] generated {len(p['train'][2])} nonzeros in {time.time() - t0:.1f}s")
    update, decay0, tasks = "Nzsl_L2(0.05)", 0.001, args.tasks or 256
    h64 = _run_to_target(D, p, update, tasks, "FP64", D.BoldDriver(decay0), None, epochs)
    target = h64[-1][1] * 1.005
    t64, e64 = next((t, k + 1) for k, (t, r) in enumerate(h64) if r <= target)
    best64 = min(range(len(h64)), key=lambda q: h64[q][1])
    log(f"[controller] FP64 plain: final rmse {h64[-1][1]:.5f} after {h64[-1][0]:.1f}s; target {target:.5f} at "
        f"epoch {e64} ({t64:.2f}s); best {h64[best64][1]:.5f} at epoch {best64 + 1}")
    res = {}
    for name, prec, rule in (("fp32_learned", "FP32", controller.LearnedDecay(decay0)),
                             ("fp32_bolddriver", "FP32", D.BoldDriver(decay0))):
        h = _run_to_target(D, p, update, tasks, prec, rule, target, epochs)
        hit = [(t, k + 1) for k, (t, r) in enumerate(h) if r <= target]
        res[name] = {"reached": bool(hit), "epochs_to_target": hit[0][1] if hit else None,
                     "seconds_to_target": hit[0][0] if hit else None, "test_rmse_reached": h[-1][1],
                     "speedup_vs_fp64_plain": (t64 / hit[0][0]) if hit else None}
        log(f"[controller] {name}: {res[name]}")
    print(json.dumps({"report": "controller", "workload": f"controller_250m: {p['m']}x{p['n']}, "
                      f"{len(p['train'][2])} train / {len(p['test'][2])} test nonzeros, rank 100, {update}, "
                      f"start eps {decay0}, {tasks}x{2 * tasks} blocks, {epochs} epochs for the plain run",
                      "target_test_rmse": target,
                      "fp64_plain": {"epochs": len(h64), "seconds_total": h64[-1][0], "final_test_rmse": h64[-1][1],
                                     "epochs_to_target": e64, "seconds_to_target": t64,
                                     "best_test_rmse": h64[best64][1], "best_epoch": best64 + 1},
                      **res}))


def bench_batch(args):
    """64 independent factorizations (shapes, ranks, lambdas drawn as in
    dsgdpp_amd.batch.batch_problems), FP32, advanced together by batched
    epochs (one persistent launch per epoch, graph replay), task counts from
    the autotuner.  value = all problems' nnz updates per second."""
    import torch
    import dsgdpp_amd as D
    from dsgdpp_amd import batch as B
    t0 = time.time()
    probs = B.batch_problems(64, seed=3, shrink=args.shrink, log=log)
    nnz = [len(p["train"][2]) for p in probs]
    log(f"[bench] generated 64 problems, {sum(nnz)} nonzeros in {time.time() - t0:.1f}s")
    rates = B.measure_rates(log=log)
    meta = [dict(m=p["m"], n=p["n"], nnz=len(p["train"][2]), rank=p["rank"]) for p in probs]
    budget = torch.cuda.get_device_properties(0).multi_processor_count   # one resident task per CU
    d = B.plan_tasks(meta, budget, rates)
    # the planner's kernel: on batch64 the streaming kernel 4, whose launch
    # leaves most of each CU's LDS to the look-ahead preps beside it
    kernel = B.batch_kernel(meta, d) if args.kernel is None else args.kernel
    log(f"[bench] batch kernel {kernel}")
    log(f"[bench] tasks per problem: {d} (sum {sum(d)}), one launch per epoch")
    t0 = time.time()
    ctxs = []
    for k, p in enumerate(probs):
        upd = f"Nzsl_L2({p['lambda']:.6g})"
        while True:   # the library has the last word on the grouped kernel's limits
            g = _ctx(D, dict(p, test=((), (), ())), upd, d[k], "FP32", seed=100 + k, kernel=kernel)
            ok, why = g.grouped_ok()
            if ok or 2 * d[k] > p["m"] or 4 * d[k] > p["n"]:
                break
            log(f"[bench] problem {k}: d={d[k]} ({why}); doubling")
            g.close()
            d[k] *= 2
        ctxs.append(g)
        if (k + 1) % 8 == 0:
            log(f"[bench] {k + 1} contexts loaded ({time.time() - t0:.1f}s)")
    # one batch: tasks beyond one per CU queue on the CUs' workgroups (the
    # library packs them, longest work first)
    parts = [list(range(len(ctxs)))]
    batches = [B.DsgdPpBatch(ctxs, use_graph=not args.no_graph)]
    log(f"[bench] contexts loaded + batch created in {time.time() - t0:.1f}s")

    class _All:
        def epoch(self, eps):
            out = {"seconds": 0.0, "sgd_seconds": 0.0}
            for part, bt in zip(parts, batches):
                st = bt.epoch([eps[k] for k in part])
                out["seconds"] += st["seconds"]
                out["sgd_seconds"] += st["sgd_seconds"]
            return out

        def loss(self):   # every problem's training loss, one wait per batch
            out = [0.0] * len(ctxs)
            for part, bt in zip(parts, batches):
                for k, v in zip(part, bt.loss(D.TRAIN)):
                    out[k] = float(v)
            return out

        def close(self):
            for bt in batches:
                bt.close()

    bt = _All()
    rules = [D.BoldDriver(0.01) for _ in ctxs]
    prev = [None] * len(ctxs)
    cur = bt.loss()

    def step():
        eps = [r(a, b) for r, a, b in zip(rules, prev, cur)]
        st = bt.epoch(eps)
        prev[:], cur[:] = cur, bt.loss()
        return st

    for _ in range(args.warmup):
        step()
    wall, kern = 0.0, 0.0
    for k in range(args.steps):
        torch.cuda.synchronize()
        a = time.perf_counter()
        st = bt.epoch([r(x, y) for r, x, y in zip(rules, prev, cur)])
        torch.cuda.synchronize()
        wall += time.perf_counter() - a
        kern += st["sgd_seconds"]
        prev[:], cur[:] = cur, bt.loss()
        log(f"[bench] batch step {k + 1}/{args.steps}: {st['seconds'] * 1e3:.1f} ms (launch {st['sgd_seconds'] * 1e3:.1f} ms)")
    total = sum(nnz)
    print(json.dumps({
        "metric": "nnz updates/s (64 batched DSGD++ solves, one launch per epoch)",
        "value": total * args.steps / wall, "unit": "nnz updates/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": wall / args.steps * 1e3, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "fp32", "data": "synthetic (uniform cells, seed 3)",
        "config": {"workload": f"batch64: 64 problems, {total} nonzeros in all, ranks "
                               f"{sorted(set(p['rank'] for p in probs))}, shrink {args.shrink}",
                   "tasks": d, "kernel": kernel, "launches_per_epoch": len(parts),
                   "launch_ms": kern / args.steps * 1e3}}))
    bt.close()
    for g in ctxs:
        g.close()


if __name__ == "__main__":
    main()
