"""Throughput of the Tsetlin engine on the 2048-clauses/class classifier workload.

Workload (one MI355X): 200k train + 20k test samples, 4096 sparse Bernoulli(0.05)
features (8192 literals, 1024-byte rows), 20 classes from planted conjunctions,
2% label noise; machine of 2048 clauses/class, target 256, specificity 10,
clause tolerance = maximum = 64, state 8 bits.  A step is one training epoch
over the 200k samples (inputs resident in HBM).  `value` is training
samples/s; predict samples/s over the 20k test rows is reported beside it.

    python bench.py [--steps K] [--warmup W] [--gpus 1] [--no-cpu]

The roofline object prices the train kernel against the integer VALU peak
(256 CU x 64 lanes x 2.4 GHz): its algorithmic operations are computed from
the kernel's own work counters (chunk-steps evaluated and feedback events).
cpu_baseline times the oracle's reference-shaped CPU path (OpenMP over
8-clause chunks, -O3 -march=native) on a bounded slice of the same workload.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
PKG = os.path.join(ROOT, "lua-santoku-tsetlin_amd")
for _p in (ROOT, PKG):
    if _p not in sys.path:
        sys.path.insert(0, _p)

METRIC = "train & predict samples/s @2048 clauses/class; search time-to-acc, surrogate vs plain"
VALU_PEAK_TOPS = 256 * 64 * 2.4e9 / 1e12      # int32 lane-ops/s
HBM_PEAK_GBS = 8000.0

MACHINE = dict(classes=20, features=4096, clauses=2048, clause_tolerance=64, clause_maximum=64,
               specificity=10.0, target=256, seed=2024)


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def algorithmic_ops(stats, wd, m=7):
    """Integer lane-ops of the train kernel from its work counters.
    eval: 8 clauses x wd words x (AND-NOT, popcount-accumulate) per chunk-step;
    a full-row feedback pass: (3(m+1)+3) ops per word; a random-literal
    decrement: 2(m+1) ops."""
    eval_ops = stats["steps"] * 16 * wd
    passes = stats["type1_true"] + stats["type1_grow"] + stats["type2_true"]
    fb_ops = passes * (3 * (m + 1) + 3) * wd + stats["draws"] * 2 * (m + 1)
    return eval_ops + fb_ops, eval_ops, fb_ops


def cpu_baseline(data, threads, budget_s=15.0):
    """Reference-shaped CPU training on a bounded slice: one epoch over the
    first `limit` samples of each thread's shuffle, all chunks."""
    import numpy as np
    from oracle import oracle as O
    try:
        so = O.build(native=True)
    except Exception as ex:  # no compiler on the host: fall back to the portable build
        log("native oracle build failed (%s); using the portable build" % ex)
        so = O.build(native=False)
    L = O.lib(so)
    p = O.params("classifier", **MACHINE)
    rows, y = data["train_rows"], data["train_y"]

    def run(limit):
        m = O.Machine(p)
        t0 = time.perf_counter()
        m.train_refshape(rows, labels=y, epochs=1, threads=threads, sample_limit=limit, L=L)
        return time.perf_counter() - t0

    run(4)
    limit = 16
    t = run(limit)
    while t < budget_s / 4 and limit < 20000:
        limit = int(limit * min(8.0, max(2.0, budget_s / 2 / max(t, 1e-3))))
        t = run(limit)
    train_sps = limit / t
    # predict: bounded test slice on the same machine size
    m = O.Machine(p)
    m.setup()
    rng = np.random.default_rng(0)
    m.actions[:] = np.packbits(rng.random(m.actions.size * 8) < 0.01, bitorder="little")
    test_rows = data["test_rows"]
    npred = min(64, test_rows.shape[0])
    while True:
        out = np.zeros(npred, np.int64)
        t0 = time.perf_counter()
        L.orc_predict_classifier_threads(O.ctypes.byref(p), O.ptr(m.actions), O.ptr(test_rows[:npred]),
                                         npred, O.ptr(out, O._I64P), threads)
        tp = time.perf_counter() - t0
        if tp >= 2.0 or npred >= test_rows.shape[0]:
            break
        npred = min(npred * 4, test_rows.shape[0])
    return dict(train_sps=train_sps, train_sample=limit, train_s=t, predict_sps=npred / tp, predict_sample=npred,
                predict_s=tp)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--n-train", type=int, default=200_000)
    ap.add_argument("--n-test", type=int, default=20_000)
    ap.add_argument("--cpu-only", action="store_true", help=argparse.SUPPRESS)
    args = ap.parse_args()
    if args.cpu_only:
        # child process of the CPU leg: no torch / HIP in this process, so the
        # oracle's OpenMP runtime is the only one loaded
        from santoku.tsetlin import data as D
        data = D.sparse_planted_task(7, n_train=args.n_train, n_test=args.n_test)
        threads = len(os.sched_getaffinity(0))
        cb = cpu_baseline(data, threads)
        cb["threads"] = threads
        print(json.dumps(cb), flush=True)
        return
    if args.gpus != 1:
        print("bench.py measures one GPU only: the engine has no multi-GPU path in this build "
              "(--gpus %d requested)" % args.gpus, file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch
    import santoku.tsetlin as tm
    from santoku.tsetlin import data as D

    t0 = time.perf_counter()
    data = D.sparse_planted_task(7, n_train=args.n_train, n_test=args.n_test)
    log("data: %d train + %d test rows of %d bytes, class histogram %s (%.1fs)" % (
        args.n_train, args.n_test, data["train_rows"].shape[1],
        np.bincount(data["train_y"], minlength=20).tolist(), time.perf_counter() - t0))
    dev = torch.device("cuda:0")
    rows = torch.from_numpy(data["train_rows"]).to(dev)
    labels = torch.from_numpy(data["train_y"]).to(dev)
    test_rows = torch.from_numpy(data["test_rows"]).to(dev)
    n, ntest = args.n_train, args.n_test

    eng = tm.classifier(dict(MACHINE, reusable=True))
    wd = (eng.info()["input_chunks"] + 3) // 4
    if args.warmup > 0:
        eng.train(samples=n, problems=rows, solutions=labels, iterations=args.warmup)
        log("warmup: %d epoch(s), kernel %.1f ms" % (args.warmup, eng.timing()[0]))
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    eng.train(samples=n, problems=rows, solutions=labels, iterations=args.steps)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    train_kernel_ms, _ = eng.timing()
    stats = eng.train_stats()
    ops, eval_ops, fb_ops = algorithmic_ops(stats, wd)
    train_sps = args.steps * n / wall
    log("train: %d epochs in %.3f s (kernel %.1f ms) -> %.0f samples/s; stats %s" % (
        args.steps, wall, train_kernel_ms, train_sps, stats))

    pred = eng.predict(test_rows, ntest)
    torch.cuda.synchronize()
    reps = 3
    t0 = time.perf_counter()
    for _ in range(reps):
        pred = eng.predict(test_rows, ntest)
    torch.cuda.synchronize()
    pwall = (time.perf_counter() - t0) / reps
    _, predict_kernel_ms = eng.timing()
    acc = float((pred.cpu().numpy() == data["test_y"]).mean())
    log("predict: %d rows in %.2f ms (kernel %.2f ms); test accuracy %.4f" % (ntest, pwall * 1e3, predict_kernel_ms, acc))

    achieved = ops / (train_kernel_ms / 1e3) / 1e12
    hbm_bytes = args.steps * (n * data["train_rows"].shape[1] + n * 8) + 2 * eng.info()["action_chunks"] * 8
    line = {
        "metric": METRIC,
        "value": round(train_sps, 1),
        "unit": "train samples/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(wall * 1e3 / args.steps, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u32",
        "data": "synthetic (seeded planted-conjunction classifier data, generated in-process)",
        "config": {"workload": "plain classifier 200k train/20k test, 4096 Bernoulli(0.05) features, 20 classes, "
                               "2048 clauses/class, target 256, specificity 10, tolerance=maximum=64",
                   "samples": n, "features": 4096, "classes": 20, "clauses": 2048},
        "predict_samples_per_s": round(ntest / pwall, 1),
        "predict_kernel_ms": round(predict_kernel_ms, 3),
        "train_kernel_ms_per_epoch": round(train_kernel_ms / args.steps, 3),
        "test_accuracy_after_steps": round(acc, 4),
        "roofline": {
            "bound": "valu",
            "achieved": round(achieved, 3),
            "peak": round(VALU_PEAK_TOPS, 2),
            "unit": "Tops/s (int32 lane-ops)",
            "frac": round(achieved / VALU_PEAK_TOPS, 4),
            "traffic": None,
            "kernel": "train_kernel<4,64,7>",
            "ops_per_launch": ops,
            "eval_ops": eval_ops,
            "feedback_ops": fb_ops,
            "hbm_algorithmic_gbs": round(hbm_bytes / (train_kernel_ms / 1e3) / 1e9, 2),
            "hbm_peak_gbs": HBM_PEAK_GBS,
        },
        "cpu_baseline": None,
    }
    if not args.no_cpu:
        import subprocess
        log("cpu baseline: child process")
        r = subprocess.run([sys.executable, os.path.abspath(__file__), "--cpu-only", "--n-train", str(n),
                            "--n-test", str(ntest)], capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError("cpu baseline failed (%d): %s" % (r.returncode, r.stderr[-2000:]))
        cb = json.loads(r.stdout.strip().splitlines()[-1])
        threads = cb["threads"]
        log("cpu baseline: %d threads, %.2f train samples/s, %.1f predict samples/s" % (
            threads, cb["train_sps"], cb["predict_sps"]))
        line["cpu_baseline"] = {
            "value": round(cb["train_sps"], 2),
            "unit": "train samples/s",
            "cores": threads,
            "kind": "port",
            "sample": "oracle reference-shaped training (OpenMP static over 5120 chunks, evaluation before the "
                      "negative-sampling test), one epoch restricted to the first %d samples of each thread's "
                      "shuffle, %.1f s" % (cb["train_sample"], cb["train_s"]),
            "predict_value": round(cb["predict_sps"], 2),
            "predict_sample": "%d test rows, %.2f s" % (cb["predict_sample"], cb["predict_s"]),
            "speedup_train": round(train_sps / cb["train_sps"], 1),
            "speedup_predict": round((ntest / pwall) / cb["predict_sps"], 1),
        }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
