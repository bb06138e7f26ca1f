"""Benchmark of the neighbour-graph step on one MI355X.

A step is one full NN-Descent run (start, iterations, sorted CSR) on a
synthetic embedding already resident in HBM. The graph is left in HBM.
Default workload: 4.06M cells x 50 PCs FP32, 150-component Gaussian mixture,
k = 15, seed 0 (polariseq_amd.datasets.mixture). After the timed steps:
  * recall@15 of the graph against the exact top-15 over all cells, for 100k
    random query cells (exact kNN on the GPU, FP64-ranked);
  * the join kernel's roofline from HIP events recorded on its stream;
  * the CPU baseline: the oracle's restatement of the reference CPU NN-Descent on
    a bounded sample of the same workload, on this host's cores.
Prints one JSON line.
"""

from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "cells/s to kNN graph at recall@15>=0.95, 4M x 50 PCs; learned+BF16 vs plain"
HBM_PEAK_GBS = 8000.0


def log(msg: str) -> None:
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def host_threads() -> int:
    env = os.environ.get("OMP_NUM_THREADS")
    if env and env.isdigit() and int(env) > 0:
        return int(env)
    try:
        return len(os.sched_getaffinity(0))
    except Exception:
        return os.cpu_count() or 1


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--n", type=int, default=4_060_000)
    ap.add_argument("--components", type=int, default=150)
    ap.add_argument("--k", type=int, default=15)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--init", default="random", choices=["random", "rp_forest"])
    ap.add_argument("--recall-queries", type=int, default=100_000)
    ap.add_argument("--cpu-cells", type=int, default=1_000_000)
    ap.add_argument("--no-cpu", action="store_true")
    args = ap.parse_args()

    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU: this build runs the neighbour-graph step on a single "
              "MI355X (no multi-GPU path); run with --gpus 1.", file=sys.stderr)
        return 2

    import torch

    import polariseq_amd as P
    from polariseq_amd import datasets as ds

    assert torch.cuda.is_available(), "bench.py needs a GPU"
    torch.cuda.set_device(0)
    n, k = args.n, args.k
    log(f"generating {n} x 50 mixture ({args.components} components, seed {args.seed})")
    x = ds.mixture(n, 50, args.components, seed=args.seed)
    xt = torch.from_numpy(x).cuda()
    torch.cuda.synchronize()

    def step():
        return P.nndescent_device(xt, k, seed=args.seed, init=args.init)

    for w in range(args.warmup):
        t = time.perf_counter()
        step()
        torch.cuda.synchronize()
        log(f"warmup {w + 1}/{args.warmup}: {time.perf_counter() - t:.3f} s")
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    out = None
    stats = []
    for s in range(args.steps):
        out = step()
        stats.append(P.last_stats())
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    ms_per_step = 1000.0 * elapsed / args.steps
    value = n * args.steps / elapsed
    log(f"timed {args.steps} steps: {ms_per_step:.1f} ms/step, {value:,.0f} cells/s")
    st = stats[-1]
    log(f"last step: iterations {st['iterations']}, ms init {st['ms_init']:.1f} "
        f"iter {st['ms_iterations']:.1f} extract {st['ms_extract']:.1f}, changed {st['changed']}")

    # roofline of the dominant kernel (the local join), over every timed launch:
    # algorithmic bytes = candidate rows gathered x d x 4 B + the two candidate lists
    # (2 m x 4 B per cell) + one threshold per candidate (4 B).
    m = st["max_candidates"]
    join_bytes = 0.0
    join_ms = 0.0
    launches = 0
    for s_ in stats:
        for it in range(s_["iterations"]):
            join_bytes += s_["cand_rows"][it] * (50 * 4 + 4) + n * (2 * m + 1) * 4
            join_ms += s_["ms_join"][it]
            launches += 1
    achieved = join_bytes / (join_ms / 1000.0) / 1e9 if join_ms > 0 else 0.0
    roofline = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None,
                "kernel": "k_join<50,1>", "launches": launches,
                "avg_launch_ms": round(join_ms / max(launches, 1), 3),
                "algorithmic_bytes_per_launch": round(join_bytes / max(launches, 1))}

    # recall@15 against exact kNN over all cells
    ip, ix, dx, nnz = out
    nq = min(args.recall_queries, n)
    q = np.sort(np.random.default_rng(12345).choice(n, nq, replace=False)).astype(np.int32)
    t = time.perf_counter()
    ei, _ = P.knn_exact_device(xt, k, torch.from_numpy(q).cuda())
    torch.cuda.synchronize()
    log(f"exact kNN for {nq} queries: {time.perf_counter() - t:.2f} s")
    g = ix[: n * k].view(n, k)[torch.from_numpy(q).cuda().long()].cpu().numpy()
    rec = ds.recall(g, ei.cpu().numpy())
    recall = float(rec.mean())
    log(f"recall@{k} = {recall:.4f} (p05 {np.percentile(rec, 5):.3f})")

    cpu = None
    if not args.no_cpu:
        import oracle
        nc = min(args.cpu_cells, n)
        threads = host_threads()
        log(f"CPU baseline: reference NN-Descent restated, {nc} cells, {threads} threads")
        xs = np.ascontiguousarray(x[:nc])
        t = time.perf_counter()
        _, pst = oracle.nndescent_port(xs, k, seed=args.seed, threads=threads)
        ct = time.perf_counter() - t
        cpu = {"value": round(nc / ct, 1), "unit": "cells/s", "cores": threads, "kind": "port",
               "sample": f"first {nc} cells of the same {n} x 50 workload, k={k}, seed {args.seed},"
                         f" auto options ({pst.n_trees} trees, {pst.iterations} iterations),"
                         f" {ct:.2f} s"}
        log(f"CPU baseline: {nc / ct:,.0f} cells/s ({ct:.2f} s)")

    line = {
        "metric": METRIC, "value": round(value, 1), "unit": "cells/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(ms_per_step, 3),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic Gaussian-mixture embedding (numpy, seeded); no dataset",
        "config": {"workload": f"plain GPU NN-Descent, {n} x 50 FP32 cells, "
                               f"{args.components}-component mixture, k={k}, init={args.init}",
                   "n_cells": n, "dims": 50, "k": k, "form": "plain", "init": args.init,
                   "seed": args.seed, "components": args.components,
                   "iterations": st["iterations"], "ms_init": round(st["ms_init"], 2),
                   "ms_iterations": round(st["ms_iterations"], 2),
                   "ms_extract": round(st["ms_extract"], 2)},
        "recall_at_15": round(recall, 5),
        "roofline": roofline,
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
