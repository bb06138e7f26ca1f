"""Benchmark of the neighbour-graph step on one MI355X.

A step is one full NN-Descent run (start, iterations, sorted CSR) on a
synthetic embedding already resident in HBM. The graph is left in HBM.
Default workload: 4.06M cells x 50 PCs FP32, 150-component Gaussian mixture,
k = 15, seed 0 (polariseq_amd.datasets.mixture). Four forms are timed, each over
the same input:
  * plain: the drop-in (FP32 join, 48-tree forest, the reference's defaults);
  * plain_r95: the same algorithm with rows of --plain-heap-k (20) while
    iterating, the smallest row size at which it clears recall@15 0.95 here;
    the speed-up of the headline is quoted against it;
  * plain_op: the plain FP32 join at the learned form's operating point (5-tree
    forest on the raw rows, rows of 26, stop at 0.07 n h, 64 slots): what the
    operating point alone gives, without BF16 or the encoder;
  * bf16_raw: the BF16 join on an RP forest of the raw rows, at the learned
    form's operating point (what the learned seeding replaces);
  * learned_bf16, the headline: in every step the encoder is trained from its
    initial weights on this embedding (exact-kNN triplet mining on a subset,
    Adam), encodes every cell, and its codes seed the BF16 form at the operating
    point --learned-trees / --learned-heap-k / --learned-delta / --learned-slots
    (5 trees on the codes, rows of 26 while iterating, early stop at 0.07 n h
    changes, 64 proposal slots per row); the encoder trains for --train-steps Adam
    steps of --train-batch triplets, hinge margin --train-margin.
After the timed steps:
  * recall@15 of every form against the exact top-15 over all cells, for 100k
    random query cells (exact kNN on the GPU, FP64-ranked);
  * the join kernel's roofline from HIP events recorded on its stream;
  * the CPU baseline: the oracle's restatement of the reference CPU NN-Descent on
    the full workload (and on the 20k-cell reference case), on this host's cores,
    with its recall on the same queries.
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
FORMS = ["plain", "plain_r95", "plain_op", "bf16_raw", "learned_bf16"]
FORM_DESC = {
    "plain": "plain GPU NN-Descent, the drop-in (FP32 join, 48-tree RP forest, the reference's "
             "defaults)",
    "plain_r95": "plain GPU NN-Descent with wider rows while iterating (FP32 join, 48-tree RP "
                 "forest, defaults otherwise)",
    "plain_op": "plain GPU NN-Descent (FP32 join) with a 5-tree RP forest on the raw rows, at the "
                "learned form's operating point",
    "bf16_raw": "BF16 MFMA join + FP32 re-rank, RP forest on the raw rows, at the learned "
                "form's operating point",
    "learned_bf16": "learned seeding (encoder trained inside the step) + BF16 MFMA join with "
                    "FP32 re-rank",
}
HBM_PEAK_GBS = 8000.0


def log(msg: str) -> None:
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def measured_traffic(kernel: str, summary: str = "join_4m_counters.md"):
    """HBM bytes per join pass of `kernel` on its workload, from the rocprofv3
    request-size counter passes summarised in profiles/<summary>."""
    path = os.path.join(ROOT, "profiles", summary)
    try:
        with open(path) as f:
            for line in f:
                parts = line.split()
                if len(parts) == 4 and parts[0] == "traffic_bytes_per_pass" and parts[1] == kernel:
                    return int(parts[3])
    except OSError:
        return None
    return None


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
    ap.add_argument("--form", default="all", choices=FORMS + ["all"],
                    help="; ".join(f"{f_} = {d_}" for f_, d_ in FORM_DESC.items()) +
                         "; all = every form, learned_bf16 (the headline) last")
    ap.add_argument("--init", default="rp_forest", choices=["random", "rp_forest"],
                    help="start of the plain and bf16 forms")
    ap.add_argument("--learned-trees", type=int, default=5)
    ap.add_argument("--learned-heap-k", type=int, default=26,
                    help="entries per row while iterating, learned form (the API default is 2k)")
    ap.add_argument("--learned-delta", type=float, default=0.07,
                    help="early-stop fraction of the learned form (the API default is 0.001)")
    ap.add_argument("--learned-slots", type=int, default=64,
                    help="proposal slots per row, learned form (0 = the API default, 128 here)")
    ap.add_argument("--plain-heap-k", type=int, default=20,
                    help="rows while iterating of the plain_r95 form (output k)")
    ap.add_argument("--train-steps", type=int, default=100)
    ap.add_argument("--train-subset", type=int, default=200_000)
    ap.add_argument("--train-anchors", type=int, default=4096)
    ap.add_argument("--train-batch", type=int, default=8192)
    ap.add_argument("--train-lr", type=float, default=3e-3)
    ap.add_argument("--train-margin", type=float, default=0.01)
    ap.add_argument("--recall-queries", type=int, default=100_000)
    ap.add_argument("--cpu-cells", type=int, default=4_060_000,
                    help="cells of the headline workload the CPU baseline runs on (all by default)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--workload", default="headline",
                    choices=["headline", "precision", "batched", "fuzzy", "scrublet", "pca"],
                    help="headline = the metric above; precision = BF16 form on 4.06M x 50 "
                         "with 5%% outliers, k=30, distance error on a 50k subset; batched = "
                         "per-sample exact kNN of 64 samples with doublets in one launch; "
                         "fuzzy = fuzzy connectivities of the headline kNN graph; scrublet = "
                         "exact kNN at Scrublet's large k over cells + simulated doublets; "
                         "pca = exact Gram PCA of 4.06M cells x 2000 genes (sparse), 50 "
                         "components")
    ap.add_argument("--pca-genes", type=int, default=2000,
                    help="genes of the pca workload (2000 = the atlas after gene selection)")
    ap.add_argument("--pca-density", type=float, default=0.08,
                    help="share of stored entries of the pca workload's matrix")
    ap.add_argument("--pca-method", default="auto", choices=["auto", "gram", "randomized"],
                    help="pca route: auto = exact Gram up to 32768 genes, randomized block Krylov above")
    ap.add_argument("--pca-n-iter", type=int, default=4,
                    help="Krylov depth of the randomized route (pp.pca passes 4)")
    ap.add_argument("--pca-oversample", type=int, default=10)
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run 5 elementwise copies of 1 GiB (y = x * 1, a kernel of known "
                         "traffic: 1 GiB read + 1 GiB written) to scale the HBM counters by")
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
    if args.copy_calibration:
        a = torch.empty(1 << 28, dtype=torch.float32, device="cuda").fill_(1.0)
        b = torch.empty_like(a)
        for _ in range(5):
            torch.mul(a, 1.0, out=b)
        torch.cuda.synchronize()
        print(json.dumps({"read_bytes_per_dispatch": a.numel() * 4,
                          "write_bytes_per_dispatch": b.numel() * 4, "dispatches": 5}), flush=True)
        return 0
    if args.workload == "precision":
        return precision_workload(args, torch, P, ds)
    if args.workload == "batched":
        return batched_workload(args, torch, P, ds)
    if args.workload == "fuzzy":
        return fuzzy_workload(args, torch, P, ds)
    if args.workload == "scrublet":
        return scrublet_workload(args, torch, P, ds)
    if args.workload == "pca":
        return pca_workload(args, torch, P, ds)
    n, k = args.n, args.k
    log(f"generating {n} x 50 mixture ({args.components} components, seed {args.seed})")
    x = ds.mixture(n, 50, args.components, seed=args.seed)
    xt = torch.from_numpy(x).cuda()
    torch.cuda.synchronize()

    forms = FORMS if args.form == "all" else [args.form]
    enc = P.Encoder(seed=args.seed) if "learned_bf16" in forms else None
    train_log = []

    def learned_step():
        # per-dataset work inside the step: the encoder is trained from its initial
        # weights on this embedding (triplet mining + Adam), then seeds the search
        enc.reset(args.seed)
        train_log.append(enc.train(xt, subset=args.train_subset, steps=args.train_steps,
                                   anchors=args.train_anchors, batch=args.train_batch,
                                   lr=args.train_lr, margin=args.train_margin, seed=args.seed))
        return P.nndescent_device(xt, k, seed=args.seed, init="learned", model=enc,
                                  precision="bf16", n_trees=args.learned_trees,
                                  heap_k=args.learned_heap_k, delta=args.learned_delta,
                                  slots=args.learned_slots)

    def make_step(form):
        if form == "plain":
            return lambda: P.nndescent_device(xt, k, seed=args.seed, init=args.init)
        if form == "plain_r95":
            return lambda: P.nndescent_device(xt, k, seed=args.seed, init=args.init,
                                              heap_k=args.plain_heap_k)
        if form == "plain_op":
            return lambda: P.nndescent_device(xt, k, seed=args.seed, init="rp_forest",
                                              n_trees=args.learned_trees, heap_k=args.learned_heap_k,
                                              delta=args.learned_delta, slots=args.learned_slots)
        if form == "bf16_raw":
            return lambda: P.nndescent_device(xt, k, seed=args.seed, init="rp_forest",
                                              precision="bf16", n_trees=args.learned_trees,
                                              heap_k=args.learned_heap_k, delta=args.learned_delta,
                                              slots=args.learned_slots)
        return learned_step

    results = {}
    for form in forms:
        step = make_step(form)
        for w in range(args.warmup):
            t = time.perf_counter()
            step()
            torch.cuda.synchronize()
            log(f"[{form}] warmup {w + 1}/{args.warmup}: {time.perf_counter() - t:.3f} s")
        train_log.clear()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        out = None
        stats = []
        for _ in range(args.steps):
            out = step()
            stats.append(P.last_stats())
        torch.cuda.synchronize()
        elapsed = time.perf_counter() - t0
        st = stats[-1]
        log(f"[{form}] {1000 * elapsed / args.steps:.1f} ms/step, {n * args.steps / elapsed:,.0f} "
            f"cells/s; iterations {st['iterations']}, ms init {st['ms_init']:.1f} iter "
            f"{st['ms_iterations']:.1f} extract {st['ms_extract']:.1f}, changed "
            f"{st['changed'][:st['iterations']]}, join ms "
            f"{[round(v, 2) for v in st['ms_join'][:st['iterations']]]}, per cell: pairs "
            f"{[round(v / n, 1) for v in st['pairs'][:st['iterations']]]} exact "
            f"{[round(v / n, 1) for v in st['exact_pairs'][:st['iterations']]]} proposals "
            f"{[round(v / n, 1) for v in st['proposals'][:st['iterations']]]} rows "
            f"{[round(v / n, 1) for v in st['cand_rows'][:st['iterations']]]}")
        results[form] = {"elapsed": elapsed, "stats": stats, "out": out,
                         "train": list(train_log) if form == "learned_bf16" else None}

    head = forms[-1]
    elapsed = results[head]["elapsed"]
    stats = results[head]["stats"]
    ms_per_step = 1000.0 * elapsed / args.steps
    value = n * args.steps / elapsed
    st = stats[-1]

    # recall@15 of each form against exact kNN over all cells (same queries)
    nq = min(args.recall_queries, n)
    q = np.sort(np.random.default_rng(12345).choice(n, nq, replace=False)).astype(np.int32)
    qt = torch.from_numpy(q).cuda()
    t = time.perf_counter()
    ei, _ = P.knn_exact_device(xt, k, qt)
    torch.cuda.synchronize()
    log(f"exact kNN for {nq} queries: {time.perf_counter() - t:.2f} s")
    ei = ei.cpu().numpy()
    recalls = {}
    for form, r in results.items():
        ix = r["out"][1]
        gq = ix[: n * k].view(n, k)[qt.long()].cpu().numpy()
        rec = ds.recall(gq, ei)
        recalls[form] = (float(rec.mean()), float(np.percentile(rec, 5)))
        log(f"[{form}] recall@{k} = {recalls[form][0]:.4f} (p05 {recalls[form][1]:.3f})")
    recall = recalls[head][0]

    # roofline of the dominant kernel of the head form (its local join), over every
    # timed join pass (one dispatch over all cells with room for 32 candidate rows,
    # one over the cells with more; HIP events on the join's stream bracket both):
    # algorithmic bytes = candidate rows gathered x (d x 4 B + 4 B threshold) + the
    # two candidate lists (2m x 4 B per cell) + the count word.
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
    kname = "k_join_bf16<50>" if "bf16" in head else "k_join<50>"
    default_workload = (n, k, args.components, args.seed, args.init) == (4_060_000, 15, 150, 0,
                                                                          "rp_forest")
    roofline = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                "traffic": measured_traffic(kname) if default_workload else None,
                "kernel": kname,
                "per": "join pass (the kernel's dispatch over every cell + its dispatch over "
                       "the cells with more than 32 candidates)",
                "passes": launches, "avg_pass_ms": round(join_ms / max(launches, 1), 3),
                "algorithmic_bytes_per_pass": round(join_bytes / max(launches, 1))}

    cpu = None
    cpu_c1 = None
    speedup_rows = []
    if not args.no_cpu:
        import oracle
        threads = host_threads()
        # the reference CPU path on C1: 20,000 cells x 50, 20 components, k = 15, seed 0
        x1 = ds.mixture(20_000, 50, 20, seed=0)
        t = time.perf_counter()
        g1, st1 = oracle.nndescent_port(x1, 15, seed=0, threads=threads)
        c1t = time.perf_counter() - t
        e1, _ = oracle.knn_exact64(x1, 15, threads=threads)
        r1 = ds.recall(g1["indices"].reshape(20_000, 15), e1)
        cpu_c1 = {"value": round(20_000 / c1t, 1), "unit": "cells/s", "cores": threads,
                  "kind": "port", "recall_at_15": round(float(r1.mean()), 5),
                  "recall_p05": round(float(np.percentile(r1, 5)), 4),
                  "sample": f"20000 x 50, 20-component mixture, k=15, seed 0, auto options "
                            f"({st1.n_trees} trees, {st1.iterations} iterations), {c1t:.2f} s"}
        log(f"CPU baseline C1: {20_000 / c1t:,.0f} cells/s, recall {r1.mean():.4f}")
        nc = min(args.cpu_cells, n)
        log(f"CPU baseline: reference NN-Descent restated, {nc} cells, {threads} threads")
        xs = x if nc == n else np.ascontiguousarray(x[:nc])
        t = time.perf_counter()
        gp, pst = oracle.nndescent_port(xs, k, seed=args.seed, threads=threads)
        ct = time.perf_counter() - t
        cpu = {"value": round(nc / ct, 1), "unit": "cells/s", "cores": threads, "kind": "port",
               "sample": (f"{'the full' if nc == n else f'first {nc} cells of the'} {n} x 50 "
                          f"workload, k={k}, seed {args.seed}, auto options ({pst.n_trees} trees, "
                          f"{pst.iterations} iterations), {ct:.2f} s")}
        if nc == n:
            pr = ds.recall(gp["indices"].reshape(n, k)[q], ei)
            cpu["recall_at_15"] = round(float(pr.mean()), 5)
            cpu["recall_p05"] = round(float(np.percentile(pr, 5)), 4)
        log(f"CPU baseline: {nc / ct:,.0f} cells/s ({ct:.2f} s), recall {cpu.get('recall_at_15')}")
        # the speed-up table: the GPU drop-in (the reference's defaults) against the CPU
        # restatement on C1, C2 and this workload, both timed from the embedding resident
        # (device / host memory) to the CSR graph resident
        speedup_rows = []
        for name, nn_, comps in (("C1", 20_000, 20), ("C2", 1_300_000, 60)):
            xc = x1 if name == "C1" else ds.mixture(nn_, 50, comps, seed=0)
            xct = torch.from_numpy(xc).cuda()
            P.nndescent_device(xct, 15, seed=0)
            torch.cuda.synchronize()
            reps = 5 if name == "C1" else 2
            t = time.perf_counter()
            for _ in range(reps):
                og = P.nndescent_device(xct, 15, seed=0)
            torch.cuda.synchronize()
            gt = (time.perf_counter() - t) / reps
            if name == "C1":
                qc, ec, cg, cs_ = np.arange(nn_, dtype=np.int32), e1, g1, c1t
            else:
                log(f"CPU restatement on C2 ({nn_} cells, {threads} threads) ...")
                t = time.perf_counter()
                cg, _ = oracle.nndescent_port(xc, 15, seed=0, threads=threads)
                cs_ = time.perf_counter() - t
                qc = np.sort(np.random.default_rng(12345).choice(nn_, 10_000, replace=False)).astype(np.int32)
                ec = P.knn_exact_device(xct, 15, torch.from_numpy(qc).cuda())[0].cpu().numpy()
            gq = og[1][: nn_ * 15].view(nn_, 15)[torch.from_numpy(qc).cuda().long()].cpu().numpy()
            speedup_rows.append({
                "workload": f"{name}: {nn_} x 50, {comps}-component mixture, k=15, seed 0",
                "gpu_ms": round(1000 * gt, 2), "cpu_ms": round(1000 * cs_, 1),
                "speedup": round(cs_ / gt, 1),
                "gpu_recall_at_15": round(float(ds.recall(gq, ec).mean()), 5),
                "cpu_recall_at_15": round(float(ds.recall(cg["indices"].reshape(nn_, 15)[qc], ec).mean()), 5),
                "queries": int(len(qc))})
            log(f"speed-up {name}: GPU {1000 * gt:.1f} ms, CPU {1000 * cs_:.0f} ms")
            del xct

    forms_out = {}
    for form, r in results.items():
        s_ = r["stats"][-1]
        forms_out[form] = {
            "what": FORM_DESC[form],
            "cells_per_s": round(n * args.steps / r["elapsed"], 1),
            "ms_per_step": round(1000.0 * r["elapsed"] / args.steps, 3),
            "recall_at_15": round(recalls[form][0], 5), "recall_p05": round(recalls[form][1], 4),
            "at_recall_bar": recalls[form][0] >= 0.95,
            "iterations": s_["iterations"], "ms_init": round(s_["ms_init"], 2),
            "ms_iterations": round(s_["ms_iterations"], 2), "ms_extract": round(s_["ms_extract"], 2),
            "n_trees": s_["n_trees"], "heap_k": s_["heap_k"], "slots": s_["slots"],
            "stop_at": s_["stop_at"]}
        if r["train"]:
            tl = r["train"]
            forms_out[form]["encoder_train_ms_mean"] = round(sum(t_["train_ms"] for t_ in tl) / len(tl), 2)
            forms_out[form]["encoder_mine_ms_mean"] = round(sum(t_["mine_ms"] for t_ in tl) / len(tl), 2)
    line = {
        "metric": METRIC, "value": round(value, 1), "unit": "cells/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(ms_per_step, 3),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "bf16" if "bf16" in head else "f32",
        "data": "synthetic Gaussian-mixture embedding (numpy, seeded); no dataset",
        "config": {"workload": f"{FORM_DESC[head]}, {n} x 50 FP32 cells, {args.components}-"
                               f"component mixture, k={k}, seed {args.seed}",
                   "n_cells": n, "dims": 50, "k": k, "form": head, "seed": args.seed,
                   "components": args.components},
        "recall_at_15": round(recall, 5),
        "forms": forms_out,
        "roofline": roofline,
        "cpu_baseline": cpu,
    }
    if cpu_c1 is not None:
        line["cpu_baseline_c1"] = cpu_c1
    if cpu is not None and "plain" in results and cpu.get("recall_at_15") is not None:
        pf = forms_out["plain"]
        speedup_rows.append({
            "workload": f"C5: {n} x 50, {args.components}-component mixture, k={k}, seed {args.seed}",
            "gpu_ms": pf["ms_per_step"], "cpu_ms": round(1000 * ct, 1),
            "speedup": round(1000 * ct / pf["ms_per_step"], 1),
            "gpu_recall_at_15": pf["recall_at_15"], "cpu_recall_at_15": cpu["recall_at_15"],
            "queries": int(nq)})
    if cpu is not None:
        line["speedup_table"] = {
            "what": ("the GPU drop-in (plain form, the reference's defaults) against the reference's "
                     f"CPU path restated (C++, {threads} host threads), same input and seed"),
            "rows": speedup_rows}
    if "plain_r95" in results and head != "plain_r95":
        line["speedup_vs_plain"] = round(forms_out[head]["cells_per_s"] /
                                         forms_out["plain_r95"]["cells_per_s"], 3)
        line["speedup_vs_plain_note"] = ("against plain_r95: the reference's algorithm (FP32 join, "
                                         "48-tree forest, its defaults) with rows of "
                                         f"{args.plain_heap_k} while iterating, the smallest row "
                                         "size that clears recall@15 0.95 here")
    if "plain" in results and head != "plain":
        line["speedup_vs_plain_defaults"] = round(forms_out[head]["cells_per_s"] /
                                                  forms_out["plain"]["cells_per_s"], 3)
    if "plain_op" in results and head != "plain_op":
        line["speedup_vs_plain_same_operating_point"] = round(
            forms_out[head]["cells_per_s"] / forms_out["plain_op"]["cells_per_s"], 3)
    if head == "learned_bf16":
        line["encoder"] = {"steps": args.train_steps, "subset": args.train_subset,
                           "anchors": args.train_anchors, "batch": args.train_batch,
                           "lr": args.train_lr, "margin": args.train_margin,
                           "note": "trained from its initial weights inside every timed step"}
    print(json.dumps(line), flush=True)
    return 0


def _bf16(a: np.ndarray) -> np.ndarray:
    """float32 -> nearest BF16 value (ties to even), as float32."""
    u = np.ascontiguousarray(a, dtype=np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32)


def precision_workload(args, torch, P, ds) -> int:
    """4.06M x 50 with 5% uniform outliers (150 components), k = 30, BF16 form
    (rows of 2k = 60, each ranked in FP32), seed 0. Reports cells/s, recall@30
    over 20k queries and the distance error against FP64 on a 50k-cell subset:
    of the output distances, and of the BF16 screen's squared distance for
    candidate pairs of the subset's cells (recomputed on the host: operands
    centred on the cell and rounded to BF16, sums in FP32)."""
    n, k = args.n, 30
    x = ds.mixture_with_outliers(n, 50, args.components, seed=args.seed, outlier_frac=0.05)
    xt = torch.from_numpy(x).cuda()
    torch.cuda.synchronize()
    step = lambda: P.nndescent_device(xt, k, seed=args.seed, precision="bf16")  # noqa: E731
    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    out = None
    for _ in range(args.steps):
        out = step()
    torch.cuda.synchronize()
    el = (time.perf_counter() - t0) / args.steps
    st = P.last_stats()
    ix = out[1][: n * k].view(n, k)
    dx = out[2][: n * k].view(n, k)
    rng = np.random.default_rng(777)
    q = np.sort(rng.choice(n, 20000, replace=False)).astype(np.int32)
    qt = torch.from_numpy(q).cuda()
    ei, _ = P.knn_exact_device(xt, k, qt)
    rec = ds.recall(ix[qt.long()].cpu().numpy(), ei.cpu().numpy())
    sub = np.sort(rng.choice(n, 50000, replace=False))
    si = ix[torch.from_numpy(sub).cuda().long()].cpu().numpy().astype(np.int64)
    sd = dx[torch.from_numpy(sub).cuda().long()].cpu().numpy().astype(np.float64)
    xs = x[sub].astype(np.float64)
    diff = x[si].astype(np.float64) - xs[:, None, :]
    d64 = np.sqrt((diff * diff).sum(-1))
    ok = d64 > 0
    rel_out = np.abs(sd[ok] - d64[ok]) / d64[ok]
    # BF16 screen, as in a join: two candidates p, q of cell c (here: c's 1st-10th
    # and 11th-20th output neighbours), both centred on x_c and rounded to BF16;
    # d~ = |p~|^2 + |q~|^2 - 2 p~.q~ with FP32 sums, against FP64 |x_p - x_q|^2
    cp = _bf16((x[si[:, :10]] - x[sub][:, None, :]).astype(np.float32))
    cq = _bf16((x[si[:, 10:20]] - x[sub][:, None, :]).astype(np.float32))
    dt = ((cp * cp).sum(-1, dtype=np.float32) + (cq * cq).sum(-1, dtype=np.float32)
          - np.float32(2) * (cp * cq).sum(-1, dtype=np.float32)).astype(np.float64)
    e = x[si[:, :10]].astype(np.float64) - x[si[:, 10:20]].astype(np.float64)
    sq64 = (e * e).sum(-1)
    okp = sq64 > 0
    rel_screen = np.abs(np.maximum(dt[okp], 0.0) - sq64[okp]) / sq64[okp]
    line = {
        "metric": "cells/s to kNN graph, BF16 form, k=30, 4.06M x 50 with 5% outliers",
        "value": round(n / el, 1), "unit": "cells/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(1000 * el, 3), "higher_is_better": True,
        "dtype": "bf16", "data": "synthetic Gaussian mixture + uniform outliers (numpy, seeded)",
        "config": {"workload": f"BF16 MFMA screen + FP32 re-rank, {n} x 50, k={k}, 5% outliers, "
                               f"{args.components} components, seed {args.seed}",
                   "heap_k": st["heap_k"], "iterations": st["iterations"]},
        "recall_at_30": round(float(rec.mean()), 5),
        "distance_error_vs_fp64_50k_subset": {
            "output_max_rel": float(rel_out.max()), "output_p99_rel": float(np.percentile(rel_out, 99)),
            "edges": int(ok.sum()),
            "bf16_screen_sq_max_rel": float(rel_screen.max()),
            "bf16_screen_sq_p99_rel": float(np.percentile(rel_screen, 99)),
            "bf16_screen_pairs": int(okp.sum())},
    }
    print(json.dumps(line), flush=True)
    return 0


def batched_workload(args, torch, P, ds) -> int:
    """64 samples, sizes lognormal(10, 1.2) clipped to 2k-200k cells, each an
    independent 20-component mixture shifted by N(0, I), plus 2x doublets; exact
    kNN (k = 15) of every point within its sample, all samples in one persistent
    launch per pass, tiles from the autotuner. Timed after one untimed call (which
    runs the autotuner)."""
    samples = ds.batched_samples(64, 50, 20, seed=args.seed)
    x = np.concatenate([s[0] for s in samples])
    off = np.zeros(len(samples) + 1, np.uint64)
    off[1:] = np.cumsum([s[0].shape[0] for s in samples])
    xt = torch.from_numpy(x).cuda()
    P.knn_exact_batched_device(xt, off, 15)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        P.knn_exact_batched_device(xt, off, 15)
    torch.cuda.synchronize()
    el = (time.perf_counter() - t0) / args.steps
    rep = P.batched_report()
    pairs = float(sum(int(off[i + 1] - off[i]) ** 2 for i in range(len(samples))))
    line = {
        "metric": "points/s, per-sample exact kNN (k=15), 64 samples with doublets, one launch",
        "value": round(x.shape[0] / el, 1), "unit": "points/s", "n_gpus": 1, "steps": args.steps,
        "ms_per_step": round(1000 * el, 3), "higher_is_better": True, "dtype": "f32",
        "data": "synthetic per-sample mixtures with doublets (numpy, seeded)",
        "config": {"workload": f"64 samples, {x.shape[0]} points in total, sizes "
                               f"{int(np.diff(off).min())}-{int(np.diff(off).max())}, k=15",
                   "autotuner": rep},
    }
    # both passes evaluate the matrix-core dot product of every pair (2 d flops;
    # the norms are per point): pass 1 over the call's pass-1 time
    p1 = rep.get("ms_pass1") or 0.0
    flop_pass = 2.0 * 50 * pairs
    line["roofline"] = {
        "bound": "mfma", "kernel": "k_bexact_pass1 (FP32 matrix-core filter, per-segment top-16)",
        "per": "2 d flops per (query, point) pair of a sample, self included",
        "achieved": round(flop_pass / (p1 / 1000.0) / 1e12, 2) if p1 > 0 else None,
        "peak": 157.3, "unit": "TFLOP/s",
        "frac": round(flop_pass / (p1 / 1000.0) / 1e12 / 157.3, 4) if p1 > 0 else None,
        "traffic": None}
    line["distance_tflops_both_passes"] = round(2 * flop_pass / el / 1e12, 2)
    print(json.dumps(line), flush=True)
    return 0


def fuzzy_workload(args, torch, P, ds) -> int:
    """Fuzzy connectivities (the kNN graph's consumer in pp.neighbors) of the
    headline graph: 4.06M x 50 mixture, plain NN-Descent k = 15, built untimed and
    left in HBM. A step is one call: strengths, transpose, fuzzy union, the
    symmetric CSR left in HBM. The roofline is the whole call's: the input CSR
    read once and the output CSR written once, over the call's time from HIP
    events on the stream the library launches on. CPU baseline: the oracle's
    restatement of umap.rs on the full graph, one thread; its output is also the
    parity check of the GPU result."""
    n, k = args.n, args.k
    x = ds.mixture(n, 50, args.components, seed=args.seed)
    xt = torch.from_numpy(x).cuda()
    ip, ix, dx, nnz = P.nndescent_device(xt, k, seed=args.seed)
    del xt
    torch.cuda.synchronize()
    cap = 2 * nnz
    out = (torch.empty(n + 1, dtype=torch.int32, device="cuda"),
           torch.empty(cap, dtype=torch.int32, device="cuda"),
           torch.empty(cap, dtype=torch.float32, device="cuda"))
    for _ in range(args.warmup):
        P.fuzzy_connectivities_device(ip, ix, dx, n, out=out)
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    t0 = time.perf_counter()
    e0.record()
    for _ in range(args.steps):
        _, _, _, cnnz = P.fuzzy_connectivities_device(ip, ix, dx, n, out=out)
    e1.record()
    torch.cuda.synchronize()
    el = (time.perf_counter() - t0) / args.steps
    gpu_s = e0.elapsed_time(e1) / 1000.0 / args.steps
    alg = 2 * (n + 1) * 4 + nnz * 8 + cnnz * 8
    line = {
        "metric": "cells/s to fuzzy connectivities of a 4.06M x 15-NN graph",
        "value": round(n / el, 1), "unit": "cells/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(1000 * el, 3), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic Gaussian mixture (numpy, seeded); its NN-Descent graph",
        "config": {"workload": f"fuzzy connectivities of the plain NN-Descent graph of {n} x 50 "
                               f"FP32 cells, {args.components}-component mixture, k={k}, "
                               f"seed {args.seed}", "edges_in": nnz, "edges_out": cnnz},
        "roofline": {"bound": "hbm", "kernel": "psq_fuzzy_connectivities_device (all kernels)",
                     "achieved": round(alg / gpu_s / 1e9, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(alg / gpu_s / 1e9 / HBM_PEAK_GBS, 4),
                     "traffic": None},
    }
    if not args.no_cpu:
        import oracle
        h_ip = ip.cpu().numpy().view(np.uint32)
        h_ix = ix.cpu().numpy().view(np.uint32)[:nnz]
        h_dx = dx.cpu().numpy()[:nnz]
        log("oracle fuzzy connectivities (one thread) ...")
        c0 = time.perf_counter()
        want = oracle.fuzzy_connectivities(h_ip, h_ix, h_dx, n)
        cs = time.perf_counter() - c0
        line["cpu_baseline"] = {"value": round(n / cs, 1), "unit": "cells/s", "cores": 1,
                                "kind": "port",
                                "sample": f"the full graph ({n} rows x {k}), one thread"}
        g_ip = out[0].cpu().numpy().view(np.uint32)
        g_ix = out[1][:cnnz].cpu().numpy().view(np.uint32)
        g_w = out[2][:cnnz].cpu().numpy()
        same = np.array_equal(g_ip, want["indptr"]) and np.array_equal(g_ix, want["indices"])
        ulps = (np.abs(g_w.view(np.int32).astype(np.int64) - want["data"].view(np.int32))
                if same else None)
        line["parity_vs_oracle"] = {
            "indptr_indices_equal": bool(same),
            "weights_max_ulp": None if ulps is None else int(ulps.max(initial=0)),
            "weights_bit_exact_frac": None if ulps is None else float((ulps == 0).mean())}
    print(json.dumps(line), flush=True)
    return 0


def scrublet_workload(args, torch, P, ds) -> int:
    """Scrublet's exact neighbour search (core/doublets.rs:1083-1090): 100k cells
    of a 30-dimensional embedding plus 200k simulated doublets (means of two
    random cells), every point's k_adj = round(0.5 sqrt(100k)) * 3 = 474 nearest,
    self excluded, ordered by (FP64 distance, index). A step is one
    psq_knn_exact_device call over all 300k points; the roofline is the FP32
    matrix-core filter's flops (two passes of 2 d flops per pair, the dot
    products' multiply-adds) over the call's HIP-event time. CPU baseline: the oracle's FP64 brute force for the first 3000
    queries on this host's cores, scaled to points/s."""
    rng = np.random.default_rng(args.seed)
    n_cells, d = 100_000, 30
    cells = ds.mixture(n_cells, d, 20, seed=args.seed)
    a, b = rng.integers(0, n_cells, (2, 2 * n_cells))
    x = np.concatenate([cells, ((cells[a] + cells[b]) / 2).astype(np.float32)])
    n = x.shape[0]
    k = int(np.round(0.5 * np.sqrt(n_cells))) * 3
    xt = torch.from_numpy(x).cuda()
    for _ in range(args.warmup):
        P.knn_exact_device(xt, k)
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    t0 = time.perf_counter()
    e0.record()
    for _ in range(args.steps):
        idx, dist = P.knn_exact_device(xt, k)
    e1.record()
    torch.cuda.synchronize()
    el = (time.perf_counter() - t0) / args.steps
    gpu_s = e0.elapsed_time(e1) / 1000.0 / args.steps
    flops = 2.0 * 2.0 * d * float(n) * float(n)
    line = {
        "metric": "points/s, exact kNN at Scrublet's k (474) over 100k cells + 200k doublets",
        "value": round(n / el, 1), "unit": "points/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(1000 * el, 3), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic mixture cells + simulated doublets (numpy, seeded)",
        "config": {"workload": f"{n_cells} cells + {2 * n_cells} doublets, d={d}, k={k}, all "
                               f"{n} points as queries"},
        "roofline": {"bound": "mfma", "kernel": "psq_knn_exact_device (FP32 filter, both passes)",
                     "achieved": round(flops / gpu_s / 1e12, 2), "peak": 157.3,
                     "unit": "TFLOP/s", "frac": round(flops / gpu_s / 1e12 / 157.3, 4),
                     "traffic": None,
                     "note": "v_mfma_f32_16x16x4f32 dot products, d padded to a multiple of 4 "
                             "(32 for d = 30), plus each pass's top-K insertion or candidate "
                             "test on the VALU"},
    }
    if not args.no_cpu:
        import oracle
        nq = 3000
        log(f"oracle FP64 exact kNN for {nq} queries ...")
        c0 = time.perf_counter()
        ei, _ = oracle.knn_exact64(x, k, np.arange(nq, dtype=np.uint32), threads=host_threads())
        cs = time.perf_counter() - c0
        line["cpu_baseline"] = {"value": round(nq / cs, 1), "unit": "points/s",
                                "cores": host_threads(), "kind": "port",
                                "sample": f"first {nq} of the {n} queries, FP64 brute force"}
        gi = idx[:nq].cpu().numpy().view(np.uint32)
        line["parity_vs_oracle_first_3000"] = bool(np.array_equal(gi, ei))
    print(json.dumps(line), flush=True)
    return 0


def pca_counts(torch, n, m, density, seed, chunk=None):
    """A synthetic log-normalised sparse expression matrix on the device with the
    structure PCA is for: ten gene programs of decaying strength (cell scores
    N(0, s_p), s = 6 ... 0.8; gene loadings N(0, 1/10)) on top of gene-specific
    count noise; entries kept with probability `density`; values
    log1p(floor(u * (1 + 4 g_j)) + 1 + max(program signal, 0)). CSR with sorted
    column indices per row."""
    chunk = chunk or max(1024, min(65536, (1 << 27) // m))  # rows of at most 2^27 draws
    gen = torch.Generator(device="cuda")
    gen.manual_seed(seed)
    gscale = 1.0 + 4.0 * torch.rand(m, generator=gen, device="cuda")
    strength = torch.tensor([6.0, 5.0, 4.0, 3.0, 2.5, 2.0, 1.5, 1.2, 1.0, 0.8], device="cuda")
    load = torch.randn((10, m), generator=gen, device="cuda") / 10 ** 0.5
    counts, idx, val = [], [], []
    for r0 in range(0, n, chunk):
        rc = min(chunk, n - r0)
        score = torch.randn((rc, 10), generator=gen, device="cuda") * strength
        keep = torch.rand((rc, m), generator=gen, device="cuda") < density
        nz = keep.nonzero()
        rows, cols = nz[:, 0], nz[:, 1]
        u = torch.rand(rows.numel(), generator=gen, device="cuda")
        sig = (score[rows] * load[:, cols].t()).sum(dim=1).clamp_min(0.0)
        v = torch.floor(u * gscale[cols]) + 1.0 + sig
        counts.append(torch.bincount(rows, minlength=rc))
        idx.append(cols.to(torch.int32))
        val.append(torch.log1p(v).to(torch.float32))
        del keep, nz, sig
    indptr = torch.zeros(n + 1, dtype=torch.int64, device="cuda")
    indptr[1:] = torch.cumsum(torch.cat(counts), 0)
    return indptr, torch.cat(idx), torch.cat(val)


def pca_workload(args, torch, P, ds) -> int:
    """pp.pca of a sparse matrix whose CSR is already in HBM, 50 centred
    components. Gram route (default: the atlas after gene selection, 4.06M
    cells x 2000 genes, 8% stored; pca.rs:487-535): a step is one
    psq_pca_csr_device call, the Gram pass (dense FP64 row chunks, rocBLAS
    DGEMM), the Chebyshev-accelerated subspace eigensolver, projection, signs;
    the roofline is the Gram product (the algorithmic n m (m + 1) flops of the
    symmetric Gram over its DGEMM launches' HIP-event time) against the FP64
    matrix peak. Randomized route (more than 32,768 genes, or --pca-method
    randomized; rsvd.rs:476-566): a step is one call, the column sums and the
    passes of the block Krylov SVD (oversample 10, n_iter 4 as pp.pca passes);
    the roofline is its transposed-product DGEMMs (2 n m flops per column). CPU
    baseline: the oracle (numpy/LAPACK or scipy, this host's cores) on the
    first 20,000 cells; the GPU's result on the same sample is its parity check."""
    from polariseq_amd import pca as gp
    n, m, k, dens = args.n, args.pca_genes, 50, args.pca_density
    method = args.pca_method
    if method == "auto":
        method = "randomized" if m > 32768 else "gram"
    over, it = args.pca_oversample, args.pca_n_iter
    log(f"generating {n} x {m} sparse matrix ({dens:.0%} stored, seed {args.seed})")
    ip, ix, dx = pca_counts(torch, n, m, dens, args.seed)
    nnz = int(ix.numel())
    out = torch.empty((n, k), dtype=torch.float32, device="cuda")
    torch.cuda.synchronize()

    def call():
        return gp.pca_csr_device(ip, ix, dx, m, k, out=out, method=method, oversample=over, n_iter=it,
                                 seed=args.seed)
    for _ in range(args.warmup):
        call()
    torch.cuda.synchronize()
    reps = []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        _, vr = call()
        reps.append(gp.pca_last_report())
        log(f"step {len(reps)}: {reps[-1]['ms_total']:.1f} ms")
    torch.cuda.synchronize()
    el = (time.perf_counter() - t0) / args.steps
    gemm_s = sum(r["ms_gemm"] for r in reps) / len(reps) / 1000.0
    rep = reps[-1]
    if method == "gram":
        title = f"exact Gram PCA (50 components) of {n / 1e6:.2f}M cells x {m} genes"
        flops = float(n) * m * (m + 1)
        kern = "Gram DGEMM (rocBLAS, FP64; algorithmic flops n m (m + 1), the product computes twice that)"
        traffic = measured_traffic("gram_dgemm", "pca_gram_hbm_traffic.md") if m == 2000 and n == 4_060_000 else None
        note = "HBM bytes per DGEMM launch (one 16710-row chunk)"
    else:
        title = (f"randomized block Krylov PCA (50 components, oversample {over}, n_iter {it}) of "
                 f"{n / 1e6:.2f}M cells x {m} genes")
        most = min(n, m)
        l = max(1, min(k + over, most))
        width = min(l * (it + 1), most)
        blocks_ = -(-width // l)
        cols = l * (blocks_ - 1) + width  # the transposed products' columns: A'^T Q_d, then B^T
        flops = 2.0 * n * m * cols
        kern = ("transposed-product DGEMMs (rocBLAS, FP64; algorithmic flops 2 n m per column; the "
                "products with A' are the sparse kernel)")
        traffic, note = None, None
    line = {
        "metric": f"cells/s, {title}",
        "value": round(n / el, 1), "unit": "cells/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(1000 * el, 3), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic log-normalised sparse counts with 10 gene programs (torch, seeded)",
        "config": {"workload": f"{n} cells x {m} genes, {nnz} stored ({dens:.0%}), {k} components, "
                               f"centred, {method} route, CSR resident in HBM", "nnz": nnz},
        "phases_ms": {kk: round(v, 3) for kk, v in rep.items() if kk.startswith("ms_")},
        "solver": {"method": rep["method"], "iterations": rep["iterations"], "block": rep["block"],
                   "chunks": rep["chunks"], "rows_per_chunk": rep["rows_per_chunk"]},
        "roofline": {"bound": "mfma", "kernel": kern,
                     "achieved": round(flops / gemm_s / 1e12, 2), "peak": 78.6, "unit": "TFLOP/s",
                     "frac": round(flops / gemm_s / 1e12 / 78.6, 4), "traffic": traffic,
                     "traffic_note": note},
        "variance_ratio_head": [round(float(v), 6) for v in vr[:5].cpu().numpy()],
    }
    if not args.no_cpu:
        from oracle import pca as opca
        ns = 20_000
        h_ip = ip[: ns + 1].cpu().numpy().astype(np.uint64)
        e = int(h_ip[-1])
        h_ix = ix[:e].cpu().numpy().astype(np.uint32)
        h_dx = dx[:e].cpu().numpy()
        comp = np.zeros((m, k), np.float32)
        if method == "gram":
            log(f"oracle Gram PCA of the first {ns} cells ...")
            c0 = time.perf_counter()
            want = opca.pca_gram(h_ip, h_ix, h_dx, ns, m, k, True)
            cs = time.perf_counter() - c0
            sample = f"first {ns} cells of the same matrix, numpy/LAPACK Gram PCA"
            r = gp.pca_sparse((h_ip, h_ix, h_dx, (ns, m)), k, True, 10, 4, 0, components=comp, method="gram")
            c = opca.centred_gram(h_ip, h_ix, h_dx, ns, m, True)[0]
            top = float(want["eigenvalues"][0])
            res = max(float(np.linalg.norm(c @ comp[:, j].astype(np.float64)
                                           - (comp[:, j] @ c @ comp[:, j]) * comp[:, j].astype(np.float64)))
                      for j in range(k)) / top
            parity = {"variance_ratio_max_abs_diff": float(np.max(np.abs(r["variance_ratio"]
                                                                         - want["variance_ratio"]))),
                      "axis_residual_max_rel": res}
        else:
            from oracle import rsvd as orsvd
            log(f"oracle randomized PCA of the first {ns} cells ...")
            c0 = time.perf_counter()
            want = orsvd.rsvd(h_ip, h_ix, h_dx, ns, m, k, over, it, args.seed, True)
            cs = time.perf_counter() - c0
            sample = f"first {ns} cells of the same matrix, numpy/scipy block Krylov SVD"
            r = gp.pca_sparse((h_ip, h_ix, h_dx, (ns, m)), k, True, over, it, args.seed, components=comp,
                              method="randomized")
            scale = np.abs(want["embedding"]).max(axis=0)
            parity = {"variance_ratio_max_rel_diff": float(np.max(np.abs(r["variance_ratio"] / want["variance_ratio"]
                                                                         - 1.0))),
                      "embedding_max_rel_diff": float(np.max(np.abs(r["embedding"] - want["embedding"])
                                                             / scale[None, :]))}
        line["cpu_baseline"] = {"value": round(ns / cs, 1), "unit": "cells/s", "cores": host_threads(),
                                "kind": "port", "sample": sample}
        line["parity_vs_oracle_sample"] = parity
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
