"""Benchmark: QPS at recall@100 >= 0.95 on 1M x 96 fp32 (one MI355X).

Workload: 1,000,000 x 96 float32 base from the seeded normalised 64-component
Gaussian mixture (sigma 0.3), 10,000 queries from the same mixture, NSG-style
graph from exact 100-NN lists with out-degree <= 50, K = 100.  A step is one
pass of the search over all 10,000 queries with inputs resident in HBM (one
persistent-kernel launch).  The operating point (M, L) is the fastest setting
whose recall@100 reaches the target, found by bisection on L per M, as the
reference's find-L driver and result selection do
(app/profile_find_L_para_single_query_top_m_search_pure_large_graph.cpp:210-260,
scripts/output_find_runtime_above_presicion.py:40-60).

Prints one JSON line.  --gpus above 1 exits with a message: the build targets
one GPU (queries shard across GPUs with no exchange; see the design notes).
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


def log(msg):
    print(f"[bench {time.strftime('%H:%M:%S')}] {msg}", file=sys.stderr, flush=True)


def algorithmic_bytes(res, dim, K, nq):
    """B_q = 4*D*n_dist + 4*(n_exp + sum of expanded degrees) + 4*D + 8*K summed
    over queries: vector gathers, CSR adjacency reads (degree word + ids), the
    query row and the K results."""
    c = res["counters"]
    return (4.0 * dim * c["distance_computations"] + 4.0 * (c["expansions"] + c["adjacency_slots"])
            + nq * (4.0 * dim + 8.0 * K))


def find_L(index, qids, K, M, gt, target, lo, hi, hash_slots):
    """Smallest L in [lo, hi] with recall@100 >= target (bisection)."""
    from parallel_anns_amd import tools

    def rec(L):
        r = index.search(qids, K, L, M, hash_slots=hash_slots)
        return tools.recall(gt, r["ids"], L)

    if rec(hi) < target:
        return None
    while lo < hi:
        mid = (lo + hi) // 2
        if rec(mid) >= target:
            hi = mid
        else:
            lo = mid + 1
    return lo


def time_plan(plan, steps, warmup):
    for _ in range(warmup):
        plan.run(sync=True)
    kms = []
    t0 = time.perf_counter()
    for _ in range(steps):
        kms.append(plan.run(sync=True))
    wall = time.perf_counter() - t0
    return wall, kms


def cpu_baseline(ds, K, L, M, init_ids, budget_s=12.0):
    """The oracle restatement of top-M search, inter-query parallel on the
    host cores (num_t = 1 per query), timed on a bounded sample."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle as po
    cores = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else os.cpu_count()
    cores = max(1, min(cores, 16))
    g = po.CsrGraph(ds["adj"], ds["deg"], ds["ep"])
    q = ds["queries"]
    n_probe = min(200, q.shape[0])
    t0 = time.perf_counter()
    po.search_batch(g, ds["base"], q[:n_probe], K, L, M, init_ids, nthreads=cores)
    per_q = (time.perf_counter() - t0) / n_probe
    n_s = int(max(n_probe, min(q.shape[0], budget_s / max(per_q, 1e-9))))
    t0 = time.perf_counter()
    ids, _, _, _ = po.search_batch(g, ds["base"], q[:n_s], K, L, M, init_ids, nthreads=cores)
    el = time.perf_counter() - t0
    return {"value": n_s / el, "unit": "QPS", "cores": cores, "kind": "port",
            "sample": f"first {n_s} of {q.shape[0]} queries, top-M M={M} L={L} K={K}, "
                      f"OpenMP inter-query with {cores} threads (oracle restatement, ids identical "
                      f"to the GPU search)"}, ids


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--config", default="c1")
    ap.add_argument("--cache-dir", default=os.environ.get("PANNS_CACHE", "/tmp/panns_cache"))
    ap.add_argument("--target-recall", type=float, default=0.95)
    ap.add_argument("--op", default=None, help="M,L: skip the operating-point search")
    ap.add_argument("--ms", default="1,2,4,8", help="M values tried by the operating-point search")
    ap.add_argument("--hash-slots", type=int, default=0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--curve", action="store_true", help="also report QPS/recall for L in 100..400")
    args = ap.parse_args()

    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; queries shard across GPUs with no "
                                   "exchange step, so N>1 is not offered", "n_gpus": args.gpus}))
        sys.exit(2)

    from parallel_anns_amd import datasets, tools
    from parallel_anns_amd.search import Index

    ds = datasets.prepare(args.config, cache_dir=args.cache_dir, gpu_knn=datasets.GpuKnn(0))
    cfg = ds["cfg"]
    K = cfg["K"]
    nq = ds["queries"].shape[0]
    dim = ds["base"].shape[1]
    index = Index(0)
    index.set_base(ds["base"])
    index.set_queries(ds["queries"])
    index.set_graph(ds["adj"], ds["deg"], ds["ep"])
    qids = np.arange(nq, dtype=np.uint32)
    gt = ds["gt_ids"]

    # operating point
    if args.op:
        M, L = (int(x) for x in args.op.split(","))
        cands = [(M, L)]
    else:
        cands = []
        for M in (int(x) for x in args.ms.split(",")):
            L = find_L(index, qids, K, M, gt, args.target_recall, K, 1024, args.hash_slots)
            log(f"find-L M={M}: L={L}")
            if L is not None:
                cands.append((M, L))
    best = None
    for M, L in cands:
        plan = index.plan(qids, K, L, M, hash_slots=args.hash_slots)
        wall, kms = time_plan(plan, 3, 1)
        qps = 3 * nq / wall
        log(f"candidate M={M} L={L}: {qps:,.0f} QPS, kernel {np.mean(kms):.3f} ms")
        if best is None or qps > best[0]:
            best = (qps, M, L)
        plan.close()
    _, M, L = best

    plan = index.plan(qids, K, L, M, hash_slots=args.hash_slots)
    desc = plan.describe()
    wall, kms = time_plan(plan, args.steps, args.warmup)
    res = plan.fetch()
    recall = tools.recall(gt, res["ids"], L)
    qps = args.steps * nq / wall
    ms_step = 1000.0 * wall / args.steps
    k_ms = float(np.mean(kms))
    B = algorithmic_bytes(res, dim, K, nq)
    achieved = B / (k_ms * 1e-3) / 1e9
    log(f"timed M={M} L={L}: {qps:,.0f} QPS recall {recall:.4f} kernel {k_ms:.3f} ms "
        f"{achieved:,.0f} GB/s algorithmic; geometry {desc}")

    curve = None
    if args.curve:
        curve = []
        for Lc in (100, 150, 200, 300, 400):
            p2 = index.plan(qids, K, Lc, M, hash_slots=args.hash_slots)
            w2, k2 = time_plan(p2, 3, 1)
            r2 = p2.fetch()
            curve.append({"L": Lc, "M": M, "qps": round(3 * nq / w2, 1),
                          "recall": round(tools.recall(gt, r2["ids"], Lc), 4),
                          "kernel_ms": round(float(np.mean(k2)), 3)})
            p2.close()

    cpu = None
    if not args.no_cpu:
        init = index.prepare_init_ids(L)
        cpu, cpu_ids = cpu_baseline(ds, K, L, M, init)
        same = bool((cpu_ids == res["ids"][:cpu_ids.shape[0]]).all())
        cpu["ids_identical_to_gpu"] = same
        log(f"cpu baseline {cpu['value']:,.0f} QPS on {cpu['cores']} cores; ids identical: {same}")

    out = {
        "metric": "QPS @ recall@100=0.95 (1M x 96 fp32): plain fixed-L search, 1 MI355X",
        "value": round(qps, 1),
        "unit": "QPS",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded normalised Gaussian mixture, generated in the run)",
        "config": {"workload": f"{args.config}: {cfg['n']:,} x {dim} fp32, {cfg['comps']}-component mixture "
                               f"sigma {cfg['sigma']}, {nq:,} queries, graph R<={cfg['R']} from exact "
                               f"{cfg['knn']}-NN, K={K}",
                   "M": M, "L": L, "stop": "fixed", "recall@100": round(recall, 4),
                   "target_recall": args.target_recall, "parallelism": "single GPU, persistent kernel",
                   "workgroups": desc["workgroups"], "lds_bytes": desc["lds_bytes"],
                   "hash_slots": desc["hash_slots"]},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None,
                     "kernel": "search_kernel<96>", "kernel_ms": round(k_ms, 4),
                     "algorithmic_bytes_per_launch": int(B)},
        "cpu_baseline": cpu,
    }
    if curve:
        out["curve"] = curve
    print(json.dumps(out), flush=True)
    plan.close()
    index.close()


if __name__ == "__main__":
    main()
