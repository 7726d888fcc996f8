"""A fixed sequence of search launches for hardware profilers.

  python -m parallel_anns_amd.profile_run --op M,L[,eighths[,rows[,overflow]]]
                                          [--learned-op M,Lmax,threshold[,eighths[,min_iter,every[,rows[,overflow]]]]]
                                          [--config c1] [--steps 5]
  python -m parallel_anns_amd.profile_run --uneven M:maxL1,maxL2,...[:concurrent[:eighths[:rows[:overflow]]]]
                                          [--config c2] [--steps 5]
  python -m parallel_anns_amd.profile_run --knn K[,method] [--config c1] [--steps 5]

Loads the prepared workload from the cache (and the stop model bench.py
saved there), then runs the fixed-L plan `steps + 1` times and, when given,
the learned-stop plan `steps + 1` times, nothing else, so that the last
`steps` dispatches of each search-kernel variant in a kernel trace or counter
run are exactly the timed launches of the benchmark's two operating points.
Prints one JSON line with the device times and the algorithmic bytes per
launch.

With --knn the workload's exact ground truth (K nearest base rows of every
query; method "mfma" (default) or "valu") is computed `steps + 1` times; the
JSON line gives each call's wall time and the products' FLOP count.

With --uneven the batch is the workload's uneven one (per-query L
log-uniform over its L range, K from its K choices, all queries) in the
given queue-size buckets (eighths 0: each bucket's visited table by its own
largest L), replayed as one hipGraph `steps + 1` times; the JSON line then
gives, per bucket, its largest L, queries, LDS bytes and algorithmic bytes
(distance rows and results exact, adjacency rows at the batch's mean read
slots per expansion), and the whole run's algorithmic bytes and device time.
"""
import argparse
import json
import os
import sys

import numpy as np


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--config", default="c1")
    ap.add_argument("--cache-dir", default=os.environ.get("PANNS_CACHE", "/tmp/panns_cache"))
    ap.add_argument("--op", default=None)
    ap.add_argument("--uneven", default=None)
    ap.add_argument("--knn", default=None)
    ap.add_argument("--learned-op", default=None)
    ap.add_argument("--steps", type=int, default=5)
    args = ap.parse_args()

    from . import datasets
    from .search import Index

    ds = datasets.prepare(args.config, cache_dir=args.cache_dir, gpu_knn=datasets.GpuKnn(0))
    K = ds["cfg"]["K"]
    nq = ds["queries"].shape[0]
    dim = ds["base"].shape[1]
    idx = Index(0)
    idx.set_base(ds["base"])
    idx.set_queries(ds["queries"])
    idx.set_graph(ds["adj"], ds["deg"], ds["ep"])
    qids = np.arange(nq, dtype=np.uint32)

    def run(M, L, eighths, rows=0, ovf=0, **kw):
        pl = idx.plan(qids, K, L, M, want_queue=False,
                      options=Index.plan_options(table_eighths=eighths, rows_per_wave=rows, overflow_slots=ovf), **kw)
        pl.run()
        ms = [pl.run() for _ in range(args.steps)]
        res = pl.fetch()
        pl.close()
        c = res["counters"]
        alg = 4.0 * dim * c["distance_computations"] + 4.0 * (c["expansions"] + c["adjacency_slots"]) \
            + nq * (4.0 * dim + 8.0 * K)
        return {"M": M, "L": L, "table_eighths": eighths, "rows_per_wave": rows, "overflow_slots": ovf,
                "kernel_ms": round(float(np.mean(ms)), 4),
                "algorithmic_bytes_per_launch": int(alg)}

    if args.knn:
        import time
        f = args.knn.split(",")
        k, method = int(f[0]), (f[1] if len(f) > 1 else "mfma")
        walls = []
        for _ in range(args.steps + 1):
            st = {}
            t0 = time.perf_counter()
            idx.exact_knn(k, source="queries", method=method, stats=st)
            walls.append(time.perf_counter() - t0)
        n = ds["base"].shape[0]
        print(json.dumps({"config": args.config, "knn": k, "method": method, "queries": nq, "rows": n,
                          "wall_s": [round(w, 4) for w in walls[1:]], "recomputed": st.get("recomputed"),
                          "product_flops": 2.0 * nq * n * dim}), flush=True)
        idx.close()
        return 0
    if args.uneven:
        print(json.dumps(uneven_run(idx, ds, args)), flush=True)
        idx.close()
        return 0
    out = {"config": args.config, "steps": args.steps}
    f = [x for x in args.op.split(",")]
    out["fixed"] = run(int(f[0]), int(f[1]), int(f[2]) if len(f) > 2 else 8, int(f[3]) if len(f) > 3 else 0,
                       int(f[4]) if len(f) > 4 else 0)
    if args.learned_op:
        g = args.learned_op.split(",")
        model = np.load(os.path.join(ds["dir"] or args.cache_dir, "stop_model.npy"))
        idx.set_stop_model(model)
        out["learned"] = run(int(g[0]), int(g[1]), int(g[3]) if len(g) > 3 else 8,
                             int(g[6]) if len(g) > 6 else 0, int(g[7]) if len(g) > 7 else 0, stop="learned",
                             stop_threshold=float(g[2]), stop_min_iter=int(g[4]) if len(g) > 4 else 0,
                             stop_every=int(g[5]) if len(g) > 5 else 1)
        out["learned"]["threshold"] = float(g[2])
    print(json.dumps(out), flush=True)
    idx.close()


def uneven_run(idx, ds, args):
    from . import datasets
    from .search import Index

    cfg = ds["cfg"]
    nq, dim = ds["queries"].shape[0], ds["base"].shape[1]
    Ls, Ks = datasets.uneven_L_K(cfg, nq)
    f = args.uneven.split(":")
    M = int(f[0])
    bounds = [int(x) for x in f[1].split(",")]
    conc = bool(int(f[2])) if len(f) > 2 else False
    eighths = int(f[3]) if len(f) > 3 else 0
    rows = int(f[4]) if len(f) > 4 else 0
    ovf = int(f[5]) if len(f) > 5 else 0
    maxL, maxK = int(Ls.max()), int(Ks.max())
    assert bounds[-1] >= maxL, "the last bucket must hold the largest L"
    qids = np.arange(nq, dtype=np.uint32)
    init = idx.prepare_init_ids(maxL)
    opt = Index.plan_options(bucket_max_L=bounds, table_eighths=eighths, use_graph=True, concurrent=conc,
                             rows_per_wave=rows, overflow_slots=ovf)
    pl = idx.plan(qids, maxK, maxL, M, init_ids=init, L_per_query=Ls, K_per_query=Ks, options=opt,
                  want_queue=False)
    pl.run()
    ms = [pl.run() for _ in range(args.steps)]
    res = pl.fetch()
    bk = pl.buckets()
    pl.close()
    c = res["counters"]
    slots = c["adjacency_slots"] / max(1, c["expansions"])  # mean read slots per expansion
    # bucket of each query (the plan may list its buckets in launch order)
    order = np.argsort([b["max_L"] for b in bk])
    member = order[np.searchsorted(np.array([bk[i]["max_L"] for i in order]), Ls)]
    nd, ne = res["ndist"].astype(np.float64), res["nexp"].astype(np.float64)
    for i, b in enumerate(bk):
        sel = member == i
        assert int(sel.sum()) == b["nq"], (i, b)
        b["algorithmic_bytes"] = int(4.0 * dim * nd[sel].sum() + 4.0 * (1.0 + slots) * ne[sel].sum()
                                     + (4.0 * dim * sel.sum() + 8.0 * Ks[sel].astype(np.float64).sum()))
        b["iterations"] = int(res["niter"][sel].astype(np.int64).sum())
    alg = 4.0 * dim * c["distance_computations"] + 4.0 * (c["expansions"] + c["adjacency_slots"]) \
        + 4.0 * dim * nq + 8.0 * float(Ks.astype(np.float64).sum())
    return {"config": args.config, "steps": args.steps, "M": M, "uneven": True,
            "options": {"bucket_max_L": bounds, "concurrent": conc, "table_eighths": eighths,
                        "rows_per_wave": rows, "overflow_slots": ovf},
            "buckets": bk, "run_ms": round(float(np.mean(ms)), 4), "algorithmic_bytes_per_run": int(alg),
            "mean_L": round(float(Ls.mean()), 1)}


if __name__ == "__main__":
    sys.exit(main())
