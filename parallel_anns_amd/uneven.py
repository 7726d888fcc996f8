"""Uneven batches: queries of different L and K in one batch.

  python -m parallel_anns_amd.uneven [--config c2] [--M 4] [--steps 3]

Workload (configuration c2): 10,000,000 x 96 float32 base from the seeded
normalised 256-component mixture (sigma 0.3), graph out-degree <= 32 from
exact 64-NN lists, 100,000 queries whose L is log-uniform in [100, 2000]
and K drawn from {10, 100}.  Three forms of the same batch are timed, with
inputs resident in HBM:

  single       one persistent launch sized for the largest L
  per_bucket   the plain form: one launch per L bucket [100,200), [200,400),
               [400,800), [800,1600), [1600,2000], one after another
  autotuned    queue-size buckets chosen by panns_plan_autotune, replayed as
               one hipGraph

All three return the same ids and distance counts (checked here; the
bucketed forms are checked against the CPU oracle by
tests/test_gpu_buckets.py on smaller inputs).  Recall is recall@K against
exact brute force.  Prints one JSON line.
"""
import argparse
import json
import os
import sys
import time

import numpy as np


def log(msg):
    print(f"[uneven {time.strftime('%H:%M:%S')}] {msg}", file=sys.stderr, flush=True)


def recall_at_k(gt, ids, K):
    """Mean over queries of |truth[:K_q] & ids[:K_q]| / K_q."""
    tot = 0.0
    for q in range(ids.shape[0]):
        k = int(K[q])
        tot += len(set(gt[q, :k].tolist()) & set(ids[q, :k].tolist())) / k
    return tot / ids.shape[0]


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--config", default="c2")
    ap.add_argument("--cache-dir", default=os.environ.get("PANNS_CACHE", "/tmp/panns_cache"))
    ap.add_argument("--M", type=int, default=4)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--queries", type=int, default=0, help="use the first n queries (0: all)")
    args = ap.parse_args()

    from . import datasets
    from .search import Index

    ds = datasets.prepare(args.config, cache_dir=args.cache_dir, gpu_knn=datasets.GpuKnn(0))
    cfg = ds["cfg"]
    nq = args.queries or ds["queries"].shape[0]
    Ls, Ks = datasets.uneven_L_K(cfg, ds["queries"].shape[0])
    Ls, Ks = Ls[:nq], Ks[:nq]
    qids = np.arange(nq, dtype=np.uint32)
    idx = Index(0)
    idx.set_base(ds["base"])
    idx.set_queries(ds["queries"])
    idx.set_graph(ds["adj"], ds["deg"], ds["ep"])
    maxL, maxK = int(Ls.max()), int(Ks.max())
    init = idx.prepare_init_ids(maxL)
    M = args.M

    def timed(options, name):
        pl = idx.plan(qids, maxK, maxL, M, init_ids=init, L_per_query=Ls, K_per_query=Ks, options=options,
                      want_queue=False)
        pl.run()
        kms = [pl.run() for _ in range(args.steps)]
        res = pl.fetch()
        info = {"kernel_ms": round(float(np.mean(kms)), 3), "qps": round(nq / (np.mean(kms) * 1e-3), 1),
                "buckets": pl.buckets()}
        pl.close()
        log(f"{name}: {info['kernel_ms']} ms, {info['qps']:,.0f} QPS, {len(info['buckets'])} launch(es)")
        return res, info

    forms = {}
    res_single, forms["single"] = timed(None, "single")
    plain = Index.plan_options(bucket_max_L=[199, 399, 799, 1599, max(1600, maxL)])
    res_plain, forms["per_bucket"] = timed(plain, "per_bucket")
    t0 = time.time()
    opt, sample_ms = idx.autotune(qids, maxK, maxL, M, init_ids=init, L_per_query=Ls, K_per_query=Ks)
    tune = {"seconds": round(time.time() - t0, 2), "sample_ms": round(sample_ms, 3),
            "n_buckets": opt.n_buckets, "bucket_max_L": [opt.bucket_max_L[i] for i in range(opt.n_buckets)],
            "concurrent": opt.concurrent, "table_eighths": opt.table_eighths, "waves": max(1, opt.waves), "use_graph": opt.use_graph,
            "rows_per_wave": opt.rows_per_wave, "overflow_slots": opt.overflow_slots}
    log(f"autotuner: {tune}")
    res_tuned, forms["autotuned"] = timed(opt, "autotuned")
    forms["autotuned"]["choice"] = tune
    for name, r in (("per_bucket", res_plain), ("autotuned", res_tuned)):
        same = bool((r["ids"] == res_single["ids"]).all() and (r["ndist"] == res_single["ndist"]).all())
        forms[name]["ids_equal_single"] = same

    gt = ds["gt_ids"][:nq]
    out = {"workload": f"{args.config}: {cfg['n']:,} x {cfg['dim']} fp32, {cfg['comps']} components, "
                       f"R<={cfg['R']}, {nq:,} queries, L log-uniform {cfg['L_range']}, K in {cfg['K_choices']}",
           "M": M, "recall_at_K": round(recall_at_k(gt, res_single["ids"], Ks), 4),
           "mean_L": round(float(Ls.mean()), 1), "ndist_mean": round(float(res_single["ndist"].mean()), 1),
           "forms": forms,
           "speedup_autotuned_vs_single": round(forms["single"]["kernel_ms"] / forms["autotuned"]["kernel_ms"], 3),
           "speedup_autotuned_vs_per_bucket": round(forms["per_bucket"]["kernel_ms"] /
                                                    forms["autotuned"]["kernel_ms"], 3)}
    print(json.dumps(out), flush=True)
    idx.close()


if __name__ == "__main__":
    main()
