"""Synthetic workloads of the benchmark configurations and their preparation.

Every workload is a seeded, L2-normalised 96-d Gaussian mixture (sigma 0.3);
seeds: centres 0, base 1, queries 2, training queries 3.  The graph is built
from exact k-NN lists (ties to the lower id) by tools.build_graph; the
ground truth is exact brute force.  Prepared inputs can be cached on disk in
the reference's own formats (fvecs, .nsg, truth).
"""
import hashlib
import json
import os
import sys
import time

import numpy as np

from . import tools

SEED_CENTERS, SEED_BASE, SEED_QUERY, SEED_TRAIN = 0, 1, 2, 3

CONFIGS = {
    # C0: 100k base, 1k queries, exact 64-NN -> R <= 32
    "c0": dict(n=100_000, nq=1_000, dim=96, comps=64, sigma=0.3, knn=64, R=32, K=100, L=100),
    # C1: 1M base, 10k queries, exact 100-NN -> R <= 50
    "c1": dict(n=1_000_000, nq=10_000, dim=96, comps=64, sigma=0.3, knn=100, R=50, K=100, L=100),
    # C2: 10M base from 256 components, 100k queries of uneven L and K; the
    # exact 64-NN lists are computed per mixture component (checked against
    # global brute force on a sample of rows)
    "c2": dict(n=10_000_000, nq=100_000, dim=96, comps=256, sigma=0.3, knn=64, R=32, K=100, L=100,
               knn_by_component=True, L_range=(100, 2000), K_choices=(10, 100)),
    # C4: 50M base (19.2 GB) from 1,024 components, 100k queries with L
    # log-uniform in [100, 2000] and K = 100; exact 64-NN per component
    "c4": dict(n=50_000_000, nq=100_000, dim=96, comps=1024, sigma=0.3, knn=64, R=32, K=100, L=100,
               knn_by_component=True, L_range=(100, 2000), K_choices=(100,)),
    # C1's base set and queries with the NSG-style construction (search-path
    # candidate pool of 40 from the navigating node, occlusion over at most
    # C = 500 candidates, reverse edges) instead of the configured k-NN ->
    # RNG prune -> DFS stitch; a secondary workload, not a configuration
    "c1nsg": dict(n=1_000_000, nq=10_000, dim=96, comps=64, sigma=0.3, knn=100, R=50, K=100, L=100,
                  L_build=40, C=500, reverse=True),
    # a 1M-row rehearsal of C4's pipeline (per-component k-NN, uneven L)
    "c4s": dict(n=1_000_000, nq=10_000, dim=96, comps=64, sigma=0.3, knn=64, R=32, K=100, L=100,
                knn_by_component=True, L_range=(100, 2000), K_choices=(100,)),
    # small cases for tests
    "t20k": dict(n=20_000, nq=200, dim=96, comps=64, sigma=0.3, knn=64, R=32, K=100, L=100),
    "t5k": dict(n=5_000, nq=64, dim=96, comps=16, sigma=0.3, knn=32, R=24, K=50, L=64),
}


def log(msg):
    print(f"[datasets {time.strftime('%H:%M:%S')}] {msg}", file=sys.stderr, flush=True)


def generate(cfg, which="base", n=None, row0=0):
    seed = {"base": SEED_BASE, "query": SEED_QUERY, "train": SEED_TRAIN}[which]
    if n is None:
        n = cfg["n"] if which == "base" else cfg["nq"]
    return tools.gen_mixture(n, cfg["dim"], cfg["comps"], cfg["sigma"], SEED_CENTERS, seed, row0=row0)


def knn_graph(base, k, nthreads=0, gpu_knn=None):
    """Exact k-NN of every base row (self excluded).  gpu_knn, when given, is a
    callable(base, k) -> ids computing the same lists on the GPU."""
    if gpu_knn is not None:
        return gpu_knn(base, k)
    ids, _ = tools.exact_knn(base, base, k, exclude_self=True, nthreads=nthreads)
    return ids


def knn_by_component(cfg, base, k, gpu_knn, check_rows=2000, verbose=True):
    """Exact k-NN lists computed within each mixture component (ids stay in
    ascending order inside a component, so ties still go to the lower id).
    On the configured mixtures a row's k nearest rows all share its
    component; `check_rows` evenly spaced rows are compared with global
    brute force and the agreement is logged (and returned)."""
    n = base.shape[0]
    labels = tools.mixture_labels(n, cfg["comps"], SEED_BASE)
    order = np.argsort(labels, kind="stable")
    counts = np.bincount(labels, minlength=cfg["comps"])
    if counts.min() <= k:
        raise ValueError("a mixture component holds no more than k rows")
    ids = np.empty((n, k), np.uint32)
    start = 0
    for c in range(cfg["comps"]):
        members = order[start:start + counts[c]]
        start += counts[c]
        local = gpu_knn(np.ascontiguousarray(base[members]), k)
        ids[members] = members[local]
        if verbose and (c + 1) % 32 == 0:
            log(f"per-component {k}-NN: {c + 1}/{cfg['comps']} components")
    rows = np.linspace(0, n - 1, min(check_rows, n)).astype(np.int64)
    g_ids, _ = gpu_knn.truth(base, np.ascontiguousarray(base[rows]), k + 1)
    agree = 0
    for r, row in zip(rows, g_ids):
        row = row[row != r][:k]
        agree += int((row == ids[r]).all())
    frac = agree / rows.shape[0]
    if verbose:
        log(f"per-component {k}-NN equals global brute force on {agree}/{rows.shape[0]} sampled rows")
    return ids, frac


class GpuKnn:
    """Exact k-NN lists and ground truth on the GPU (search.Index.exact_knn)."""

    def __init__(self, device=0):
        self.device = device

    def __call__(self, base, k):
        from .search import Index
        idx = Index(self.device)
        try:
            idx.set_base(base)
            ids, _ = idx.exact_knn(k, source="base")
        finally:
            idx.close()
        return ids

    def truth(self, base, queries, k):
        from .search import Index
        idx = Index(self.device)
        try:
            idx.set_base(base)
            idx.set_queries(queries)
            return idx.exact_knn(k, source="queries")
        finally:
            idx.close()


def cfg_key(cfg):
    return hashlib.sha256(json.dumps(cfg, sort_keys=True).encode()).hexdigest()[:12]


def uneven_L_K(cfg, nq, seed=11):
    """Per-query L log-uniform in cfg["L_range"] and K drawn from
    cfg["K_choices"] (K <= L), seeded."""
    rng = np.random.default_rng(seed)
    lo, hi = cfg["L_range"]
    L = np.exp(rng.uniform(np.log(lo), np.log(hi), nq)).astype(np.uint32)
    L = np.clip(L, lo, hi)
    K = rng.choice(np.asarray(cfg["K_choices"], np.uint32), nq)
    return L, np.minimum(K, L).astype(np.uint32)


def prepare(name_or_cfg, cache_dir=None, nthreads=0, gpu_knn=None, truth_k=100, verbose=True):
    """Return dict(base, queries, adj, deg, ep, gt_ids, gt_dists, cfg)."""
    cfg = CONFIGS[name_or_cfg] if isinstance(name_or_cfg, str) else dict(name_or_cfg)
    d = None
    if cache_dir:
        d = os.path.join(cache_dir, f"panns_{cfg_key(cfg)}")
        os.makedirs(d, exist_ok=True)
        paths = {k: os.path.join(d, f) for k, f in
                 (("base", "base.fvecs"), ("query", "query.fvecs"), ("graph", "graph.nsg"),
                  ("truth", "truth.bin"))}
        if all(os.path.exists(p) for p in paths.values()):
            if verbose:
                log(f"cache hit {d}")
            base = tools.fvecs_read(paths["base"])
            queries = tools.fvecs_read(paths["query"])
            adj, deg, ep, _ = tools.nsg_read(paths["graph"])
            gt_ids, gt_d = tools.truth_read(paths["truth"])
            return dict(base=base, queries=queries, adj=adj, deg=deg, ep=ep, gt_ids=gt_ids,
                        gt_dists=gt_d, cfg=cfg, dir=d)
    t0 = time.time()
    base = generate(cfg, "base")
    queries = generate(cfg, "query")
    if verbose:
        log(f"generated {base.shape} + {queries.shape} in {time.time() - t0:.1f}s")
    t0 = time.time()
    if cfg.get("knn_by_component"):
        knn, _ = knn_by_component(cfg, base, cfg["knn"], gpu_knn, verbose=verbose)
    else:
        knn = knn_graph(base, cfg["knn"], nthreads, gpu_knn)
    if verbose:
        log(f"exact {cfg['knn']}-NN in {time.time() - t0:.1f}s")
    t0 = time.time()
    # default: the configured construction, exact k-NN -> RNG prune -> DFS
    # stitch; "L_build" > 0 / "reverse" add NSG's search-path pool and
    # reverse edges
    adj, deg, ep = tools.build_graph(base, knn, cfg["R"], L_build=cfg.get("L_build", 0),
                                     C=cfg.get("C", 500), reverse_edges=cfg.get("reverse", False),
                                     nthreads=nthreads)
    del knn
    if verbose:
        log(f"graph R<={cfg['R']} avg deg {deg.mean():.2f} max {deg.max()} ep {ep} in {time.time() - t0:.1f}s")
    t0 = time.time()
    if gpu_knn is not None and hasattr(gpu_knn, "truth"):
        gt_ids, gt_d = gpu_knn.truth(base, queries, truth_k)
    else:
        gt_ids, gt_d = tools.exact_knn(base, queries, truth_k, nthreads=nthreads)
    if verbose:
        log(f"ground truth {truth_k}-NN in {time.time() - t0:.1f}s")
    out = dict(base=base, queries=queries, adj=adj, deg=deg, ep=ep, gt_ids=gt_ids, gt_dists=gt_d,
               cfg=cfg, dir=d)
    if d:
        tools.fvecs_write(paths["base"], base)
        tools.fvecs_write(paths["query"], queries)
        tools.nsg_write(paths["graph"], adj, deg, ep)
        tools.truth_write(paths["truth"], gt_ids, gt_d)
    return out
