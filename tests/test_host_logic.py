"""Host-side logic around the search, checked on the CPU: the learned-stop
training labels and schedule, the per-query stop bound, the threshold
bisection, the uneven-batch draws, and the benchmark's traffic lookup and
recall-deficit summary."""
import importlib.util
import itertools
import json
import os

import numpy as np
import pytest

from parallel_anns_amd import datasets, learned

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _bench():
    spec = importlib.util.spec_from_file_location("bench_mod", os.path.join(ROOT, "bench.py"))
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


def test_samples_from_trace_labels_the_final_top_k():
    # query 0: hashes 5, 7, 7, 7 over 4 iterations; query 1: 3, 9 over 2 of 4 slots
    feat = np.arange(2 * 4 * 8, dtype=np.float32).reshape(2, 4, 8)
    hsh = np.array([[5, 7, 7, 7], [3, 9, 0, 0]], np.uint32)
    niter = np.array([4, 2], np.uint32)
    X, Y = learned.samples_from_trace(feat, hsh, niter)
    assert X.shape == (6, 8)
    assert list(Y) == [0, 1, 1, 1, 0, 1]
    assert (X[4] == feat[1, 0]).all()


def test_first_final_iteration_quantile():
    # final run starts at iteration 2, 3, 5 and 1 (1-based)
    hsh = np.array([[1, 2, 2, 2, 2], [1, 1, 3, 3, 3], [4, 4, 4, 4, 6], [8, 8, 8, 8, 8]], np.uint32)
    niter = np.array([5, 5, 5, 5], np.uint32)
    assert learned.first_final_iteration(hsh, niter, 0.0) == 1
    assert learned.first_final_iteration(hsh, niter, 1.0) == 5
    assert learned.first_final_iteration(hsh, niter, 0.5) == 2  # quantile of [2, 3, 5, 1], floored


def test_oracle_bound_equals_brute_force_on_tiny_grids():
    """The Lagrangian bound against an exhaustive search over every choice of
    one column per query: its choice is feasible (so it costs at least the
    optimum) and the dual value at its price is at most the optimum."""
    rng = np.random.default_rng(0)
    for trial in range(30):
        nq, nc = 4, 3
        N = np.sort(rng.uniform(100, 1000, (nq, nc)), axis=1)
        R = np.sort(rng.uniform(0, 1, (nq, nc)), axis=1)
        target = float(rng.uniform(R[:, 0].mean(), R[:, -1].mean()))
        best = None
        for choice in itertools.product(range(nc), repeat=nq):
            rec = R[np.arange(nq), choice].mean()
            nd = N[np.arange(nq), choice].mean()
            if rec >= target - 1e-12 and (best is None or nd < best):
                best = nd
        b = learned.oracle_bound(R, N, target)
        assert b is not None and b[1] >= target - 1e-9
        # feasible, so no better than the optimum; and the Lagrangian dual at
        # its price bounds the optimum from below
        lam = b[2]
        dual = (N - lam * R).min(axis=1).mean() + lam * target
        assert dual - 1e-6 <= best <= b[0] + 1e-9
    assert learned.oracle_bound(np.zeros((3, 2)), np.ones((3, 2)), 0.5) is None


def test_calibrate_bisects_a_monotone_recall(monkeypatch):
    """calibrate finds the smallest threshold (log scale) whose recall meets
    the target, given a recall that rises with the threshold."""
    def fake(index, qids, gt, K, L, M, threshold, runs=1, L_per_query=None, options=None, **kw):
        return min(1.0, 0.8 + 0.2 * np.log10(threshold / 1e-4) / 4.0), 0.0, 0.0
    monkeypatch.setattr(learned, "recall_work", fake)
    thr = learned.calibrate(None, None, None, 100, 100, 4, 0.9)
    assert abs(fake(None, None, None, 0, 0, 0, thr)[0] - 0.9) < 1e-3
    assert fake(None, None, None, 0, 0, 0, thr * 0.98)[0] < 0.9
    assert learned.calibrate(None, None, None, 100, 100, 4, 1.01) is None


def test_calibrate_refuses_a_threshold_of_one(monkeypatch):
    """The kernel's sigmoid can round to exactly 1.0, so a threshold of 1.0
    may still stop queries: when only the upper end reaches the target there
    is no usable threshold, and every caller's fallback is the one sentinel."""
    def fake(index, qids, gt, K, L, M, threshold, runs=1, L_per_query=None, options=None, **kw):
        return (0.95 if threshold >= 1.0 else 0.9), 0.0, 0.0
    monkeypatch.setattr(learned, "recall_work", fake)
    assert learned.calibrate(None, None, None, 100, 100, 4, 0.93) is None
    assert learned.is_fallback(None) and learned.is_fallback(learned.FIXED_FALLBACK)
    assert not learned.is_fallback(1.0) and not learned.is_fallback(0.4)


def test_uneven_draws_stay_in_range_and_are_seeded():
    cfg = datasets.CONFIGS["c2"]
    L, K = datasets.uneven_L_K(cfg, 20000)
    assert L.min() >= 100 and L.max() <= 2000 and set(np.unique(K)) <= {10, 100}
    assert (K <= L).all()
    L2, K2 = datasets.uneven_L_K(cfg, 20000)
    assert (L == L2).all() and (K == K2).all()
    # log-uniform: about half the draws below the geometric mean
    assert abs((L < np.sqrt(100 * 2000)).mean() - 0.5) < 0.02


def test_measured_traffic_lookup(tmp_path, monkeypatch):
    b = _bench()
    f = tmp_path / "t.json"
    f.write_text(json.dumps({"records": [
        {"config": "c1", "kernel": "search_kernel<96, true>", "M": 4, "L": 289, "traffic_bytes": 900,
         "algorithmic_bytes": 2000, "threshold": 0.5, "command": "cmd"},
        {"config": "c1", "kernel": "search_kernel<96, false>", "M": 4, "L": 289, "traffic_bytes": 1000,
         "algorithmic_bytes": 2000, "command": "cmd"},
        # a record of an uneven launch (buckets, no single L) is skipped
        {"config": "c1", "kernel": "search_kernel<96, false>", "M": 4, "uneven": True, "buckets": [417, 1999],
         "traffic_bytes": 5, "algorithmic_bytes": 7, "command": "uneven"}]}))
    monkeypatch.setattr(b, "TRAFFIC_FILE", str(f))
    t, src = b.measured_traffic("c1", "search_kernel<96, true>", {"M": 4, "L_max": 289, "threshold": 0.5})
    assert t == 900 and "cmd" in src
    # another threshold of the same kernel: scaled by traffic per algorithmic byte
    t, src = b.measured_traffic("c1", "search_kernel<96, true>",
                                {"M": 4, "L_max": 289, "threshold": 0.4, "algorithmic_bytes": 3000})
    assert t == 1350 and src.startswith("scaled")
    assert b.measured_traffic("c1", "search_kernel<96, false>", {"M": 8, "L": 289}) == (None, None)
    t, src = b.measured_traffic("c1", "search_kernel<96, false>", {"M": 4, "L": 289})
    assert t == 1000 and "uneven" not in src


def test_deficit_distribution():
    b = _bench()
    d = b.deficit_distribution([0.0, 0.0, 0.01, 0.03, 0.2])
    assert d["frac_queries_with_loss"] == 0.6
    assert d["histogram"]["queries"] == [2, 0, 1, 1, 0, 0, 1]
    assert d["max"] == 0.2
