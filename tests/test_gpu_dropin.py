"""The C++ host drop-in (include/panns/Searching.h) end to end.

The find-L driver built from parallel_anns_amd/csrc/apps/find_L_top_m.cpp
(command line of the reference's
app/profile_find_L_para_single_query_top_m_search_pure_large_graph.cpp) reads
the base set, queries, graph and ground truth from files in the reference's
formats, searches through PANNS::Searching, and writes set_K lists in the
format of include/utils.h:146-163 save_result.  Its ids must equal the
oracle's and its printed P@100 the recall of those ids.
"""
import os
import re
import subprocess

import numpy as np
import pytest

import pyoracle as po
from parallel_anns_amd import build, tools

pytestmark = pytest.mark.gpu

DRIVER = os.path.join(build.BIN, "find_L_top_m")


def _write_inputs(d, ds, nq):
    p = {k: os.path.join(d, f) for k, f in (("base", "base.fvecs"), ("query", "query.fvecs"),
                                            ("graph", "graph.nsg"), ("truth", "truth.bin"))}
    tools.fvecs_write(p["base"], ds["base"])
    tools.fvecs_write(p["query"], ds["queries"][:nq])
    tools.nsg_write(p["graph"], ds["adj"], ds["deg"], ds["ep"])
    # the reference's load_true_NN wants at least 100 neighbours per query
    gt_ids, gt_d = tools.exact_knn(ds["base"], ds["queries"][:nq], 100)
    tools.truth_write(p["truth"], gt_ids, gt_d)
    return p, gt_ids


def _read_result(path):
    raw = np.fromfile(path, np.uint32)
    rows, i = [], 0
    while i < raw.size:
        k = int(raw[i])
        rows.append(raw[i + 1:i + 1 + k])
        i += 1 + k
    return np.stack(rows)


@pytest.mark.parametrize("how,M", [("per_query", 4), ("per_query", 1), ("batch", 8), ("pss", 1), ("opt", 4),
                                   ("opt_pss", 1)])
def test_find_L_driver(tiny_set, tmp_path, how, M):
    assert os.path.exists(DRIVER), "build the drop-in driver first (python -m parallel_anns_amd.build)"
    ds = tiny_set
    nq, K, L = 48, 100, 128
    print("case drop-in driver", how, "M", M, flush=True)
    p, gt_ids = _write_inputs(str(tmp_path), ds, nq)
    res_path = os.path.join(str(tmp_path), "result.bin")
    env = dict(os.environ)
    env["PANNS_BATCH"] = "1" if how == "batch" else "0"
    env["PANNS_PSS"] = "1" if how in ("pss", "opt_pss") else "0"
    # the opt-graph family: load_nsg_graph, norm - 2 v.q distances
    env["PANNS_OPT_GRAPH"] = "1" if how.startswith("opt") else "0"
    # L_lower = L_upper = L: exactly one search at L
    cmd = [DRIVER, p["base"], p["query"], p["graph"], str(L), str(K), res_path, str(M), p["truth"], "1",
           str(L), "0.5"]
    r = subprocess.run(cmd, capture_output=True, text=True, timeout=120, env=env)
    assert r.returncode == 0, r.stderr
    got = _read_result(res_path)
    assert got.shape == (nq, K)
    g = po.CsrGraph(ds["adj"], ds["deg"], ds["ep"], base=ds["base"] if how.startswith("opt") else None)
    init = po.prepare_init_ids(g, L)
    for q in range(nq):
        ref = po.search_one(g, ds["base"], ds["queries"][q], K, L, M, init)
        assert (got[q] == ref["ids"]).all(), f"query {q}"
    final = [ln for ln in r.stdout.splitlines() if ln.startswith("P_dest")][0]
    p100 = float(re.search(r"P@100: ([0-9.]+)", final).group(1))
    assert abs(p100 - tools.recall(gt_ids, got, L)) < 2e-6


def test_driver_rejects_threads(tiny_set, tmp_path):
    """num_threads_ > 1 is refused with the reference's error style."""
    ds = tiny_set
    p, _ = _write_inputs(str(tmp_path), ds, 8)
    cmd = [DRIVER, p["base"], p["query"], p["graph"], "64", "50", os.path.join(str(tmp_path), "r.bin"), "2",
           p["truth"], "4", "64", "0.5"]
    r = subprocess.run(cmd, capture_output=True, text=True, timeout=120)
    assert r.returncode != 0
    assert "Error:" in r.stderr and "num_threads_" in r.stderr


def test_driver_rejects_mismatched_graph(tiny_set, tmp_path):
    """A graph whose vertex count differs from the base set is an error."""
    ds = tiny_set
    p, _ = _write_inputs(str(tmp_path), ds, 8)
    n = 1000
    adj = np.minimum(ds["adj"][:n], n - 1).astype(np.uint32)
    tools.nsg_write(p["graph"], adj, ds["deg"][:n], 0)
    cmd = [DRIVER, p["base"], p["query"], p["graph"], "64", "50", os.path.join(str(tmp_path), "r.bin"), "2",
           p["truth"], "1", "64", "0.5"]
    r = subprocess.run(cmd, capture_output=True, text=True, timeout=120)
    assert r.returncode != 0 and "Error:" in r.stderr


def test_find_L_bisection_against_a_full_scan(small_set, tmp_path):
    """The driver's find-L (app/profile_find_L_para_single_query_top_m_search_
    pure_large_graph.cpp:210-260) over L in [100, 400] on the 20k set: a full
    oracle scan gives recall@100 at every L (a staircase: plateaus between
    the L where new components' init ids arrive); for each target the
    driver reports the L and P@100 that the reference's bisection reaches on
    that curve -- the smallest L reaching the target where the curve never
    decreases, an exact hit where the target is a plateau value, and the
    upper end with its recall where nothing reaches it."""
    from parallel_anns_amd import operating_point
    ds = small_set
    nq, K, M, lo, hi = ds["queries"].shape[0], 100, 4, 100, 400
    p, gt = _write_inputs(str(tmp_path), ds, nq)
    g = po.CsrGraph(ds["adj"], ds["deg"], ds["ep"])
    print("case find-L full scan", lo, hi, flush=True)
    init = po.prepare_init_ids(g, hi)
    rec = {}
    for L in range(lo, hi + 1):
        ids, _, _, _ = po.search_batch(g, ds["base"], ds["queries"], K, L, M, init[:L], nthreads=16)
        # the reference's count: hits / (100 * queries), in double
        hits = sum(len(np.intersect1d(gt[q, :100], ids[q, :100])) for q in range(nq))
        rec[L] = hits / (100.0 * nq)
    vals = [rec[L] for L in range(lo, hi + 1)]
    plateaus = sum(1 for a, b in zip(vals, vals[1:]) if a == b)
    assert plateaus > (hi - lo) // 2  # mostly flat steps: a staircase
    monotone = all(b >= a for a, b in zip(vals, vals[1:]))
    mid = vals[(hi - lo) // 2]
    targets = [0.5 * (vals[0] + vals[-1]), mid, max(vals) - 1e-4, vals[-1] + 0.01]
    res_path = os.path.join(str(tmp_path), "result.bin")
    env = dict(os.environ, PANNS_BATCH="1", PANNS_PSS="0", PANNS_OPT_GRAPH="0")
    cmd = [DRIVER, p["base"], p["query"], p["graph"], str(lo), str(K), res_path, str(M), p["truth"], "1",
           str(hi)] + [repr(t) for t in targets]
    r = subprocess.run(cmd, capture_output=True, text=True, timeout=300, env=env)
    assert r.returncode == 0, r.stderr
    finals = [ln for ln in r.stdout.splitlines() if ln.startswith("P_dest")]
    assert len(finals) == len(targets)
    for t, line in zip(targets, finals):
        L_drv = int(re.search(r" L: ([0-9]+)", line).group(1))
        p_drv = float(re.search(r"P@100: ([0-9.]+)", line).group(1))
        L_exp, r_exp, _ = operating_point.find_L(rec.__getitem__, lo, hi, t)
        assert L_drv == L_exp, (t, L_drv, L_exp)
        assert abs(p_drv - rec[L_drv]) < 2e-6
        reach = [L for L in range(lo, hi + 1) if rec[L] >= t]
        if not reach:
            assert L_drv == hi
        elif monotone and r_exp != t:
            assert L_drv == reach[0]
        else:
            assert rec[L_drv] >= t


@pytest.mark.parametrize("T,X", [(2, 1), (4, 4), (8, 2)])
def test_driver_pss_v5_worker_queues(tiny_set, tmp_path, T, X):
    """para_search_PSS_v5_large_graph_dist_thresh_profiling through the
    drop-in with num_threads_ = T queues and local_queue_capacity = L (the
    reference driver's setting): ids equal the oracle's lockstep PSS_v5."""
    ds = tiny_set
    nq, K, L = 32, 50, 96
    print("case drop-in pss workers T", T, "X", X, flush=True)
    p, _ = _write_inputs(str(tmp_path), ds, nq)
    res_path = os.path.join(str(tmp_path), "result.bin")
    env = dict(os.environ, PANNS_BATCH="0", PANNS_PSS="1", PANNS_OPT_GRAPH="0", PANNS_PSS_X=str(X))
    cmd = [DRIVER, p["base"], p["query"], p["graph"], str(L), str(K), res_path, "1", p["truth"], str(T),
           str(L), "0.5"]
    r = subprocess.run(cmd, capture_output=True, text=True, timeout=120, env=env)
    assert r.returncode == 0, r.stderr
    got = _read_result(res_path)
    g = po.CsrGraph(ds["adj"], ds["deg"], ds["ep"])
    init = po.prepare_init_ids(g, L)
    for q in range(nq):
        ref = po.search_pss_v5(g, ds["base"], ds["queries"][q], K, L, X, T, init)
        assert (got[q] == ref["ids"]).all(), f"query {q}"
