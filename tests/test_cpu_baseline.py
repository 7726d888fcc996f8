"""The CPU baseline's two modes (bench.py, run in a child process):

- the reference design, PSS_v5 with T OpenMP threads on one query
  (po_search_pss_v5_par): at T = 1 it is plain best-first search, equal in
  ids and distance counts to top-M with M = 1 for every X (what the
  reference's own T = 1 runs show); at T > 1 the schedule decides ties of
  the shared visited set, so it is checked by properties: K distinct ids
  per query, at least L distance computations, recall within 0.02 of the
  lockstep oracle's;
- the child end to end on a small cached set: its throughput-mode ids are
  the oracle's top-M ids, the X sweep is L/T .. L/T + 64 in steps of 32,
  and the host report holds the core count, CPU model and compiler.
"""
import numpy as np
import pytest

import pyoracle as po
from parallel_anns_amd import datasets, tools


@pytest.fixture(scope="module")
def t5k(tmp_path_factory):
    return datasets.prepare("t5k", cache_dir=str(tmp_path_factory.mktemp("cache")), verbose=False)


@pytest.mark.parametrize("L", [50, 120])
def test_pss_par_single_thread_is_top_m_1(t5k, L):
    g = po.CsrGraph(t5k["adj"], t5k["deg"], t5k["ep"])
    q = t5k["queries"][:60]
    init = po.prepare_init_ids(g, L)
    ids, _, nd, _ = po.search_batch(g, t5k["base"], q, 20, L, 1, init, nthreads=2)
    for X in (1, 8, 50):
        pid, pnd = po.search_pss_v5_par(g, t5k["base"], q, 20, L, X, 1)
        assert (pid == ids).all() and (pnd == nd).all(), X


@pytest.mark.parametrize("T,X", [(2, 4), (4, 8), (4, 40)])
def test_pss_par_threads(t5k, T, X):
    g = po.CsrGraph(t5k["adj"], t5k["deg"], t5k["ep"])
    K, L = 20, 100
    q = t5k["queries"][:80]
    init = po.prepare_init_ids(g, L)
    pid, pnd = po.search_pss_v5_par(g, t5k["base"], q, K, L, X, T)
    assert all(np.unique(r).shape[0] == K for r in pid)
    assert (pnd >= L).all()
    lock = np.stack([po.search_pss_v5(g, t5k["base"], x, K, L, X, T, init)["ids"] for x in q])
    gt = t5k["gt_ids"][:80]
    assert abs(tools.recall(gt, pid, K) - tools.recall(gt, lock, K)) <= 0.02


def test_cpu_child_end_to_end(t5k):
    import bench
    child = bench.start_cpu_child(2)
    out, ids = bench.cpu_baseline(child, 2, t5k, 20, 64, 2, budget_s=1.0)
    g = po.CsrGraph(t5k["adj"], t5k["deg"], t5k["ep"])
    want, _, _, _ = po.search_batch(g, t5k["base"], t5k["queries"][:ids.shape[0]], 20, 64, 2,
                                    po.prepare_init_ids(g, 64), nthreads=1)
    assert (ids == want).all()
    assert out["cores"] == 2 and out["unit"] == "QPS" and out["value"] > 0 and out["kind"] == "port"
    assert [r["X"] for r in out["reference_design"]["rows"]] == [32, 64, 96]
    h = out["host"]
    assert h["omp_proc_bind"] == "close" and h["omp_places"] == "cores"
    assert h["nproc"] >= 1 and "lscpu_model_name" in h and h["gcc"].startswith("gcc")
    assert "gcc" in out["compiler"]
