"""GPU parity of PSS_v5 with T > 1 queues (search_kernel_pss) against the
oracle's lockstep restatement (po_search_pss_v5, PSS5LG.cpp:1376-1724).

Bar: result ids, the master queue's ids, distance bits and checked flags,
distance computations, expansions and merge rounds equal the oracle's for
every query."""
import numpy as np
import pytest

import pyoracle as po
from parallel_anns_amd import tools
from parallel_anns_amd.search import Index, SearchError

pytestmark = pytest.mark.gpu


def _case(**kw):
    print("case", " ".join(f"{k}={v}" for k, v in kw.items()), flush=True)


def _check(res, refs, qids):
    for i, q in enumerate(qids):
        r = refs[i]
        assert (res["ids"][i] == r["ids"]).all(), f"ids differ for query {q}"
        assert (res["queue_ids"][i] == r["queue"]["id"]).all(), f"queue ids differ for query {q}"
        assert (res["queue_dists"][i].view(np.uint32) == r["queue"]["dist"].view(np.uint32)).all(), q
        assert (res["queue_checked"][i] == r["queue"]["checked"]).all(), f"checked flags differ for query {q}"
        assert res["ndist"][i] == r["ndist"], f"distance count differs for query {q}"
        assert res["nexp"][i] == r["nexp"], f"expansions differ for query {q}"
        assert res["niter"][i] == r["merges"], f"merge rounds differ for query {q}"
    assert res["counters"]["merges"] == sum(r["merges"] for r in refs)
    assert res["counters"]["distance_computations"] == sum(r["ndist"] for r in refs)


@pytest.mark.parametrize("T,X,L", [(2, 1, 100), (2, 4, 100), (3, 2, 64), (4, 4, 100), (4, 8, 200), (8, 2, 100),
                                   (8, 8, 150), (16, 4, 100), (5, 3, 300)])
def test_pss_v5_workers_bit_exact(small_set, gpu_index, T, X, L):
    ds = small_set
    K = min(100, L)
    qids = np.arange(48, dtype=np.uint32)
    _case(test="pss_workers", T=T, X=X, L=L)
    init = gpu_index.prepare_init_ids(L)
    res = gpu_index.search(qids, K, L, X, mode="pss_v5", num_threads=T, init_ids=init, want_queue=True)
    g = po.CsrGraph(ds["adj"], ds["deg"], ds["ep"])
    refs = [po.search_pss_v5(g, ds["base"], ds["queries"][q], K, L, X, T, init) for q in qids]
    _check(res, refs, qids)


@pytest.mark.parametrize("bs", [8, 4, 2])
def test_pss_v5_workers_with_spills(small_set, gpu_index, bs):
    """A 64-slot visited table (buckets of 8, 4 or 2 slots) sends most claims
    to the global bitmap, whose ordered atomics must keep the worker order
    of ownership."""
    ds = small_set
    K, L, T, X = 100, 120, 4, 4
    qids = np.arange(24, dtype=np.uint32)
    _case(test="pss_workers_spills", T=T, X=X, L=L, bucket_slots=bs)
    init = gpu_index.prepare_init_ids(L)
    res = gpu_index.search(qids, K, L, X, mode="pss_v5", num_threads=T, init_ids=init, want_queue=True,
                           hash_slots=64, bucket_slots=bs)
    g = po.CsrGraph(ds["adj"], ds["deg"], ds["ep"])
    refs = [po.search_pss_v5(g, ds["base"], ds["queries"][q], K, L, X, T, init) for q in qids]
    _check(res, refs, qids)
    assert res["counters"]["visited_spills"] > 0


def test_pss_v5_workers_norm_distance(tiny_set):
    """The opt-graph distance (norm - 2 v.q, PSS5.cpp:1211-1462 with its
    worker queues): negative keys through the deal, steps and merges."""
    ds = tiny_set
    idx = Index(0)
    idx.set_base(ds["base"])
    idx.set_queries(ds["queries"])
    idx.set_graph(ds["adj"], ds["deg"], ds["ep"])
    idx.set_distance("norm_dot")
    K, L, T, X = 50, 80, 4, 2
    qids = np.arange(24, dtype=np.uint32)
    _case(test="pss_workers_norm", T=T, X=X, L=L)
    init = idx.prepare_init_ids(L)
    res = idx.search(qids, K, L, X, mode="pss_v5", num_threads=T, init_ids=init, want_queue=True)
    g = po.CsrGraph(ds["adj"], ds["deg"], ds["ep"], base=ds["base"])
    refs = [po.search_pss_v5(g, ds["base"], ds["queries"][q], K, L, X, T, init) for q in qids]
    _check(res, refs, qids)
    idx.close()


def test_pss_v5_workers_uneven_buckets(small_set, gpu_index):
    """Per-query L in queue-size buckets replayed as a graph: each query
    equals the oracle at its own L."""
    ds = small_set
    rng = np.random.default_rng(3)
    nq = 40
    qids = np.arange(nq, dtype=np.uint32)
    Ls = rng.integers(60, 260, nq).astype(np.uint32)
    Ks = np.minimum(rng.choice(np.array([10, 50], np.uint32), nq), Ls)
    T, X = 4, 3
    _case(test="pss_workers_buckets", T=T, X=X)
    init = gpu_index.prepare_init_ids(int(Ls.max()))
    pl = gpu_index.plan(qids, 50, int(Ls.max()), X, mode="pss_v5", num_threads=T, init_ids=init, L_per_query=Ls,
                        K_per_query=Ks, options=Index.plan_options(buckets=2, use_graph=True))
    pl.run()
    pl.run()
    res = pl.fetch(want_queue=True)
    pl.close()
    g = po.CsrGraph(ds["adj"], ds["deg"], ds["ep"])
    for i in range(nq):
        L, K = int(Ls[i]), int(Ks[i])
        r = po.search_pss_v5(g, ds["base"], ds["queries"][i], K, L, X, T, init[:L])
        assert (res["ids"][i, :K] == r["ids"]).all(), i
        assert (res["queue_ids"][i, :L] == r["queue"]["id"]).all(), i
        assert res["ndist"][i] == r["ndist"] and res["niter"][i] == r["merges"], i


def test_pss_v5_workers_recall_against_truth(small_set, gpu_index):
    """Recall of the T-queue search over the whole query set, against the
    exact ground truth, is at least the one-queue search's at the same L."""
    ds = small_set
    q = np.arange(ds["queries"].shape[0], dtype=np.uint32)
    L = 120
    r1 = tools.recall(ds["gt_ids"], gpu_index.search(q, 100, L, 4, mode="pss_v5")["ids"], L)
    for T in (4, 8):
        rt = tools.recall(ds["gt_ids"], gpu_index.search(q, 100, L, 4, mode="pss_v5", num_threads=T)["ids"], L)
        assert rt >= r1 - 0.01, (T, rt, r1)


def test_pss_v5_workers_refusals(small_set, gpu_index):
    q = np.arange(2, dtype=np.uint32)
    init = gpu_index.prepare_init_ids(100)
    for kw in (dict(num_threads=17), dict(num_threads=0), dict(num_threads=4, waves=2),
               dict(num_threads=4, stop="learned", stop_threshold=0.5)):
        with pytest.raises(SearchError):
            gpu_index.search(q, 10, 100, 2, mode="pss_v5", init_ids=init, **kw)
    with pytest.raises(SearchError):
        gpu_index.search(q, 10, 100, 2, mode="top_m", num_threads=4, init_ids=init)
