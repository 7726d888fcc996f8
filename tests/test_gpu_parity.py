"""GPU parity: the HIP search (through the C-ABI) against the CPU oracle.

Bar: result ids, queue ids and distances bit-exact; distance-computation
counts equal (the visited sets are identical).
"""
import numpy as np
import pytest

import pyoracle as po
from parallel_anns_amd import tools
from parallel_anns_amd.search import Index, SearchError

pytestmark = pytest.mark.gpu


def oracle_run(ds, qids, K, L, M, init_ids, mode="top_m", X=1):
    g = po.CsrGraph(ds["adj"], ds["deg"], ds["ep"])
    out = [po.search_one(g, ds["base"], ds["queries"][q], K, L, M, init_ids, mode=mode, X=X) for q in qids]
    return out


def _print_case(**kw):
    print("case", " ".join(f"{k}={v}" for k, v in kw.items()), flush=True)


@pytest.mark.parametrize("L,M", [(100, 1), (100, 2), (100, 4), (100, 8), (200, 1), (200, 8), (64, 3)])
def test_top_m_bit_exact(small_set, gpu_index, L, M):
    ds = small_set
    K = min(100, L)
    qids = np.arange(64, dtype=np.uint32)
    _print_case(test="top_m", L=L, M=M, K=K)
    init = gpu_index.prepare_init_ids(L)
    g = po.CsrGraph(ds["adj"], ds["deg"], ds["ep"])
    assert (init == po.prepare_init_ids(g, L)).all()
    res = gpu_index.search(qids, K, L, M, init_ids=init, want_queue=True)
    ref = oracle_run(ds, qids, K, L, M, init)
    for i, q in enumerate(qids):
        r = ref[i]
        assert (res["ids"][i] == r["ids"]).all(), f"ids differ for query {q}"
        assert (res["queue_ids"][i] == r["queue"]["id"]).all(), f"queue ids differ for query {q}"
        assert (res["queue_dists"][i].view(np.uint32) == r["queue"]["dist"].view(np.uint32)).all()
        assert (res["queue_checked"][i] == r["queue"]["checked"]).all()
        assert res["ndist"][i] == r["ndist"], f"distance count differs for query {q}"
        assert res["nexp"][i] == r["nexp"]
        assert res["niter"][i] == r["merges"]
    assert res["counters"]["distance_computations"] == sum(r["ndist"] for r in ref)
    assert res["counters"]["merges"] == sum(r["merges"] for r in ref)


@pytest.mark.parametrize("X", [1, 4, 8])
def test_pss_v5_t1_bit_exact(small_set, gpu_index, X):
    ds = small_set
    K, L = 100, 100
    qids = np.arange(32, dtype=np.uint32)
    _print_case(test="pss_v5", X=X)
    init = gpu_index.prepare_init_ids(L)
    res = gpu_index.search(qids, K, L, X, mode="pss_v5", init_ids=init)
    ref = oracle_run(ds, qids, K, L, 1, init, mode="pss_v5", X=X)
    for i in range(len(qids)):
        assert (res["ids"][i] == ref[i]["ids"]).all()
        assert res["ndist"][i] == ref[i]["ndist"]
    assert res["counters"]["merges"] == sum(r["merges"] for r in ref)


def test_all_queries_recall_and_exact(small_set, gpu_index):
    ds = small_set
    K, L, M = 100, 100, 4
    qids = np.arange(ds["queries"].shape[0], dtype=np.uint32)
    _print_case(test="all_queries", L=L, M=M)
    init = gpu_index.prepare_init_ids(L)
    res = gpu_index.search(qids, K, L, M, init_ids=init)
    g = po.CsrGraph(ds["adj"], ds["deg"], ds["ep"])
    ids, dists, nd, total = po.search_batch(g, ds["base"], ds["queries"], K, L, M, init)
    assert (res["ids"] == ids).all()
    assert (res["dists"].view(np.uint32) == dists.view(np.uint32)).all()
    assert (res["ndist"] == nd).all()
    rec = tools.recall(ds["gt_ids"], res["ids"], L)
    assert rec > 0.7


@pytest.mark.parametrize("slots", [64, 192, 1536, 4864])
def test_small_tables_spill_and_keep_results(small_set, gpu_index, slots):
    """Small LDS tables (power of two or not, so the bucket index is a true
    remainder) force visited ids into second buckets and the global bitmap;
    the results stay those of the oracle."""
    ds = small_set
    K, L, M = 100, 100, 2
    qids = np.arange(48, dtype=np.uint32)
    _print_case(test="spill", hash_slots=slots)
    init = gpu_index.prepare_init_ids(L)
    res = gpu_index.search(qids, K, L, M, init_ids=init, hash_slots=slots)
    if slots < 1536:
        assert res["counters"]["visited_spills"] > 0
    ref = oracle_run(ds, qids, K, L, M, init)
    for i in range(qids.shape[0]):
        assert (res["ids"][i] == ref[i]["ids"]).all(), f"query {i}"
        assert res["ndist"][i] == ref[i]["ndist"]
    # a second run on the same scratch must start from clean bitmaps
    again = gpu_index.search(qids, K, L, M, init_ids=init, hash_slots=slots)
    assert (again["ids"] == res["ids"]).all() and (again["ndist"] == res["ndist"]).all()


@pytest.mark.parametrize("bs", [4, 2])
@pytest.mark.parametrize("slots", [64, 192, 1536, 8192])
def test_narrow_buckets_keep_results(small_set, gpu_index, slots, bs):
    """Visited tables of 4-slot buckets (the automatic choice for id spaces
    above 2^25 at the default size) and of 2-slot buckets, forced here:
    second buckets and spills, results those of the oracle."""
    ds = small_set
    K, L, M = 100, 160, 4
    qids = np.arange(48, dtype=np.uint32)
    _print_case(test="narrow_buckets", hash_slots=slots, bucket_slots=bs)
    init = gpu_index.prepare_init_ids(L)
    res = gpu_index.search(qids, K, L, M, init_ids=init, hash_slots=slots, bucket_slots=bs, want_queue=True)
    if slots < 1536:
        assert res["counters"]["visited_spills"] > 0
    ref = oracle_run(ds, qids, K, L, M, init)
    for i in range(qids.shape[0]):
        assert (res["ids"][i] == ref[i]["ids"]).all(), f"query {i}"
        assert (res["queue_ids"][i] == ref[i]["queue"]["id"]).all()
        assert res["ndist"][i] == ref[i]["ndist"]
    for w in (2, 4):
        mw = gpu_index.search(qids, K, L, M, init_ids=init, hash_slots=slots, bucket_slots=bs, waves=w)
        assert (mw["ids"] == res["ids"]).all() and (mw["ndist"] == res["ndist"]).all()
    with pytest.raises(SearchError):
        gpu_index.search(qids[:2], K, L, M, init_ids=init, bucket_slots=6)


@pytest.mark.parametrize("slots,ovf,bs", [(64, 256, 8), (192, 64, 8), (1536, 1024, 8), (128, 512, 4),
                                         (128, 256, 2)])
def test_overflow_area_keeps_results(small_set, gpu_index, slots, ovf, bs):
    """Ids whose buckets are full go to the LDS overflow entries, then to the
    bitmap; the results stay the oracle's and fewer ids reach the bitmap."""
    ds = small_set
    K, L, M = 100, 160, 4
    qids = np.arange(48, dtype=np.uint32)
    _print_case(test="overflow", hash_slots=slots, overflow=ovf, bucket_slots=bs)
    init = gpu_index.prepare_init_ids(L)
    base = gpu_index.search(qids, K, L, M, init_ids=init, hash_slots=slots, bucket_slots=bs)
    pl = gpu_index.plan(qids, K, L, M, init_ids=init, hash_slots=slots, bucket_slots=bs, want_queue=True,
                        options=Index.plan_options(overflow_slots=ovf))
    pl.run()
    res = pl.fetch(want_queue=True)
    pl.close()
    sp0, sp1 = base["counters"]["visited_spills"], res["counters"]["visited_spills"]
    assert sp1 < sp0 or sp0 == sp1 == 0
    ref = oracle_run(ds, qids, K, L, M, init)
    for i in range(qids.shape[0]):
        assert (res["ids"][i] == ref[i]["ids"]).all(), f"query {i}"
        assert (res["queue_ids"][i] == ref[i]["queue"]["id"]).all()
        assert res["ndist"][i] == ref[i]["ndist"]
    for w in (2, 4):
        pm = gpu_index.plan(qids, K, L, M, init_ids=init, hash_slots=slots, bucket_slots=bs, want_queue=False,
                            options=Index.plan_options(overflow_slots=ovf, waves=w))
        pm.run()
        r2 = pm.fetch()
        pm.close()
        assert (r2["ids"] == res["ids"]).all() and (r2["ndist"] == res["ndist"]).all()
    with pytest.raises(SearchError):
        gpu_index.plan(qids[:2], K, L, M, init_ids=init, options=Index.plan_options(overflow_slots=100))


def test_uneven_plan_matches_uniform(small_set, gpu_index):
    ds = small_set
    rng = np.random.default_rng(5)
    nq = 96
    qids = np.arange(nq, dtype=np.uint32)
    Ls = rng.choice([64, 100, 150, 220], size=nq).astype(np.uint32)
    Ks = np.minimum(rng.choice([10, 50, 100], size=nq), Ls).astype(np.uint32)
    _print_case(test="uneven", nq=nq)
    maxL = int(Ls.max())
    init = gpu_index.prepare_init_ids(maxL)
    plan = gpu_index.plan(qids, 100, maxL, 2, init_ids=init, L_per_query=Ls, K_per_query=Ks)
    plan.run()
    out = plan.fetch()
    plan.close()
    for L in np.unique(Ls):
        sel = np.nonzero(Ls == L)[0]
        r = gpu_index.search(qids[sel], 100 if L >= 100 else int(L), int(L), 2, init_ids=init[:L])
        for j, q in enumerate(sel):
            k = int(Ks[q])
            assert (out["ids"][q, :k] == r["ids"][j, :k]).all()
            assert (out["ids"][q, k:] == 0xFFFFFFFF).all()


def test_subsearch_with_top_m(small_set, gpu_index):
    ds = small_set
    L = 120
    g = po.CsrGraph(ds["adj"], ds["deg"], ds["ep"])
    init = po.prepare_init_ids(g, L)
    for q, mmax in ((0, 1), (1, 4), (2, 16), (3, 3)):
        _print_case(test="subsearch", q=q, M_max=mmax)
        set_L, visited, _ = po.initialize_set_L(ds["base"], ds["queries"][q], L, init, ds["base"].shape[0])
        ref_L, ref_vis = set_L.copy(), visited.copy()
        nd_ref = po.subsearch_with_top_m(g, ds["base"], ds["queries"][q], mmax, L, ref_L, ref_vis)
        qi = set_L["id"].copy()
        qd = set_L["dist"].copy()
        qc = set_L["checked"].copy()
        vis = visited.copy()
        nd = gpu_index.subsearch_with_top_m(mmax, q, L, qi, qd, qc, vis)
        assert nd == nd_ref
        assert (qi == ref_L["id"]).all()
        assert (qd.view(np.uint32) == ref_L["dist"].view(np.uint32)).all()
        assert (qc == ref_L["checked"]).all()
        assert (vis == ref_vis).all()


def test_edge_cases_tiny_graph():
    """L == N (the queue holds every vertex: exact), K = 1, M > L, duplicate init ids."""
    rng = np.random.default_rng(11)
    n, d = 200, 32
    base = tools.gen_mixture(n, d, 4, 0.3, 0, 7)
    queries = tools.gen_mixture(8, d, 4, 0.3, 0, 8)
    knn, _ = tools.exact_knn(base, base, 16, exclude_self=True)
    adj, deg, ep = tools.build_graph(base, knn, 12)
    idx = Index(0)
    idx.set_base(base)
    idx.set_queries(queries)
    idx.set_graph(adj, deg, ep)
    g = po.CsrGraph(adj, deg, ep)
    qids = np.arange(8, dtype=np.uint32)
    bf_ids, _ = po.brute_force_knn(base, queries, n)
    cases = [(n, n, 1), (n, 10, 64), (50, 1, 1), (50, 50, 7), (64, 10, 1000), (33, 5, 2)]
    for L, K, M in cases:
        _print_case(test="edge", L=L, K=K, M=M)
        init = idx.prepare_init_ids(L)
        res = idx.search(qids, K, L, M, init_ids=init, want_queue=True)
        for i, q in enumerate(qids):
            r = po.search_one(g, base, queries[q], K, L, M, init)
            assert (res["ids"][i] == r["ids"]).all()
            assert (res["queue_ids"][i] == r["queue"]["id"]).all()
            assert res["ndist"][i] == r["ndist"]
            if L == n:
                assert (res["queue_ids"][i] == bf_ids[q]).all()
    # duplicate init ids (a caller-built list): the reference keeps both copies
    init = np.concatenate([idx.prepare_init_ids(20), idx.prepare_init_ids(20)[:10]]).astype(np.uint32)
    _print_case(test="edge_dup_init")
    res = idx.search(qids, 10, 30, 2, init_ids=init, want_queue=True)
    for i, q in enumerate(qids):
        r = po.search_one(g, base, queries[q], 10, 30, 2, init)
        assert (res["queue_ids"][i] == r["queue"]["id"]).all()
        assert res["ndist"][i] == r["ndist"]
    idx.close()


def _fixtures():
    import json
    import os
    return json.load(open(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "fixtures.json")))


def test_golden_distance_bits_on_gpu():
    """The fixture rows v of one dimension form the base set and their partners
    q the queries; with L = N the final queue holds every row with the kernel's
    distance, and row i's distance to query i must carry the fixture's bits."""
    fx = _fixtures()
    for dim in sorted({r["dim"] for r in fx["distance"]}):
        if dim not in (16, 32, 40, 64, 96, 128):
            continue
        recs = [r for r in fx["distance"] if r["dim"] == dim]
        base = np.stack([np.asarray(r["v"], np.uint32).view(np.float32) for r in recs])
        queries = np.stack([np.asarray(r["q"], np.uint32).view(np.float32) for r in recs])
        n = base.shape[0]
        adj = np.array([[j for j in range(n) if j != i] for i in range(n)], np.uint32)
        deg = np.full(n, n - 1, np.uint32)
        idx = Index(0)
        idx.set_base(base)
        idx.set_queries(queries)
        idx.set_graph(adj, deg, 0)
        _print_case(test="golden_distance", dim=dim)
        res = idx.search(np.arange(n, dtype=np.uint32), 1, n, 1, want_queue=True)
        for i in range(n):
            pos = int(np.nonzero(res["queue_ids"][i] == i)[0][0])
            assert int(res["queue_dists"][i][pos].view(np.uint32)) == recs[i]["d_bits"], (dim, i)
        idx.close()


def test_golden_tiny_search_on_gpu():
    t = _fixtures()["tiny_search"]
    base = np.asarray(t["base"], np.uint32).view(np.float32).reshape(t["n"], t["dim"])
    queries = np.asarray(t["queries"], np.uint32).view(np.float32).reshape(-1, t["dim"])
    adj = np.asarray(t["adj"], np.uint32)
    deg = np.full(t["n"], adj.shape[1], np.uint32)
    idx = Index(0)
    idx.set_base(base)
    idx.set_queries(queries)
    idx.set_graph(adj, deg, t["ep"])
    init = idx.prepare_init_ids(t["L"])
    assert init.tolist() == t["init_ids"]
    for c in t["cases"]:
        _print_case(test="golden_tiny", M=c["M"], q=c["query"])
        r = idx.search(np.array([c["query"]], np.uint32), t["K"], t["L"], c["M"], init_ids=init, want_queue=True)
        assert r["queue_ids"][0].tolist() == c["queue_ids"]
        assert r["queue_dists"][0].view(np.uint32).tolist() == c["queue_dist_bits"]
        assert int(r["ndist"][0]) == c["ndist"]
    idx.close()


@pytest.mark.parametrize("dim", [16, 32, 40, 64, 128])
def test_other_dimensions(dim):
    n = 3000
    base = tools.gen_mixture(n, dim, 8, 0.3, 0, 21)
    queries = tools.gen_mixture(16, dim, 8, 0.3, 0, 22)
    knn, _ = tools.exact_knn(base, base, 24, exclude_self=True)
    adj, deg, ep = tools.build_graph(base, knn, 16)
    idx = Index(0)
    idx.set_base(base)
    idx.set_queries(queries)
    idx.set_graph(adj, deg, ep)
    g = po.CsrGraph(adj, deg, ep)
    _print_case(test="dims", dim=dim)
    init = idx.prepare_init_ids(80)
    res = idx.search(np.arange(16, dtype=np.uint32), 40, 80, 2, init_ids=init)
    for q in range(16):
        r = po.search_one(g, base, queries[q], 40, 80, 2, init)
        assert (res["ids"][q] == r["ids"]).all()
        assert res["ndist"][q] == r["ndist"]
    idx.close()


def test_error_paths(small_set, gpu_index):
    q = np.arange(2, dtype=np.uint32)
    init = gpu_index.prepare_init_ids(100)
    bad = [dict(K=101, L=100, M=1), dict(K=10, L=100, M=0), dict(K=10, L=30000, M=1),
           dict(K=10, L=100, M=1, mode="pss_v5", num_threads=17), dict(K=10, L=100, M=1, num_threads=4)]
    for kw in bad:
        with pytest.raises(SearchError):
            gpu_index.search(q, init_ids=init, **kw)
    with pytest.raises(SearchError):
        gpu_index.search(np.array([10**6], dtype=np.uint32), 10, 100, 1, init_ids=init)
    with pytest.raises(SearchError):
        gpu_index.search(q, 10, 100, 1, init_ids=init[:50])
    idx = Index(0)
    with pytest.raises(SearchError):
        idx.search(q, 10, 100, 1, init_ids=init)  # nothing loaded
    with pytest.raises(SearchError):
        idx.set_base(np.zeros((10, 12), np.float32))  # D not a multiple of 8
    idx.close()


def test_one_shot_buffers_across_sizes(small_set, gpu_index):
    """panns_search lends its device buffers from the handle to each call:
    calls that grow, shrink, change L and turn the queue outputs on and off
    return what a resident plan (its own buffers) returns for the same ids."""
    ds = small_set
    q = np.arange(ds["queries"].shape[0], dtype=np.uint32)
    for n, L, wq in ((3, 160, False), (200, 160, True), (50, 120, True), (1, 200, False), (200, 100, False)):
        r = gpu_index.search(q[:n], 100, L, 4, want_queue=wq)
        pl = gpu_index.plan(q[:n], 100, L, 4, want_queue=wq)
        pl.run()
        e = pl.fetch(want_queue=wq)
        pl.close()
        for f in ("ids", "dists", "ndist", "nexp", "niter"):
            assert (r[f] == e[f]).all(), (n, L, wq, f)
        if wq:
            for f in ("queue_ids", "queue_dists", "queue_checked"):
                assert (r[f] == e[f]).all(), (n, L, wq, f)
