"""The opt-graph search family on the GPU: load_nsg_graph (PSS5.cpp:56-132)
with the distance norm(v) - 2 v.q (compute_norm PSS5.cpp:293-334,
compute_distance_with_norm :336-390), searched by para_search_PSS_v5_profiling
(PSS5.cpp:1211-1462), para_search_with_top_m_pure (TOPM.cpp:1881-2070) and
seq_search_with_top_m_scale_m (TOPM.cpp:1273-1380) at one thread.

Bar: norms, ids, queue ids, distance bits and checked flags, and every
counter equal the oracle's opt-graph restatement.  On the normalised
mixture norm - 2 v.q is negative for every close neighbour, so the key
order of negative distances is exercised throughout.
"""
import os
import tempfile

import numpy as np
import pytest

import pyoracle as po
from parallel_anns_amd import tools
from parallel_anns_amd.search import Index, SearchError

pytestmark = pytest.mark.gpu


def _case(**kw):
    print("case", " ".join(f"{k}={v}" for k, v in kw.items()), flush=True)


@pytest.fixture(scope="module")
def norm_index(small_set):
    ds = small_set
    idx = Index(0)
    idx.set_base(ds["base"])
    idx.set_queries(ds["queries"])
    idx.set_graph(ds["adj"], ds["deg"], ds["ep"])
    idx.set_distance("norm_dot")
    g = po.CsrGraph(ds["adj"], ds["deg"], ds["ep"], base=ds["base"])
    yield idx, g
    idx.close()


def _same_as_oracle(res, refs, with_queue=True):
    for i, r in enumerate(refs):
        assert (res["ids"][i] == r["ids"]).all(), f"ids differ, row {i}"
        assert res["ndist"][i] == r["ndist"], f"distance count differs, row {i}"
        assert res["nexp"][i] == r["nexp"]
        if with_queue:
            assert (res["queue_ids"][i] == r["queue"]["id"]).all()
            assert (res["queue_dists"][i].view(np.uint32) == r["queue"]["dist"].view(np.uint32)).all()
            assert (res["queue_checked"][i] == r["queue"]["checked"]).all()


def test_norms_bit_exact(small_set, norm_index):
    idx, g = norm_index
    _case(test="norms", n=small_set["base"].shape[0])
    assert (idx.norms().view(np.uint32) == g.norms.view(np.uint32)).all()


@pytest.mark.parametrize("L,M", [(100, 1), (100, 4), (200, 8), (64, 3)])
def test_top_m_bit_exact(small_set, norm_index, L, M):
    ds = small_set
    idx, g = norm_index
    K = min(100, L)
    qids = np.arange(64, dtype=np.uint32)
    _case(test="norm_top_m", L=L, M=M)
    init = idx.prepare_init_ids(L)
    res = idx.search(qids, K, L, M, init_ids=init, want_queue=True)
    refs = [po.search_one(g, ds["base"], ds["queries"][q], K, L, M, init) for q in qids]
    _same_as_oracle(res, refs)
    assert (res["dists"] < 0).any()
    assert (res["niter"] == np.array([r["merges"] for r in refs])).all()
    assert res["counters"]["merges"] == sum(r["merges"] for r in refs)


@pytest.mark.parametrize("X", [1, 4])
def test_pss_v5_bit_exact(small_set, norm_index, X):
    ds = small_set
    idx, g = norm_index
    K, L = 100, 100
    qids = np.arange(32, dtype=np.uint32)
    _case(test="norm_pss_v5", X=X)
    init = idx.prepare_init_ids(L)
    res = idx.search(qids, K, L, X, mode="pss_v5", init_ids=init, want_queue=True)
    refs = [po.search_one(g, ds["base"], ds["queries"][q], K, L, 1, init, mode="pss_v5", X=X) for q in qids]
    _same_as_oracle(res, refs)
    assert res["counters"]["merges"] == sum(r["merges"] for r in refs)


def test_scale_m_and_subsearch(small_set, norm_index):
    ds = small_set
    idx, g = norm_index
    L, n = 120, ds["base"].shape[0]
    init = po.prepare_init_ids(g, L)
    for q, mmax in ((0, 1), (1, 4), (2, 16)):
        _case(test="norm_subsearch", q=q, M_max=mmax)
        set_L, visited, _ = po.initialize_set_L(ds["base"], ds["queries"][q], L, init, n, norms=g.norms)
        assert (set_L["dist"] < 0).any()
        ref_L, ref_vis = set_L.copy(), visited.copy()
        nd_ref = po.subsearch_with_top_m(g, ds["base"], ds["queries"][q], mmax, L, ref_L, ref_vis)
        qi, qd, qc, vis = set_L["id"].copy(), set_L["dist"].copy(), set_L["checked"].copy(), visited.copy()
        nd = idx.subsearch_with_top_m(mmax, q, L, qi, qd, qc, vis)
        assert nd == nd_ref
        assert (qi == ref_L["id"]).all() and (qc == ref_L["checked"]).all()
        assert (qd.view(np.uint32) == ref_L["dist"].view(np.uint32)).all()
        assert (vis == ref_vis).all()
    # scale-M from the initial queue (seq_search_with_top_m_scale_m)
    qids = np.arange(16, dtype=np.uint32)
    _case(test="norm_scale_m")
    res = idx.search(qids, 50, L, 8, mode="scale_m", init_ids=init, want_queue=True)
    for i, q in enumerate(qids):
        set_L, visited, nd0 = po.initialize_set_L(ds["base"], ds["queries"][q], L, init, n, norms=g.norms)
        nd = po.subsearch_with_top_m(g, ds["base"], ds["queries"][q], 8, L, set_L, visited)
        assert (res["queue_ids"][i] == set_L["id"]).all()
        assert (res["queue_dists"][i].view(np.uint32) == set_L["dist"].view(np.uint32)).all()
        assert res["ndist"][i] == nd0 + nd


@pytest.mark.parametrize("waves", [2, 4])
def test_multiwave_equals_single_wave(small_set, norm_index, waves):
    idx, _ = norm_index
    qids = np.arange(64, dtype=np.uint32)
    init = idx.prepare_init_ids(160)
    _case(test="norm_multiwave", waves=waves)
    a = idx.search(qids, 100, 160, 8, init_ids=init, want_queue=True)
    b = idx.search(qids, 100, 160, 8, init_ids=init, want_queue=True, waves=waves)
    for k in ("ids", "dists", "queue_ids", "queue_dists", "queue_checked", "ndist", "nexp", "niter"):
        x, y = a[k], b[k]
        if x.dtype == np.float32:
            x, y = x.view(np.uint32), y.view(np.uint32)
        assert (x == y).all(), k


def test_large_batch_in_locality_order(small_set, norm_index):
    """A batch above the reordering size equals the oracle per query."""
    ds = small_set
    idx, g = norm_index
    q = tools.gen_mixture(700, ds["base"].shape[1], ds["cfg"]["comps"], ds["cfg"]["sigma"], 0, 21)
    idx2 = Index(0)
    try:
        idx2.set_base(ds["base"])
        idx2.set_queries(q)
        idx2.set_graph(ds["adj"], ds["deg"], ds["ep"])
        idx2.set_distance("norm_dot")
        init = idx2.prepare_init_ids(100)
        _case(test="norm_locality", nq=700)
        res = idx2.search(np.arange(700, dtype=np.uint32), 100, 100, 4, init_ids=init)
        for i in range(0, 700, 37):
            r = po.search_one(g, ds["base"], q[i], 100, 100, 4, init)
            assert (res["ids"][i] == r["ids"]).all() and res["ndist"][i] == r["ndist"]
    finally:
        idx2.close()


@pytest.mark.parametrize("dim", [16, 40, 128])
def test_other_dimensions(dim):
    n = 3000
    base = tools.gen_mixture(n, dim, 8, 0.3, 0, 70 + dim)
    queries = tools.gen_mixture(24, dim, 8, 0.3, 0, 80 + dim)
    knn, _ = tools.exact_knn(base, base, 20, exclude_self=True)
    adj, deg, ep = tools.build_graph(base, knn, 16)
    g = po.CsrGraph(adj, deg, ep, base=base)
    idx = Index(0)
    try:
        idx.set_base(base)
        idx.set_queries(queries)
        idx.set_graph(adj, deg, ep)
        idx.set_distance("norm_dot")
        _case(test="norm_dims", dim=dim)
        assert (idx.norms().view(np.uint32) == g.norms.view(np.uint32)).all()
        init = idx.prepare_init_ids(80)
        res = idx.search(np.arange(24, dtype=np.uint32), 40, 80, 4, init_ids=init, want_queue=True)
        refs = [po.search_one(g, base, queries[i], 40, 80, 4, init) for i in range(24)]
        _same_as_oracle(res, refs)
    finally:
        idx.close()


def test_integer_data_equals_l2_search():
    """Size-independent property: on small-integer rows every distance is
    exact and norm - 2 v.q = |v - q|^2 - |q|^2, so the two families return
    the same ids and counts, distances shifted by |q|^2."""
    n, dim = 4000, 32
    base = np.round(tools.gen_mixture(n, dim, 6, 0.3, 0, 91) * 10).astype(np.float32)
    queries = np.round(tools.gen_mixture(32, dim, 6, 0.3, 0, 92) * 10).astype(np.float32)
    knn, _ = tools.exact_knn(base, base, 20, exclude_self=True)
    adj, deg, ep = tools.build_graph(base, knn, 16)
    idx = Index(0)
    try:
        idx.set_base(base)
        idx.set_queries(queries)
        idx.set_graph(adj, deg, ep)
        init = idx.prepare_init_ids(90)
        qids = np.arange(32, dtype=np.uint32)
        _case(test="norm_integer")
        a = idx.search(qids, 50, 90, 4, init_ids=init, want_queue=True)
        idx.set_distance("norm_dot")
        b = idx.search(qids, 50, 90, 4, init_ids=init, want_queue=True)
        assert (a["queue_ids"] == b["queue_ids"]).all() and (a["ndist"] == b["ndist"]).all()
        assert (a["niter"] == b["niter"]).all()
        q2 = (queries.astype(np.int64) ** 2).sum(axis=1).astype(np.float32)
        assert (b["queue_dists"] == a["queue_dists"] - q2[:, None]).all()
        idx.set_distance("l2")
        c = idx.search(qids, 50, 90, 4, init_ids=init)
        assert (c["ids"] == a["ids"]).all() and (c["dists"].view(np.uint32) == a["dists"].view(np.uint32)).all()
    finally:
        idx.close()


def test_load_nsg_graph_and_rejections(small_set):
    ds = small_set
    d = tempfile.mkdtemp(prefix="panns_nsg_")
    path = os.path.join(d, "g.nsg")
    tools.nsg_write(path, ds["adj"], ds["deg"], ds["ep"])
    idx = Index(0)
    try:
        with pytest.raises(SearchError):
            idx.load_nsg_graph(path)  # the base set comes first, as in the reference
        idx.set_base(ds["base"])
        idx.set_queries(ds["queries"])
        idx.load_nsg_graph(path)
        assert idx.distance == "norm_dot"
        _case(test="norm_load_nsg")
        init = idx.prepare_init_ids(100)
        res = idx.search(np.arange(8, dtype=np.uint32), 100, 100, 2, init_ids=init)
        g = po.CsrGraph(ds["adj"], ds["deg"], ds["ep"], base=ds["base"])
        for i in range(8):
            r = po.search_one(g, ds["base"], ds["queries"][i], 100, 100, 2, init)
            assert (res["ids"][i] == r["ids"]).all()
        with pytest.raises(SearchError):
            idx.search(np.arange(4, dtype=np.uint32), 10, 100, 2, init_ids=init, stop="learned",
                       stop_threshold=0.5)
        with pytest.raises(SearchError):
            idx.set_distance("cosine")
        idx.load_common_nsg_graph(path)
        assert idx.distance == "l2"
    finally:
        idx.close()
