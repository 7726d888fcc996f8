"""Pins of the oracle's queue and search (PSS5LG / TOPMLG restatement).

The search is checked against an independent statement of its semantics:
each iteration takes the first M unchecked queue entries, marks them
checked, marks their unvisited neighbours visited, and the queue becomes the
L smallest (distance, id) of old queue plus new neighbours.  That model is
written here with Python sets and sorted lists, not the oracle's
insert-and-shift code; a wrong bound test, index update or tie rule in the
oracle changes its queue, distance count or iteration count.
"""
import numpy as np
import pytest

import pyoracle as po
from parallel_anns_amd import tools


# ----------------------------------------------------------------------------
# add_into_queue (PSS5LG.cpp:565-608)
def _queue_model(items, cand, cap):
    """Reference behaviour: reject duplicates of an entry with the same (dist, id)
    position, keep the cap smallest by (dist, id)."""
    key = (np.float32(cand[1]), cand[0])
    keys = [(np.float32(d), i) for i, d in items]
    # lower_bound position
    pos = 0
    while pos < len(keys) and keys[pos] < key:
        pos += 1
    if pos < len(keys) and keys[pos][1] == cand[0]:
        return cap, items
    if pos == len(keys) and len(keys) >= cap:
        return cap, items
    new = items[:pos] + [cand] + items[pos:]
    return pos, new[:cap]


def test_add_into_queue_matches_model():
    rng = np.random.default_rng(0)
    for trial in range(300):
        cap = int(rng.integers(1, 9))
        q = np.zeros(cap + 1, po.CAND_DT)
        size = 0
        items = []
        for _ in range(20):
            cid = int(rng.integers(0, 12))
            d = float(rng.choice([0.5, 1.0, 1.5, 2.0, float(rng.random())]))
            exp_ret, items2 = _queue_model(items, (cid, d), cap)
            ret, size = po.add_into_queue(q, size, cap, cid, d)
            assert ret == exp_ret
            items = items2
            assert size == len(items)
            assert [int(x) for x in q["id"][:size]] == [i for i, _ in items]


# ----------------------------------------------------------------------------
def small_graph(n=300, dim=16, R=8, seed=0, comps=6):
    base = tools.gen_mixture(n, dim, comps, 0.3, 0, 100 + seed)
    knn, _ = tools.exact_knn(base, base, 12, exclude_self=True)
    adj, deg, ep = tools.build_graph(base, knn, R, L_build=16)
    queries = tools.gen_mixture(12, dim, comps, 0.3, 0, 200 + seed)
    return base, queries, adj, deg, ep


def set_model(adj, deg, base, q, L, M, init_ids, scale_M=None, queue=None, visited=None, norms=None):
    """The search's semantics with sets; returns (queue [(d, id, checked)],
    ndist, iters, visited).  norms: the opt-graph distance norm - 2 v.q."""
    if norms is None:
        dist = lambda v: np.float32(po.distance(base[v], q))
    else:
        dist = lambda v: np.float32(po.distance_with_norm(base[v], q, norms[v]))
    if queue is None:
        visited = set(int(v) for v in init_ids)
        Q = sorted([(dist(v), int(v), False) for v in init_ids[:L]], key=lambda e: (e[0], e[1]))
        ndist = L
    else:
        Q = list(queue)
        ndist = 0
    iters = 0
    m = 1 if scale_M else M
    while True:
        pick = [i for i, e in enumerate(Q) if not e[2]][:m]
        if not pick:
            break
        iters += 1
        for i in pick:
            Q[i] = (Q[i][0], Q[i][1], True)
        new = []
        for i in pick:
            c = Q[i][1]
            for n in adj[c, :deg[c]]:
                n = int(n)
                if n not in visited:
                    visited.add(n)
                    new.append(n)
        ndist += len(new)
        Q = sorted(Q + [(dist(n), n, False) for n in new], key=lambda e: (e[0], e[1]))[:L]
        if scale_M and m < scale_M:
            m <<= 1
    return Q, ndist, iters, visited


@pytest.mark.parametrize("M", [1, 2, 3, 8])
@pytest.mark.parametrize("seed", [0, 1])
def test_top_m_equals_set_semantics(M, seed):
    base, queries, adj, deg, ep = small_graph(seed=seed)
    g = po.CsrGraph(adj, deg, ep)
    L, K = 40, 20
    init = po.prepare_init_ids(g, L)
    for qi in range(queries.shape[0]):
        r = po.search_one(g, base, queries[qi], K, L, M, init)
        Q, nd, it, vis = set_model(adj, deg, base, queries[qi], L, M, init)
        assert [e[1] for e in Q] == list(r["queue"]["id"])
        assert all(e[2] for e in Q) and r["queue"]["checked"].all()
        assert nd == r["ndist"]
        assert it == r["merges"]
        assert nd == L + len(vis - set(int(x) for x in init))
        assert r["visited_clear"]
        assert list(r["ids"]) == [e[1] for e in Q[:K]]


@pytest.mark.parametrize("X", [1, 2, 5, 16])
def test_pss_v5_t1_equals_top_m_1(X):
    base, queries, adj, deg, ep = small_graph(seed=3)
    g = po.CsrGraph(adj, deg, ep)
    L, K = 50, 30
    init = po.prepare_init_ids(g, L)
    for qi in range(queries.shape[0]):
        a = po.search_one(g, base, queries[qi], K, L, 1, init)
        b = po.search_one(g, base, queries[qi], K, L, 1, init, mode="pss_v5", X=X)
        assert (a["queue"]["id"] == b["queue"]["id"]).all()
        assert a["ndist"] == b["ndist"]
        assert a["nexp"] == b["nexp"]
        # one sequential expansion, then rounds of up to X (PSS5LG.cpp:1478-1673)
        assert b["merges"] == -(-(b["nexp"] - 1) // X)


@pytest.mark.parametrize("mmax", [1, 3, 4, 16])
def test_scale_m_subsearch_equals_set_semantics(mmax):
    base, queries, adj, deg, ep = small_graph(seed=4)
    g = po.CsrGraph(adj, deg, ep)
    L = 36
    init = po.prepare_init_ids(g, L)
    n = base.shape[0]
    for qi in range(6):
        set_L, visited, _ = po.initialize_set_L(base, queries[qi], L, init, n)
        queue = [(np.float32(d), int(i), bool(c)) for i, d, c in zip(set_L["id"], set_L["dist"], set_L["checked"])]
        vis_set = set(int(x) for x in init)
        nd = po.subsearch_with_top_m(g, base, queries[qi], mmax, L, set_L, visited)
        Q, nd_m, _, vis_m = set_model(adj, deg, base, queries[qi], L, None, None, scale_M=mmax,
                                      queue=queue, visited=vis_set)
        assert [e[1] for e in Q] == list(set_L["id"])
        assert nd == nd_m
        bits = np.unpackbits(visited.view(np.uint8), bitorder="little")[:n]
        assert set(np.nonzero(bits)[0].tolist()) == vis_m


# ----------------------------------------------------------------------------
# opt-graph family (PSS5 / TOPM): the same searches with norm - 2 v.q
@pytest.mark.parametrize("M", [1, 4])
def test_norm_dot_top_m_equals_set_semantics(M):
    base, queries, adj, deg, ep = small_graph(seed=6)
    g = po.CsrGraph(adj, deg, ep, base=base)
    L, K = 40, 20
    init = po.prepare_init_ids(g, L)
    negative = 0
    for qi in range(queries.shape[0]):
        r = po.search_one(g, base, queries[qi], K, L, M, init)
        Q, nd, it, _ = set_model(adj, deg, base, queries[qi], L, M, init, norms=g.norms)
        assert [e[1] for e in Q] == list(r["queue"]["id"])
        assert [e[0] for e in Q] == list(r["queue"]["dist"])
        assert nd == r["ndist"] and it == r["merges"]
        negative += int((r["queue"]["dist"] < 0).sum())
    assert negative > 0  # the ordering of negative distances is exercised


def test_norm_dot_scale_m_equals_set_semantics():
    base, queries, adj, deg, ep = small_graph(seed=7)
    g = po.CsrGraph(adj, deg, ep, base=base)
    L, n = 36, base.shape[0]
    init = po.prepare_init_ids(g, L)
    for qi in range(6):
        set_L, visited, _ = po.initialize_set_L(base, queries[qi], L, init, n, norms=g.norms)
        queue = [(np.float32(d), int(i), bool(c)) for i, d, c in zip(set_L["id"], set_L["dist"], set_L["checked"])]
        nd = po.subsearch_with_top_m(g, base, queries[qi], 8, L, set_L, visited)
        Q, nd_m, _, _ = set_model(adj, deg, base, queries[qi], L, None, None, scale_M=8, queue=queue,
                                  visited=set(int(x) for x in init), norms=g.norms)
        assert [e[1] for e in Q] == list(set_L["id"]) and nd == nd_m


@pytest.mark.parametrize("mode,M", [("top_m", 1), ("top_m", 4), ("pss_v5", 1)])
def test_norm_dot_search_on_integer_data_equals_l2_search(mode, M):
    """On small-integer vectors every distance is exact, and norm - 2 v.q is
    |v - q|^2 - |q|^2: the same order (ties included) as squared L2, so both
    families must return the same ids, counts and iterations, with distances
    shifted by |q|^2 exactly."""
    base, queries, adj, deg, ep = small_graph(seed=8)
    base = np.round(base * 12).astype(np.float32)
    queries = np.round(queries * 12).astype(np.float32)
    g2 = po.CsrGraph(adj, deg, ep)
    gn = po.CsrGraph(adj, deg, ep, base=base)
    L, K = 48, 16
    init = po.prepare_init_ids(g2, L)
    for qi in range(queries.shape[0]):
        a = po.search_one(g2, base, queries[qi], K, L, M, init, mode=mode, X=3)
        b = po.search_one(gn, base, queries[qi], K, L, M, init, mode=mode, X=3)
        assert (a["queue"]["id"] == b["queue"]["id"]).all()
        assert (a["ids"] == b["ids"]).all()
        assert (a["ndist"], a["merges"], a["nexp"]) == (b["ndist"], b["merges"], b["nexp"])
        q2 = float((queries[qi].astype(np.int64) ** 2).sum())
        assert (b["queue"]["dist"] == a["queue"]["dist"] - np.float32(q2)).all()
        assert (b["queue"]["dist"] < 0).any()


def test_L_equals_N_is_exact_brute_force():
    base, queries, adj, deg, ep = small_graph(n=120, seed=5)
    g = po.CsrGraph(adj, deg, ep)
    n = base.shape[0]
    init = po.prepare_init_ids(g, n)
    assert sorted(init.tolist()) == list(range(n))
    bf, _ = po.brute_force_knn(base, queries, n)
    for qi in range(queries.shape[0]):
        r = po.search_one(g, base, queries[qi], 10, n, 2, init)
        assert (r["queue"]["id"] == bf[qi]).all()


def test_complete_graph_is_exact_after_one_expansion():
    n, dim = 60, 16
    base = tools.gen_mixture(n, dim, 3, 0.3, 0, 50)
    queries = tools.gen_mixture(5, dim, 3, 0.3, 0, 51)
    adj = np.array([[j for j in range(n) if j != i] for i in range(n)], np.uint32)
    deg = np.full(n, n - 1, np.uint32)
    g = po.CsrGraph(adj, deg, 0)
    bf, _ = po.brute_force_knn(base, queries, 20)
    init = po.prepare_init_ids(g, 20)
    for qi in range(5):
        r = po.search_one(g, base, queries[qi], 20, 20, 1, init)
        assert (r["ids"] == bf[qi]).all()
        assert r["ndist"] == n  # every vertex exactly once


def test_prepare_init_ids_known_answers():
    # ep = 3 with neighbours [5, 1, 5, 7] (a duplicate), N = 9
    n = 9
    adj = np.zeros((n, 4), np.uint32)
    deg = np.zeros(n, np.uint32)
    adj[3] = [5, 1, 5, 7]
    deg[3] = 4
    g = po.CsrGraph(adj, deg, 3)
    assert list(po.prepare_init_ids(g, 2)) == [5, 1]
    assert list(po.prepare_init_ids(g, 3)) == [5, 1, 7]
    # then ep+1, ep+2, ... wrapping at N, skipping chosen ids
    assert list(po.prepare_init_ids(g, 9)) == [5, 1, 7, 4, 6, 8, 0, 2, 3]


# ----------------------------------------------------------------------------
# learned-stop MLP oracle: its backward pass against finite differences
def test_mlp_oracle_gradient_finite_differences():
    import mlp_oracle as mo
    rng = np.random.default_rng(2)
    P = mo.init_params(3)[:mo.N_PARAMS].astype(np.float64)
    P[512:576] = rng.standard_normal(64) * 0.1
    X = rng.standard_normal((40, 8))
    Y = (rng.random(40) < 0.5).astype(np.float64)
    _, _, g = mo.forward_backward(P, X, Y)
    for i in rng.choice(mo.N_PARAMS, 60, replace=False).tolist() + [4800, 0, 600, 4740]:
        e = np.zeros_like(P)
        e[i] = 1e-6
        fd = (mo.loss_only(P + e, X, Y) - mo.loss_only(P - e, X, Y)) / 2e-6
        assert abs(fd - g[i]) <= 1e-6 + 1e-4 * abs(fd), i


def test_mlp_oracle_bf16_rounding():
    import mlp_oracle as mo
    # 1 + 2^-8 is halfway between 1 and 1 + 2^-7: ties go to the even 1.0;
    # 1 + 3*2^-8 is halfway between 1 + 2^-7 (odd) and 1 + 2^-6 (even)
    x = np.array([1.0, 1.00390625, 1.01171875, 1.0 + 2 ** -8 + 2 ** -12, -3.0e-5], np.float32)
    r = mo.bf16(x)
    assert r[0] == 1.0 and r[1] == 1.0 and r[2] == 1.015625 and r[3] == 1.0078125
    assert (np.asarray(r, np.float32).view(np.uint32) & 0xFFFF == 0).all()
