"""Pins of the oracle's PSS_v5 with T > 1 queues (PSS5LG.cpp:648-947,
1376-1724): the two queue merges, the dealing of candidates to workers, the
X-step expansions and the merge back into the master.

The reference runs its T expansion loops as racing OpenMP threads; the
oracle runs them in lockstep in a fixed order (step s expands the s-th
candidate of worker 0, 1, ..., then of the master).  The model below states
that schedule with Python lists and set semantics -- a queue after an
expansion is the capacity smallest (distance, id) of old entries plus the
accepted new ones, a merge is a sorted union with duplicates folded -- not
with the oracle's lower-bound-and-shift code, so a wrong index update,
bound or duplicate rule in the oracle changes its queue or its counts.
"""
import numpy as np
import pytest

import pyoracle as po
from test_oracle_search import small_graph


def _key(e):
    return (np.float32(e[0]), e[1])


# ----------------------------------------------------------------------------
# merge_two_queues_into_1st_queue_seq_fixed / _incr (PSS5LG.cpp:648-805)
def _lower_bound(q, e):
    i = 0
    while i < len(q) and _key(q[i]) < _key(e):
        i += 1
    return i


def merge_fixed_model(q1, q2):
    """queue1 keeps its length; a duplicate keeps queue1's copy, unchecked if
    queue2's copy is; queue2's first element landing on queue1's last index
    replaces it outright (PSS5LG.cpp:665-667)."""
    L = len(q1)
    idx = _lower_bound(q1, q2[0])
    out = [list(e) for e in q1]
    if idx == L:
        return idx, out
    if idx == L - 1:
        out[L - 1] = list(q2[0])
        return idx, out
    pos = {e[1]: i for i, e in enumerate(out)}
    extra = []
    for e in q2:
        if e[1] in pos:
            out[pos[e[1]]][2] = out[pos[e[1]]][2] and e[2]
        else:
            extra.append(list(e))
    return idx, sorted(out + extra, key=_key)[:L]


def merge_incr_model(q1, cap, q2):
    idx = _lower_bound(q1, q2[0])
    out = [list(e) for e in q1]
    pos = {e[1]: i for i, e in enumerate(out)}
    extra = []
    for e in q2:
        if e[1] in pos:
            out[pos[e[1]]][2] = out[pos[e[1]]][2] and e[2]
        else:
            extra.append(list(e))
    return idx, sorted(out + extra, key=_key)[:cap]


def _rand_queues(rng, n1, n2):
    """Two sorted queues over a small id pool whose distances tie often; a
    shared id has the same distance in both (the same vertex and query)."""
    pool = 14
    dist = rng.choice([0.25, 0.5, 0.5, 1.0, 1.5], size=pool).astype(np.float32)
    ids1 = rng.choice(pool, size=n1, replace=False)
    ids2 = rng.choice(pool, size=n2, replace=False)
    q1 = sorted([[float(dist[i]), int(i), bool(rng.random() < 0.5)] for i in ids1], key=_key)
    q2 = sorted([[float(dist[i]), int(i), bool(rng.random() < 0.5)] for i in ids2], key=_key)
    return q1, q2


def _arr(q, cap=None):
    a = np.zeros(cap if cap is not None else len(q), po.CAND_DT)
    for i, e in enumerate(q):
        a[i]["dist"], a[i]["id"], a[i]["checked"] = e[0], e[1], 1 if e[2] else 0
    return a


def _lst(a, n=None):
    n = a.shape[0] if n is None else n
    return [[float(a[i]["dist"]), int(a[i]["id"]), bool(a[i]["checked"])] for i in range(n)]


def test_merge_fixed_matches_model():
    rng = np.random.default_rng(1)
    for _ in range(3000):
        q1, q2 = _rand_queues(rng, int(rng.integers(1, 9)), int(rng.integers(1, 9)))
        a1, a2 = _arr(q1), _arr(q2)
        r = po.merge_two_queues_fixed(a1, a2)
        idx, want = merge_fixed_model(q1, q2)
        assert r == idx
        assert _lst(a1) == want


def test_merge_fixed_hand_cases():
    # duplicate of a checked master entry from an unchecked worker copy:
    # the master entry is unchecked, nothing else moves
    q1 = [[0.1, 5, True], [0.2, 3, True], [0.3, 9, False], [0.4, 1, False]]
    a = _arr(q1)
    assert po.merge_two_queues_fixed(a, _arr([[0.2, 3, False]])) == 1
    assert _lst(a) == [[np.float32(0.1), 5, True], [np.float32(0.2), 3, False], [np.float32(0.3), 9, False],
                       [np.float32(0.4), 1, False]]
    # a checked worker copy never re-checks an unchecked master entry
    a = _arr(q1)
    po.merge_two_queues_fixed(a, _arr([[0.3, 9, True]]))
    assert [e[2] for e in _lst(a)] == [True, True, False, False]
    # everything past the end: nothing merged, index = size
    a = _arr(q1)
    assert po.merge_two_queues_fixed(a, _arr([[0.5, 2, False], [0.6, 7, False]])) == 4
    assert _lst(a)[3][1] == 1
    # lands on the last index: replaces it, the rest of queue2 is ignored
    a = _arr(q1)
    assert po.merge_two_queues_fixed(a, _arr([[0.35, 8, True], [0.36, 6, False]])) == 3
    assert [e[1] for e in _lst(a)] == [5, 3, 9, 8]
    # equal distance, the lower id first (Candidate::operator<)
    a = _arr(q1)
    assert po.merge_two_queues_fixed(a, _arr([[0.2, 2, False], [0.3, 10, False]])) == 1
    assert [e[1] for e in _lst(a)] == [5, 2, 3, 9]


def test_merge_incr_matches_model():
    rng = np.random.default_rng(2)
    for _ in range(3000):
        cap = int(rng.integers(1, 11))
        n1 = int(rng.integers(1, cap + 1))
        q1, q2 = _rand_queues(rng, n1, int(rng.integers(1, 9)))
        a1, a2 = _arr(q1, cap), _arr(q2)
        r, size = po.merge_two_queues_incr(a1, n1, a2)
        idx, want = merge_incr_model(q1, cap, q2)
        assert r == idx
        assert size == len(want)
        assert _lst(a1, size) == want


# ----------------------------------------------------------------------------
# the lockstep PSS_v5 schedule with T queues
def pss_model(adj, deg, base, q, L, X, T, init_ids, local_L=None):
    """Returns (master [(d, id, checked)], ndist, merges, expansions, visited)."""
    local_L = L if local_L is None else local_L
    dist = lambda v: np.float32(po.distance(base[v], q))
    visited = set(int(v) for v in init_ids[:L])
    master = sorted([[dist(int(v)), int(v), False] for v in init_ids[:L]], key=_key)
    stats = {"ndist": L, "exp": 0, "merges": 0}

    def expand(queue, cap, c):
        """One expansion of vertex c into queue (capacity cap): neighbours not
        yet visited are visited and measured; those not farther than the
        master's last entry join the queue, which keeps its cap smallest."""
        stats["exp"] += 1
        bound = master[-1][0]
        new = []
        for n in adj[c, :deg[c]]:
            n = int(n)
            if n in visited:
                continue
            visited.add(n)
            stats["ndist"] += 1
            d = dist(n)
            if d <= bound:
                new.append([d, n, False])
        queue[:] = sorted(queue + new, key=_key)[:cap]

    def first_unchecked(queue):
        return next((i for i, e in enumerate(queue) if not e[2]), None)

    # sequential start: the master's first candidate
    master[0][2] = True
    expand(master, L, master[0][1])
    workers = [[] for _ in range(T - 1)]
    while True:
        unc = [i for i, e in enumerate(master) if not e[2]]
        if not unc:
            break
        # deal: the j-th unchecked entry to worker j mod T, every T-th stays
        for j, i in enumerate(unc):
            w = j % T
            if w == T - 1:
                continue
            workers[w].append(list(master[i]))
            master[i][2] = True
            if len(workers[w]) == local_L:
                break
        for _ in range(X):
            any_step = False
            for w in range(T):
                queue, cap = (workers[w], local_L) if w < T - 1 else (master, L)
                i = first_unchecked(queue)
                if i is None:
                    continue
                queue[i][2] = True
                expand(queue, cap, queue[i][1])
                any_step = True
            if not any_step:
                break
        stats["merges"] += 1
        # merge: each worker's entries folded into the master in worker order
        for w in range(T - 1):
            if workers[w]:
                _, master[:] = merge_fixed_model(master, workers[w])
                workers[w] = []
    return master, stats["ndist"], stats["merges"], stats["exp"], visited


@pytest.mark.parametrize("X", [1, 2, 5, 16])
def test_pss_v5_one_queue_is_the_t1_search(X):
    base, queries, adj, deg, ep = small_graph(seed=3)
    g = po.CsrGraph(adj, deg, ep)
    L, K = 50, 30
    init = po.prepare_init_ids(g, L)
    for qi in range(queries.shape[0]):
        a = po.search_one(g, base, queries[qi], K, L, 1, init, mode="pss_v5", X=X)
        b = po.search_pss_v5(g, base, queries[qi], K, L, X, 1, init)
        assert (a["queue"]["id"] == b["queue"]["id"]).all()
        assert (a["queue"]["dist"].view(np.uint32) == b["queue"]["dist"].view(np.uint32)).all()
        assert (a["queue"]["checked"] == b["queue"]["checked"]).all()
        assert (a["ndist"], a["merges"], a["nexp"]) == (b["ndist"], b["merges"], b["nexp"])


@pytest.mark.parametrize("T,X", [(2, 1), (2, 4), (3, 2), (4, 1), (4, 4), (8, 2), (8, 8)])
@pytest.mark.parametrize("seed", [0, 5])
def test_pss_v5_workers_equal_the_lockstep_model(T, X, seed):
    base, queries, adj, deg, ep = small_graph(seed=seed)
    g = po.CsrGraph(adj, deg, ep)
    L, K = 40, 20
    init = po.prepare_init_ids(g, L)
    for qi in range(queries.shape[0]):
        r = po.search_pss_v5(g, base, queries[qi], K, L, X, T, init)
        Q, nd, merges, nexp, vis = pss_model(adj, deg, base, queries[qi], L, X, T, init)
        assert [e[1] for e in Q] == list(r["queue"]["id"])
        assert [np.float32(e[0]) for e in Q] == list(r["queue"]["dist"])
        assert [e[2] for e in Q] == [bool(c) for c in r["queue"]["checked"]]
        assert (nd, merges, nexp) == (r["ndist"], r["merges"], r["nexp"])
        # invariants: everything checked at the end, one distance per
        # visited vertex, no id twice, visited bits cleared on return
        assert r["queue"]["checked"].all()
        assert nd == L + len(vis - set(int(x) for x in init))
        assert len(set(r["queue"]["id"].tolist())) == L
        assert r["visited_clear"]
        assert list(r["ids"]) == [e[1] for e in Q[:K]]


def test_pss_v5_workers_small_local_queues():
    """local_L < L (APP_PSS's L_local parameter): worker queues fill and drop
    their last entries, and dealing stops at a full worker queue."""
    base, queries, adj, deg, ep = small_graph(seed=2)
    g = po.CsrGraph(adj, deg, ep)
    L, K = 60, 20
    init = po.prepare_init_ids(g, L)
    for T, X, lL in ((2, 4, 8), (4, 3, 5), (3, 6, 12)):
        for qi in range(queries.shape[0]):
            r = po.search_pss_v5(g, base, queries[qi], K, L, X, T, init, local_L=lL)
            Q, nd, merges, nexp, _ = pss_model(adj, deg, base, queries[qi], L, X, T, init, local_L=lL)
            assert [e[1] for e in Q] == list(r["queue"]["id"])
            assert [e[2] for e in Q] == [bool(c) for c in r["queue"]["checked"]]
            assert (nd, merges, nexp) == (r["ndist"], r["merges"], r["nexp"])


def test_pss_v5_workers_recall_parity_with_top_m(small_set):
    """More queues expand more candidates per round; recall stays at least
    the one-queue search's (the reference's claim for its parallel form is
    the same recall at lower latency)."""
    ds = small_set
    g = po.CsrGraph(ds["adj"], ds["deg"], ds["ep"])
    L, K = 120, 100
    init = po.prepare_init_ids(g, L)
    nq = 60
    from parallel_anns_amd import tools
    base_ids = np.stack([po.search_one(g, ds["base"], ds["queries"][i], K, L, 1, init)["ids"] for i in range(nq)])
    r1 = tools.recall(ds["gt_ids"][:nq], base_ids, L)
    for T, X in ((4, 4), (8, 2)):
        ids = np.stack([po.search_pss_v5(g, ds["base"], ds["queries"][i], K, L, X, T, init)["ids"]
                        for i in range(nq)])
        rt = tools.recall(ds["gt_ids"][:nq], ids, L)
        assert rt >= r1 - 0.01, (T, X, rt, r1)
