"""Pins of the graph builder (pt_build_graph) against an independent
statement of its rules, written here with Python lists and sets:

1. candidate pool of v: the vertices expanded by a best-first search for v
   over the k-NN graph from ep with an L_build queue (when L_build > 0),
   plus v's k-NN list; sorted by (distance, id), duplicates and v removed;
2. RNG occlusion (sync_prune, include/efanna2e/index_nsg.cpp:233-283): walk
   the pool in order, at most its first C entries, keep p unless a kept r has
   d(r, p) < d(v, p) (strictly), stop at R kept;
3. reverse edges (the role of InterInsert, :285-352, done in one pass):
   v's pool gains every s that kept v, pruned again only when above R;
4. stitching (tree_grow / DFS / findroot, :638-718, deterministic and
   degree-capped): the lowest-id unreached vertex u gets an edge from the
   first reached vertex with spare degree in u's k-NN list, else from the
   lowest-id such vertex, and everything u reaches is reached.

Distances are the oracle's (po_distance, pinned on its own).  Inputs with
many exactly equal distances (small integer coordinates) make a `<` / `<=`
slip or a different tie order change the graph; the caps are exercised by
small R and C."""
import numpy as np
import pytest

import pyoracle as po
from parallel_anns_amd import tools

SENT = 0xFFFFFFFF


def model_build(base, knn, R, L_build, C, ep, reverse_edges, strict=True):
    n = base.shape[0]
    dist = lambda a, b: np.float32(po.distance(base[a], base[b]))  # d(row a, query b)
    key = lambda e: (e[0], e[1])

    def search(v):
        """Expanded vertices of a best-first search for v from ep."""
        visited = {ep}
        Q = [[dist(ep, v), ep, False]]
        expanded = []
        while True:
            i = next((j for j, e in enumerate(Q) if not e[2]), None)
            if i is None:
                return expanded
            Q[i][2] = True
            c = Q[i][1]
            expanded.append((Q[i][0], c))
            for nb in knn[c]:
                nb = int(nb)
                if nb == SENT or nb in visited:
                    continue
                visited.add(nb)
                Q.append([dist(nb, v), nb, False])
            Q = sorted(Q, key=key)[:L_build]

    def prune(v, pool, cap):
        kept = []
        for i, (d, p) in enumerate(pool):
            if i >= cap or len(kept) >= R:
                break
            occluded = False
            for _, r in kept:
                if r == p:
                    occluded = True
                    break
                drp = dist(r, p)
                if (drp < d) if strict else (drp <= d):
                    occluded = True
                    break
            if not occluded:
                kept.append((d, p))
        return kept

    out = []
    for v in range(n):
        pool = set(search(v)) if L_build > 0 else set()
        pool |= {(dist(int(nb), v), int(nb)) for nb in knn[v] if nb != SENT}
        pool = sorted((e for e in pool if e[1] != v), key=key)
        out.append(prune(v, pool, C))
    adj = []
    for v in range(n):
        pool = set(out[v])
        if reverse_edges:
            pool |= {(dist(s, v), s) for s in range(n) if s != v and any(p == v for _, p in out[s])}
        pool = sorted(pool, key=key)
        adj.append([p for _, p in (pool if len(pool) <= R else prune(v, pool, len(pool)))])
    # stitching
    reached = set()

    def dfs(root):
        if root in reached:
            return
        reached.add(root)
        stack = [root]
        while stack:
            u = stack.pop()
            for w in adj[u]:
                if w not in reached:
                    reached.add(w)
                    stack.append(w)

    dfs(ep)
    for u in range(n):
        if u in reached:
            continue
        spare = lambda c: c in reached and len(adj[c]) < R
        root = next((int(c) for c in knn[u] if c != SENT and spare(int(c))), None)
        if root is None:
            root = min(c for c in reached if len(adj[c]) < R)
        adj[root].append(u)
        dfs(u)
    return adj


def _grid_points(n, dim, seed):
    """Small integer coordinates: many pairs at exactly equal distances."""
    rng = np.random.default_rng(seed)
    pts = set()
    while len(pts) < n:
        pts.add(tuple(int(x) for x in rng.integers(0, 3, dim)))
    return np.array(sorted(pts), np.float32)


def _check(base, k, R, L_build, C, reverse, ep=0):
    knn, _ = tools.exact_knn(base, base, k, exclude_self=True)
    adj, deg, ep_out = tools.build_graph(base, knn, R, L_build=L_build, C=C, ep=ep, reverse_edges=reverse)
    want = model_build(base, knn, R, L_build, C, ep, reverse)
    for v in range(base.shape[0]):
        assert adj[v, :deg[v]].tolist() == want[v], (v, R, L_build, C, reverse)
    assert tools.reachable(adj, deg, ep) == base.shape[0]
    return knn, want


@pytest.mark.parametrize("R,C", [(3, 4), (4, 500), (5, 6), (8, 500)])
@pytest.mark.parametrize("L_build,reverse", [(0, False), (0, True), (10, False), (16, True)])
def test_builder_equals_model_on_ties(R, C, L_build, reverse):
    base = _grid_points(60, 8, 1)
    _check(base, 12, R, L_build, C, reverse, ep=7)


@pytest.mark.parametrize("L_build,reverse", [(0, False), (24, True)])
def test_builder_equals_model_on_mixture(L_build, reverse):
    base = tools.gen_mixture(300, 16, 6, 0.3, 0, 31)
    ep = tools.navigating_node(base)
    _check(base, 16, 6, L_build, 40, reverse, ep=ep)


def test_tie_data_tells_strict_from_non_strict_occlusion():
    """The tie inputs above would catch a `<=` in the occlusion test: the
    model with `<=` builds a different graph from the same k-NN lists."""
    base = _grid_points(60, 8, 1)
    knn, _ = tools.exact_knn(base, base, 12, exclude_self=True)
    a = model_build(base, knn, 5, 0, 500, 7, False)
    b = model_build(base, knn, 5, 0, 500, 7, False, strict=False)
    assert a != b


def test_stitch_uses_the_lowest_id_spare_vertex():
    """Two far-apart clusters whose k-NN lists stay inside each cluster: the
    second is joined by one edge, from the lowest-id reached vertex with
    spare degree (every k-NN neighbour of the first unreached vertex is
    unreached)."""
    rng = np.random.default_rng(4)
    a = rng.standard_normal((20, 8)).astype(np.float32)
    b = (rng.standard_normal((20, 8)) + 100.0).astype(np.float32)
    base = np.concatenate([a, b])
    knn, want = _check(base, 6, 4, 0, 500, False, ep=3)
    assert not set(knn[20:].ravel().tolist()) & set(range(20))
    into_b = [(v, w) for v in range(20) for w in want[v] if w >= 20]
    assert len(into_b) == 1 and into_b[0][1] == 20
    assert into_b[0][0] == min(v for v in range(20) if len([w for w in want[v] if w < 20]) < 4)


def test_tie_data_exercises_the_candidate_cap():
    """C bounds how far down the pool the prune looks: on the tie inputs a
    cap of 4 keeps fewer neighbours than no cap for some vertices."""
    base = _grid_points(60, 8, 1)
    knn, _ = tools.exact_knn(base, base, 12, exclude_self=True)
    capped = model_build(base, knn, 5, 0, 4, 7, False)
    free = model_build(base, knn, 5, 0, 500, 7, False)
    assert capped != free
