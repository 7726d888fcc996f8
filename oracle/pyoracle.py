"""ctypes wrapper of the CPU oracle (liboracle.so) -- TEST INFRASTRUCTURE.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg use this
module, and only as the checker / the timed CPU baseline.  The product
package never imports it.
"""
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
SO = os.path.join(HERE, "liboracle.so")

u32, u64, cint, cf, cd, vp = (ctypes.c_uint32, ctypes.c_uint64, ctypes.c_int, ctypes.c_float,
                              ctypes.c_double, ctypes.c_void_p)


class Cand(ctypes.Structure):
    _fields_ = [("id", u32), ("dist", cf), ("checked", ctypes.c_uint8), ("pad_", ctypes.c_uint8 * 3)]


CAND_DT = np.dtype([("id", "<u4"), ("dist", "<f4"), ("checked", "u1"), ("pad", "u1", (3,))])


class Graph(ctypes.Structure):
    _fields_ = [("num_v", u32), ("width", u32), ("ep", u32), ("num_e", u64),
                ("vertex_base", vp), ("deg_ngbrs", vp), ("norms", vp)]


class Counters(ctypes.Structure):
    _fields_ = [("distance_computations", u64), ("merges", u64), ("expansions", u64)]


_API = {
    "po_distance": (cf, [vp, vp, u32]),
    "po_norm": (cf, [vp, u32]),
    "po_distance_with_norm": (cf, [vp, vp, u32, cf]),
    "po_add_into_queue": (u32, [vp, u32, ctypes.POINTER(u32), u32, Cand]),
    "po_prepare_init_ids": (cint, [ctypes.POINTER(Graph), u32, vp]),
    "po_search_top_m": (None, [ctypes.POINTER(Graph), vp, u32, vp, u32, u32, u32, vp, vp, vp, vp,
                               ctypes.POINTER(Counters)]),
    "po_search_pss_v5_t1": (None, [ctypes.POINTER(Graph), vp, u32, vp, u32, u32, u32, vp, vp, vp, vp,
                                   ctypes.POINTER(Counters)]),
    "po_subsearch_with_top_m": (None, [ctypes.POINTER(Graph), vp, u32, vp, u32, u32, vp, u32,
                                       ctypes.POINTER(u32), vp, ctypes.POINTER(u64)]),
    "po_initialize_set_L": (None, [vp, u32, vp, u32, vp, u32, ctypes.POINTER(u32), vp, vp,
                                   ctypes.POINTER(u64), vp]),
    "po_merge_two_queues_fixed": (u32, [vp, u32, u32, vp, u32, u32]),
    "po_merge_two_queues_incr": (u32, [vp, u32, ctypes.POINTER(u32), u32, vp, u32, u32]),
    "po_search_pss_v5": (None, [ctypes.POINTER(Graph), vp, u32, vp, u32, u32, u32, u32, u32, vp, vp, vp, vp,
                                ctypes.POINTER(Counters)]),
    "po_search_batch": (u64, [ctypes.POINTER(Graph), vp, u32, vp, u32, u32, u32, u32, cint, vp, vp,
                              vp, vp, cint]),
    "po_search_pss_v5_par": (u64, [ctypes.POINTER(Graph), vp, u32, vp, u32, u32, u32, u32, u32, u32, vp, vp]),
    "po_build_info": (ctypes.c_char_p, []),
    "po_brute_force_knn": (None, [vp, u32, u32, vp, u32, u32, vp, vp, cint]),
    "po_recall": (cd, [vp, u32, vp, u32, u32, u32, u32]),
}

_lib = None


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(SO):
            build()
        _lib = ctypes.CDLL(SO)
        for n, (r, a) in _API.items():
            f = getattr(_lib, n)
            f.restype = r
            f.argtypes = a
    return _lib


def _p(a):
    return None if a is None else a.ctypes.data_as(vp)


class CsrGraph:
    """The reference's CSR (vertex_base + [deg, ids...]) built from padded rows.

    With `base` given, the graph also carries every vertex's norm (po_norm)
    and the searches use the opt-graph distance norm - 2 v.q (the PSS5 / TOPM
    family, load_nsg_graph PSS5.cpp:56-132); without, squared L2."""

    def __init__(self, adj, deg, ep, width=None, base=None):
        adj = np.asarray(adj, dtype=np.uint32)
        deg = np.asarray(deg, dtype=np.uint32)
        n = adj.shape[0]
        # vertex v's record [deg, ids...] starts at v + (sum of earlier degrees)
        lens = deg.astype(np.uint64) + 1
        self.vertex_base = np.zeros(n, dtype=np.uint64)
        if n:
            self.vertex_base[1:] = np.cumsum(lens)[:-1]
        total = int(n + deg.sum())
        self.deg_ngbrs = np.empty(total, dtype=np.uint32)
        self.deg_ngbrs[self.vertex_base.astype(np.int64)] = deg
        if adj.shape[1]:
            slot = np.arange(adj.shape[1], dtype=np.uint32)
            real = slot[None, :] < deg[:, None]
            pos = (self.vertex_base.astype(np.int64) + 1)[:, None] + slot[None, :].astype(np.int64)
            self.deg_ngbrs[pos[real]] = adj[real]
        self.norms = None if base is None else norms(base)
        self.g = Graph(n, int(width if width is not None else (deg.max() if n else 0)), int(ep),
                       int(deg.sum()), _p(self.vertex_base), _p(self.deg_ngbrs), _p(self.norms))
        self.n = n


def distance(v, q):
    v = np.ascontiguousarray(v, np.float32)
    q = np.ascontiguousarray(q, np.float32)
    return float(lib().po_distance(_p(v), _p(q), v.shape[-1]))


def norm(v):
    v = np.ascontiguousarray(v, np.float32)
    return float(lib().po_norm(_p(v), v.shape[-1]))


def norms(base):
    base = np.ascontiguousarray(base, np.float32)
    return np.array([lib().po_norm(_p(base[i]), base.shape[1]) for i in range(base.shape[0])], np.float32)


def distance_with_norm(v, q, nrm):
    v = np.ascontiguousarray(v, np.float32)
    q = np.ascontiguousarray(q, np.float32)
    return float(lib().po_distance_with_norm(_p(v), _p(q), v.shape[-1], float(np.float32(nrm))))


def prepare_init_ids(graph, L):
    out = np.empty(L, np.uint32)
    rc = lib().po_prepare_init_ids(ctypes.byref(graph.g), L, _p(out))
    if rc:
        raise ValueError("prepare_init_ids failed")
    return out


def search_one(graph, base, query, K, L, M, init_ids, mode="top_m", X=1):
    base = np.ascontiguousarray(base, np.float32)
    query = np.ascontiguousarray(query, np.float32)
    init_ids = np.ascontiguousarray(init_ids, np.uint32)
    set_L = np.zeros(L, CAND_DT)
    set_K = np.zeros(K, np.uint32)
    visited = np.zeros((graph.n + 63) // 64, np.uint64)
    c = Counters()
    f = lib().po_search_pss_v5_t1 if mode == "pss_v5" else lib().po_search_top_m
    f(ctypes.byref(graph.g), _p(base), base.shape[1], _p(query), K, L, X if mode == "pss_v5" else M,
      _p(init_ids), _p(set_L), _p(set_K), _p(visited), ctypes.byref(c))
    return {"ids": set_K, "queue": set_L, "ndist": c.distance_computations, "merges": c.merges,
            "nexp": c.expansions, "visited_clear": not visited.any()}


def search_pss_v5(graph, base, query, K, L, X, T, init_ids, local_L=None):
    """PSS_v5 with T queues in the oracle's lockstep order; returns the
    master queue and counters as search_one does, plus the whole set_L."""
    base = np.ascontiguousarray(base, np.float32)
    query = np.ascontiguousarray(query, np.float32)
    init_ids = np.ascontiguousarray(init_ids, np.uint32)
    local_L = L if local_L is None else local_L
    set_L = np.zeros((T - 1) * local_L + L, CAND_DT)
    set_K = np.zeros(K, np.uint32)
    visited = np.zeros((graph.n + 63) // 64, np.uint64)
    c = Counters()
    lib().po_search_pss_v5(ctypes.byref(graph.g), _p(base), base.shape[1], _p(query), K, L, X, T, local_L,
                           _p(init_ids), _p(set_L), _p(set_K), _p(visited), ctypes.byref(c))
    ms = (T - 1) * local_L
    return {"ids": set_K, "queue": set_L[ms:ms + L].copy(), "ndist": c.distance_computations, "merges": c.merges,
            "nexp": c.expansions, "visited_clear": not visited.any()}


def merge_two_queues_fixed(q1, q2):
    """q1, q2: CAND_DT arrays (q1 in place, its length fixed); returns the index."""
    return int(lib().po_merge_two_queues_fixed(_p(q1), 0, q1.shape[0], _p(q2), 0, q2.shape[0]))


def merge_two_queues_incr(q1, size1, q2):
    """q1: CAND_DT array of capacity len(q1) holding size1 entries (in place);
    returns (index, new size)."""
    s = u32(size1)
    r = lib().po_merge_two_queues_incr(_p(q1), 0, ctypes.byref(s), q1.shape[0], _p(q2), 0, q2.shape[0])
    return int(r), int(s.value)


def initialize_set_L(base, query, L, init_ids, n, norms=None):
    base = np.ascontiguousarray(base, np.float32)
    query = np.ascontiguousarray(query, np.float32)
    init_ids = np.ascontiguousarray(init_ids, np.uint32)
    set_L = np.zeros(L, CAND_DT)
    size = u32(0)
    visited = np.zeros((n + 63) // 64, np.uint64)
    cnt = u64(0)
    lib().po_initialize_set_L(_p(base), base.shape[1], _p(query), L, _p(set_L), 0, ctypes.byref(size),
                              _p(init_ids), _p(visited), ctypes.byref(cnt), _p(norms))
    return set_L, visited, cnt.value


def subsearch_with_top_m(graph, base, query, M_max, local_L, set_L, visited):
    """In place on set_L (CAND_DT, local_L entries) and visited (uint64 words)."""
    base = np.ascontiguousarray(base, np.float32)
    query = np.ascontiguousarray(query, np.float32)
    size = u32(local_L)
    cnt = u64(0)
    lib().po_subsearch_with_top_m(ctypes.byref(graph.g), _p(base), base.shape[1], _p(query), M_max,
                                  local_L, _p(set_L), 0, ctypes.byref(size), _p(visited),
                                  ctypes.byref(cnt))
    return cnt.value


def search_batch(graph, base, queries, K, L, M, init_ids, mode="top_m", nthreads=0):
    base = np.ascontiguousarray(base, np.float32)
    queries = np.ascontiguousarray(queries, np.float32)
    init_ids = np.ascontiguousarray(init_ids, np.uint32)
    nq = queries.shape[0]
    ids = np.empty((nq, K), np.uint32)
    dists = np.empty((nq, K), np.float32)
    nd = np.empty(nq, np.uint32)
    total = lib().po_search_batch(ctypes.byref(graph.g), _p(base), base.shape[1], _p(queries), nq, K, L, M,
                                  1 if mode == "pss_v5" else 0, _p(init_ids), _p(ids), _p(dists), _p(nd),
                                  nthreads)
    return ids, dists, nd, int(total)


def search_pss_v5_par(graph, base, queries, K, L, X, T, local_L=None):
    """The reference design for timing: PSS_v5 with T OpenMP threads per
    query, queries one after another (po_search_pss_v5_par).  Returns
    (ids [nq, K], distance counts [nq])."""
    base = np.ascontiguousarray(base, np.float32)
    queries = np.ascontiguousarray(queries, np.float32)
    nq = queries.shape[0]
    ids = np.empty((nq, K), np.uint32)
    nd = np.empty(nq, np.uint32)
    lib().po_search_pss_v5_par(ctypes.byref(graph.g), _p(base), base.shape[1], _p(queries), nq, K, L, X, T,
                               L if local_L is None else local_L, _p(ids), _p(nd))
    return ids, nd


def build_info():
    """The compiler string the oracle library was built with."""
    return lib().po_build_info().decode()


def brute_force_knn(base, queries, K, nthreads=0):
    base = np.ascontiguousarray(base, np.float32)
    queries = np.ascontiguousarray(queries, np.float32)
    nq = queries.shape[0]
    ids = np.empty((nq, K), np.uint32)
    dists = np.empty((nq, K), np.float32)
    lib().po_brute_force_knn(_p(base), base.shape[0], base.shape[1], _p(queries), nq, K, _p(ids), _p(dists),
                             nthreads)
    return ids, dists


def add_into_queue(queue, size, capacity, cand_id, cand_dist):
    """queue: CAND_DT array of length >= capacity; returns (ret, new_size)."""
    s = u32(size)
    c = Cand(cand_id, cand_dist, 0)
    r = lib().po_add_into_queue(_p(queue), 0, ctypes.byref(s), capacity, c)
    return int(r), int(s.value)
