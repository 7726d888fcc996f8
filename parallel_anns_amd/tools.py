"""Host tools around the search path: synthetic inputs, file formats, exact
brute-force k-NN (the recall judge) and NSG-style graph construction.  Thin
numpy wrappers over include/panns_tools.h."""
import numpy as np

from . import _native
from ._native import ptr

U32_NONE = 0xFFFFFFFF


class ToolsError(RuntimeError):
    pass


def _chk(rc, what):
    if rc != 0:
        msg = _native.tools().pt_last_error().decode(errors="replace")
        raise ToolsError(f"{what} failed ({rc}): {msg}")


def _f32(a):
    return np.ascontiguousarray(a, dtype=np.float32)


def _u32(a):
    return np.ascontiguousarray(a, dtype=np.uint32)


def l2(v, q):
    v, q = _f32(v), _f32(q)
    return float(_native.tools().pt_l2(ptr(v), ptr(q), v.shape[-1]))


def gen_mixture(n, dim, n_comp, sigma, seed_centers, seed_rows, row0=0, normalize=True, nthreads=0):
    """Rows [row0, row0+n) of the seeded normalised Gaussian mixture."""
    out = np.empty((n, dim), dtype=np.float32)
    _chk(_native.tools().pt_gen_mixture(seed_centers, seed_rows, row0, n, dim, n_comp, sigma,
                                        1 if normalize else 0, ptr(out), nthreads), "gen_mixture")
    return out


def mixture_labels(n, n_comp, seed_rows, row0=0, nthreads=0):
    """Mixture component of rows [row0, row0+n) of gen_mixture's stream."""
    out = np.empty(n, dtype=np.uint32)
    _chk(_native.tools().pt_mixture_labels(seed_rows, row0, n, n_comp, ptr(out), nthreads), "mixture_labels")
    return out


def fvecs_write(path, data):
    data = _f32(data)
    _chk(_native.tools().pt_fvecs_write(str(path).encode(), ptr(data), data.shape[0], data.shape[1]),
         "fvecs_write")


def fvecs_read(path):
    t = _native.tools()
    n, d = _native.c_u32(), _native.c_u32()
    _chk(t.pt_fvecs_info(str(path).encode(), n, d), "fvecs_info")
    out = np.empty((n.value, d.value), dtype=np.float32)
    _chk(t.pt_fvecs_read(str(path).encode(), ptr(out), n.value, d.value), "fvecs_read")
    return out


def nsg_write(path, adj, deg, ep, width=None):
    adj, deg = _u32(adj), _u32(deg)
    if width is None:
        width = int(deg.max()) if deg.size else 0
    _chk(_native.tools().pt_nsg_write(str(path).encode(), width, ep, adj.shape[0], adj.shape[1],
                                      ptr(adj), ptr(deg)), "nsg_write")


def nsg_read(path):
    """Returns (adj [n, max_deg] padded with 0xFFFFFFFF, deg, ep, width)."""
    t = _native.tools()
    w, ep, n, md = _native.c_u32(), _native.c_u32(), _native.c_u32(), _native.c_u32()
    ne = _native.c_u64()
    _chk(t.pt_nsg_info(str(path).encode(), w, ep, n, ne, md), "nsg_info")
    stride = max(md.value, 1)
    adj = np.full((n.value, stride), U32_NONE, dtype=np.uint32)
    deg = np.zeros(n.value, dtype=np.uint32)
    _chk(t.pt_nsg_read(str(path).encode(), n.value, stride, ptr(adj), ptr(deg)), "nsg_read")
    return adj, deg, ep.value, w.value


def truth_write(path, ids, dists):
    ids, dists = _u32(ids), _f32(dists)
    _chk(_native.tools().pt_truth_write(str(path).encode(), ids.shape[0], ids.shape[1], ptr(ids),
                                        ptr(dists)), "truth_write")


def truth_read(path):
    t = _native.tools()
    nq, k = _native.c_u32(), _native.c_u32()
    _chk(t.pt_truth_info(str(path).encode(), nq, k), "truth_info")
    ids = np.empty((nq.value, k.value), dtype=np.uint32)
    dists = np.empty((nq.value, k.value), dtype=np.float32)
    _chk(t.pt_truth_read(str(path).encode(), nq.value, k.value, ptr(ids), ptr(dists)), "truth_read")
    return ids, dists


def result_write(path, ids):
    ids = _u32(ids)
    _chk(_native.tools().pt_result_write(str(path).encode(), ids.shape[0], ids.shape[1], ptr(ids)),
         "result_write")


def exact_knn(base, queries, k, exclude_self=False, nthreads=0):
    """Exact k-NN with the reference distance; ties go to the lower id."""
    base, queries = _f32(base), _f32(queries)
    nq = queries.shape[0]
    ids = np.empty((nq, k), dtype=np.uint32)
    dists = np.empty((nq, k), dtype=np.float32)
    _chk(_native.tools().pt_exact_knn(ptr(base), base.shape[0], base.shape[1], ptr(queries), nq, k,
                                      1 if exclude_self else 0, ptr(ids), ptr(dists), nthreads),
         "exact_knn")
    return ids, dists


def recall(true_ids, res_ids, L, R=100):
    true_ids, res_ids = _u32(true_ids), _u32(res_ids)
    return float(_native.tools().pt_recall(ptr(true_ids), true_ids.shape[1], ptr(res_ids),
                                           res_ids.shape[1], res_ids.shape[0], L, R))


def recall_per_query(true_ids, res_ids, L, R=100):
    true_ids, res_ids = _u32(true_ids), _u32(res_ids)
    out = np.empty(res_ids.shape[0], dtype=np.float32)
    _chk(_native.tools().pt_recall_per_query(ptr(true_ids), true_ids.shape[1], ptr(res_ids),
                                             res_ids.shape[1], res_ids.shape[0], L, R, ptr(out)),
         "recall_per_query")
    return out


def navigating_node(base, nthreads=0):
    base = _f32(base)
    ep = _native.c_u32()
    _chk(_native.tools().pt_navigating_node(ptr(base), base.shape[0], base.shape[1], ep, nthreads),
         "navigating_node")
    return ep.value


def build_graph(base, knn_ids, R, L_build=40, C=500, ep=None, reverse_edges=True, nthreads=0):
    """NSG-style graph (kNN + search-path pool, RNG prune, reverse edges,
    DFS stitching with out-degree <= R).  L_build=0 and reverse_edges=False
    give the plain k-NN -> RNG prune -> DFS stitch graph.
    Returns (adj [n, R], deg, ep)."""
    base, knn_ids = _f32(base), _u32(knn_ids)
    n, dim = base.shape
    if ep is None:
        ep = navigating_node(base, nthreads)
    adj = np.full((n, R), U32_NONE, dtype=np.uint32)
    deg = np.zeros(n, dtype=np.uint32)
    _chk(_native.tools().pt_build_graph(ptr(base), n, dim, ptr(knn_ids), knn_ids.shape[1], R, L_build,
                                        C, ep, 1 if reverse_edges else 0, ptr(adj), ptr(deg), nthreads),
         "build_graph")
    return adj, deg, ep


def reachable(adj, deg, ep):
    adj, deg = _u32(adj), _u32(deg)
    r = _native.c_u32()
    _chk(_native.tools().pt_reachable(ptr(adj), ptr(deg), adj.shape[0], adj.shape[1], ep, r), "reachable")
    return r.value
