"""Python face of the MI355X search C-ABI (include/panns_amd.h).

Index mirrors the search-path methods of the reference's PANNS::Searching
(core/Searching.202011241615.top_m_pure_large_graph.h:142-352 and
core/Searching.202102031939.PSS_v5.large_graph.dist_thresh.profiling.h:30-268)
over ctypes; every call runs the HIP kernels and raises SearchError on a
non-zero status.  There is no CPU fallback.
"""
import numpy as np

from . import _native
from ._native import PannsCounters, PannsOutputs, PannsParams, ptr

MODES = {"top_m": 0, "pss_v5": 1, "scale_m": 2}
STOPS = {"fixed": 0, "learned": 1}


class SearchError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"panns error {code}: {msg}")
        self.code = code


def _chk(rc):
    if rc != 0:
        raise SearchError(rc, _native.amd().panns_last_error().decode(errors="replace"))


def make_params(K, L, M=1, mode="top_m", num_threads=1, stop="fixed", stop_threshold=0.5,
                hash_slots=0, waves=1, stop_min_iter=0, stop_every=1, bucket_slots=0):
    p = PannsParams()
    p.K, p.L, p.M = K, L, M
    p.mode = MODES[mode] if isinstance(mode, str) else int(mode)
    p.num_threads = num_threads
    p.stop = STOPS[stop] if isinstance(stop, str) else int(stop)
    p.stop_threshold = stop_threshold
    p.hash_slots = hash_slots
    p.waves = waves
    p.stop_min_iter, p.stop_every = stop_min_iter, stop_every
    p.bucket_slots = bucket_slots
    return p


def counters_dict(c):
    return {f: getattr(c, f) for f, _ in PannsCounters._fields_}


class Plan:
    """A resident batch: inputs in HBM, re-runnable, results fetched on demand."""

    def __init__(self, index, handle, nq, max_k, max_l, keeps_queue=True):
        self.index, self.h, self.nq, self.max_k, self.max_l = index, handle, nq, max_k, max_l
        self.keeps_queue = keeps_queue

    def run(self, sync=True):
        ms = _native.c_float(0.0)
        _chk(_native.amd().panns_plan_run(self.h, 1 if sync else 0, ms))
        return ms.value

    def describe(self):
        g, lds, hs, ml = (_native.c_u32() for _ in range(4))
        _chk(_native.amd().panns_plan_describe(self.h, g, lds, hs, ml))
        return {"workgroups": g.value, "lds_bytes": lds.value, "hash_slots": hs.value, "max_L": ml.value}

    def buckets(self):
        """Per queue-size bucket: largest L, queries, workgroups, LDS bytes,
        visited-table slots."""
        n = _native.c_u32()
        arrs = [np.zeros(_native.MAX_BUCKETS, np.uint32) for _ in range(5)]
        _chk(_native.amd().panns_plan_describe_buckets(self.h, n, *(ptr(a) for a in arrs)))
        keys = ("max_L", "nq", "workgroups", "lds_bytes", "hash_slots")
        return [{k: int(a[b]) for k, a in zip(keys, arrs)} for b in range(n.value)]

    def fetch(self, want_queue=False):
        if want_queue and not self.keeps_queue:
            raise ValueError("this plan was created with want_queue=False")
        nq, K, L = self.nq, self.max_k, self.max_l
        res = {"ids": np.empty((nq, K), np.uint32), "dists": np.empty((nq, K), np.float32),
               "ndist": np.empty(nq, np.uint32), "nexp": np.empty(nq, np.uint32),
               "niter": np.empty(nq, np.uint32)}
        if want_queue:
            res["queue_ids"] = np.empty((nq, L), np.uint32)
            res["queue_dists"] = np.empty((nq, L), np.float32)
            res["queue_checked"] = np.empty((nq, L), np.uint8)
        o = PannsOutputs()
        for f, _ in PannsOutputs._fields_:
            setattr(o, f, ptr(res[f]) if f in res else None)
        c = PannsCounters()
        _chk(_native.amd().panns_plan_fetch(self.h, o, c))
        res["counters"] = counters_dict(c)
        return res

    PHASES = ("init", "pick", "claim", "distance", "merge", "stop", "output", "idle")

    def phase_cycles(self):
        """Cycles per search phase of the last run (timer builds only)."""
        out = np.zeros(8, np.uint64)
        _chk(_native.amd().panns_plan_phase_cycles(self.h, ptr(out)))
        return dict(zip(self.PHASES, (int(x) for x in out)))

    def close(self):
        if self.h:
            _native.amd().panns_plan_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Index:
    def __init__(self, device=0):
        lib = _native.amd()
        h = _native.vp()
        _chk(lib.panns_create(device, h))
        self.h = h
        self.lib = lib

    def close(self):
        if getattr(self, "h", None):
            self.lib.panns_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # loaders (PSS5LG.cpp:14-49, :138-251)
    def load_data_load(self, path):
        _chk(self.lib.panns_load_data_load(self.h, str(path).encode()))

    def load_queries_load(self, path):
        _chk(self.lib.panns_load_queries_load(self.h, str(path).encode()))

    def load_common_nsg_graph(self, path):
        _chk(self.lib.panns_load_common_nsg_graph(self.h, str(path).encode()))

    # the opt-graph family (PSS5.cpp:56-132): same file, norm - 2 v.q distances
    def load_nsg_graph(self, path):
        _chk(self.lib.panns_load_nsg_graph(self.h, str(path).encode()))

    DISTANCES = {"l2": 0, "norm_dot": 1}

    def set_distance(self, name):
        """"l2" (the large-graph searches) or "norm_dot" (the opt-graph
        searches: norm(v) - 2 v.q, PSS5.cpp:336-390)."""
        if name not in self.DISTANCES:
            raise SearchError(-1, f"unknown distance {name!r}")
        _chk(self.lib.panns_set_distance(self.h, self.DISTANCES[name]))

    @property
    def distance(self):
        v = _native.c_int()
        _chk(self.lib.panns_get_distance(self.h, v))
        return {0: "l2", 1: "norm_dot"}[v.value]

    def norms(self):
        """Every base row's norm, compute_norm (PSS5.cpp:293-334)."""
        n = self.info()["num_v"]
        out = np.empty(n, dtype=np.float32)
        _chk(self.lib.panns_get_norms(self.h, ptr(out)))
        return out

    def set_base(self, data):
        data = np.ascontiguousarray(data, dtype=np.float32)
        _chk(self.lib.panns_set_base(self.h, ptr(data), data.shape[0], data.shape[1]))

    def set_queries(self, data):
        data = np.ascontiguousarray(data, dtype=np.float32)
        _chk(self.lib.panns_set_queries(self.h, ptr(data), data.shape[0], data.shape[1]))

    def set_graph(self, adj, deg, ep, width=None):
        adj = np.ascontiguousarray(adj, dtype=np.uint32)
        deg = np.ascontiguousarray(deg, dtype=np.uint32)
        if width is None:
            width = int(deg.max())
        _chk(self.lib.panns_set_graph(self.h, width, ep, adj.shape[0], adj.shape[1], ptr(adj), ptr(deg)))

    def info(self):
        v = [_native.c_u32() for _ in range(5)]
        ne = _native.c_u64()
        _chk(self.lib.panns_get_info(self.h, *v, ne))
        return {"num_v": v[0].value, "dim": v[1].value, "num_queries": v[2].value,
                "width": v[3].value, "ep": v[4].value, "num_e": ne.value}

    def prepare_init_ids(self, L):
        out = np.empty(L, dtype=np.uint32)
        _chk(self.lib.panns_prepare_init_ids(self.h, L, ptr(out)))
        return out

    def search(self, query_ids, K, L, M=1, mode="top_m", init_ids=None, want_queue=False,
               num_threads=1, stop="fixed", stop_threshold=0.5, hash_slots=0, waves=1, stop_min_iter=0,
               stop_every=1, bucket_slots=0, no_alloc=False):
        """Batch form of para_search_with_top_m_pure_large_graph /
        para_search_PSS_v5_large_graph_dist_thresh_profiling; mode "pss_v5"
        takes M as X and num_threads as the reference's num_threads_ (its
        T queues run in lockstep in a fixed order).  no_alloc: the
        allocation-free entry point (panns_search_noalloc), whose device
        buffers come from the library's fixed scratch array."""
        qids = np.ascontiguousarray(np.atleast_1d(query_ids), dtype=np.uint32)
        nq = qids.shape[0]
        if init_ids is None:
            init_ids = self.prepare_init_ids(L)
        init_ids = np.ascontiguousarray(init_ids, dtype=np.uint32)
        p = make_params(K, L, M, mode, num_threads, stop, stop_threshold, hash_slots, waves, stop_min_iter,
                        stop_every, bucket_slots)
        res = {"ids": np.empty((nq, K), np.uint32), "dists": np.empty((nq, K), np.float32),
               "ndist": np.empty(nq, np.uint32), "nexp": np.empty(nq, np.uint32),
               "niter": np.empty(nq, np.uint32)}
        if want_queue:
            res["queue_ids"] = np.empty((nq, L), np.uint32)
            res["queue_dists"] = np.empty((nq, L), np.float32)
            res["queue_checked"] = np.empty((nq, L), np.uint8)
        o = PannsOutputs()
        for f, _ in PannsOutputs._fields_:
            setattr(o, f, ptr(res[f]) if f in res else None)
        c = PannsCounters()
        f = self.lib.panns_search_noalloc if no_alloc else self.lib.panns_search
        _chk(f(self.h, p, ptr(qids), nq, ptr(init_ids), init_ids.shape[0], o, c))
        res["counters"] = counters_dict(c)
        return res

    @staticmethod
    def _batch_args(query_ids, K, L, L_per_query, K_per_query):
        qids = np.ascontiguousarray(np.atleast_1d(query_ids), dtype=np.uint32)
        Lq = None if L_per_query is None else np.ascontiguousarray(L_per_query, dtype=np.uint32)
        Kq = None if K_per_query is None else np.ascontiguousarray(K_per_query, dtype=np.uint32)
        maxL = int(Lq.max()) if Lq is not None else L
        maxK = int(Kq.max()) if Kq is not None else K
        return qids, Lq, Kq, maxL, maxK

    @staticmethod
    def plan_options(buckets=1, bucket_max_L=None, table_eighths=8, use_graph=False, concurrent=False,
                     want_queue=False, waves=0, rows_per_wave=0, overflow_slots=0, deferred_merge=0,
                     visited_table=0):
        o = _native.PannsPlanOptions()
        o.n_buckets = len(bucket_max_L) if bucket_max_L else buckets
        for i, v in enumerate(bucket_max_L or ()):
            o.bucket_max_L[i] = int(v)
        o.table_eighths, o.use_graph, o.concurrent = table_eighths, int(use_graph), int(concurrent)
        o.want_queue = int(want_queue)
        o.waves = waves
        o.rows_per_wave = rows_per_wave
        o.overflow_slots = overflow_slots
        o.deferred_merge = deferred_merge
        o.visited_table = visited_table
        return o

    def plan(self, query_ids, K, L, M=1, mode="top_m", init_ids=None, L_per_query=None,
             K_per_query=None, stop="fixed", stop_threshold=0.5, hash_slots=0, options=None,
             want_queue=True, waves=1, stop_min_iter=0, stop_every=1, bucket_slots=0, num_threads=1):
        """A resident batch (panns_plan_create_ex).  options: a
        PannsPlanOptions (plan_options() or autotune()) for queue-size
        buckets, graph replay and side-by-side buckets; want_queue keeps the
        final queues for fetch(want_queue=True)."""
        qids, Lq, Kq, maxL, maxK = self._batch_args(query_ids, K, L, L_per_query, K_per_query)
        if init_ids is None:
            init_ids = self.prepare_init_ids(maxL)
        init_ids = np.ascontiguousarray(init_ids, dtype=np.uint32)
        p = make_params(K, L, M, mode, num_threads, stop, stop_threshold, hash_slots, waves, stop_min_iter,
                        stop_every, bucket_slots)
        o = _native.PannsPlanOptions() if options is None else options
        o.want_queue = int(want_queue)
        h = _native.vp()
        _chk(self.lib.panns_plan_create_ex(self.h, p, ptr(qids), qids.shape[0], ptr(Lq), ptr(Kq),
                                           ptr(init_ids), init_ids.shape[0], o, h))
        return Plan(self, h, qids.shape[0], maxK, maxL, keeps_queue=bool(want_queue))

    def autotune(self, query_ids, K, L, M=1, mode="top_m", init_ids=None, L_per_query=None,
                 K_per_query=None, stop="fixed", stop_threshold=0.5, sample_stride=0, stop_min_iter=0,
                 stop_every=1, num_threads=1):
        """The fastest bucketing for this batch (panns_plan_autotune): returns
        (PannsPlanOptions, device ms on the sample)."""
        qids, Lq, Kq, maxL, maxK = self._batch_args(query_ids, K, L, L_per_query, K_per_query)
        if init_ids is None:
            init_ids = self.prepare_init_ids(maxL)
        init_ids = np.ascontiguousarray(init_ids, dtype=np.uint32)
        p = make_params(K, L, M, mode, num_threads, stop, stop_threshold, 0, 1, stop_min_iter, stop_every)
        best = _native.PannsPlanOptions()
        ms = _native.c_float(0.0)
        _chk(self.lib.panns_plan_autotune(self.h, p, ptr(qids), qids.shape[0], ptr(Lq), ptr(Kq),
                                          ptr(init_ids), init_ids.shape[0], sample_stride, best, ms))
        return best, ms.value

    def search_trace(self, query_ids, K, L, M=1, mode="top_m", init_ids=None, trace_cap=256, gt_ids=None):
        """Fixed-L search that also records the stop features and the top-K
        hash after every iteration; returns (feat [nq, cap, 8], topk_hash
        [nq, cap], niter [nq], ids [nq, K]).  With gt_ids (the loaded
        queries' true neighbours, one row per query id) it also returns
        hits [nq, cap], recall@100's count after every iteration, and ndist
        [nq, cap], the distance computations so far, as fifth and sixth
        elements."""
        qids = np.ascontiguousarray(np.atleast_1d(query_ids), dtype=np.uint32)
        nq = qids.shape[0]
        if init_ids is None:
            init_ids = self.prepare_init_ids(L)
        init_ids = np.ascontiguousarray(init_ids, dtype=np.uint32)
        p = make_params(K, L, M, mode)
        feat = np.zeros((nq, trace_cap, 8), np.float32)
        hsh = np.zeros((nq, trace_cap), np.uint32)
        niter = np.zeros(nq, np.uint32)
        ids = np.zeros((nq, K), np.uint32)
        if gt_ids is None:
            _chk(self.lib.panns_search_trace(self.h, p, ptr(qids), nq, ptr(init_ids), init_ids.shape[0],
                                             trace_cap, ptr(feat), ptr(hsh), ptr(niter), ptr(ids)))
            return feat, hsh, niter, ids
        gt = np.ascontiguousarray(gt_ids, dtype=np.uint32)
        if gt.shape[0] != self.info()["num_queries"]:
            raise ValueError("gt_ids needs one row per loaded query")
        hits = np.zeros((nq, trace_cap), np.uint16)
        nd = np.zeros((nq, trace_cap), np.uint32)
        _chk(self.lib.panns_search_trace_hits(self.h, p, ptr(qids), nq, ptr(init_ids), init_ids.shape[0],
                                              trace_cap, ptr(gt), gt.shape[1], ptr(feat), ptr(hsh), ptr(hits),
                                              ptr(nd), ptr(niter), ptr(ids)))
        return feat, hsh, niter, ids, hits, nd

    def set_stop_model(self, params):
        params = np.ascontiguousarray(params, dtype=np.float32).ravel()
        _chk(self.lib.panns_set_stop_model(self.h, ptr(params), params.shape[0]))

    def mlp_train(self, X, Y, params, epochs=4, batch=4096, lr=2e-3, seed=0):
        """Trains the stop model on the GPU (BF16 MFMA kernels); returns
        (trained params, per-epoch mean loss)."""
        X = np.ascontiguousarray(X, np.float32)
        Y = np.ascontiguousarray(Y, np.float32)
        P = np.array(params, np.float32, copy=True)
        losses = np.zeros(epochs, np.float32)
        _chk(self.lib.panns_mlp_train(self.h, ptr(X), ptr(Y), X.shape[0], ptr(P), epochs, batch, lr, seed,
                                      ptr(losses)))
        return P, losses

    def mlp_grad(self, params, X, Y):
        X = np.ascontiguousarray(X, np.float32)
        Y = np.ascontiguousarray(Y, np.float32)
        P = np.ascontiguousarray(params, np.float32)
        g = np.zeros(4801, np.float32)
        loss = _native.c_float()
        _chk(self.lib.panns_mlp_grad(self.h, ptr(P), ptr(X), ptr(Y), X.shape[0], ptr(g), loss))
        return g, loss.value

    def mfma_probe(self, A, B):
        A = np.ascontiguousarray(A, np.float32)
        B = np.ascontiguousarray(B, np.float32)
        C = np.zeros((16, 16), np.float32)
        _chk(self.lib.panns_mfma_probe(self.h, ptr(A), ptr(B), ptr(C)))
        return C

    def device_copy(self, nbytes):
        """Device time (ms) of a plain streaming copy of nbytes."""
        ms = _native.c_float()
        _chk(self.lib.panns_device_copy(self.h, nbytes, ms))
        return ms.value

    def exact_knn(self, k, source="queries", method="auto", stats=None):
        """Exact k-NN on the GPU (reference distance, ties to the lower id).
        source="queries": of the loaded queries (ground truth); "base": of
        every base row with itself excluded (graph k-NN lists).  method
        "valu": the exact chain for every pair; "mfma": an FP32 MFMA filter
        bounded by the exact k-th distance over a sample of the rows, then
        the exact chain on the candidates; "auto": mfma from 8,192 base
        rows.  stats (a dict) receives the number of queries recomputed with
        the exact chain over every row (candidate lists that overflowed)."""
        info = self.info()
        nq = info["num_queries"] if source == "queries" else info["num_v"]
        ids = np.empty((nq, k), np.uint32)
        dists = np.empty((nq, k), np.float32)
        redo = _native.c_u32(0)
        _chk(self.lib.panns_exact_knn_ex(self.h, 1 if source == "queries" else 0, k,
                                         {"auto": 0, "valu": 1, "mfma": 2}[method], ptr(ids), ptr(dists), redo))
        if stats is not None:
            stats["recomputed"] = redo.value
        return ids, dists

    def mfma_f32_probe(self, A, B):
        """C (16x16) = A (16x4) @ B (4x16) by one FP32 MFMA tile."""
        A = np.ascontiguousarray(A, np.float32)
        B = np.ascontiguousarray(B, np.float32)
        C = np.zeros((16, 16), np.float32)
        _chk(self.lib.panns_mfma_f32_probe(self.h, ptr(A), ptr(B), ptr(C)))
        return C

    def subsearch_with_top_m(self, M_max, query_id, local_L, queue_ids, queue_dists, queue_checked,
                             visited_words):
        """TOPMLG.cpp:1023-1071 on a caller-held full queue; arrays are updated in place."""
        for a, dt in ((queue_ids, np.uint32), (queue_dists, np.float32), (queue_checked, np.uint8),
                      (visited_words, np.uint64)):
            if a.dtype != dt or not a.flags["C_CONTIGUOUS"]:
                raise TypeError("subsearch arrays must be C-contiguous with the documented dtypes")
        size = _native.c_u32(local_L)
        cnt = _native.c_u64(0)
        _chk(self.lib.panns_subsearch_with_top_m(self.h, M_max, query_id, local_L, ptr(queue_ids),
                                                 ptr(queue_dists), ptr(queue_checked), size,
                                                 ptr(visited_words), cnt))
        return cnt.value
