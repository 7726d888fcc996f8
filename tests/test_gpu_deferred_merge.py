"""The deferred merge (panns_plan_options.deferred_merge): survivors wait in
a sorted 256-key buffer and are merged into the queue when it fills or a
pick could rank outside the top L.  Results must be those of the search
that merges every iteration -- ids, distance bits, final queue with checked
flags, distance counts, expansions and iterations -- and the oracle's."""
import numpy as np
import pytest

import pyoracle as po
from parallel_anns_amd.search import Index

pytestmark = pytest.mark.gpu

FIELDS = ("ids", "dists", "ndist", "nexp", "niter", "queue_ids", "queue_dists", "queue_checked")


def _run(idx, q, K, L, M, defer, **kw):
    pl = idx.plan(q, K, L, M, options=Index.plan_options(deferred_merge=defer), **kw)
    pl.run()
    r = pl.fetch(want_queue=True)
    pl.close()
    return r


def _same(a, b):
    for f in FIELDS:
        x, y = a[f], b[f]
        if x.dtype == np.float32:
            x, y = x.view(np.uint32), y.view(np.uint32)
        assert (x == y).all(), f


@pytest.mark.parametrize("L,M", [(100, 1), (100, 4), (200, 2), (200, 8), (600, 4), (600, 64), (1000, 16)])
def test_deferred_equals_every_step_merge(small_set, gpu_index, L, M):
    ds = small_set
    q = np.arange(ds["queries"].shape[0], dtype=np.uint32)
    K = min(100, L)
    print("case deferred", L, M, flush=True)
    a = _run(gpu_index, q, K, L, M, 2)
    b = _run(gpu_index, q, K, L, M, 1)
    _same(a, b)
    g = po.CsrGraph(ds["adj"], ds["deg"], ds["ep"])
    init = gpu_index.prepare_init_ids(L)
    for i in range(0, q.shape[0], 25):
        r = po.search_one(g, ds["base"], ds["queries"][i], K, L, M, init)
        assert (a["queue_ids"][i] == r["queue"]["id"]).all()
        assert (a["queue_checked"][i] == r["queue"]["checked"]).all()
        assert a["ndist"][i] == r["ndist"] and a["niter"][i] == r["merges"]


def test_deferred_with_spills_modes_and_buckets(small_set, gpu_index):
    ds = small_set
    q = np.arange(ds["queries"].shape[0], dtype=np.uint32)
    print("case deferred spills", flush=True)
    _same(_run(gpu_index, q, 100, 300, 4, 2, hash_slots=64), _run(gpu_index, q, 100, 300, 4, 1, hash_slots=64))
    print("case deferred scale_m / pss_v5", flush=True)
    _same(_run(gpu_index, q, 50, 250, 32, 2, mode="scale_m"), _run(gpu_index, q, 50, 250, 32, 1, mode="scale_m"))
    _same(_run(gpu_index, q, 50, 250, 3, 2, mode="pss_v5"), _run(gpu_index, q, 50, 250, 3, 1, mode="pss_v5"))
    # uneven L in buckets, automatic choice (deferred above L = 1,000)
    rng = np.random.default_rng(5)
    Ls = rng.integers(100, 2001, q.shape[0]).astype(np.uint32)
    init = gpu_index.prepare_init_ids(int(Ls.max()))
    res = []
    for defer in (0, 1):
        pl = gpu_index.plan(q, 100, int(Ls.max()), 4, init_ids=init, L_per_query=Ls,
                            options=Index.plan_options(buckets=3, use_graph=True, deferred_merge=defer))
        pl.run()
        pl.run()
        res.append(pl.fetch())
        pl.close()
    for f in ("ids", "ndist", "niter"):
        assert (res[0][f] == res[1][f]).all(), f


def test_deferred_learned_stop_equals_every_step_merge(small_set, gpu_index):
    """The learned-stop search with the deferred merge: the pending keys are
    merged before every iteration where the model runs, so its features, its
    stops and all results equal those of the search that merges every
    iteration -- at every evaluation stride and with spills."""
    from parallel_anns_amd import learned
    ds = small_set
    q = np.arange(ds["queries"].shape[0], dtype=np.uint32)
    K, L, M = 100, 600, 4
    feat, hsh, niter, _ = gpu_index.search_trace(q, K, L, M, trace_cap=256)
    X, Y = learned.samples_from_trace(feat, hsh, niter)
    P, _ = learned.train(gpu_index, X, Y, epochs=2, batch=1024)
    gpu_index.set_stop_model(P)
    stops = 0
    for thr, every, mi, hs in ((0.3, 1, 0, 0), (0.5, 4, 20, 0), (0.5, 3, 7, 64), (0.7, 2, 0, 0)):
        print("case deferred learned", thr, every, mi, hs, flush=True)
        kw = dict(stop="learned", stop_threshold=thr, stop_every=every, stop_min_iter=mi, hash_slots=hs)
        a = _run(gpu_index, q, K, L, M, 2, **kw)
        b = _run(gpu_index, q, K, L, M, 1, **kw)
        _same(a, b)
        stops += int((a["niter"] < niter).sum())
    assert stops > 0  # the model stopped some queries early
