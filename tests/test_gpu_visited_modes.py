"""Visited-set forms (panns_plan_options.visited_table): with no LDS table
every visited id is claimed in the per-workgroup bitmap (logged words
cleared between queries, the whole bitmap past the log's capacity).  The
results must be those of the table form and the oracle's: ids, distance
bits, final queues with checked flags, distance counts, iterations."""
import numpy as np
import pytest

import pyoracle as po
from parallel_anns_amd.search import Index

pytestmark = pytest.mark.gpu

FIELDS = ("ids", "dists", "ndist", "nexp", "niter", "queue_ids", "queue_dists", "queue_checked")


def _run(idx, q, K, L, M, table, defer=0, **kw):
    pl = idx.plan(q, K, L, M, options=Index.plan_options(visited_table=table, deferred_merge=defer), **kw)
    assert (pl.describe()["hash_slots"] == 0) == (table == 2)
    pl.run()
    pl.run()  # the second query on each workgroup's bitmap starts from the first's clear
    r = pl.fetch(want_queue=True)
    pl.close()
    return r


def _same(a, b):
    for f in FIELDS:
        x, y = a[f], b[f]
        if x.dtype == np.float32:
            x, y = x.view(np.uint32), y.view(np.uint32)
        assert (x == y).all(), f


@pytest.mark.parametrize("L,M,defer", [(100, 1, 0), (300, 4, 0), (900, 4, 2), (2000, 4, 2), (2000, 16, 1)])
def test_bitmap_only_equals_table_and_oracle(small_set, gpu_index, L, M, defer):
    ds = small_set
    q = np.arange(ds["queries"].shape[0], dtype=np.uint32)
    K = 100
    print("case bitmap-only", L, M, defer, flush=True)
    a = _run(gpu_index, q, K, L, M, 2, defer)
    b = _run(gpu_index, q, K, L, M, 1, defer)
    _same(a, b)
    g = po.CsrGraph(ds["adj"], ds["deg"], ds["ep"])
    init = gpu_index.prepare_init_ids(L)
    for i in range(0, q.shape[0], 40):
        r = po.search_one(g, ds["base"], ds["queries"][i], K, L, M, init)
        assert (a["ids"][i] == r["ids"]).all(), i
        assert (a["queue_ids"][i] == r["queue"]["id"]).all(), i
        assert (a["queue_checked"][i] == r["queue"]["checked"]).all(), i
        assert a["ndist"][i] == r["ndist"] and a["niter"][i] == r["merges"], i


def test_bitmap_only_uneven_buckets(small_set, gpu_index):
    ds = small_set
    q = np.arange(ds["queries"].shape[0], dtype=np.uint32)
    rng = np.random.default_rng(9)
    Ls = rng.integers(100, 2001, q.shape[0]).astype(np.uint32)
    Ks = rng.choice(np.array([10, 100], np.uint32), q.shape[0])
    init = gpu_index.prepare_init_ids(int(Ls.max()))
    res = []
    for table in (2, 1, 0):
        print("case bitmap-only uneven", table, flush=True)
        pl = gpu_index.plan(q, 100, int(Ls.max()), 4, init_ids=init, L_per_query=Ls, K_per_query=Ks,
                            options=Index.plan_options(buckets=3, use_graph=True, concurrent=True,
                                                       visited_table=table))
        assert all((b["hash_slots"] == 0) == (table == 2) for b in pl.buckets())
        pl.run()
        pl.run()
        res.append(pl.fetch())
        pl.close()
    for r in res[1:]:
        for f in ("ids", "ndist", "niter", "nexp"):
            assert (r[f] == res[0][f]).all(), f
        assert (r["dists"].view(np.uint32) == res[0]["dists"].view(np.uint32)).all()


@pytest.mark.parametrize("waves,mode", [(2, "top_m"), (4, "scale_m")])
def test_bitmap_only_multi_wave(small_set, gpu_index, waves, mode):
    q = np.arange(small_set["queries"].shape[0], dtype=np.uint32)
    print("case bitmap-only waves", waves, mode, flush=True)
    _same(_run(gpu_index, q, 100, 600, 16, 2, waves=waves, mode=mode),
          _run(gpu_index, q, 100, 600, 16, 1, waves=waves, mode=mode))
