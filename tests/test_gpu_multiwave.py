"""Several waves per query (the deterministic form of the reference's
multi-thread search and its merge of worker queues into the master queue,
PSS5LG.cpp:811-842 and :1543-1673): the results, final queues and counters
equal the single-wave search, which equals the oracle."""
import numpy as np
import pytest

import pyoracle as po
from parallel_anns_amd.search import Index, SearchError

pytestmark = pytest.mark.gpu


def _same(a, b, keys):
    for k in keys:
        x, y = a[k], b[k]
        if x.dtype == np.float32:
            x, y = x.view(np.uint32), y.view(np.uint32)
        assert (x == y).all(), k


@pytest.mark.parametrize("waves", [2, 4])
@pytest.mark.parametrize("L,M,mode", [(100, 4, "top_m"), (160, 8, "top_m"), (200, 16, "top_m"),
                                      (128, 3, "top_m"), (150, 8, "scale_m")])
def test_multiwave_equals_single_wave(small_set, gpu_index, waves, L, M, mode):
    ds = small_set
    K = min(100, L)
    qids = np.arange(96, dtype=np.uint32)
    print("case multiwave", waves, L, M, mode, flush=True)
    init = gpu_index.prepare_init_ids(L)
    one = gpu_index.search(qids, K, L, M, mode=mode, init_ids=init, want_queue=True)
    many = gpu_index.search(qids, K, L, M, mode=mode, init_ids=init, want_queue=True, waves=waves)
    _same(many, one, ("ids", "dists", "queue_ids", "queue_dists", "queue_checked", "ndist", "nexp", "niter"))
    assert many["counters"]["distance_computations"] == one["counters"]["distance_computations"]
    assert many["counters"]["adjacency_slots"] == one["counters"]["adjacency_slots"]
    if mode == "top_m":
        g = po.CsrGraph(ds["adj"], ds["deg"], ds["ep"])
        for i in range(0, qids.shape[0], 12):
            r = po.search_one(g, ds["base"], ds["queries"][qids[i]], K, L, M, init)
            assert (many["ids"][i] == r["ids"]).all() and many["ndist"][i] == r["ndist"]


@pytest.mark.parametrize("waves", [2, 4])
def test_multiwave_spills_and_buckets(small_set, gpu_index, waves):
    """Small tables (spills logged per wave) and an uneven bucketed batch."""
    rng = np.random.default_rng(3)
    nq = small_set["queries"].shape[0]
    qids = np.arange(nq, dtype=np.uint32)
    init = gpu_index.prepare_init_ids(300)
    a = gpu_index.search(qids[:48], 100, 120, 8, init_ids=init, hash_slots=128)
    b = gpu_index.search(qids[:48], 100, 120, 8, init_ids=init, hash_slots=128, waves=waves)
    assert b["counters"]["visited_spills"] > 0
    _same(b, a, ("ids", "dists", "ndist", "nexp", "niter"))
    Ls = rng.integers(100, 301, nq).astype(np.uint32)
    Ks = np.minimum(rng.choice(np.array([10, 100], np.uint32), nq), Ls)
    ref = gpu_index.plan(qids, 100, 300, 4, init_ids=init, L_per_query=Ls, K_per_query=Ks, want_queue=False)
    ref.run()
    r0 = ref.fetch()
    ref.close()
    pl = gpu_index.plan(qids, 100, 300, 4, init_ids=init, L_per_query=Ls, K_per_query=Ks, want_queue=False,
                        options=Index.plan_options(buckets=3, use_graph=True, concurrent=True, waves=waves))
    for _ in range(2):
        pl.run()
        _same(pl.fetch(), r0, ("ids", "dists", "ndist", "nexp", "niter"))
    pl.close()


def test_multiwave_rejections(gpu_index):
    q = np.arange(4, dtype=np.uint32)
    with pytest.raises(SearchError):
        gpu_index.search(q, 10, 50, 4, waves=3)
    with pytest.raises(SearchError):
        gpu_index.search(q, 10, 50, 2, mode="pss_v5", waves=2)
    with pytest.raises(SearchError):
        gpu_index.search(q, 10, 50, 4, stop="learned", stop_threshold=0.5, waves=2)
