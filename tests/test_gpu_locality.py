"""Batches of 512 queries or more run in locality order (queries sorted by
their nearest pivot row, per-XCD work ranges); every query's results must
be those of running it in a small batch (caller order) and the oracle's."""
import numpy as np
import pytest

import pyoracle as po
from parallel_anns_amd import tools
from parallel_anns_amd.search import Index

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def big_batch(small_set):
    ds = small_set
    q = tools.gen_mixture(1500, ds["base"].shape[1], ds["cfg"]["comps"], ds["cfg"]["sigma"], 0, 7)
    idx = Index(0)
    idx.set_base(ds["base"])
    idx.set_queries(q)
    idx.set_graph(ds["adj"], ds["deg"], ds["ep"])
    yield idx, q
    idx.close()


def test_reordered_batch_equals_small_batches(small_set, big_batch):
    idx, q = big_batch
    nq = q.shape[0]
    K, L, M = 100, 120, 4
    init = idx.prepare_init_ids(L)
    rng = np.random.default_rng(1)
    qids = rng.permutation(nq).astype(np.uint32)  # any caller order
    print("case locality batch", nq, flush=True)
    big = idx.search(qids, K, L, M, init_ids=init, want_queue=True)
    for s in range(0, nq, 100):
        part = idx.search(qids[s:s + 100], K, L, M, init_ids=init, want_queue=True)
        for k in ("ids", "dists", "queue_ids", "queue_dists", "queue_checked", "ndist", "nexp", "niter"):
            a, b = big[k][s:s + 100], part[k]
            if a.dtype == np.float32:
                a, b = a.view(np.uint32), b.view(np.uint32)
            assert (a == b).all(), (k, s)
    ds = small_set
    g = po.CsrGraph(ds["adj"], ds["deg"], ds["ep"])
    for i in range(0, nq, 97):
        r = po.search_one(g, ds["base"], q[qids[i]], K, L, M, init)
        assert (big["ids"][i] == r["ids"]).all() and big["ndist"][i] == r["ndist"]


def test_reordered_uneven_buckets(big_batch):
    idx, q = big_batch
    nq = q.shape[0]
    rng = np.random.default_rng(2)
    qids = np.arange(nq, dtype=np.uint32)
    Ls = rng.integers(100, 400, nq).astype(np.uint32)
    Ks = np.minimum(rng.choice(np.array([10, 100], np.uint32), nq), Ls)
    init = idx.prepare_init_ids(int(Ls.max()))
    pl = idx.plan(qids, 100, int(Ls.max()), 2, init_ids=init, L_per_query=Ls, K_per_query=Ks, want_queue=False,
                  options=Index.plan_options(buckets=3, use_graph=True, concurrent=True))
    pl.run()
    res = pl.fetch()
    pl.close()
    for s in range(0, nq, 250):
        sl = slice(s, s + 250)
        # small plans (under the reordering size) in caller order
        p2 = idx.plan(qids[sl], 100, int(Ls[sl].max()), 2, init_ids=init, L_per_query=Ls[sl], K_per_query=Ks[sl],
                      want_queue=False)
        p2.run()
        r2 = p2.fetch()
        p2.close()
        assert (res["ids"][sl] == r2["ids"]).all() and (res["ndist"][sl] == r2["ndist"]).all(), s
