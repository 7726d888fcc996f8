"""Uneven batches: queue-size buckets, hipGraph replay, side-by-side buckets
and the autotuner give exactly the single-plan results (and the oracle's)
for queries of different L and K."""
import numpy as np
import pytest

import pyoracle as po
from parallel_anns_amd.search import Index, SearchError

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def uneven(small_set):
    rng = np.random.default_rng(7)
    nq = small_set["queries"].shape[0]
    qids = np.arange(nq, dtype=np.uint32)
    Ls = np.exp(rng.uniform(np.log(100), np.log(600), nq)).astype(np.uint32)
    Ks = rng.choice(np.array([10, 100], np.uint32), nq)
    Ks = np.minimum(Ks, Ls)
    return qids, Ls, Ks


def _same(a, b, keys=("ids", "dists", "ndist", "nexp", "niter")):
    for k in keys:
        x, y = a[k], b[k]
        if x.dtype == np.float32:
            x, y = x.view(np.uint32), y.view(np.uint32)
        assert (x == y).all(), k


def test_buckets_equal_single_plan_and_oracle(small_set, gpu_index, uneven):
    ds = small_set
    qids, Ls, Ks = uneven
    M = 4
    init = gpu_index.prepare_init_ids(int(Ls.max()))
    single = gpu_index.plan(qids, 100, int(Ls.max()), M, init_ids=init, L_per_query=Ls, K_per_query=Ks)
    single.run()
    ref = single.fetch(want_queue=True)
    single.close()
    g = po.CsrGraph(ds["adj"], ds["deg"], ds["ep"])
    for i in range(0, qids.shape[0], 9):
        o = po.search_one(g, ds["base"], ds["queries"][qids[i]], int(Ks[i]), int(Ls[i]), M, init[:Ls[i]])
        assert (ref["ids"][i, :Ks[i]] == o["ids"]).all(), f"query {i}"
        assert (ref["ids"][i, Ks[i]:] == 0xFFFFFFFF).all()
        assert ref["ndist"][i] == o["ndist"]
    for opts in (dict(buckets=3), dict(buckets=3, use_graph=True), dict(buckets=4, concurrent=True),
                 dict(buckets=2, use_graph=True, concurrent=True, table_eighths=5),
                 dict(bucket_max_L=[150, 300, 600], use_graph=True)):
        print("case buckets", opts, flush=True)
        pl = gpu_index.plan(qids, 100, int(Ls.max()), M, init_ids=init, L_per_query=Ls, K_per_query=Ks,
                            options=Index.plan_options(**opts, want_queue=True), want_queue=True)
        b = pl.buckets()
        assert sum(x["nq"] for x in b) == qids.shape[0]
        assert all(x["max_L"] <= y["max_L"] for x, y in zip(b[1:], b[:-1]))  # largest L first
        pl.run()
        res = pl.fetch(want_queue=True)
        _same(res, ref, keys=("ids", "dists", "ndist", "nexp", "niter", "queue_ids", "queue_dists",
                              "queue_checked"))
        assert res["counters"]["distance_computations"] == ref["counters"]["distance_computations"]
        pl.close()


def test_graph_replays_with_eager_work_between(small_set, gpu_index, uneven):
    qids, Ls, Ks = uneven
    init = gpu_index.prepare_init_ids(int(Ls.max()))
    pl = gpu_index.plan(qids, 100, int(Ls.max()), 2, init_ids=init, L_per_query=Ls, K_per_query=Ks,
                        options=Index.plan_options(buckets=3, use_graph=True, concurrent=True))
    first = None
    for r in range(4):
        print("case graph replay", r, flush=True)
        pl.run()
        res = pl.fetch()
        if first is None:
            first = res
        else:
            _same(res, first)
        # eager work on the same handle between replays
        e = gpu_index.search(qids[:16], 50, 120, 2)
        assert e["ids"].shape == (16, 50)
    pl.close()


def test_autotune_returns_a_valid_bucketing(small_set, gpu_index, uneven):
    qids, Ls, Ks = uneven
    init = gpu_index.prepare_init_ids(int(Ls.max()))
    opt, ms = gpu_index.autotune(qids, 100, int(Ls.max()), 4, init_ids=init, L_per_query=Ls, K_per_query=Ks)
    assert 1 <= opt.n_buckets <= 8 and ms > 0
    bounds = [opt.bucket_max_L[i] for i in range(opt.n_buckets)]
    assert bounds == sorted(bounds) and bounds[-1] >= Ls.max()
    pl = gpu_index.plan(qids, 100, int(Ls.max()), 4, init_ids=init, L_per_query=Ls, K_per_query=Ks, options=opt,
                        want_queue=False)
    pl.run()
    res = pl.fetch()
    pl.close()
    ref = gpu_index.plan(qids, 100, int(Ls.max()), 4, init_ids=init, L_per_query=Ls, K_per_query=Ks,
                         want_queue=False)
    ref.run()
    _same(res, ref.fetch())
    ref.close()


def test_bucket_options_are_checked(gpu_index, uneven):
    qids, Ls, Ks = uneven
    init = gpu_index.prepare_init_ids(int(Ls.max()))
    with pytest.raises(SearchError):
        gpu_index.plan(qids, 100, int(Ls.max()), 2, init_ids=init, L_per_query=Ls,
                       options=Index.plan_options(bucket_max_L=[300, 200, 700]))
    with pytest.raises(SearchError):
        gpu_index.plan(qids, 100, int(Ls.max()), 2, init_ids=init, L_per_query=Ls,
                       options=Index.plan_options(bucket_max_L=[150, int(Ls.max()) - 1]))
    with pytest.raises(SearchError):
        gpu_index.plan(qids, 100, int(Ls.max()), 2, init_ids=init, L_per_query=Ls,
                       options=Index.plan_options(buckets=9))
    pl = gpu_index.plan(qids[:8], 50, 100, 2, want_queue=False)
    pl.run()
    with pytest.raises(ValueError):
        pl.fetch(want_queue=True)
    pl.close()


@pytest.mark.parametrize("rows", [32, 64, 96])
def test_rows_per_wave_keep_results(small_set, gpu_index, uneven, rows):
    """Neighbour rows gathered per wave and distance step (an autotuner
    dimension) change the launch shape only: buckets with one or two waves
    per query give the default's results."""
    qids, Ls, Ks = uneven
    print("case rows_per_wave", rows, flush=True)
    init = gpu_index.prepare_init_ids(int(Ls.max()))
    ref = gpu_index.plan(qids, 100, int(Ls.max()), 4, init_ids=init, L_per_query=Ls, K_per_query=Ks)
    ref.run()
    r0 = ref.fetch(want_queue=True)
    ref.close()
    for waves in (0, 2):
        pl = gpu_index.plan(qids, 100, int(Ls.max()), 4, init_ids=init, L_per_query=Ls, K_per_query=Ks,
                            options=Index.plan_options(buckets=2, rows_per_wave=rows, waves=waves))
        pl.run()
        _same(pl.fetch(want_queue=True), r0, ("ids", "dists", "ndist", "nexp", "niter", "queue_ids", "queue_dists"))
        pl.close()
    with pytest.raises(SearchError):
        gpu_index.plan(qids[:4], 100, 200, 4, init_ids=init, options=Index.plan_options(rows_per_wave=48))


def test_plans_are_refused_after_inputs_change(small_set, gpu_index, uneven):
    """A plan holds the device addresses of the queries, base rows and graph
    of its build: after set_queries (or set_base, or a graph load) running
    it is an error (PANNS_ERR_STATE, no kernel launched) until it is created
    again; results of an earlier run can still be fetched.  A graph plan is
    covered the same way."""
    qids, Ls, Ks = uneven
    init = gpu_index.prepare_init_ids(int(Ls.max()))
    plain = gpu_index.plan(qids[:16], 50, 100, 2, want_queue=False)
    graph = gpu_index.plan(qids, 100, int(Ls.max()), 2, init_ids=init, L_per_query=Ls, K_per_query=Ks,
                           options=Index.plan_options(buckets=2, use_graph=True), want_queue=False)
    plain.run()
    graph.run()
    r0, g0 = plain.fetch(), graph.fetch()
    gpu_index.set_queries(small_set["queries"])
    for pl in (plain, graph):
        with pytest.raises(SearchError) as e:
            pl.run()
        assert e.value.code == -9
    _same(plain.fetch(), r0)
    _same(graph.fetch(), g0)
    plain.close()
    graph.close()
    again = gpu_index.plan(qids[:16], 50, 100, 2, want_queue=False)
    again.run()
    _same(again.fetch(), r0)
    again.close()


def test_plans_outliving_their_index(small_set, uneven):
    """An index destroyed before its plans detaches them: running or
    fetching such a plan is an error (PANNS_ERR_STATE) and destroying it
    frees only its own resources (it touched the freed index before)."""
    qids, Ls, Ks = uneven
    idx = Index(0)
    idx.set_base(small_set["base"])
    idx.set_queries(small_set["queries"])
    idx.set_graph(small_set["adj"], small_set["deg"], small_set["ep"])
    init = idx.prepare_init_ids(int(Ls.max()))
    plain = idx.plan(qids[:16], 50, 100, 2, want_queue=False)
    graph = idx.plan(qids, 100, int(Ls.max()), 2, init_ids=init, L_per_query=Ls, K_per_query=Ks,
                     options=Index.plan_options(buckets=2, use_graph=True, concurrent=True), want_queue=False)
    plain.run()
    graph.run()
    idx.close()
    for pl in (plain, graph):
        with pytest.raises(SearchError) as e:
            pl.run()
        assert e.value.code == -9
        with pytest.raises(SearchError):
            pl.fetch()
        pl.close()
