"""The allocation-free search (panns_search_noalloc) through the C-ABI:
its results equal panns_search's for the same batch, on calls in a row that
grow and shrink the batch, change the mode and push claims into the
visited bitmaps, and a batch whose buffers do not fit in the library's
fixed array is refused before any launch."""
import numpy as np
import pytest

from parallel_anns_amd.search import Index, SearchError

pytestmark = pytest.mark.gpu

FIELDS = ("ids", "dists", "ndist", "nexp", "niter")
QFIELDS = ("queue_ids", "queue_dists", "queue_checked")
COUNTERS = ("distance_computations", "merges", "expansions", "adjacency_slots", "visited_spills")


def _same(a, b, fields):
    for f in fields:
        x, y = a[f], b[f]
        if x.dtype == np.float32:
            x, y = x.view(np.uint32), y.view(np.uint32)
        assert (x == y).all(), f


@pytest.fixture(scope="module")
def index(small_set):
    idx = Index(0)
    idx.set_base(small_set["base"])
    idx.set_queries(small_set["queries"])
    idx.set_graph(small_set["adj"], small_set["deg"], small_set["ep"])
    yield idx
    idx.close()


def test_noalloc_equals_search_on_calls_in_a_row(small_set, index):
    nq = small_set["queries"].shape[0]
    q = np.arange(nq, dtype=np.uint32)
    # the second call is smaller; later ones change mode and table size
    calls = (dict(n=nq, K=100, L=160, M=4, want_queue=True),
             dict(n=37, K=50, L=100, M=2, want_queue=False),
             dict(n=nq, K=100, L=200, M=4, want_queue=True, hash_slots=64),
             dict(n=64, K=100, L=120, M=3, want_queue=True, mode="pss_v5", num_threads=4),
             dict(n=5, K=10, L=64, M=1, want_queue=False, mode="scale_m"))
    for c in calls:
        c = dict(c)
        n = c.pop("n")
        print("case", n, c, flush=True)
        a = index.search(q[:n], no_alloc=True, **c)
        b = index.search(q[:n], **c)
        _same(a, b, FIELDS + (QFIELDS if c["want_queue"] else ()))
        for k in COUNTERS:
            assert a["counters"][k] == b["counters"][k], k
        if c.get("hash_slots") == 64:
            assert a["counters"]["visited_spills"] > 0


def test_noalloc_grid_limited_by_the_array(small_set, index):
    """A batch whose outputs take most of the array leaves room for fewer
    workgroups' bitmaps than the occupancy allows; results are unchanged."""
    nq = small_set["queries"].shape[0]
    q = np.tile(np.arange(nq, dtype=np.uint32), 60000 // nq)
    print("case grid-limited", q.shape[0], flush=True)
    a = index.search(q, 120, 120, 4, no_alloc=True)
    b = index.search(q, 120, 120, 4)
    _same(a, b, FIELDS)


def test_noalloc_refuses_what_does_not_fit(index):
    q = np.zeros(100_000, np.uint32)
    with pytest.raises(SearchError) as e:
        index.search(q, 100, 100, 4, no_alloc=True)
    assert e.value.code == -1
    # the handle still works afterwards
    r = index.search(q[:4], 10, 64, 2, no_alloc=True)
    assert (r["ids"][0] == r["ids"][3]).all()
