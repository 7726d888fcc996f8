"""GPU exact k-NN (graph lists and ground truth) against the host brute force
and the oracle's get_true_NN restatement: ids and distance bits equal."""
import numpy as np
import pytest

import pyoracle as po
from parallel_anns_amd import tools
from parallel_anns_amd.search import Index

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("n,dim,k", [(5000, 96, 100), (3000, 32, 64), (777, 128, 128), (130, 16, 5)])
def test_exact_knn_matches_cpu(n, dim, k):
    print(f"case knn n={n} dim={dim} k={k}", flush=True)
    base = tools.gen_mixture(n, dim, 8, 0.3, 0, 31)
    queries = tools.gen_mixture(70, dim, 8, 0.3, 0, 32)
    idx = Index(0)
    idx.set_base(base)
    idx.set_queries(queries)
    gi, gd = idx.exact_knn(k, source="queries")
    ci, cd = tools.exact_knn(base, queries, k)
    assert (gi == ci).all()
    assert (gd.view(np.uint32) == cd.view(np.uint32)).all()
    oi, od = po.brute_force_knn(base, queries, k)
    assert (gi == oi).all()
    bi, _ = idx.exact_knn(min(k, n - 1), source="base")
    cbi, _ = tools.exact_knn(base, base, min(k, n - 1), exclude_self=True)
    assert (bi == cbi).all()
    idx.close()


def test_exact_knn_ties_go_to_lower_id():
    """Duplicated rows give equal distances; the lower id must come first."""
    dim = 32
    rng = np.random.default_rng(3)
    uniq = rng.standard_normal((50, dim)).astype(np.float32)
    base = np.concatenate([uniq, uniq, uniq])  # ids i, i+50, i+100 identical
    queries = uniq[:10] + 0.01
    idx = Index(0)
    idx.set_base(base)
    idx.set_queries(queries.astype(np.float32))
    gi, gd = idx.exact_knn(30, source="queries")
    oi, od = po.brute_force_knn(base, queries.astype(np.float32), 30)
    assert (gi == oi).all()
    for q in range(10):
        assert list(gi[q, :3]) == [q, q + 50, q + 100]
    idx.close()
