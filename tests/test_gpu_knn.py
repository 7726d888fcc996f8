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


def test_mfma_f32_layout():
    """The fragment layout the MFMA k-NN kernel assumes: C = A (16x4) @ B (4x16)."""
    rng = np.random.default_rng(5)
    A = rng.standard_normal((16, 4)).astype(np.float32)
    B = rng.standard_normal((4, 16)).astype(np.float32)
    idx = Index(0)
    print("case mfma f32 probe", flush=True)
    C = idx.mfma_f32_probe(A, B)
    idx.close()
    ref = A.astype(np.float64) @ B.astype(np.float64)
    assert np.abs(C - ref).max() <= 1e-5 * (np.abs(A).max() * np.abs(B).max() * 4)


@pytest.mark.parametrize("n,dim,k,src", [(20000, 96, 100, "queries"), (20000, 96, 64, "base"),
                                         (6000, 40, 50, "queries"), (4000, 128, 120, "base"),
                                         (3000, 16, 10, "queries")])
def test_mfma_candidates_equal_exact_chain(n, dim, k, src):
    """The MFMA filter bounded by a sample's exact k-th distance + the exact
    re-rank of its candidates give the exact kernel's ids and distance bits."""
    print(f"case knn mfma n={n} dim={dim} k={k} {src}", flush=True)
    base = tools.gen_mixture(n, dim, 16, 0.3, 0, 41)
    queries = tools.gen_mixture(300, dim, 16, 0.3, 0, 42)
    idx = Index(0)
    idx.set_base(base)
    idx.set_queries(queries)
    st = {}
    a_i, a_d = idx.exact_knn(k, source=src, method="mfma", stats=st)
    v_i, v_d = idx.exact_knn(k, source=src, method="valu")
    idx.close()
    assert (a_i == v_i).all()
    assert (a_d.view(np.uint32) == v_d.view(np.uint32)).all()
    nq = queries.shape[0] if src == "queries" else n
    assert st["recomputed"] < nq  # the bound proves most queries


def test_mfma_candidates_with_exact_ties():
    """Every row repeated 100 times: every copy of a row within the bound is a
    candidate, and the exact re-rank orders equal distances by id, so the ids
    equal the oracle's (ties to the lower id).  Integer rows keep every
    distance exact."""
    rng = np.random.default_rng(9)
    uniq = rng.integers(-2, 3, (30, 32)).astype(np.float32)
    base = np.tile(uniq, (100, 1))
    queries = rng.integers(-2, 3, (64, 32)).astype(np.float32)
    idx = Index(0)
    idx.set_base(base)
    idx.set_queries(queries)
    st = {}
    print("case knn mfma ties", flush=True)
    gi, gd = idx.exact_knn(50, source="queries", method="mfma", stats=st)
    idx.close()
    oi, od = po.brute_force_knn(base, queries, 50)
    assert (gi == oi).all() and (gd.view(np.uint32) == od.view(np.uint32)).all()


def test_mfma_overflowing_lists_fall_back_to_exact_chain():
    """3,000 copies of each of 4 rows: every copy of the nearest row passes
    the filter, more than a candidate list holds, so those queries are
    recomputed with the exact chain over every row; ids and distances still
    equal the oracle's."""
    rng = np.random.default_rng(13)
    uniq = rng.integers(-3, 4, (4, 32)).astype(np.float32)
    base = np.tile(uniq, (3000, 1))
    queries = rng.integers(-3, 4, (40, 32)).astype(np.float32)
    idx = Index(0)
    idx.set_base(base)
    idx.set_queries(queries)
    st = {}
    print("case knn mfma overflow", flush=True)
    gi, gd = idx.exact_knn(10, source="queries", method="mfma", stats=st)
    idx.close()
    oi, od = po.brute_force_knn(base, queries, 10)
    assert (gi == oi).all() and (gd.view(np.uint32) == od.view(np.uint32)).all()
    assert st["recomputed"] > 0
