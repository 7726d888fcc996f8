"""Learned stop on the GPU: MFMA fragment layout, BF16-MFMA gradients against
the FP64 oracle, training, and the in-loop stop's semantics."""
import numpy as np
import pytest

import mlp_oracle as mo
from parallel_anns_amd import learned, tools
from parallel_anns_amd.search import Index

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def idx():
    i = Index(0)
    i.set_base(np.zeros((8, 96), np.float32))
    yield i
    i.close()


def test_mfma_fragment_layout(idx):
    rng = np.random.default_rng(0)
    A = mo.bf16(rng.standard_normal((16, 16))).astype(np.float32)
    B = mo.bf16(rng.standard_normal((16, 16))).astype(np.float32)
    C = idx.mfma_probe(A, B)
    ref = A.astype(np.float64) @ B.astype(np.float64)
    assert np.allclose(C, ref, rtol=1e-5, atol=1e-5)
    # a transposed operand layout would fail this
    assert not np.allclose(C, A.T.astype(np.float64) @ B, rtol=1e-3, atol=1e-3)


@pytest.mark.parametrize("n", [64, 200, 1000])
def test_mlp_gradient_matches_oracle(idx, n):
    rng = np.random.default_rng(n)
    P = mo.init_params(1)
    P[512:576] = rng.standard_normal(64) * 0.1
    P[4672:4736] = rng.standard_normal(64) * 0.1
    P[4800] = 0.2
    X = rng.standard_normal((n, 8)).astype(np.float32) * 3 + 1
    P[4804:4812] = X.mean(0)
    P[4812:4820] = X.std(0)
    Y = (rng.random(n) < 0.4).astype(np.float32)
    g, loss = idx.mlp_grad(P, X, Y)
    Xn = mo.normalise(X, P)
    l_ref, _, g_ref = mo.forward_backward(P[:mo.N_PARAMS], Xn, Y, bf16_operands=True)
    assert abs(loss - l_ref) < 1e-4 * max(1.0, abs(l_ref))
    scale = np.abs(g_ref).max()
    assert np.abs(g - g_ref).max() <= 2e-4 * scale
    # and the plain fp64 gradient agrees to bf16 precision
    _, _, g64 = mo.forward_backward(P[:mo.N_PARAMS], Xn, Y)
    assert np.abs(g - g64).max() <= 3e-2 * scale


def test_training_reduces_loss(idx):
    rng = np.random.default_rng(5)
    n = 20000
    X = rng.standard_normal((n, 8)).astype(np.float32)
    Y = ((X[:, 0] + 0.5 * X[:, 3] > 0.2) ^ (X[:, 5] > 1.5)).astype(np.float32)
    P, losses = learned.train(idx, X, Y, epochs=6, batch=2048, lr=3e-3)
    assert losses[-1] < 0.6 * losses[0]
    l_final, p, _ = mo.forward_backward(P[:mo.N_PARAMS], mo.normalise(X, P), Y)
    assert np.mean((p > 0.5) == (Y > 0.5)) > 0.9


def test_learned_stop_semantics(small_set, gpu_index):
    ds = small_set
    K, L, M = 100, 160, 4
    q = np.arange(ds["queries"].shape[0], dtype=np.uint32)
    feat, hsh, niter, ids_t = gpu_index.search_trace(q, K, L, M, trace_cap=256)
    fixed = gpu_index.search(q, K, L, M)
    assert (ids_t == fixed["ids"]).all()
    assert (niter == fixed["niter"]).all()
    X, Y = learned.samples_from_trace(feat, hsh, niter)
    assert X.shape[0] == int(np.minimum(niter, 256).sum())
    # iteration feature counts 1, 2, ...; the last step of each query is positive
    assert feat[0, 0, 0] == 1.0 and feat[0, 1, 0] == 2.0
    assert Y.mean() > 0
    P, _ = learned.train(gpu_index, X, Y, epochs=2, batch=1024)
    gpu_index.set_stop_model(P)
    # a threshold the sigmoid never reaches leaves the fixed-L search unchanged
    never = gpu_index.search(q, K, L, M, stop="learned", stop_threshold=2.0)
    assert (never["ids"] == fixed["ids"]).all() and (never["ndist"] == fixed["ndist"]).all()
    # threshold 0 stops every query after its first iteration
    first = gpu_index.search(q, K, L, M, stop="learned", stop_threshold=0.0)
    assert (first["niter"] == 1).all()
    mid = gpu_index.search(q, K, L, M, stop="learned", stop_threshold=0.5)
    assert (mid["ndist"] <= fixed["ndist"]).all()
    # stopping early never reorders what was found: the partial queue is sorted
    assert (np.diff(mid["dists"], axis=1) >= 0).all()
    assert tools.recall(ds["gt_ids"], mid["ids"], L) <= tools.recall(ds["gt_ids"], fixed["ids"], L) + 1e-9


def test_in_loop_model_matches_oracle(small_set, gpu_index):
    """Each query stops at the first iteration where the FP64 oracle of the
    network (the kernel evaluates it in fp32), evaluated on the traced
    features, reaches the threshold; queries with a
    step within 1e-3 of the threshold before that are not judged (the kernel
    scales features by a precomputed reciprocal of sd in fp32 and sums in
    fp32, the oracle divides and sums in fp64)."""
    ds = small_set
    K, L, M = 100, 160, 4
    q = np.arange(ds["queries"].shape[0], dtype=np.uint32)
    feat, hsh, niter, _ = gpu_index.search_trace(q, K, L, M, trace_cap=256)
    assert int(niter.max()) <= 256
    X, Y = learned.samples_from_trace(feat, hsh, niter)
    P, _ = learned.train(gpu_index, X, Y, epochs=2, batch=1024)
    gpu_index.set_stop_model(P)
    Pb = np.array(P, np.float64)
    for thr in (0.3, 0.5, 0.7):
        print("case in-loop model threshold", thr, flush=True)
        res = gpu_index.search(q, K, L, M, stop="learned", stop_threshold=thr)
        judged = 0
        for i in range(q.shape[0]):
            n = int(niter[i])
            _, p, _ = mo.forward_backward(Pb[:mo.N_PARAMS], mo.normalise(feat[i, :n], P), np.zeros(n))
            fire = np.nonzero(p >= thr)[0]
            expect = int(fire[0]) + 1 if fire.size else n
            if np.any(np.abs(p[:expect] - thr) < 1e-3):
                continue
            judged += 1
            assert int(res["niter"][i]) == expect, f"query {i} threshold {thr}"
        assert judged >= 0.9 * q.shape[0]


def test_stop_schedule(small_set, gpu_index):
    """stop_min_iter and stop_every only restrict where the model may stop a
    query: never before min_iter, only at multiples of stop_every."""
    ds = small_set
    K, L, M = 100, 160, 4
    q = np.arange(ds["queries"].shape[0], dtype=np.uint32)
    feat, hsh, niter, _ = gpu_index.search_trace(q, K, L, M, trace_cap=256)
    X, Y = learned.samples_from_trace(feat, hsh, niter)
    P, _ = learned.train(gpu_index, X, Y, epochs=2, batch=1024)
    gpu_index.set_stop_model(P)
    fixed = gpu_index.search(q, K, L, M)
    late = gpu_index.search(q, K, L, M, stop="learned", stop_threshold=0.0, stop_min_iter=int(niter.max()) + 1)
    assert (late["ids"] == fixed["ids"]).all() and (late["niter"] == fixed["niter"]).all()
    five = gpu_index.search(q, K, L, M, stop="learned", stop_threshold=0.0, stop_min_iter=5)
    assert ((five["niter"] == 5) | ((five["niter"] < 5) & (five["niter"] == fixed["niter"]))).all()
    # both restrictions: the first evaluation is the first multiple of
    # stop_every at or after max(1, stop_min_iter)
    for mi, ev, first in ((5, 2, 6), (6, 3, 6), (0, 4, 4), (7, 1, 7)):
        both = gpu_index.search(q, K, L, M, stop="learned", stop_threshold=0.0, stop_min_iter=mi, stop_every=ev)
        assert ((both["niter"] == first) | ((both["niter"] < first) & (both["niter"] == fixed["niter"]))).all(), \
            (mi, ev)
    every = gpu_index.search(q, K, L, M, stop="learned", stop_threshold=0.4, stop_every=3)
    stopped_early = every["niter"] < fixed["niter"]
    assert stopped_early.any()
    assert (every["niter"][stopped_early] % 3 == 0).all()
    m1 = learned.first_final_iteration(hsh, niter, 0.01)
    assert 1 <= m1 <= int(niter.max())


def test_trained_network_gains_over_fixed_l(small_set, gpu_index):
    """What the trained network gains over the method without it: with the
    model trained on traces of 3,000 seed-disjoint training queries, some
    threshold of the learned stop (L_max = 200) reaches at least the recall of
    a fixed-L search with fewer distances per query than the cheapest fixed L
    that reaches that recall.  On this set the best such point saves about
    9 % of the distances (recall 0.87 at 436 distances against L = 140's 0.90
    at 478); the assertion keeps a margin for training noise."""
    ds = small_set
    cfg = ds["cfg"]
    K, M, Lmax = 100, 4, 200
    from parallel_anns_amd import datasets
    tq = datasets.generate(cfg, "train", n=3000)
    gpu_index.set_queries(tq)
    try:
        feat, hsh, niter, _ = gpu_index.search_trace(np.arange(3000, dtype=np.uint32), K, 300, M, trace_cap=256)
        X, Y = learned.samples_from_trace(feat, hsh, niter)
        P, _ = learned.train(gpu_index, X, Y, epochs=3, batch=2048)
    finally:
        gpu_index.set_queries(ds["queries"])
    gpu_index.set_stop_model(P)
    q = np.arange(ds["queries"].shape[0], dtype=np.uint32)
    gt = ds["gt_ids"]
    fixed = []
    for L in range(100, Lmax + 1, 10):
        r = gpu_index.search(q, K, L, M)
        fixed.append((tools.recall(gt, r["ids"], L), float(r["ndist"].mean()), L))
    best = (0.0, None)
    for thr in (0.2, 0.3, 0.4, 0.5, 0.6, 0.7, 0.8, 0.9):
        r = gpu_index.search(q, K, Lmax, M, stop="learned", stop_threshold=thr)
        rec, nd = tools.recall(gt, r["ids"], Lmax), float(r["ndist"].mean())
        cheaper = [f for f in fixed if f[0] >= rec]
        if cheaper:
            f = min(cheaper, key=lambda x: x[1])
            print(f"threshold {thr}: recall {rec:.4f} at {nd:.1f} distances; fixed L={f[2]}: "
                  f"recall {f[0]:.4f} at {f[1]:.1f}", flush=True)
            if f[1] / nd > best[0]:
                best = (f[1] / nd, thr)
    assert best[0] >= 1.03, best


def test_trace_with_hits(small_set, gpu_index):
    """panns_search_trace_hits: the per-iteration recall count ends at the
    final result's count against the same truth, and the per-iteration
    distance count is non-decreasing up to the search's own count; the
    features and hashes equal the plain trace's."""
    ds = small_set
    K, L, M = 100, 160, 4
    q = np.arange(ds["queries"].shape[0], dtype=np.uint32)
    gt = ds["gt_ids"]
    feat, hsh, niter, ids, hits, nd = gpu_index.search_trace(q, K, L, M, trace_cap=256, gt_ids=gt)
    f0, h0, n0, i0 = gpu_index.search_trace(q, K, L, M, trace_cap=256)
    assert (feat == f0).all() and (hsh == h0).all() and (niter == n0).all() and (ids == i0).all()
    fixed = gpu_index.search(q, K, L, M)
    for i in range(q.shape[0]):
        n = int(niter[i])
        assert int(hits[i, n - 1]) == len(np.intersect1d(gt[i, :100], fixed["ids"][i, :100]))
        assert (np.diff(nd[i, :n].astype(np.int64)) >= 0).all() and int(nd[i, n - 1]) == int(fixed["ndist"][i])
        assert int(nd[i, 0]) >= L
