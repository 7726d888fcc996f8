"""Learned early termination for the top-M search.

Pipeline (all device work runs through include/panns_amd.h):
  1. traces: the plain fixed-L search (search.Index.search_trace) records after
     every iteration the 8 stop features and a hash of the top-K id set;
  2. labels: a step is positive when its top-K set already equals the final
     one (the set the search ends with);
  3. training: the 8-64-64-1 MLP on the GPU with the BF16-MFMA forward and
     backward kernels (search.Index.mlp_train), features normalised by the
     training set's mean and deviation;
  4. calibration: the stop threshold is the smallest-work threshold whose
     recall on held-out validation queries reaches the target;
  5. search: panns_params.stop = LEARNED runs the MLP after every iteration in
     the search kernel and stops the query once its output reaches the
     threshold.

The reference has no learned component; its hand-tuned stopping studies are
core/Searching.202011301508.top_m_pure_large_graph.num_unchecked.* and
core/Searching.202012041559.top_m_pure_large_graph.last_update_step.*.
"""
import numpy as np

from . import tools

FEATURES = ("iteration", "top1_dist", "kth_dist", "kth_over_top1", "inserted_last_step",
            "best_initial_dist", "query_norm", "unchecked")
MODEL_FLOATS = 4820


def init_model(seed=0):
    """He-initialised weights, zero biases, identity normalisation."""
    rng = np.random.default_rng(seed)
    P = np.zeros(MODEL_FLOATS, np.float32)
    P[0:512] = rng.standard_normal(512) * np.sqrt(2 / 8)
    P[576:4672] = rng.standard_normal(4096) * np.sqrt(2 / 64)
    P[4736:4800] = rng.standard_normal(64) * np.sqrt(1 / 64)
    P[4812:4820] = 1.0
    return P


def samples_from_trace(feat, topk_hash, niter, hits=None):
    """(X [n, 8], Y [n]) with Y = 1 where the step's top-K set is the final
    one; with hits (recall@100's count per step, search_trace with gt_ids)
    Y = 1 where the step's recall already equals the final recall."""
    cap = feat.shape[1]
    key = topk_hash if hits is None else hits
    X, Y = [], []
    for q in range(feat.shape[0]):
        n = int(min(niter[q], cap))
        if n == 0:
            continue
        final = key[q, n - 1]
        X.append(feat[q, :n])
        Y.append((key[q, :n] == final).astype(np.float32))
    return np.concatenate(X), np.concatenate(Y)


def set_normalisation(P, X):
    P = P.copy()
    P[4804:4812] = X.mean(0)
    P[4812:4820] = X.std(0) + 1e-6
    return P


def train(index, X, Y, epochs=4, batch=4096, lr=2e-3, seed=0):
    P = set_normalisation(init_model(seed), X)
    P, losses = index.mlp_train(X, Y, P, epochs=epochs, batch=batch, lr=lr, seed=seed)
    return P, losses


def first_final_iteration(topk_hash, niter, quantile=0.01):
    """The iteration (1-based) by which `quantile` of the traced queries have
    reached their final top-K set (or, given per-step recall counts instead
    of hashes, their final recall): the stop model need not run earlier."""
    firsts = []
    for q in range(topk_hash.shape[0]):
        n = int(min(niter[q], topk_hash.shape[1]))
        if n == 0:
            continue
        same = topk_hash[q, :n] == topk_hash[q, n - 1]
        # first step of the final run of equal hashes
        j = n - 1
        while j > 0 and same[j - 1]:
            j -= 1
        firsts.append(j + 1)
    return int(np.quantile(np.asarray(firsts), quantile)) if firsts else 0


def recall_work(index, qids, gt, K, L, M, threshold, runs=1, L_per_query=None, options=None, **stop_kw):
    """(recall@K, mean distance computations, kernel ms -- the least of
    `runs` launches) of a learned-stop search; stop_kw: stop_min_iter,
    stop_every.  threshold None runs the fixed-L search.  L_per_query: each
    query's own L (then L is their maximum and recall is counted over the
    first min(K, smallest L) ids); options: plan options (buckets)."""
    kw = dict(stop="learned", stop_threshold=threshold, **stop_kw) if threshold is not None else {}
    plan = index.plan(qids, K, L, M, want_queue=False, L_per_query=L_per_query, options=options, **kw)
    ms = min(plan.run() for _ in range(runs))
    res = plan.fetch()
    plan.close()
    Lr = L if L_per_query is None else int(np.min(L_per_query))
    return tools.recall(gt, res["ids"], Lr, R=min(K, gt.shape[1])), float(res["ndist"].mean()), ms


# The stop threshold that means "the plain fixed-L search": the kernel's
# sigmoid output never exceeds 1.0f (it can round to exactly 1.0f, so a
# threshold of 1.0 may still stop a query), so any threshold above 1 never
# fires.  Every caller uses this one value for the fallback.
FIXED_FALLBACK = 2.0


def is_fallback(threshold):
    """True when a calibrated threshold means "run the fixed-L search"."""
    return threshold is None or threshold > 1.0


def calibrate(index, qids, gt, K, L, M, target, lo=1e-4, hi=1.0, steps=14, **stop_kw):
    """Largest-saving threshold whose recall on (qids, gt) reaches target: the
    recall rises with the threshold, so bisect on it (log scale).  None when
    no threshold below `hi` reaches the target (only `hi` itself, where the
    stop is all but the fixed search, or not even that).  stop_kw also
    passes L_per_query and options through to recall_work."""
    r_hi = recall_work(index, qids, gt, K, L, M, hi, **stop_kw)[0]
    if r_hi < target:
        return None
    a, b = np.log(lo), np.log(hi)
    if recall_work(index, qids, gt, K, L, M, lo, **stop_kw)[0] >= target:
        return lo
    passed = False
    for _ in range(steps):
        m = 0.5 * (a + b)
        if recall_work(index, qids, gt, K, L, M, float(np.exp(m)), **stop_kw)[0] >= target:
            b = m
            passed = True
        else:
            a = m
    return float(np.exp(b)) if passed else None


def oracle_bound(per_query_recall, per_query_ndist, target):
    """Least mean distance computations at mean recall >= target when every
    query may take its own L from the measured grid (columns): a bound for
    any per-query stop rule that stops at one of those L.

    Lagrangian relaxation: at price lam each query takes the column that
    minimises ndist - lam * recall, and the mean recall of that choice rises
    with lam; bisect for the smallest lam that reaches target.  This is
    exact at the vertices of the convex hull of the (recall, ndist) choices,
    which is where the bisection ends up.  Returns (ndist, recall, lam)."""
    R = np.asarray(per_query_recall, np.float64)
    N = np.asarray(per_query_ndist, np.float64)
    rows = np.arange(R.shape[0])

    def choose(lam):
        c = np.argmin(N - lam * R, axis=1)
        return float(N[rows, c].mean()), float(R[rows, c].mean())

    lo, hi = 0.0, 1.0
    while choose(hi)[1] < target and hi < 1e12:
        hi *= 4.0
    nd, rec = choose(hi)
    if rec < target:
        return None
    for _ in range(80):
        mid = 0.5 * (lo + hi)
        if choose(mid)[1] >= target:
            hi = mid
        else:
            lo = mid
    nd, rec = choose(hi)
    return (nd, rec, hi)
