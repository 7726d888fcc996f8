"""Operating points: the find-L bisection and the "fastest setting at or
above a recall target" selection, as the reference's harness does them.

find_L follows the loop of
app/profile_find_L_para_single_query_top_m_search_pure_large_graph.cpp:210-260
(also the C++ driver csrc/apps/find_L_top_m.cpp): start at L_upper, move
L_lower above an L whose recall is below the target, L_upper below one
whose recall is above it, stop at an exact hit, bisect; when the last L
probed is below the target, report the last L that was above it.  For a
recall that never decreases with L the result is the smallest L in the
range reaching the target; for any recall curve the reported L reaches the
target and L - 1 was probed below it (or L is L_lower).

fastest_above follows scripts/output_find_runtime_above_presicion.py:40-60:
among rows whose precision reaches the target, the one with the least
runtime, the first such row on ties; None when no row reaches it.
"""


def find_L(recall_at, L_lower, L_upper, target):
    """recall_at(L) -> recall@100.  Returns (L, recall, probes) where probes
    lists the (L, recall) pairs in the order they were evaluated."""
    probes = []
    lo, hi = L_lower, L_upper
    L = hi
    cur = None
    last = None
    while lo <= hi:
        cur = recall_at(L)
        probes.append((L, cur))
        if cur < target:
            lo = L + 1
        elif cur > target:
            hi = L - 1
            last = (L, cur)
        else:
            break
        if lo <= hi:
            L = (lo + hi) // 2
    if cur is not None and cur < target and L < L_upper and last is not None:
        L, cur = last
    return L, cur, probes


def fastest_above(rows, target, runtime=lambda r: r["runtime"], precision=lambda r: r["recall"]):
    """The row of least runtime among those with precision >= target (the
    first of equal ones), or None."""
    best, best_rt = None, float("inf")
    for r in rows:
        rt = runtime(r)
        if precision(r) >= target and rt < best_rt:
            best, best_rt = r, rt
    return best
