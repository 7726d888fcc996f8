"""The find-L bisection and the fastest-above-target selection
(parallel_anns_amd.operating_point), checked against what they must give,
not against a second copy of their loops:

- recall that never decreases with L (steps and plateaus, as the
  configured graphs give): the smallest L in range reaching the target;
- an exact hit: an L whose recall equals the target;
- any recall curve: the reported L reaches the target and L - 1 was probed
  below it (or L is the range's lower end); a target nothing reaches gives
  the upper end with its recall;
- selection (scripts/output_find_runtime_above_presicion.py:40-60): the
  least runtime among rows at or above the target, the first on ties.
"""
import numpy as np

from parallel_anns_amd import operating_point as op


def _staircase(rng, lo, hi):
    L = np.arange(lo, hi + 1)
    steps = np.sort(rng.choice(L, size=int(rng.integers(1, 8)), replace=False))
    r = np.zeros(L.shape[0])
    level = float(rng.uniform(0.3, 0.6))
    for s in steps:
        level = min(1.0, level + float(rng.uniform(0.01, 0.2)))
        r[L >= s] = level
    r[L < steps[0]] = float(rng.uniform(0.1, 0.3))
    return {int(l): float(x) for l, x in zip(L, r)}


def test_monotone_staircases_give_the_smallest_L():
    rng = np.random.default_rng(0)
    for _ in range(400):
        lo, hi = int(rng.integers(10, 100)), int(rng.integers(120, 700))
        rec = _staircase(rng, lo, hi)
        for target in (float(rng.uniform(0.1, 1.0)), float(rng.choice(list(rec.values())))):
            L, r, probes = op.find_L(rec.__getitem__, lo, hi, target)
            reach = [l for l in range(lo, hi + 1) if rec[l] >= target]
            if reach:
                if r == target:
                    # an exact hit stops the search at a probe on the plateau
                    assert rec[L] == target
                else:
                    assert L == reach[0] and r == rec[L]
            else:
                assert L == hi and r == rec[hi] < target
            # every probe lies in range and the first is the upper end
            assert probes[0][0] == hi and all(lo <= p <= hi for p, _ in probes)


def test_exact_hits_stop_the_search():
    rec = {L: (0.5 if L < 40 else 0.8 if L < 90 else 0.95) for L in range(10, 201)}
    L, r, probes = op.find_L(rec.__getitem__, 10, 200, 0.8)
    assert r == 0.8 and 40 <= L < 90
    assert probes[-1] == (L, 0.8)


def test_any_curve_gives_a_locally_minimal_L():
    rng = np.random.default_rng(1)
    for _ in range(400):
        lo, hi = int(rng.integers(10, 60)), int(rng.integers(80, 400))
        rec = {l: float(x) for l, x in zip(range(lo, hi + 1), rng.random(hi - lo + 1))}
        target = float(rng.uniform(0.2, 0.95))
        L, r, probes = op.find_L(rec.__getitem__, lo, hi, target)
        seen = dict(probes)
        if r >= target:
            assert rec[L] == r
            assert L == lo or (L - 1 in seen and seen[L - 1] < target)
        else:
            assert L == hi and not any(x >= target for x in seen.values())


def test_fastest_above_follows_the_script():
    rows = [dict(runtime=3.0, recall=0.96, tag="a"), dict(runtime=2.0, recall=0.94, tag="b"),
            dict(runtime=2.5, recall=0.95, tag="c"), dict(runtime=2.5, recall=0.99, tag="d"),
            dict(runtime=1.0, recall=0.90, tag="e")]
    assert op.fastest_above(rows, 0.95)["tag"] == "c"  # equal runtimes: the first row
    assert op.fastest_above(rows, 0.97)["tag"] == "d"
    assert op.fastest_above(rows, 0.90)["tag"] == "e"
    assert op.fastest_above(rows, 0.995) is None
