"""Launch geometry under gfx950's LDS allocation granule: a workgroup's LDS
is allocated in 1,280-byte granules (160 KiB of LDS per CU), so the
persistent grid holds only the workgroups whose rounded-up LDS fits, and a
visited table the plan sizes itself grows into what that occupancy leaves
free (whole groups of 8 buckets) unless the plan is a bucket launched side
by side with others.  The results never depend on the table size: ids,
distance bits, distance counts and iterations equal those of a plan with a
given table size, and the oracle's."""
import numpy as np
import pytest

import pyoracle as po
from parallel_anns_amd.search import Index

pytestmark = pytest.mark.gpu

GRANULE = 1280
STEP = 2 * 64  # one group of 8 eight-slot buckets, 2 bytes per slot


def _device_attr(attr):
    """hipDeviceGetAttribute on device 0 (the HIP runtime the library has
    loaded); enum values of ROCm 7's hip_runtime_api.h."""
    import ctypes
    hip = ctypes.CDLL("libamdhip64.so")
    v = ctypes.c_int(0)
    assert hip.hipDeviceGetAttribute(ctypes.byref(v), attr, 0) == 0
    return v.value


def _num_cu():
    return _device_attr(63)  # hipDeviceAttributeMultiprocessorCount


def _lds_per_cu():
    return _device_attr(10002)  # hipDeviceAttributeMaxSharedMemoryPerMultiprocessor


def _ru(x):
    return -(-x // GRANULE) * GRANULE


@pytest.mark.parametrize("L,M,defer", [(100, 4, 0), (289, 4, 0), (600, 4, 0), (1000, 4, 2), (1500, 4, 2),
                                       (2000, 16, 0)])
def test_grid_and_grown_table_fit_the_granule(small_set, gpu_index, L, M, defer):
    ds = small_set
    num_cu, LDS_PER_CU = _num_cu(), _lds_per_cu()
    assert LDS_PER_CU == 160 * 1024
    K = min(100, L)
    # more work items than any grid, so the grid is the resident count
    q = np.tile(np.arange(ds["queries"].shape[0], dtype=np.uint32), 40)
    print("case granule", L, M, defer, flush=True)
    pl = gpu_index.plan(q, K, L, M, options=Index.plan_options(deferred_merge=defer))
    d = pl.describe()
    assert d["workgroups"] % num_cu == 0, d
    per_cu = d["workgroups"] // num_cu
    assert per_cu >= 1 and _ru(d["lds_bytes"]) * per_cu <= LDS_PER_CU, d
    budget = LDS_PER_CU // per_cu // GRANULE * GRANULE
    assert d["lds_bytes"] <= budget < d["lds_bytes"] + STEP, (d, budget)
    pl.run()
    a = pl.fetch()
    pl.close()
    # the same search with a given (not grown) table
    pl = gpu_index.plan(q, K, L, M, options=Index.plan_options(deferred_merge=defer), hash_slots=4096)
    assert pl.describe()["hash_slots"] == 4096
    pl.run()
    b = pl.fetch()
    pl.close()
    for f in ("ids", "ndist", "niter", "nexp"):
        assert (a[f] == b[f]).all(), f
    assert (a["dists"].view(np.uint32) == b["dists"].view(np.uint32)).all()
    g = po.CsrGraph(ds["adj"], ds["deg"], ds["ep"])
    init = gpu_index.prepare_init_ids(L)
    for i in range(0, ds["queries"].shape[0], 50):
        r = po.search_one(g, ds["base"], ds["queries"][i], K, L, M, init)
        assert (a["ids"][i] == r["ids"]).all(), i
        assert a["ndist"][i] == r["ndist"] and a["niter"][i] == r["merges"], i


def test_side_by_side_buckets_keep_their_tables(small_set, gpu_index):
    ds = small_set
    num_cu, LDS_PER_CU = _num_cu(), _lds_per_cu()
    rng = np.random.default_rng(5)
    q = np.tile(np.arange(ds["queries"].shape[0], dtype=np.uint32), 40)
    Ls = np.exp(rng.uniform(np.log(100), np.log(2000), q.size)).astype(np.uint32)
    Ks = rng.choice(np.array([10, 100], np.uint32), q.size)
    init = gpu_index.prepare_init_ids(int(Ls.max()))
    bounds = [300, 800, 1400, int(Ls.max())]
    out = {}
    for conc in (True, False):
        print("case granule buckets, side by side", conc, flush=True)
        pl = gpu_index.plan(q, 100, int(Ls.max()), 4, init_ids=init, L_per_query=Ls, K_per_query=Ks,
                            options=Index.plan_options(bucket_max_L=bounds, concurrent=conc, table_eighths=0))
        bk = pl.buckets()
        for b in bk:
            assert _ru(b["lds_bytes"]) * (b["workgroups"] // num_cu) <= LDS_PER_CU, b
        pl.run()
        out[conc] = (pl.fetch(), {b["max_L"]: b for b in bk})
        pl.close()
    (a, side), (b, alone) = out[True], out[False]
    for L, s in side.items():
        # alone, a bucket's table may grow; side by side it keeps its size
        assert s["hash_slots"] <= alone[L]["hash_slots"], (s, alone[L])
    assert any(s["hash_slots"] < alone[L]["hash_slots"] for L, s in side.items())
    for f in ("ids", "ndist", "niter", "nexp"):
        assert (a[f] == b[f]).all(), f
    assert (a["dists"].view(np.uint32) == b["dists"].view(np.uint32)).all()
