"""Pins of the oracle's distances: squared L2 (PSS5LG.cpp:525-560 as
compiled) and the opt-graph norm and norm - 2 v.q (PSS5.cpp:293-390).

The reference itself cannot be built here (it needs Boost), so its compiled
summation order is pinned by an independent route: the same 8-wide AVX
intrinsic structure compiled by GCC with the contraction the reference's
-O3 -march=native build gets must equal the oracle bit for bit, while the
build without contraction must differ on some inputs (so the comparison
can tell the two orders apart).  Closed forms and fp64 closeness cover the
rest.
"""
import ctypes
import os
import subprocess
import tempfile

import numpy as np
import pytest

import pyoracle as po
from parallel_anns_amd import tools

HERE = os.path.dirname(os.path.abspath(__file__))


def _build_probe(extra):
    out = os.path.join(tempfile.mkdtemp(prefix="panns_probe_"), "probe.so")
    subprocess.run(["gcc", "-O3", "-mavx2", "-mfma", *extra, "-fPIC", "-shared",
                    os.path.join(HERE, "native", "avx_l2_probe.c"), "-o", out], check=True)
    lib = ctypes.CDLL(out)
    lib.probe_l2.restype = ctypes.c_float
    lib.probe_l2.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_uint]
    lib.probe_norm.restype = ctypes.c_float
    lib.probe_norm.argtypes = [ctypes.c_void_p, ctypes.c_uint]
    lib.probe_dot_norm.restype = ctypes.c_float
    lib.probe_dot_norm.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_uint, ctypes.c_float]
    return lib


@pytest.fixture(scope="module")
def probes():
    return _build_probe([]), _build_probe(["-ffp-contract=off"])


def _bits(x):
    return np.float32(x).view(np.uint32)


def _pairs(dim, n, seed):
    rng = np.random.default_rng(seed)
    # mixed magnitudes make rounding order visible
    v = (rng.standard_normal((n, dim)) * rng.choice([1e-3, 1.0, 30.0], size=(n, dim))).astype(np.float32)
    q = (rng.standard_normal((n, dim)) * rng.choice([1e-3, 1.0, 30.0], size=(n, dim))).astype(np.float32)
    return v, q


@pytest.mark.parametrize("dim", [8, 16, 24, 96, 104, 128])
def test_oracle_equals_contracted_avx_build(probes, dim):
    fused, _ = probes
    v, q = _pairs(dim, 3000, dim)
    for i in range(v.shape[0]):
        a = po.distance(v[i], q[i])
        b = fused.probe_l2(v[i].ctypes.data, q[i].ctypes.data, dim)
        assert _bits(a) == _bits(b), (dim, i)


def test_unfused_order_is_distinguishable(probes):
    _, plain = probes
    dim = 96
    v, q = _pairs(dim, 3000, 7)
    diff = sum(_bits(po.distance(v[i], q[i])) != _bits(plain.probe_l2(v[i].ctypes.data, q[i].ctypes.data, dim))
               for i in range(v.shape[0]))
    assert diff > 0


def test_order_matters_for_a_shifted_chain():
    """A mistake in the block order (tail block last instead of first) changes bits."""
    dim = 104
    v, q = _pairs(dim, 2000, 9)

    def wrong_order(a, b):
        acc = np.zeros(8, np.float32)
        for blk in range(dim // 8):
            t = (a[blk * 8:blk * 8 + 8] - b[blk * 8:blk * 8 + 8]).astype(np.float64)
            acc = (t * t + acc.astype(np.float64)).astype(np.float32)  # exact fma then round
        r = np.float32(0)
        for x in acc:
            r = np.float32(r + x)
        return r

    diff = sum(_bits(po.distance(v[i], q[i])) != _bits(wrong_order(v[i], q[i])) for i in range(v.shape[0]))
    assert diff > 0


def test_closed_forms():
    dim = 96
    z = np.zeros(dim, np.float32)
    x = np.random.default_rng(1).standard_normal(dim).astype(np.float32)
    assert _bits(po.distance(x, x)) == 0  # +0.0
    for i in range(dim):
        e = z.copy()
        e[i] = 1.0
        assert po.distance(e, z) == 1.0
        f = z.copy()
        f[(i + 5) % dim] = 2.0
        assert po.distance(f, e) == 5.0


def test_fp64_closeness():
    dim = 96
    rng = np.random.default_rng(3)
    v = rng.standard_normal((500, dim)).astype(np.float32)
    q = rng.standard_normal((500, dim)).astype(np.float32)
    for i in range(500):
        exact = float(((v[i].astype(np.float64) - q[i].astype(np.float64)) ** 2).sum())
        assert abs(po.distance(v[i], q[i]) - exact) <= 1e-6 * exact


def test_package_distance_equals_oracle():
    v, q = _pairs(96, 500, 5)
    for i in range(v.shape[0]):
        assert _bits(tools.l2(v[i], q[i])) == _bits(po.distance(v[i], q[i]))


# -- opt-graph family: compute_norm and compute_distance_with_norm ----------

@pytest.mark.parametrize("dim", [8, 16, 24, 96, 104, 128])
def test_norm_and_dot_equal_contracted_avx_build(probes, dim):
    fused, _ = probes
    v, q = _pairs(dim, 2000, 100 + dim)
    for i in range(v.shape[0]):
        n = po.norm(v[i])
        assert _bits(n) == _bits(fused.probe_norm(v[i].ctypes.data, dim)), (dim, i)
        a = po.distance_with_norm(v[i], q[i], n)
        b = fused.probe_dot_norm(v[i].ctypes.data, q[i].ctypes.data, dim, n)
        assert _bits(a) == _bits(b), (dim, i)


def test_norm_dot_unfused_order_is_distinguishable(probes):
    _, plain = probes
    dim = 96
    v, q = _pairs(dim, 2000, 11)
    dn = sum(_bits(po.norm(v[i])) != _bits(plain.probe_norm(v[i].ctypes.data, dim)) for i in range(v.shape[0]))
    dd = sum(_bits(po.distance_with_norm(v[i], q[i], 1.0)) !=
             _bits(plain.probe_dot_norm(v[i].ctypes.data, q[i].ctypes.data, dim, 1.0)) for i in range(v.shape[0]))
    assert dn > 0 and dd > 0


def test_norm_dot_integer_closed_form():
    """Small integers keep every product and partial sum exact, so
    norm - 2 v.q must equal |v - q|^2 - |q|^2 exactly (negative values too)."""
    rng = np.random.default_rng(4)
    for dim in (16, 40, 96, 104):
        for _ in range(200):
            v = rng.integers(-6, 7, dim).astype(np.float32)
            q = rng.integers(-6, 7, dim).astype(np.float32)
            n = po.norm(v)
            assert n == float((v.astype(np.int64) ** 2).sum())
            d = po.distance_with_norm(v, q, n)
            assert d == po.distance(v, q) - float((q.astype(np.int64) ** 2).sum())


def test_norm_dot_fp64_closeness():
    dim = 96
    rng = np.random.default_rng(6)
    v = rng.standard_normal((500, dim)).astype(np.float32)
    q = rng.standard_normal((500, dim)).astype(np.float32)
    for i in range(500):
        v64, q64 = v[i].astype(np.float64), q[i].astype(np.float64)
        n64 = float((v64 * v64).sum())
        n = po.norm(v[i])
        assert abs(n - n64) <= 1e-6 * n64
        dot = float((v64 * q64).sum())
        d = po.distance_with_norm(v[i], q[i], n)
        assert abs(d - (n64 - 2 * dot)) <= 1e-6 * (n64 + 2 * abs(dot))
