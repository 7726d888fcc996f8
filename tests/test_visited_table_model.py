"""CPU model of the search kernel's visited set (the LDS tag table with two
buckets per id and a bitmap for the rest), written from its arithmetic:
scramble, divide by the bucket count with a multiply-high reciprocal, second
bucket at a tag-derived offset.  Membership must equal a plain set for any
insertion sequence, table size and id space -- the property that keeps the
GPU's distance counts equal to the reference's boost::dynamic_bitset ones.
"""
import numpy as np
import pytest

EMPTY = 0xFFFF


class TableModel:
    def __init__(self, slots, n, bucket_slots=0):
        self.id_bits = max(1, int(np.ceil(np.log2(n)))) if n > 1 else 1
        min_nb = ((1 << self.id_bits) - 1) // 0x7FFE + 1
        # 8-slot buckets, or 4 when an 8-slot table this size has too few,
        # or 2 for a table of at most half the least 4-slot one
        nb_min = (min_nb + 7) // 8 * 8
        bs = bucket_slots or (8 if slots // 8 >= min_nb else 2 if 2 * slots <= 4 * nb_min else 4)
        nb = max(slots // bs, nb_min, 8)
        self.nb, self.bucket_slots = nb, bs
        self.magic = (1 << 32) // nb
        self.idmask = (1 << self.id_bits) - 1
        self.xshift = self.id_bits // 2 + 1
        self.buckets = [[EMPTY] * bs for _ in range(nb)]
        self.bitmap = set()

    def scramble(self, i):
        h = (i * 0x9E3779B1) & 0xFFFFFFFF & self.idmask
        return h ^ (h >> self.xshift)

    def _bucket(self, b, tag):
        s = self.buckets[b]
        if tag in s:
            return 0
        if EMPTY in s:
            s[s.index(EMPTY)] = tag
            return 1
        return 2

    def claim(self, i):
        """True when i was not visited before (and is now)."""
        h = self.scramble(i)
        tag = (h * self.magic) >> 32
        b1 = h - tag * self.nb
        if b1 >= self.nb:
            b1 -= self.nb
            tag += 1
        assert tag < 0x8000
        r = self._bucket(b1, tag)
        if r == 2 and tag != 0x7FFF:
            b2 = b1 + 1 + ((((tag * 0x85EBCA77) & 0xFFFFFFFF) * (self.nb - 1)) >> 32)
            if b2 >= self.nb:
                b2 -= self.nb
            r = self._bucket(b2, tag | 0x8000)
        if r == 2:
            new = i not in self.bitmap
            self.bitmap.add(i)
            return new
        return r == 1


def test_scramble_is_a_bijection():
    for n in (1000, 20000, 1 << 16):
        t = TableModel(64, n)
        vals = {t.scramble(i) for i in range(1 << t.id_bits)}
        assert len(vals) == 1 << t.id_bits


@pytest.mark.parametrize("slots,n,seed", [(64, 20000, 0), (192, 20000, 1), (1536, 1_000_000, 2),
                                          (4864, 1_000_000, 3), (8192, 10_000_000, 4), (64, 50_000_000, 5)])
def test_membership_equals_a_set(slots, n, seed):
    rng = np.random.default_rng(seed)
    t = TableModel(slots, n)
    seen = set()
    # repeats and clustered ids, as a search offers them
    ids = np.concatenate([rng.integers(0, n, 6000), rng.integers(0, n, 200).repeat(4),
                          rng.integers(max(0, n - 3000), n, 1500)])
    rng.shuffle(ids)
    for i in ids.tolist():
        assert t.claim(i) == (i not in seen), i
        seen.add(i)
    stored = sum(x != EMPTY for s in t.buckets for x in s)
    assert stored + len(t.bitmap) == len(seen)


def test_second_buckets_take_the_overflow():
    """At load 0.7 the second bucket keeps nearly every id out of the bitmap,
    where one bucket per id overflows several percent."""
    rng = np.random.default_rng(9)
    n, slots = 1_000_000, 8192
    ids = rng.choice(n, 5700, replace=False).tolist()
    t = TableModel(slots, n)
    for i in ids:
        t.claim(i)
    assert len(t.bitmap) < 0.01 * len(ids)


@pytest.mark.parametrize("bs", [4, 2])
@pytest.mark.parametrize("slots,n,seed", [(64, 20000, 6), (1536, 1_000_000, 7), (8192, 50_000_000, 8),
                                          (4096, 10_000_000, 9), (4096, 50_000_000, 10)])
def test_narrow_buckets_membership_equals_a_set(slots, n, seed, bs):
    rng = np.random.default_rng(seed)
    t = TableModel(slots, n, bucket_slots=bs)
    assert t.bucket_slots == bs
    seen = set()
    ids = np.concatenate([rng.integers(0, n, 6000), rng.integers(0, n, 200).repeat(4)])
    rng.shuffle(ids)
    for i in ids.tolist():
        assert t.claim(i) == (i not in seen), i
        seen.add(i)


def test_automatic_bucket_width_keeps_the_table_size():
    """Above 2^25 ids an 8-slot table of 8,192 slots has too few buckets for
    15-bit tags; 4-slot buckets keep it at about 8,192 slots instead of
    doubling it."""
    assert TableModel(8192, 1_000_000).bucket_slots == 8
    assert TableModel(8192, 10_000_000).bucket_slots == 8
    t = TableModel(8192, 50_000_000)
    assert t.bucket_slots == 4 and t.nb * 4 <= 8192 + 64
    t8 = TableModel(8192, 50_000_000, bucket_slots=8)
    assert t8.nb * 8 > 16000
    # 2-slot buckets: the least bucket count (2,056 at 26-bit ids) in half the
    # slots, chosen for a table of at most half the least 4-slot one
    t2 = TableModel(4096, 50_000_000, bucket_slots=2)
    assert t2.nb == 2056 and t2.nb * 2 == 4112
    assert TableModel(3840, 50_000_000).bucket_slots == 2
    assert TableModel(4160, 50_000_000).bucket_slots == 4
    assert TableModel(3840, 10_000_000).bucket_slots == 4  # 480 eight-slot buckets < 513
    t = TableModel(1024, 10_000_000)
    assert t.bucket_slots == 2 and t.nb == 520
