"""Rebuilds the inputs of the reference's Masstree unit tests from the committed draws.

Follows test/unittest_masstree.cu in hyoungjook/IndexinGPU:
  * make_input (:104-124): key i of length L has slice s = ((i >> (L-1-s)) + 1) * 2 + s
    (u32 arithmetic, shift taken modulo 32 as the x86 shift instruction does);
    "not-exist" keys add 1 to every slice; values = i + 1, update values = i + 7 in
    every slice;
  * duplicate (:125-139): the second half repeats the first half;
  * sort (:149-184): lexicographic order of the slices;
  * shuffle (:185-215) and the ConcurrentMix shuffle (:593-602) use std::shuffle with
    std::mt19937(0), whose output is implementation-defined, so the permutations come
    from tests/golden/unittest_inputs.npz (made by tests/golden/make_unittest_fixture.py);
  * range queries (:230-279); ConcurrentMix (:571-708).
"""
import os

import numpy as np

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden",
                      "unittest_inputs.npz")
CONFIGS = [(1, 1, 1, 1), (1, 200, 1, 1), (1, 1, 1, 200), (1, 200, 1, 200)]
NUM_KEYS = 1024

_fix = None


def fixture():
    global _fix
    if _fix is None:
        _fix = dict(np.load(GOLDEN))
    return _fix


def key_checksum(keys, lens, maxk):
    h = 1469598103934665603
    for i in range(len(lens)):
        for s in range(int(lens[i])):
            h = ((h ^ int(keys[i * maxk + s])) * 1099511628211) & ((1 << 64) - 1)
    return h


class TestingInput:
    """testing_input (test/unittest_masstree.cu:82-230) for config index c."""

    def __init__(self, c, duplicate=False):
        kmin, kmax, vmin, vmax = CONFIGS[c]
        f = fixture()
        n = NUM_KEYS
        self.num_keys = n
        self.max_key_length = kmax
        self.max_value_length = vmax
        klen = f[f"cfg{c}_key_lengths"].astype(np.uint32)
        vlen = f[f"cfg{c}_value_lengths"].astype(np.uint32)
        keys = np.zeros((n, kmax), np.uint32)
        i = np.arange(n, dtype=np.uint64)
        for s in range(kmax):
            live = klen > s
            sh = ((klen.astype(np.int64) - 1 - s) & 31).astype(np.uint64)
            eff = (i >> sh).astype(np.uint64)
            sl = (((eff + 1) & 0xFFFFFFFF) * 2 + s) & 0xFFFFFFFF
            keys[:, s] = np.where(live, sl, 0).astype(np.uint32)
        vals = np.zeros((n, vmax), np.uint32)
        vals2 = np.zeros((n, vmax), np.uint32)
        for s in range(vmax):
            live = vlen > s
            vals[:, s] = np.where(live, (i + 1).astype(np.uint32), 0)
            vals2[:, s] = np.where(live, (i + 7).astype(np.uint32), 0)
        kne = np.where(np.arange(kmax)[None, :] < klen[:, None], keys + np.uint32(1), 0)
        self.keys = keys
        self.keys_not_exist = kne.astype(np.uint32)
        self.lengths = klen
        self.values = vals
        self.values2 = vals2
        self.value_lengths = vlen
        if duplicate:
            h = (n + 1) // 2
            for arr in (self.keys, self.keys_not_exist, self.values, self.values2):
                arr[h:] = arr[:n - h]
            self.lengths[h:] = self.lengths[:n - h]
            self.value_lengths[h:] = self.value_lengths[:n - h]

    def flat(self, name):
        return np.ascontiguousarray(getattr(self, name)).reshape(-1)

    def reorder(self, order):
        for name in ("keys", "keys_not_exist", "values", "values2"):
            setattr(self, name, np.ascontiguousarray(getattr(self, name)[order]))
        self.lengths = self.lengths[order].copy()
        self.value_lengths = self.value_lengths[order].copy()

    def sort(self):
        order = sorted(range(self.num_keys),
                       key=lambda j: tuple(self.keys[j, :self.lengths[j]].tolist()))
        self.reorder(np.array(order))

    def shuffle(self):
        self.reorder(fixture()["input_shuffle_order"].astype(np.int64))


class RangeInput:
    """testing_range_input (test/unittest_masstree.cu:230-303); sorts `inp`."""

    def __init__(self, inp, num_queries, max_count):
        inp.sort()
        n, mk, mv = inp.num_keys, inp.max_key_length, inp.max_value_length
        self.num_queries = num_queries
        self.max_count = max_count
        self.lower = np.zeros((num_queries, mk), np.uint32)
        self.upper = np.zeros((num_queries, mk), np.uint32)
        self.lower_len = np.zeros(num_queries, np.uint32)
        self.upper_len = np.zeros(num_queries, np.uint32)
        self.counts = np.zeros(num_queries, np.uint32)
        self.begin = np.zeros(num_queries, np.int64)
        for q in range(num_queries):
            begin = q % n
            count = 1 + (q % (2 * max_count))
            end = min(begin + count - 1, n - 1)
            count = end - begin + 1
            self.lower[q] = inp.keys[begin]
            self.lower_len[q] = inp.lengths[begin]
            self.upper[q] = inp.keys[end]
            self.upper_len[q] = inp.lengths[end]
            self.counts[q] = min(count, max_count)
            self.begin[q] = begin


class ConcurrentMix:
    """The mixed batch of test_concurrentmix (test/unittest_masstree.cu:571-626)."""

    def __init__(self, inp):
        inp.shuffle()
        n = inp.num_keys
        ks = n // 4
        self.num_keyset = ks
        m = 10 * ks
        mk, mv = inp.max_key_length, inp.max_value_length
        order = fixture()["mix_shuffle_order"].astype(np.int64)
        assert len(order) == m
        self.types = np.zeros(m, np.uint8)
        self.keys = np.zeros((m, mk), np.uint32)
        self.lens = np.zeros(m, np.uint32)
        self.values = np.zeros((m, mv), np.uint32)
        self.vlens = np.zeros(m, np.uint32)
        self.dst_of = np.zeros(m, np.int64)  # shuffle_order_inverse
        plan = [(0, 0, 0), (ks, 2 * ks, 0), (2 * ks, 0, 3), (3 * ks, 3 * ks, 3),
                (4 * ks, 0, 2), (5 * ks, ks, 2), (6 * ks, 2 * ks, 2), (7 * ks, 3 * ks, 2),
                (8 * ks, ks, 1), (9 * ks, 3 * ks, 1)]
        for dst_begin, src_begin, t in plan:
            for i in range(ks):
                src, dst = src_begin + i, dst_begin + i
                sd = order[dst]
                self.types[sd] = t
                self.keys[sd] = inp.keys[src]
                self.lens[sd] = inp.lengths[src]
                if t in (0, 3):
                    self.values[sd] = inp.values[src] if t == 0 else inp.values2[src]
                    self.vlens[sd] = inp.value_lengths[src]
                self.dst_of[sd] = dst
