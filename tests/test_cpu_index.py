"""The multi-threaded CPU index (the timed CPU baseline) against the oracle: inserts,
upserts, finds, erases and adapter-shaped scans on var-length keys, 4 worker threads."""
import numpy as np

from indexingpu_amd import workloads as W
from oracle.cpu_index import CpuIndex
from oracle.oracle import Oracle


def test_cpu_index_matches_oracle():
    rng = np.random.default_rng(4)
    keys, lens = W.unique_random_keys(rng, 20000, 1, 6)
    mk, mv = 6, 3
    n = len(lens)
    vals = rng.integers(0, 1 << 32, size=(n, mv), dtype=np.uint64).astype(np.uint32)
    vl = rng.integers(1, mv + 1, size=n).astype(np.uint32)
    # a few keys share a first slice region so several shards and one shard both fill
    keys[: n // 4, 0] &= 0x00FFFFFF
    h = W.key_hash64(keys, lens)
    _, first = np.unique(h, return_index=True)
    keys, lens, vals, vl = keys[first], lens[first], vals[first], vl[first]
    n = len(lens)
    c, o = CpuIndex(shard_bits=6, threads=4), Oracle()
    half = n // 2
    flat = lambda a: np.ascontiguousarray(a).reshape(-1)
    assert np.array_equal(c.insert(flat(keys[:half]), mk, lens[:half], flat(vals[:half]), mv,
                                   vl[:half], half),
                          o.insert(flat(keys[:half]), mk, lens[:half], flat(vals[:half]), mv,
                                   vl[:half], half))
    # upsert over everything: the first half is overwritten, the second inserted
    v2 = vals ^ np.uint32(0xA5A5A5A5)
    assert np.array_equal(c.insert(flat(keys), mk, lens, flat(v2), mv, vl, n, True),
                          o.insert(flat(keys), mk, lens, flat(v2), mv, vl, n, update_if_exists=True))
    gv, gl = c.find(flat(keys), mk, lens, mv, n)
    ev, el = o.find(flat(keys), mk, lens, mv, n)
    assert np.array_equal(gl, el) and np.array_equal(gv, ev)
    er = rng.permutation(n)[: n // 3]
    assert np.array_equal(c.erase(flat(keys[er]), mk, lens[er], len(er)),
                          o.erase(flat(keys[er]), mk, lens[er], len(er)))
    assert len(c) == len(o)
    nq = 500
    lo = keys[rng.integers(0, n, size=nq)]
    lol = lens[rng.integers(0, n, size=nq)]
    cnt, row = c.scan(flat(lo), mk, lol, 20, nq)
    ec, evv, evl, ek, ekl = o.scan(flat(lo), lol, mk, 20, nq, None, None, 1)
    assert np.array_equal(cnt, ec)
    row = row.reshape(nq, 20)
    evv = evv.reshape(nq, 20)
    for q in range(nq):
        assert np.array_equal(row[q, :cnt[q]], evv[q, :cnt[q]])
        assert (row[q, cnt[q]:] == 0xFFFFFFFF).all()


def test_adapter_conventions_and_result_line():
    """The reference CPU adapter's conventions on its own CPU workload shape (2-slice
    fixed-length keys, one-slice values = the key's hash, then finds of 50% stored and
    50% fresh keys): insert overwrites, an absent key reads 0xFFFFFFFF, key order is the
    order of the big-endian bytes (scans ascend like the oracle's), and the result line
    parses the way the reference's evaluation script reads it."""
    from oracle.cpu_index import result_line
    rng = np.random.default_rng(8)
    n = 20000
    keys = rng.integers(0, 1 << 32, size=(2 * n, 2), dtype=np.uint64).astype(np.uint32)
    _, first = np.unique(keys, axis=0, return_index=True)
    keys = keys[np.sort(first)][: 2 * n]
    stored, fresh = keys[:n], keys[n:2 * n]
    vals = (W.key_hash64(stored, np.full(n, 2, np.uint32)) & 0xFFFFFFFE).astype(np.uint32)
    c, o = CpuIndex(shard_bits=6, threads=4), Oracle()
    c.put1(stored.reshape(-1), 2, None, vals, n)
    c.put1(stored[: n // 2].reshape(-1), 2, None, vals[: n // 2] ^ np.uint32(1), n // 2)  # overwrite
    exp = vals.copy()
    exp[: n // 2] ^= np.uint32(1)
    o.insert(stored.reshape(-1).copy(), 2, None, exp.copy(), 1, None, n)
    q = np.concatenate([stored, fresh])[rng.permutation(2 * n)]
    got = c.get1(q.reshape(-1), 2, None, 2 * n)
    ev, el = o.find(q.reshape(-1).copy(), 2, None, 1, 2 * n)
    assert np.array_equal(got, np.where(el > 0, ev, np.uint32(0xFFFFFFFF)))
    assert int((got == 0xFFFFFFFF).sum()) == n
    cnt, row = c.scan(q[:200].reshape(-1), 2, None, 10, 200)
    ec, evv, *_ = o.scan(q[:200].reshape(-1).copy(), None, 2, 10, 200, None, None, 1)
    assert np.array_equal(cnt, ec)
    assert np.array_equal(row.reshape(200, 10)[ec == 10], evv.reshape(200, 10)[ec == 10])
    line = result_line("find", [12.5, 13.0])
    tok = line.split(" ")  # evaluation/scripts/evaluate.py:244-255
    assert tok[0][:-1] == "find" and float(tok[1]) == 12.75 and tok[2] == "Mop/s"
    assert int(tok[3].strip("(;")) == 2 and [float(t) for t in tok[4:6]] == [12.5, 13.0]
