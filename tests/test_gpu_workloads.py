"""GPU parity on the benchmark workload shapes, at sizes the oracle replays in seconds,
plus edge cases.  Every comparison is bit-exact against the oracle on the same inputs."""
import numpy as np
import pytest

import scenarios as S
from indexingpu_amd import workloads as W

pytestmark = pytest.mark.gpu


def _eq(a, b):
    assert np.array_equal(np.asarray(a), np.asarray(b))


@pytest.mark.parametrize("hints", [1, 0xFFFFFFFF], ids=["write_hints", "from_root"])
@pytest.mark.parametrize("tile", [16, 64])
def test_plain_batch_varlen(tile, hints):
    """Plain batch path shape: var-len keys of 2..8 slices, 2-slice values, one find
    batch with 50% hits, one erase batch, then a find of everything; writes start at
    border nodes found by a read-only pass, or at the root."""
    d = W.plain_batch_workload(200_000, seed=5)
    mk, mv = d["max_key_length"], d["max_value_length"]
    g, o = S.GpuBackend(tile, pool_bytes=256 << 20), S.OracleBackend()
    g.t.set_write_hints(hints)
    _eq(g.insert(d["keys"], d["lens"], d["values"], None, mk, mv),
        o.insert(d["keys"], d["lens"], d["values"], None, mk, mv))
    for a, b in zip(g.find(d["find_keys"], d["find_lens"], mk, mv),
                    o.find(d["find_keys"], d["find_lens"], mk, mv)):
        _eq(a, b)
    _eq(g.erase(d["erase_keys"], d["erase_lens"], mk), o.erase(d["erase_keys"], d["erase_lens"], mk))
    for a, b in zip(g.find(d["keys"], d["lens"], mk, mv, concurrent=True),
                    o.find(d["keys"], d["lens"], mk, mv)):
        _eq(a, b)
    assert g.items() == o.items()
    ok, *_, msg = g.validate()
    assert ok, msg


def test_cpu_path_shape_fixed_length():
    """CPU-runnable case shape: fixed 2-slice keys (key_lengths = null), 1-slice values."""
    d = W.cpu_path_workload(100_000, seed=7)
    g, o = S.GpuBackend(16, pool_bytes=128 << 20), S.OracleBackend()
    n = len(d["keys"])
    import torch
    keys = g._d(d["keys"])
    vals = g._d(d["values"])
    g.t.insert(keys, 2, None, vals, 1, None, n)
    o.o.insert(d["keys"].reshape(-1).copy(), 2, None, d["values"].reshape(-1).copy(), 1, None, n)
    fv = g._zeros(n)
    fl = g._zeros(n)
    g.t.find(g._d(d["find_keys"]), 2, None, fv, 1, fl, n)
    ev, el = o.o.find(d["find_keys"].reshape(-1).copy(), 2, None, 1, n)
    _eq(g._h(fv)[:n], ev)
    _eq(g._h(fl)[:n], el)
    g.t.check_device_error()
    assert int((el > 0).sum()) == n // 2


@pytest.mark.parametrize("split", [0, 0xFFFFFFFF], ids=["finds_pass", "one_launch"])
@pytest.mark.parametrize("tile", [16, 32])
def test_mixed_batches_and_scans(tile, split):
    """Mixed shape: Zipf(0.99) finds / fresh inserts / erases of unread keys, then range
    scans with random lower bounds and per-query limits in [1, 100]; mixed batches run as
    one launch or as a read-only pass over their finds followed by the other requests."""
    wl = W.MixedWorkload(200_000, key_len=4, seed=13)
    mk, mv = 4, 2
    g, o = S.GpuBackend(tile, pool_bytes=256 << 20), S.OracleBackend()
    g.t.set_mixed_split(split)
    _eq(g.insert(wl.keys, wl.lens, wl.values, None, mk, mv),
        o.insert(wl.keys, wl.lens, wl.values, None, mk, mv))
    for _ in range(4):
        bt = wl.batch(50_000)
        vl = np.full(len(bt["types"]), mv, np.uint32)
        for a, b in zip(g.mixed(bt["types"], bt["keys"], bt["lens"], bt["values"], vl, mk, mv),
                        o.mixed(bt["types"], bt["keys"], bt["lens"], bt["values"], vl, mk, mv)):
            _eq(a, b)
    sc = wl.scans(4096, max_scan=100)
    nq = len(sc["limits"])
    lens = np.full(nq, mk, np.uint32)
    for a, b in zip(g.scan(sc["lower"], lens, mk, 100, maxV=mv, limits=sc["limits"]),
                    o.scan(sc["lower"], lens, mk, 100, maxV=mv, limits=sc["limits"])):
        _eq(a, b)
    assert g.items() == o.items()


@pytest.mark.parametrize("hints", [1, 0xFFFFFFFF], ids=["write_hints", "from_root"])
def test_layers_push_down_and_long_keys(hints):
    """Keys sharing long prefixes (trie layers, suffix push-downs) and keys longer than
    a tile (global key reads, multi-slab records), with scans across layers."""
    rng = np.random.default_rng(21)
    n, mk, mv = 20_000, 80, 40
    lens = rng.integers(1, mk + 1, size=n).astype(np.uint32)
    keys = rng.integers(0, 4, size=(n, mk)).astype(np.uint32)
    keys[:, :40] = (keys[:, :40] % 2)  # long shared prefixes
    keys[np.arange(mk)[None, :] >= lens[:, None]] = 0
    h = W.key_hash64(keys, lens)
    _, first = np.unique(h, return_index=True)
    first = np.sort(first)
    keys, lens = np.ascontiguousarray(keys[first]), lens[first]
    n = len(lens)
    vlens = rng.integers(0, mv + 1, size=n).astype(np.uint32)
    vals = rng.integers(0, 1 << 32, size=(n, mv), dtype=np.uint64).astype(np.uint32)
    for tile in (16, 64):
        g, o = S.GpuBackend(tile, pool_bytes=512 << 20), S.OracleBackend()
        g.t.set_write_hints(hints)
        half = n // 2
        _eq(g.insert(keys[:half], lens[:half], vals[:half], vlens[:half], mk, mv),
            o.insert(keys[:half], lens[:half], vals[:half], vlens[:half], mk, mv))
        _eq(g.insert(keys, lens, vals, vlens, mk, mv), o.insert(keys, lens, vals, vlens, mk, mv))
        for a, b in zip(g.find(keys, lens, mk, mv), o.find(keys, lens, mk, mv)):
            _eq(a, b)
        q = rng.choice(n, 500, replace=False)
        u = rng.choice(n, 500, replace=False)
        for a, b in zip(g.scan(keys[q], lens[q], mk, 30, keys[u], lens[u], maxV=mv),
                        o.scan(keys[q], lens[q], mk, 30, keys[u], lens[u], maxV=mv)):
            _eq(a, b)
        e = rng.choice(n, n // 3, replace=False)
        _eq(g.erase(keys[e], lens[e], mk), o.erase(keys[e], lens[e], mk))
        assert g.items() == o.items()
        ok, *_, msg = g.validate()
        assert ok, msg


def test_upsert_update_and_zero_length_values():
    """insert<update_if_exists=true>, update, and values of length 0 (stored, read back
    as length 0, but the key still exists for insert/update/erase)."""
    rng = np.random.default_rng(3)
    keys, lens = W.unique_random_keys(rng, 30_000, 1, 6)
    mk, mv = 6, 3
    vl = rng.integers(0, mv + 1, size=len(lens)).astype(np.uint32)
    vals = rng.integers(0, 1 << 32, size=(len(lens), mv), dtype=np.uint64).astype(np.uint32)
    g, o = S.GpuBackend(16, pool_bytes=128 << 20), S.OracleBackend()
    _eq(g.insert(keys, lens, vals, vl, mk, mv), o.insert(keys, lens, vals, vl, mk, mv))
    vl2 = rng.integers(0, mv + 1, size=len(lens)).astype(np.uint32)
    vals2 = vals ^ np.uint32(0x5A5A5A5A)
    half = len(lens) // 2
    _eq(g.insert(keys[:half], lens[:half], vals2[:half], vl2[:half], mk, mv, upsert=True),
        o.insert(keys[:half], lens[:half], vals2[:half], vl2[:half], mk, mv, upsert=True))
    _eq(g.update(keys[half:], lens[half:], vals2[half:], vl2[half:], mk, mv),
        o.update(keys[half:], lens[half:], vals2[half:], vl2[half:], mk, mv))
    _eq(g.insert(keys, lens, vals, vl, mk, mv), o.insert(keys, lens, vals, vl, mk, mv))
    for a, b in zip(g.find(keys, lens, mk, mv), o.find(keys, lens, mk, mv)):
        _eq(a, b)
    assert g.items() == o.items()


def test_empty_batches_and_tiny_pool_error():
    g = S.GpuBackend(16, pool_bytes=1 << 16)  # 512 slabs
    import torch
    e = g._zeros(1)
    g.t.find(e, 1, None, e, 1, e, 0)
    g.t.insert(e, 1, None, e, 1, None, 0)
    g.t.check_device_error()
    rng = np.random.default_rng(0)
    keys, lens = W.unique_random_keys(rng, 20_000, 3, 3)
    vals = np.zeros((len(lens), 2), np.uint32)
    r = g._zeros(len(lens), np.uint8)
    g.t.insert(g._d(keys), 3, g._d(lens), g._d(vals), 2, None, len(lens), results=r)
    assert g.t.device_error() == 1  # pool exhausted is reported, not a hang


def test_tiny_pool_keeps_working_after_exhaustion():
    """A pool of 512 slabs runs out during an insert batch: some inserts report false,
    splits may leave right halves without a parent slot, records whose slabs ran out part
    way are freed.  The index stays valid and usable: the successful inserts equal the
    oracle fed only them, finds of every key agree, erases return slabs, and the failed
    keys then go in."""
    g, o = S.GpuBackend(16, pool_bytes=1 << 16), S.OracleBackend()
    rng = np.random.default_rng(3)
    keys, lens = W.unique_random_keys(rng, 6000, 2, 3)
    keys[:, 0] %= 64  # shared first slices: suffix records, push-downs, layers
    _, first = np.unique(keys, axis=0, return_index=True)
    keys, lens = keys[np.sort(first)], lens[np.sort(first)]
    vals = rng.integers(0, 1 << 32, size=(len(lens), 2), dtype=np.uint64).astype(np.uint32)
    mk, mv = 3, 2
    n = len(lens)
    r = g._zeros(n, np.uint8)
    g.t.insert(g._d(keys), mk, g._d(lens), g._d(vals), mv, None, n, results=r)
    ok = g._h(r, np.uint8)[:n].astype(bool)
    assert g.t.device_error() == 1  # out of slabs
    assert 0 < ok.sum() < n
    idx = np.nonzero(ok)[0]
    o.insert(keys[idx], lens[idx], vals[idx], None, mk, mv)
    good, *_, msg = g.validate()
    assert good, msg
    assert g.items() == o.items()
    for a, b in zip(g.find(keys, lens, mk, mv), o.find(keys, lens, mk, mv)):
        _eq(a, b)
    # erase most of what is stored (records go back to the pool), then insert the rest
    er = idx[: max(1, len(idx) * 3 // 4)]
    _eq(g.erase(keys[er], lens[er], mk), o.erase(keys[er], lens[er], mk))
    rest = np.nonzero(~ok)[0][: len(er) // 2]
    r2 = g._zeros(len(rest), np.uint8)
    g.t.insert(g._d(keys[rest]), mk, g._d(lens[rest]), g._d(vals[rest]), mv, None, len(rest),
               results=r2)
    ok2 = g._h(r2, np.uint8)[:len(rest)].astype(bool)
    g.t.device_error()  # may run out again; results say which went in
    assert ok2.sum() > 0
    idx2 = rest[ok2]
    o.insert(keys[idx2], lens[idx2], vals[idx2], None, mk, mv)
    good, *_, msg = g.validate()
    assert good, msg
    assert g.items() == o.items()


@pytest.mark.parametrize("order", ["ascending", "descending", "two_ranges"])
def test_key_ordered_insert_batches(order):
    """Insert batches in key order (and one whose second half lies past everything in
    the first) grow the tree through the scattered growth stages; every insert succeeds
    once and the entries equal the oracle's."""
    rng = np.random.default_rng(77)
    keys, lens = W.unique_random_keys(rng, 300_000, 1, 4)
    mk, mv = 4, 2
    tup = [tuple(keys[i, :lens[i]].tolist()) for i in range(len(lens))]
    idx = np.array(sorted(range(len(lens)), key=lambda i: tup[i]))
    if order == "descending":
        idx = idx[::-1]
    elif order == "two_ranges":
        half = len(idx) // 2
        idx = np.concatenate([rng.permutation(idx[:half]), rng.permutation(idx[half:])])
    keys, lens = np.ascontiguousarray(keys[idx]), lens[idx].copy()
    vals = W.make_values(keys, lens, mv)
    g, o = S.GpuBackend(16, pool_bytes=128 << 20), S.OracleBackend()
    r = g.insert(keys, lens, vals, None, mk, mv)
    assert r.all()
    assert np.array_equal(r, o.insert(keys, lens, vals, None, mk, mv))
    assert g.items() == o.items()
    ok, *_ = g.t.validate()
    assert ok


@pytest.mark.parametrize("find_tile", [1, 2, 4])
def test_mixed_finds_pass_long_keys_and_values(find_tile):
    """Mixed batches whose finds run as a read-only pass, over keys of 1..40 slices that
    share prefixes (trie layers, suffix records) and values of 0..70 slices (records of
    several slabs): finds of present and absent keys, inserts of fresh keys with long
    values, erases of present keys none of the finds read."""
    rng = np.random.default_rng(31)
    mk, mv = 40, 70
    base, blen = W.unique_random_keys(rng, 6000, 1, mk)
    keys, lens = base.copy(), blen.copy()
    donor = rng.integers(0, len(lens), size=len(lens))
    share = rng.integers(0, mk, size=len(lens))
    for i in range(0, len(lens), 3):
        p = min(share[i], lens[i] - 1, lens[donor[i]])
        keys[i, :p] = base[donor[i], :p]
    _, first = np.unique(W.key_hash64(keys, lens), return_index=True)
    first = np.sort(first)
    keys, lens = keys[first], lens[first]
    n = len(lens)
    vals = rng.integers(0, 1 << 32, size=(n, mv), dtype=np.uint64).astype(np.uint32)
    vl = rng.choice([0, 1, 2, 5, 30, 31, 62, 70], size=n).astype(np.uint32)
    g, o = S.GpuBackend(16, pool_bytes=256 << 20), S.OracleBackend()
    g.t.set_mixed_split(0)
    g.t.set_find_tile(find_tile)
    pre = n // 2
    _eq(g.insert(keys[:pre], lens[:pre], vals[:pre], vl[:pre], mk, mv),
        o.insert(keys[:pre], lens[:pre], vals[:pre], vl[:pre], mk, mv))
    perm_pre = rng.permutation(pre)
    perm_new = pre + rng.permutation(n - pre)
    finds = np.concatenate([perm_pre[: pre // 2], perm_new[: (n - pre) // 3]])
    erases = perm_pre[pre // 2:]
    inserts = perm_new[(n - pre) // 3:]
    idx = np.concatenate([finds, erases, inserts])
    types = np.concatenate([np.full(len(finds), 2), np.full(len(erases), 1),
                            np.full(len(inserts), 0)]).astype(np.uint8)
    order = rng.permutation(len(idx))
    idx, types = idx[order], types[order]
    bk, bl, bv, bvl = keys[idx], lens[idx], vals[idx], vl[idx]
    for a, b in zip(g.mixed(types, bk, bl, bv, bvl, mk, mv), o.mixed(types, bk, bl, bv, bvl, mk, mv)):
        _eq(a, b)
    assert g.items() == o.items()


def test_keys_of_2_15_slices_or_more():
    """max_key_length >= 2^15: the two-pass find parks keystate and layer only and its
    record pass reads the key lengths (shorter maximum lengths pack the tail length into the
    park word instead).  Short and very long keys sharing first slices (suffix records with
    tails of tens of thousands of slices), found present and absent, against the oracle."""
    rng = np.random.default_rng(41)
    n, mk, mv = 96, 40_000, 3
    lens = np.concatenate([rng.integers(1, 7, size=32), rng.integers(32_768, mk + 1, size=64)])
    lens = lens.astype(np.uint32)
    keys = rng.integers(0, 1 << 32, size=(n, mk), dtype=np.uint64).astype(np.uint32)
    keys[:, 0] %= 4  # shared first slices: suffix records, push-downs
    keys[np.arange(mk)[None, :] >= lens[:, None]] = 0
    h = W.key_hash64(keys, lens)
    _, first = np.unique(h, return_index=True)
    first = np.sort(first)
    keys, lens = np.ascontiguousarray(keys[first]), lens[first]
    n = len(lens)
    vals = rng.integers(0, 1 << 32, size=(n, mv), dtype=np.uint64).astype(np.uint32)
    vlens = np.full(n, mv, np.uint32)
    g, o = S.GpuBackend(16, pool_bytes=512 << 20), S.OracleBackend()
    _eq(g.insert(keys, lens, vals, vlens, mk, mv), o.insert(keys, lens, vals, vlens, mk, mv))
    # absent keys: the last slice of each key changed (same length, same tail length)
    absent = keys.copy()
    absent[np.arange(n), lens - 1] ^= 0x5A5A5A5A
    for q in (keys, absent):
        for a, b in zip(g.find(q, lens, mk, mv), o.find(q, lens, mk, mv)):
            _eq(a, b)
    ok, *_, msg = g.validate()
    assert ok, msg


@pytest.mark.parametrize("tile", [16, 32])
def test_churn_keeps_the_pool_bounded(tile):
    """Rounds that erase every key under half of the live prefixes and insert keys under
    as many new prefixes (plain erase and insert batches, and mixed batches): the layers
    of a retired prefix empty out and are unlinked, and their nodes are retired with the
    records, so the slabs in use stay bounded over the rounds (kept, every round would
    leave a retired prefix's empty layers behind).  Every result and the final entries
    equal the oracle's, and the structure validates."""
    rng = np.random.default_rng(55)
    n_pre, per, mk, mv = 2000, 6, 6, 2

    def keys_under(prefixes):
        n = len(prefixes) * per
        k = np.zeros((n, mk), np.uint32)
        k[:, :3] = np.repeat(prefixes, per, axis=0)
        k[:, 3:] = rng.integers(0, 1 << 32, size=(n, 3), dtype=np.uint64).astype(np.uint32)
        k[:, 3] %= 3  # a layer under each prefix slice, and under each of its key[3] values
        lens = rng.integers(4, mk + 1, size=n).astype(np.uint32)
        k[np.arange(mk)[None, :] >= lens[:, None]] = 0
        _, first = np.unique(k, axis=0, return_index=True)
        first = np.sort(first)
        owner = np.repeat(np.arange(len(prefixes)), per)[first]
        return k[first], lens[first], owner

    g, o = S.GpuBackend(tile, pool_bytes=256 << 20), S.OracleBackend()
    pre = rng.integers(0, 1 << 32, size=(n_pre, 3), dtype=np.uint64).astype(np.uint32)
    keys, lens, owner = keys_under(pre)
    vals = rng.integers(0, 1 << 32, size=(len(lens), mv), dtype=np.uint64).astype(np.uint32)
    _eq(g.insert(keys, lens, vals, None, mk, mv), o.insert(keys, lens, vals, None, mk, mv))
    used = []
    for rnd in range(8):
        old = rng.permutation(n_pre)[: n_pre // 2]  # prefixes whose keys all go
        gone = np.isin(owner, old)
        e = np.nonzero(gone)[0]
        newp = rng.integers(0, 1 << 32, size=(len(old), 3), dtype=np.uint64).astype(np.uint32)
        nk, nl, nown = keys_under(newp)
        nv = rng.integers(0, 1 << 32, size=(len(nl), mv), dtype=np.uint64).astype(np.uint32)
        if rnd % 2 == 0:
            _eq(g.erase(keys[e], lens[e], mk), o.erase(keys[e], lens[e], mk))
            _eq(g.insert(nk, nl, nv, None, mk, mv), o.insert(nk, nl, nv, None, mk, mv))
        else:
            types = np.concatenate([np.ones(len(e), np.uint8), np.zeros(len(nl), np.uint8)])
            perm = rng.permutation(len(types))
            mkeys = np.concatenate([keys[e], nk])[perm]
            mlens = np.concatenate([lens[e], nl])[perm]
            mvals = np.concatenate([vals[e], nv])[perm]
            vl = np.full(len(types), mv, np.uint32)
            for a, b in zip(g.mixed(types[perm], mkeys, mlens, mvals, vl, mk, mv),
                            o.mixed(types[perm], mkeys, mlens, mvals, vl, mk, mv)):
                _eq(a, b)
        # surviving prefixes keep their slots; the new ones take the retired slots
        pre[old] = newp
        keys = np.concatenate([keys[~gone], nk])
        lens = np.concatenate([lens[~gone], nl])
        vals = np.concatenate([vals[~gone], nv])
        owner = np.concatenate([owner[~gone], old[nown]])
        used.append(g.t.slabs_allocated())
    assert g.items() == o.items()
    ok, *_, msg = g.validate()
    assert ok, msg
    assert max(used[4:]) <= 1.1 * max(used[:2]), used


@pytest.mark.parametrize("hints", [1, 0xFFFFFFFF], ids=["write_hints", "from_root"])
@pytest.mark.parametrize("tile", [16, 64])
def test_erase_merges_empty_border_nodes(tile, hints):
    """Erases that empty border nodes fold them into a neighbour and retire them: after
    erasing a key range and a random share of the rest, the structure validates with
    fewer than half the nodes; then rounds of mixed batches (re-inserts, erases, finds of
    other keys) over the same ranges, and every key erased at the end.  Results equal the
    oracle's; the slabs in use stay bounded over the rounds."""
    rng = np.random.default_rng(81)
    n, mk, mv = 200_000, 2, 1
    keys, lens = W.unique_random_keys(rng, n, 1, 2)
    keys[:, 0] %= 100_000  # ~2 keys per first slice: layers below, many border nodes
    k = np.where(np.arange(mk)[None, :] < lens[:, None], keys, 0).astype(np.uint32)
    _, first = np.unique(k, axis=0, return_index=True)
    keys, lens = k[np.sort(first)], lens[np.sort(first)]
    n = len(lens)
    vals = rng.integers(0, 1 << 32, size=(n, mv), dtype=np.uint64).astype(np.uint32)
    g, o = S.GpuBackend(tile, pool_bytes=256 << 20), S.OracleBackend()
    g.t.set_write_hints(hints)
    _eq(g.insert(keys, lens, vals, None, mk, mv), o.insert(keys, lens, vals, None, mk, mv))
    ok, _, full, _, msg = g.validate()
    assert ok, msg
    gone = ((keys[:, 0] >= 20_000) & (keys[:, 0] < 80_000)) | (rng.random(n) < 0.3)
    e = np.nonzero(gone)[0]
    _eq(g.erase(keys[e], lens[e], mk), o.erase(keys[e], lens[e], mk))
    assert g.items() == o.items()
    ok, _, after, _, msg = g.validate()
    assert ok, msg
    assert after < full // 2, (full, after)
    used = []
    for rnd in range(6):
        back = rng.permutation(e)[: len(e) // 2]
        rest = np.setdiff1d(np.arange(n), e)
        out = rng.permutation(rest)[: len(rest) // 2]
        probe = np.setdiff1d(rng.permutation(n)[:20_000], np.concatenate([back, out]))
        types = np.concatenate([np.zeros(len(back), np.uint8), np.ones(len(out), np.uint8),
                                np.full(len(probe), 2, np.uint8)])
        idx = np.concatenate([back, out, probe])
        p = rng.permutation(len(idx))
        types, idx = types[p], idx[p]
        v = vals[idx].copy()
        vl = np.full(len(idx), mv, np.uint32)
        for a, b in zip(g.mixed(types, keys[idx], lens[idx], v, vl, mk, mv),
                        o.mixed(types, keys[idx], lens[idx], v.copy(), vl, mk, mv)):
            _eq(a, b)
        e = np.union1d(np.setdiff1d(e, back), out)
        used.append(g.t.slabs_allocated())
    assert g.items() == o.items()
    ok, *_, msg = g.validate()
    assert ok, msg
    assert max(used[3:]) <= 1.1 * max(used[:2]), used
    _eq(g.erase(keys, lens, mk), o.erase(keys, lens, mk))
    assert g.items() == []
    ok, _, left, _, msg = g.validate()
    assert ok, msg
    assert left < full // 10, (full, left)


@pytest.mark.parametrize("fanout", [(3, 4), (8, 6)], ids=["b3i4", "b8i6"])
@pytest.mark.parametrize("tile", [16, 64])
def test_small_fanout(tile, fanout):
    """Node fanout below the layout's maximum (the autotuner's fanout knob): inserts of
    var-len keys into an empty tree split nearly every border node and carry splits up
    several interior levels; then mixed batches (split into a read-only pass and the
    writes), scans, erases that empty and merge nodes, and finds of every key.  Results
    equal the oracle's, the structure validates, and the fanout then returns to the
    maximum on the same tree for more mixed batches."""
    wl = W.MixedWorkload(100_000, key_len=4, seed=19)
    mk, mv = 4, 2
    g, o = S.GpuBackend(tile, pool_bytes=256 << 20, fanout=fanout), S.OracleBackend()
    _eq(g.insert(wl.keys, wl.lens, wl.values, None, mk, mv),
        o.insert(wl.keys, wl.lens, wl.values, None, mk, mv))
    ok, _, nodes_small, _, msg = g.validate()
    assert ok, msg
    for rnd in range(4):
        if rnd == 2:
            g.t.set_fanout(0, 0)
        bt = wl.batch(30_000)
        vl = np.full(len(bt["types"]), mv, np.uint32)
        for a, b in zip(g.mixed(bt["types"], bt["keys"], bt["lens"], bt["values"], vl, mk, mv),
                        o.mixed(bt["types"], bt["keys"], bt["lens"], bt["values"], vl, mk, mv)):
            _eq(a, b)
    sc = wl.scans(2048, max_scan=100)
    lens = np.full(len(sc["limits"]), mk, np.uint32)
    for a, b in zip(g.scan(sc["lower"], lens, mk, 100, maxV=mv, limits=sc["limits"]),
                    o.scan(sc["lower"], lens, mk, 100, maxV=mv, limits=sc["limits"])):
        _eq(a, b)
    e = np.arange(0, len(wl.lens), 2)
    _eq(g.erase(wl.keys[e], wl.lens[e], mk), o.erase(wl.keys[e], wl.lens[e], mk))
    for a, b in zip(g.find(wl.keys, wl.lens, mk, mv), o.find(wl.keys, wl.lens, mk, mv)):
        _eq(a, b)
    assert g.items() == o.items()
    ok, *_, msg = g.validate()
    assert ok, msg
    full = S.GpuBackend(tile, pool_bytes=256 << 20)
    full.insert(wl.keys, wl.lens, wl.values, None, mk, mv)
    assert nodes_small > 2 * full.validate()[2]


@pytest.mark.parametrize("find_tile", [2, 8])
def test_find_shortcut_table(find_tile):
    """Read-only descents of a tree of more than 2^18 keys start at the layer-0 shortcut
    table (level-1 nodes by first-slice prefix; find-only descents at its border-level twin
    while no node has merged since it was built).  Finds of present and absent keys, a split
    mixed batch (its read-only pass and write hints use the table), finds after its inserts
    split border nodes, erases of short ranges whose merges rewrite the table, erases of a
    long range that retire it, then inserts that split nodes while the table stays as built
    (stale), and finds again: every result equals the oracle's, and equals the same calls
    with the table off."""
    rng = np.random.default_rng(23)
    n, mk, mv = 1_300_000, 4, 2
    keys, lens = W.unique_random_keys(rng, n, 1, mk)
    keys = np.where(np.arange(mk)[None, :] < lens[:, None], keys, 0).astype(np.uint32)
    vals = rng.integers(0, 1 << 32, size=(n, mv), dtype=np.uint64).astype(np.uint32)
    g, o = S.GpuBackend(16, pool_bytes=1 << 30), S.OracleBackend()
    g.t.set_find_tile(find_tile)
    base = n - 200_000
    _eq(g.insert(keys[:base], lens[:base], vals[:base], None, mk, mv),
        o.insert(keys[:base], lens[:base], vals[:base], None, mk, mv))
    probe = np.concatenate([rng.permutation(n)[:300_000]])
    for a, b in zip(g.find(keys[probe], lens[probe], mk, mv), o.find(keys[probe], lens[probe], mk, mv)):
        _eq(a, b)
    assert g.t.find_shortcut_bits() > 0
    # a split mixed batch: finds of present keys, inserts of the rest, erases of others
    ins = np.arange(base, n)
    era = rng.permutation(base)[:150_000]
    fnd = np.setdiff1d(rng.permutation(base)[:100_000], era)
    types = np.concatenate([np.zeros(len(ins), np.uint8), np.ones(len(era), np.uint8),
                            np.full(len(fnd), 2, np.uint8)])
    idx = np.concatenate([ins, era, fnd])
    p = rng.permutation(len(idx))
    types, idx = types[p], idx[p]
    v = vals[idx].copy()
    vl = np.full(len(idx), mv, np.uint32)
    for a, b in zip(g.mixed(types, keys[idx], lens[idx], v, vl, mk, mv),
                    o.mixed(types, keys[idx], lens[idx], v.copy(), vl, mk, mv)):
        _eq(a, b)
    # finds through the border-level table after those inserts split its nodes
    for a, b in zip(g.find(keys, lens, mk, mv), o.find(keys, lens, mk, mv)):
        _eq(a, b)
    # erase the keys of 256 short first-slice ranges (2^16 wide): their border nodes empty
    # and merge one or two at a time, each merge rewriting the few table entries that named
    # the node it removed, so finds keep starting at the border-level table
    small = np.nonzero((keys[:, 0] & ((1 << 24) - 1)) < (1 << 16))[0]
    _eq(g.erase(keys[small], lens[small], mk), o.erase(keys[small], lens[small], mk))
    for a, b in zip(g.find(keys, lens, mk, mv), o.find(keys, lens, mk, mv)):
        _eq(a, b)
    _eq(g.insert(keys[small], lens[small], vals[small], None, mk, mv),
        o.insert(keys[small], lens[small], vals[small], None, mk, mv))
    # erase every key whose first slice is below 2^28 (a contiguous range: its border
    # nodes empty and merge into long runs of table entries, which retires the
    # border-level table; finds fall back to the level-1 one)
    low = np.nonzero(keys[:, 0] < (1 << 28))[0]
    _eq(g.erase(keys[low], lens[low], mk), o.erase(keys[low], lens[low], mk))
    for a, b in zip(g.find(keys, lens, mk, mv), o.find(keys, lens, mk, mv)):
        _eq(a, b)
    _eq(g.insert(keys[low], lens[low], vals[low], None, mk, mv),
        o.insert(keys[low], lens[low], vals[low], None, mk, mv))
    # more erases and inserts, then finds through the table as built before them
    _eq(g.erase(keys[era[:50_000]], lens[era[:50_000]], mk), o.erase(keys[era[:50_000]], lens[era[:50_000]], mk))
    _eq(g.insert(keys[era], lens[era], vals[era], None, mk, mv), o.insert(keys[era], lens[era], vals[era], None, mk, mv))
    with_table = g.find(keys, lens, mk, mv)
    for a, b in zip(with_table, o.find(keys, lens, mk, mv)):
        _eq(a, b)
    g.t.set_find_shortcut(False)
    assert g.t.find_shortcut_bits() == 0
    for a, b in zip(g.find(keys, lens, mk, mv), with_table):
        _eq(a, b)
    ok, *_, msg = g.validate()
    assert ok, msg


def test_shortcut_tables_belong_to_the_owning_instance():
    """Only the 16-lane instance (the one that owns the tree) builds shortcut tables: inserts
    and finds at 32 lanes build none, finds at 16 lanes build them, erases at 32 lanes (their
    launches have no table, so their layer-0 merges are counted and retire the border-level
    table) leave finds at 16 lanes equal to the oracle's."""
    rng = np.random.default_rng(29)
    n, mk, mv = 400_000, 2, 1
    keys, lens = W.unique_random_keys(rng, n, 1, mk)
    keys = np.where(np.arange(mk)[None, :] < lens[:, None], keys, 0).astype(np.uint32)
    vals = rng.integers(0, 1 << 32, size=(n, mv), dtype=np.uint64).astype(np.uint32)
    g, o = S.GpuBackend(32, pool_bytes=1 << 30), S.OracleBackend()
    _eq(g.insert(keys, lens, vals, None, mk, mv), o.insert(keys, lens, vals, None, mk, mv))
    for a, b in zip(g.find(keys, lens, mk, mv), o.find(keys, lens, mk, mv)):
        _eq(a, b)
    assert g.t.find_shortcut_bits() == 0
    g.t.tile_size = 16
    for a, b in zip(g.find(keys, lens, mk, mv), o.find(keys, lens, mk, mv)):
        _eq(a, b)
    assert g.t.find_shortcut_bits() > 0
    g.t.tile_size = 32
    low = np.nonzero(keys[:, 0] < (1 << 28))[0]
    _eq(g.erase(keys[low], lens[low], mk), o.erase(keys[low], lens[low], mk))
    g.t.tile_size = 16
    for a, b in zip(g.find(keys, lens, mk, mv), o.find(keys, lens, mk, mv)):
        _eq(a, b)
    ok, *_, msg = g.validate()
    assert ok, msg
