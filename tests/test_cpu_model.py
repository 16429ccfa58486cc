"""The batch engine's state machine (include/indexingpu_amd/tree_ops.hpp) run on the host
with the CPU tile policy, against the oracle: reference scenarios with seeded random
interleavings of many tiles, and real concurrency with 8 host threads.  Parity is exact:
find outputs, per-op results and the full ordered dump."""
import numpy as np
import pytest

import refinputs as R
import scenarios as S
from indexingpu_amd import workloads as W

PLAIN = [S.sc_find_exist, S.sc_find_notexist, S.sc_update, S.sc_erase_all, S.sc_erase_none,
         S.sc_erase_all_insert_all]


def _same(a, b, skip=()):
    assert a.keys() == b.keys()
    for k in a:
        if k in skip:
            continue
        if isinstance(a[k], list):
            assert a[k] == b[k], k
        else:
            assert np.array_equal(a[k], b[k]), k


@pytest.mark.parametrize("nthreads", [0, 8])
@pytest.mark.parametrize("c", range(4))
def test_reference_scenarios(c, nthreads):
    for sc in PLAIN:
        inp = R.TestingInput(c)
        _same(sc(S.CpuModelBackend(nthreads=nthreads), inp), sc(S.OracleBackend(), inp))
    inp = R.TestingInput(c)
    mix = R.ConcurrentMix(inp)
    got = S.sc_concurrent_mix(S.CpuModelBackend(nthreads=nthreads), inp, mix)
    inp = R.TestingInput(c)
    mix = R.ConcurrentMix(inp)
    exp = S.sc_concurrent_mix(S.OracleBackend(), inp, mix)
    # finds of keys written in the same batch are order-dependent (reference :657-676)
    ks = mix.num_keyset
    det = (mix.dst_of < 4 * ks) | (mix.dst_of >= 7 * ks)
    assert np.array_equal(got["mix_results"][det], exp["mix_results"][det])
    assert np.array_equal(got["mix_lens"][det], exp["mix_lens"][det])
    _same({k: got[k] for k in ("find_values", "find_lens", "items")},
          {k: exp[k] for k in ("find_values", "find_lens", "items")})


@pytest.mark.parametrize("nthreads", [0, 8])
def test_duplicates_in_batch(nthreads):
    for c in range(4):
        for sc in (S.sc_insert_twice_erase_all, S.sc_erase_all_twice):
            got = sc(S.CpuModelBackend(nthreads=nthreads), R.TestingInput(c, duplicate=True))
            exp = sc(S.OracleBackend(), R.TestingInput(c, duplicate=True))
            # which copy of a duplicated key wins is order-dependent; totals are not
            for k in ("insert_dup", "erase_dup"):
                if k in got:
                    assert got[k].sum() == exp[k].sum()
            _same({k: got[k] for k in ("find_values", "find_lens", "items")},
                  {k: exp[k] for k in ("find_values", "find_lens", "items")})


@pytest.mark.parametrize("hinted", [False, True], ids=["from_root", "start_hints"])
@pytest.mark.parametrize("nthreads", [0, 8])
def test_random_mixed_batches(nthreads, hinted):
    """Mixed batches from the root, and with every write begun at the start hint of a
    read-only walk made before the batch (as the GPU write passes begin them): hints go
    stale while the batch splits nodes and pushes keys down a layer."""
    wl = W.MixedWorkload(20000, key_len=3, seed=11)
    # shared first slices force layers and suffix push-downs
    wl.keys[:, 0] %= 64
    wl.hashes = W.key_hash64(wl.keys, wl.lens)
    b = S.CpuModelBackend(nthreads=nthreads, hinted=hinted)
    o = S.OracleBackend()
    mk, mv = 3, 2
    assert np.array_equal(b.insert(wl.keys, wl.lens, wl.values, None, mk, mv),
                          o.insert(wl.keys, wl.lens, wl.values, None, mk, mv))
    for _ in range(3):
        bt = wl.batch(4000)
        vl = np.full(len(bt["types"]), mv, np.uint32)
        got = b.mixed(bt["types"], bt["keys"], bt["lens"], bt["values"], vl, mk, mv)
        exp = o.mixed(bt["types"], bt["keys"], bt["lens"], bt["values"], vl, mk, mv)
        for g, e in zip(got, exp):
            assert np.array_equal(g, e)
    assert b.items() == o.items()


@pytest.mark.parametrize("nthreads", [0, 8])
def test_hinted_growth_inserts(nthreads):
    """Insert batches up to 4x the tree's size with start hints from the tree before the
    batch (the GPU growth stages): the hinted writes move right past nodes that split after
    the walk and find their parents from stale level-1 hints; the result equals the oracle's."""
    rng = np.random.default_rng(5)
    keys, lens = W.unique_random_keys(rng, 60000, 2, 6)
    keys[:, 0] %= 512  # shared first slices: layers and push-downs under the hints
    _, first = np.unique(keys, axis=0, return_index=True)
    keys = keys[np.sort(first)]
    lens = lens[np.sort(first)]
    vals = rng.integers(0, 1 << 32, size=(len(lens), 2), dtype=np.uint64).astype(np.uint32)
    b = S.CpuModelBackend(nthreads=nthreads, hinted=True)
    o = S.OracleBackend()
    mk, mv = keys.shape[1], 2
    done, step = 0, 256
    while done < len(lens):
        sl = slice(done, min(done + step, len(lens)))
        assert np.array_equal(b.insert(keys[sl], lens[sl], vals[sl], None, mk, mv),
                              o.insert(keys[sl], lens[sl], vals[sl], None, mk, mv))
        done = sl.stop
        step *= 4
    assert b.m.error == 0
    assert b.items() == o.items()


@pytest.mark.parametrize("nthreads", [0, 8])
def test_split_parent_out_of_memory(nthreads):
    """Every allocation of a split inserting into its parent fails (a parent that must
    split cannot): the new right halves keep no parent slot and are reached through
    sibling links only.  Later splits of such nodes link them from the slot of their left
    neighbour instead of searching for a slot that does not exist (which never ends),
    inserts keep succeeding, and once slabs are available again every new right half is
    linked.  Results equal the oracle's throughout."""
    rng = np.random.default_rng(21)
    keys, lens = W.unique_random_keys(rng, 40000, 1, 3)
    vals = rng.integers(0, 1 << 32, size=(len(lens), 2), dtype=np.uint64).astype(np.uint32)
    b = S.CpuModelBackend(nthreads=nthreads, allow_error=True)
    o = S.OracleBackend()
    mk, mv = keys.shape[1], 2
    b.m.set_fail(parent_every=1)
    half = len(lens) // 2
    for lo in range(0, half, 2000):
        sl = slice(lo, min(lo + 2000, half))
        assert np.array_equal(b.insert(keys[sl], lens[sl], vals[sl], None, mk, mv),
                              o.insert(keys[sl], lens[sl], vals[sl], None, mk, mv))
    assert b.m.error == 1  # kErrOutOfMemory
    orphans = b.m.unlinked
    assert orphans > 100  # the root could not split: long sibling chains
    assert b.items() == o.items()
    fk = keys[:half]
    got = b.find(fk, lens[:half], mk, mv)
    exp = o.find(fk, lens[:half], mk, mv)
    for g, e in zip(got, exp):
        assert np.array_equal(g, e)
    b.m.set_fail()
    for lo in range(half, len(lens), 2000):
        sl = slice(lo, min(lo + 2000, len(lens)))
        assert np.array_equal(b.insert(keys[sl], lens[sl], vals[sl], None, mk, mv),
                              o.insert(keys[sl], lens[sl], vals[sl], None, mk, mv))
    assert b.m.error == 0
    # new right halves, including those of unlinked nodes, all got parent slots
    assert b.m.unlinked <= orphans
    assert b.items() == o.items()


@pytest.mark.parametrize("hinted", [False, True], ids=["from_root", "start_hints"])
@pytest.mark.parametrize("nthreads", [0, 8])
def test_random_out_of_memory(nthreads, hinted):
    """One allocation in seven fails, wherever it is made (records, border splits, layer
    push-downs, parent splits): every insert either succeeds or reports false and leaves
    no trace; the tree stays valid and equals the oracle fed only the successful inserts."""
    rng = np.random.default_rng(22)
    keys, lens = W.unique_random_keys(rng, 20000, 1, 5)
    keys[:, 0] %= 256  # shared first slices: layers, suffix records and push-downs
    _, first = np.unique(keys, axis=0, return_index=True)
    keys, lens = keys[np.sort(first)], lens[np.sort(first)]
    vl = rng.integers(1, 4, size=len(lens)).astype(np.uint32)
    vals = rng.integers(0, 1 << 32, size=(len(lens), 3), dtype=np.uint64).astype(np.uint32)
    b = S.CpuModelBackend(nthreads=nthreads, allow_error=True, hinted=hinted)
    o = S.OracleBackend()
    mk, mv = keys.shape[1], 3
    b.m.set_fail(any_every=7)
    ok_all = np.zeros(len(lens), bool)
    for lo in range(0, len(lens), 2500):
        sl = slice(lo, min(lo + 2500, len(lens)))
        ok = b.insert(keys[sl], lens[sl], vals[sl], vl[sl], mk, mv).astype(bool)
        ok_all[sl] = ok
        idx = np.nonzero(ok)[0] + lo
        o.insert(keys[idx], lens[idx], vals[idx], vl[idx], mk, mv)
    assert 0 < ok_all.sum() < len(lens)
    assert b.items() == o.items()
    b.m.set_fail()
    # the failed ones go in now
    idx = np.nonzero(~ok_all)[0]
    assert np.array_equal(b.insert(keys[idx], lens[idx], vals[idx], vl[idx], mk, mv),
                          o.insert(keys[idx], lens[idx], vals[idx], vl[idx], mk, mv))
    assert b.m.error == 0
    assert b.items() == o.items()


@pytest.mark.parametrize("nthreads", [0, 8])
def test_inserts_without_suffix_records(nthreads):
    """insert<..., enable_suffix = false>: a new key's remaining slices become one layer
    each.  Finds, erases, updates and later suffix-record inserts over such layers give
    the oracle's results, with writes begun from the root and from start hints."""
    rng = np.random.default_rng(31)
    keys, lens = W.unique_random_keys(rng, 12000, 1, 6)
    keys[:, 0] %= 32
    keys[:, 1] %= 4  # long shared prefixes: deep layer chains that later inserts split
    _, first = np.unique(keys, axis=0, return_index=True)
    keys, lens = keys[np.sort(first)], lens[np.sort(first)]
    vl = rng.integers(0, 4, size=len(lens)).astype(np.uint32)
    vals = rng.integers(0, 1 << 32, size=(len(lens), 3), dtype=np.uint64).astype(np.uint32)
    mk, mv = keys.shape[1], 3
    for hinted in (False, True):
        b = S.CpuModelBackend(nthreads=nthreads, no_suffix=True, hinted=hinted)
        o = S.OracleBackend()
        half = len(lens) // 2
        _eq = np.array_equal
        assert _eq(b.insert(keys[:half], lens[:half], vals[:half], vl[:half], mk, mv),
                   o.insert(keys[:half], lens[:half], vals[:half], vl[:half], mk, mv))
        b.m.set_no_suffix(False)  # the rest with suffix records, over the layer chains
        assert _eq(b.insert(keys, lens, vals, vl, mk, mv), o.insert(keys, lens, vals, vl, mk, mv))
        for g, e in zip(b.find(keys, lens, mk, mv), o.find(keys, lens, mk, mv)):
            assert _eq(g, e)
        e = rng.choice(len(lens), len(lens) // 3, replace=False)
        assert _eq(b.erase(keys[e], lens[e], mk), o.erase(keys[e], lens[e], mk))
        b.m.set_no_suffix(True)
        assert _eq(b.insert(keys[e], lens[e], vals[e], vl[e], mk, mv, upsert=True),
                   o.insert(keys[e], lens[e], vals[e], vl[e], mk, mv, upsert=True))
        assert b.items() == o.items()


@pytest.mark.parametrize("hinted", [False, True], ids=["from_root", "start_hints"])
@pytest.mark.parametrize("no_suffix", [False, True], ids=["suffix_records", "layer_chains"])
@pytest.mark.parametrize("nthreads", [0, 8])
def test_erase_removes_empty_layers(nthreads, no_suffix, hinted):
    """Erasing every key under a shared prefix empties the layers below it: each emptied
    layer root is unlinked from the layer above and retired (cascading up through layer
    chains), while inserts into the same prefixes run in the same batches.  Results equal
    the oracle's; once every key is gone only the top layer is left.  (Layers here stay
    single nodes: a layer that grew an interior root keeps its empty nodes.)"""
    rng = np.random.default_rng(41)
    n_pre, per = 64, 12
    pre = rng.integers(0, 1 << 32, size=(n_pre, 3), dtype=np.uint64).astype(np.uint32)
    keys = np.zeros((n_pre * per, 6), np.uint32)
    keys[:, :3] = np.repeat(pre, per, axis=0)
    keys[:, 3:] = rng.integers(0, 1 << 32, size=(n_pre * per, 3), dtype=np.uint64).astype(np.uint32)
    keys[:, 3] %= 4  # a layer below each prefix, and layers below those
    lens = rng.integers(4, 7, size=len(keys)).astype(np.uint32)
    tup = [tuple(keys[i, :lens[i]].tolist()) for i in range(len(keys))]
    _, first = np.unique(np.array([hash(t) for t in tup]), return_index=True)
    keys, lens = keys[np.sort(first)], lens[np.sort(first)]
    pid = np.array([np.nonzero((pre == k[:3]).all(1))[0][0] for k in keys])
    vals = rng.integers(0, 1 << 32, size=(len(lens), 2), dtype=np.uint64).astype(np.uint32)
    mk, mv = 6, 2
    b = S.CpuModelBackend(nthreads=nthreads, no_suffix=no_suffix, hinted=hinted)
    o = S.OracleBackend()
    _eq = np.array_equal
    assert _eq(b.insert(keys, lens, vals, None, mk, mv), o.insert(keys, lens, vals, None, mk, mv))
    full = b.m.layers
    gone = np.isin(pid, np.arange(0, n_pre, 2))  # erase every key of the even prefixes
    e = np.nonzero(gone)[0]
    assert _eq(b.erase(keys[e], lens[e], mk), o.erase(keys[e], lens[e], mk))
    assert b.items() == o.items()
    assert b.m.layers < full
    # erase the rest while re-inserting the even prefixes' keys, as one mixed batch
    k = np.nonzero(~gone)[0]
    types = np.concatenate([np.ones(len(k), np.uint8), np.zeros(len(e), np.uint8)])
    idx = np.concatenate([k, e])
    p = rng.permutation(len(idx))
    types, idx = types[p], idx[p]
    vl = np.full(len(idx), mv, np.uint32)
    for g, x in zip(b.mixed(types, keys[idx], lens[idx], vals[idx], vl, mk, mv),
                    o.mixed(types, keys[idx], lens[idx], vals[idx], vl, mk, mv)):
        assert _eq(g, x)
    assert b.items() == o.items()
    for g, x in zip(b.find(keys, lens, mk, mv), o.find(keys, lens, mk, mv)):
        assert _eq(g, x)
    # everything erased: only the top layer is left
    assert _eq(b.erase(keys, lens, mk), o.erase(keys, lens, mk))
    assert b.items() == [] and b.m.layers == 1


@pytest.mark.parametrize("nthreads", [0, 8])
def test_hinted_writes_on_few_hot_nodes(nthreads):
    """Hinted writes (the GPU write passes' fast path: load, compare-and-swap, publish as
    separate steps) crowded onto a few hundred keys: every kind of write, upserts and
    updates replacing values of every length (inline, record, suffix record), erases,
    re-inserts, with most lock attempts failing and going round again.  Results equal the
    oracle's batch after batch."""
    rng = np.random.default_rng(61)
    keys, lens = W.unique_random_keys(rng, 600, 1, 3)
    keys[:, 0] %= 40  # ~40 first slices: a few border nodes, suffix records, layers
    _, first = np.unique(keys, axis=0, return_index=True)
    keys, lens = keys[np.sort(first)], lens[np.sort(first)]
    n, mk, mv = len(lens), 3, 3
    b = S.CpuModelBackend(nthreads=nthreads, ntiles=64, hinted=True)
    o = S.OracleBackend()
    for rnd in range(12):
        # each key at most once per batch, every kind
        idx = rng.permutation(n)[: n // 2]
        types = rng.choice([0, 1, 3, 2, 0], size=len(idx)).astype(np.uint8)
        types[types == 2] = 4 if rnd % 2 else 0  # upserts (kind 4) in odd rounds
        vals = rng.integers(0, 1 << 32, size=(len(idx), mv), dtype=np.uint64).astype(np.uint32)
        vl = rng.integers(0, mv + 1, size=len(idx)).astype(np.uint32)
        for kind in (0, 4, 3, 1):
            sel = idx[types == kind]
            if not len(sel):
                continue
            v, l = vals[types == kind], vl[types == kind]
            if kind == 0:
                got = b.insert(keys[sel], lens[sel], v, l, mk, mv)
                exp = o.insert(keys[sel], lens[sel], v, l, mk, mv)
            elif kind == 4:
                got = b.insert(keys[sel], lens[sel], v, l, mk, mv, upsert=True)
                exp = o.insert(keys[sel], lens[sel], v, l, mk, mv, upsert=True)
            elif kind == 3:
                got = b.update(keys[sel], lens[sel], v, l, mk, mv)
                exp = o.update(keys[sel], lens[sel], v, l, mk, mv)
            else:
                got = b.erase(keys[sel], lens[sel], mk)
                exp = o.erase(keys[sel], lens[sel], mk)
            assert np.array_equal(got, exp), (rnd, kind)
        assert b.items() == o.items()


@pytest.mark.parametrize("nthreads,ntiles", [(0, 1), (0, 64), (8, 64)],
                         ids=["one_tile", "interleaved", "threads"])
def test_merges_rewrite_shortcut_table(nthreads, ntiles):
    """A border-level shortcut table built on the tree stays a valid start for every prefix
    while erases empty and merge layer-0 border nodes and inserts split them, batch after
    batch: each merge rewrites the entries that named the node it removed, and results
    equal the oracle's.  One tile at a time no merge is ever counted (which would retire the
    table); with tiles running together a merge that finds the node it rewrites to busy is
    counted, after which the table is rebuilt, as the index's next rebuild would."""
    rng = np.random.default_rng(75)
    n, mk, mv = 6000, 2, 1
    keys, lens = W.unique_random_keys(rng, n, 1, mk)
    vals = rng.integers(0, 1 << 32, size=(n, mv), dtype=np.uint64).astype(np.uint32)
    b = S.CpuModelBackend(nthreads=nthreads, ntiles=ntiles)
    o = S.OracleBackend()
    _eq = np.array_equal
    assert _eq(b.insert(keys, lens, vals, None, mk, mv), o.insert(keys, lens, vals, None, mk, mv))
    full_border = b.m.nodes(1)
    b.m.build_shortcut(10)
    assert b.m.check_shortcut() == 0
    checked = 0

    def settle(tag, counted_before):
        nonlocal checked
        if b.m.nodes(5) == counted_before:
            assert b.m.check_shortcut() == 0, tag
            checked += 1
        else:
            b.m.build_shortcut(10)

    counted = b.m.nodes(5)
    gone = ((keys[:, 0] >= (1 << 30)) & (keys[:, 0] < (1 << 31))) | (rng.random(n) < 0.3)
    e = np.nonzero(gone)[0]
    assert _eq(b.erase(keys[e], lens[e], mk), o.erase(keys[e], lens[e], mk))
    assert b.items() == o.items()
    assert b.m.nodes(1) < full_border - full_border // 8  # layer-0 border nodes merged away
    settle("first erase", counted)
    for rnd in range(3):
        back = rng.permutation(e)[: len(e) // 3]
        rest = np.setdiff1d(np.arange(n), e)
        out = rng.permutation(rest)[: len(rest) // 3]
        assert _eq(b.insert(keys[back], lens[back], vals[back], None, mk, mv),
                   o.insert(keys[back], lens[back], vals[back], None, mk, mv)), rnd
        assert b.m.check_shortcut() == 0, rnd  # splits leave every entry a valid start
        counted = b.m.nodes(5)
        assert _eq(b.erase(keys[out], lens[out], mk), o.erase(keys[out], lens[out], mk)), rnd
        e = np.union1d(np.setdiff1d(e, back), out)
        assert b.items() == o.items(), rnd
        settle(rnd, counted)
    if ntiles == 1:
        assert b.m.nodes(5) == 0 and checked == 4


@pytest.mark.parametrize("nthreads,ntiles", [(0, 1), (0, 64), (8, 64)],
                         ids=["one_tile", "interleaved", "threads"])
def test_shortcut_table_through_split_and_merge_churn(nthreads, ntiles):
    """Rounds of inserts of fresh keys (splits make nodes that no entry names) and erases of
    short contiguous ranges (runs of neighbouring nodes empty and merge, new nodes into old
    ones and old into new): after every batch in which no merge had to retire the border-
    level table, every entry is a live layer-0 border node at or left of its prefix's node.
    One tile at a time, no merge ever finds the node it rewrites to busy, so the table is
    never retired; with tiles running together a busy one retires it, and it is rebuilt."""
    rng = np.random.default_rng(79)
    mk, mv = 1, 1
    live = set()
    b = S.CpuModelBackend(nthreads=nthreads, ntiles=ntiles)
    o = S.OracleBackend()
    _eq = np.array_equal

    def insert(k):
        k = np.array(sorted(set(k.tolist()) - live), np.uint32)
        if len(k) == 0:
            return
        keys, lens = k.reshape(-1, 1), np.ones(len(k), np.uint32)
        vals = rng.integers(0, 1 << 32, size=(len(k), mv), dtype=np.uint64).astype(np.uint32)
        assert _eq(b.insert(keys, lens, vals, None, mk, mv), o.insert(keys, lens, vals, None, mk, mv))
        live.update(k.tolist())

    insert(rng.integers(0, 1 << 32, size=3000, dtype=np.uint64).astype(np.uint32))
    b.m.build_shortcut(10)
    assert b.m.check_shortcut() == 0
    checked = 0
    for rnd in range(16):
        insert(rng.integers(0, 1 << 32, size=800, dtype=np.uint64).astype(np.uint32))
        assert b.m.check_shortcut() == 0, rnd  # splits leave every entry a valid start
        retired = b.m.nodes(5)
        lo = int(rng.integers(0, (1 << 32) - (1 << 26)))
        k = np.array(sorted(x for x in live if lo <= x < lo + (1 << 26)), np.uint32)
        if len(k):
            keys, lens = k.reshape(-1, 1), np.ones(len(k), np.uint32)
            assert _eq(b.erase(keys, lens, mk), o.erase(keys, lens, mk)), rnd
            live.difference_update(k.tolist())
        assert b.items() == o.items(), rnd
        if b.m.nodes(5) == retired:
            assert b.m.check_shortcut() == 0, rnd
            checked += 1
        else:
            b.m.build_shortcut(10)  # retired: a fresh table, as the next rebuild gives
    if ntiles == 1:
        assert checked == 16
    else:
        assert checked >= 4


def test_long_table_run_retires_shortcut_table():
    """When the entries naming the node a merge removes are a run longer than the merge
    rewrites (here the tree's last border node names every prefix above its keys), the
    merge is counted instead, which retires the table."""
    rng = np.random.default_rng(77)
    n, mk, mv = 600, 1, 1
    keys = np.sort(rng.choice(1 << 22, size=n, replace=False)).astype(np.uint32).reshape(n, 1)
    lens = np.ones(n, np.uint32)
    vals = rng.integers(0, 1 << 32, size=(n, mv), dtype=np.uint64).astype(np.uint32)
    b = S.CpuModelBackend(nthreads=0, ntiles=8)
    o = S.OracleBackend()
    _eq = np.array_equal
    assert _eq(b.insert(keys, lens, vals, None, mk, mv), o.insert(keys, lens, vals, None, mk, mv))
    b.m.build_shortcut(12)
    assert b.m.check_shortcut() == 0
    top = np.arange(n - 40, n)  # the largest keys: the last border nodes empty and merge
    assert _eq(b.erase(keys[top], lens[top], mk), o.erase(keys[top], lens[top], mk))
    assert b.items() == o.items()
    assert b.m.nodes(5) > 0


@pytest.mark.parametrize("nthreads", [0, 8])
def test_merge_count_is_layer_zero_only(nthreads):
    """The merge count a border-level shortcut table is checked against (layer-0 border nodes
    only: the table holds no other node) stays put when erases empty and merge border nodes
    of a deeper layer: every key shares its first slice, so every border node that merges
    lies in layer 1.  Results equal the oracle's and the node count falls."""
    rng = np.random.default_rng(73)
    n, mk, mv = 4000, 2, 1
    keys = np.zeros((n, mk), np.uint32)
    keys[:, 0] = 0x12345678
    keys[:, 1] = rng.permutation(1 << 20)[:n].astype(np.uint32)
    lens = np.full(n, 2, np.uint32)
    vals = rng.integers(0, 1 << 32, size=(n, mv), dtype=np.uint64).astype(np.uint32)
    b = S.CpuModelBackend(nthreads=nthreads, ntiles=64)
    o = S.OracleBackend()
    _eq = np.array_equal
    assert _eq(b.insert(keys, lens, vals, None, mk, mv), o.insert(keys, lens, vals, None, mk, mv))
    full_border = b.m.nodes(1)
    e = np.nonzero(keys[:, 1] < (1 << 19))[0]  # a contiguous half of layer 1
    assert _eq(b.erase(keys[e], lens[e], mk), o.erase(keys[e], lens[e], mk))
    assert b.items() == o.items()
    assert b.m.nodes(1) < full_border * 3 // 4  # layer-1 border nodes merged away
    assert b.m.nodes(5) == 0


@pytest.mark.parametrize("hinted", [False, True], ids=["from_root", "start_hints"])
@pytest.mark.parametrize("merge", [True, False], ids=["merge", "no_merge"])
@pytest.mark.parametrize("nthreads", [0, 8])
def test_erase_merges_empty_border_nodes(nthreads, hinted, merge):
    """Erases that empty non-root border nodes fold them into their left sibling (the
    sibling takes the node's key range and sibling link, the parent drops its slot, the
    node is retired), while inserts, finds and other erases run in the same batches over
    the same ranges.  Results equal the oracle's batch after batch; the walk finds the
    structure valid and no removed node reachable; the node count falls with the keys.
    With erase<..., do_merge = false> (and do_remove_empty_root = false, which the
    reference requires with it) every node stays, with the same results."""
    rng = np.random.default_rng(71)
    n, mk, mv = 6000, 2, 1
    keys, lens = W.unique_random_keys(rng, n, 1, 2)
    keys[:, 0] %= 3000  # ~2 keys per first slice: a few layers below, many border nodes
    _, first = np.unique(keys, axis=0, return_index=True)
    keys, lens = keys[np.sort(first)], lens[np.sort(first)]
    n = len(lens)
    vals = rng.integers(0, 1 << 32, size=(n, mv), dtype=np.uint64).astype(np.uint32)
    b = S.CpuModelBackend(nthreads=nthreads, ntiles=64, hinted=hinted)
    if not merge:
        b.m.set_erase_flags(4 | 2)
    o = S.OracleBackend()
    _eq = np.array_equal
    assert _eq(b.insert(keys, lens, vals, None, mk, mv), o.insert(keys, lens, vals, None, mk, mv))
    full, full_border = b.m.nodes(0), b.m.nodes(1)
    # erase whole key ranges (every key with first slice in [lo, hi)) and a random share
    # elsewhere; finds of erased and kept keys and re-inserts of other erased keys mixed in
    gone = ((keys[:, 0] >= 500) & (keys[:, 0] < 2500)) | (rng.random(n) < 0.3)
    e = np.nonzero(gone)[0]
    assert _eq(b.erase(keys[e], lens[e], mk), o.erase(keys[e], lens[e], mk))
    assert b.items() == o.items()
    if merge:
        assert b.m.nodes(0) < full // 2 and b.m.nodes(1) < full_border // 2
        # an emptied node stays only as its parent's one child (or when its neighbour was busy)
        assert b.m.nodes(2) <= b.m.nodes(0) - b.m.nodes(1)
        # layer-0 merges advanced the count a border-level shortcut table is checked against
        # (the rest of the fall in border nodes is layers removed whole and deeper merges)
        assert b.m.nodes(5) > 0
    else:
        assert b.m.nodes(0) == full and b.m.nodes(2) > 0 and b.m.nodes(5) == 0
    for rnd in range(3):
        back = rng.permutation(e)[: len(e) // 4]
        rest = np.setdiff1d(np.arange(n), e)
        out = rng.permutation(rest)[: len(rest) // 3]
        probe = rng.permutation(n)[:500]
        probe = np.setdiff1d(probe, np.concatenate([back, out]))
        types = np.concatenate([np.zeros(len(back), np.uint8), np.ones(len(out), np.uint8),
                                np.full(len(probe), 2, np.uint8)])
        idx = np.concatenate([back, out, probe])
        p = rng.permutation(len(idx))
        types, idx = types[p], idx[p]
        v = vals[idx].copy()
        vl = np.full(len(idx), mv, np.uint32)
        for g, x in zip(b.mixed(types, keys[idx], lens[idx], v, vl, mk, mv),
                        o.mixed(types, keys[idx], lens[idx], v.copy(), vl, mk, mv)):
            assert _eq(g, x), rnd
        assert b.items() == o.items(), rnd
        e = np.union1d(np.setdiff1d(e, back), out)
    for g, x in zip(b.find(keys, lens, mk, mv), o.find(keys, lens, mk, mv)):
        assert _eq(g, x)
    # everything erased: one border node per layer root left, plus the few nodes a
    # layer that grew interior levels keeps
    assert _eq(b.erase(keys, lens, mk), o.erase(keys, lens, mk))
    assert b.items() == []
    assert b.m.nodes(0) < full // 10 if merge else b.m.nodes(0) >= full


def test_guard_trip_record():
    """The livelock guard, forced from the host: a border node whose lock is never
    released (set directly, as a holder that never finishes would leave it) and a guard
    time limit of one tick.  The insert into it is abandoned with the step-limit error and
    the guard records the operation it abandoned: its phase, the node it waited on (with
    that node's locked header), its layer, kind, key length and key slice.  Once the lock
    is gone the same insert succeeds and the structure validates."""
    from cpu_model.model import CpuModel
    m = CpuModel(1 << 12)
    keys = np.array([[5], [9], [7]], np.uint32)
    lens = np.ones(3, np.uint32)
    vals = np.array([[50], [90], [70]], np.uint32)
    assert m.batch(0, keys[:2], 1, lens[:2], vals[:2], 1, None, n=2, ntiles=1)[2].tolist() == [1, 1]
    assert not m.trip().any()
    root = m.root
    m.set_lock(root, True)
    m.set_max_op_time(1)
    r = m.batch(0, keys[2:], 1, lens[2:], vals[2:], 1, None, n=1, ntiles=1)[2]
    assert r.tolist() == [0] and m.error == 2
    t = m.trip()
    assert t[0] == 1
    assert t[1] in (0, 2)           # waiting to read the node (descend) or to lock it
    assert t[2] == root and t[3] == 0 and t[4] == root
    assert t[10] & (1 << 7)         # the node's header still carries the lock bit
    assert t[11] > (1 << 16)        # steps taken before the clock started and ran out
    assert t[12] == 0 and t[13] == 1 and t[14] == 7
    m.set_lock(root, False)
    m.set_max_op_time(1 << 21)
    m.set_fail()  # clears the error word
    assert m.batch(0, keys[2:], 1, lens[2:], vals[2:], 1, None, n=1, ntiles=1)[2].tolist() == [1]
    assert m.items() == [((5,), (50,)), ((7,), (70,)), ((9,), (90,))]


@pytest.mark.parametrize("fanout", [(2, 3), (3, 4), (6, 5)], ids=["b2i3", "b3i4", "b6i5"])
@pytest.mark.parametrize("nthreads", [0, 8])
def test_small_fanout(nthreads, fanout):
    """Node fanout below the layout's maximum (border nodes split at 2, 3 or 6 entries,
    interior nodes at 3, 4 or 5 children): every insert of a growing tree splits and most
    splits go up several levels, while mixed batches, hinted writes and erases that empty
    and merge nodes run over the same keys.  Results equal the oracle's, the walk finds the
    structure valid with no node above the fanout, and the tree is much deeper than at full
    fanout.  The fanout then changes back to the maximum on the same tree."""
    wl = W.MixedWorkload(6000, key_len=3, seed=17)
    wl.keys[:, 0] %= 300
    wl.hashes = W.key_hash64(wl.keys, wl.lens)
    mk, mv = 3, 2
    b = S.CpuModelBackend(nthreads=nthreads, hinted=True, fanout=fanout)
    o = S.OracleBackend()
    assert np.array_equal(b.insert(wl.keys, wl.lens, wl.values, None, mk, mv),
                          o.insert(wl.keys, wl.lens, wl.values, None, mk, mv))
    assert b.items() == o.items()
    small = b.m.nodes(0)
    assert b.m.nodes(3) <= fanout[0] and b.m.nodes(4) <= fanout[1]
    for rnd in range(4):
        if rnd == 2:
            b.m.set_fanout(0, 0)  # back to full nodes: later splits fill them up
        bt = wl.batch(1500)
        vl = np.full(len(bt["types"]), mv, np.uint32)
        got = b.mixed(bt["types"], bt["keys"], bt["lens"], bt["values"], vl, mk, mv)
        exp = o.mixed(bt["types"], bt["keys"], bt["lens"], bt["values"], vl, mk, mv)
        for g, e in zip(got, exp):
            assert np.array_equal(g, e), rnd
        assert b.items() == o.items(), rnd
    full = S.CpuModelBackend(nthreads=0)
    full.insert(wl.keys, wl.lens, wl.values, None, mk, mv)
    assert small > 2 * full.m.nodes(0)
    e = np.arange(0, len(wl.lens), 3)
    assert np.array_equal(b.erase(wl.keys[e], wl.lens[e], mk), o.erase(wl.keys[e], wl.lens[e], mk))
    assert b.items() == o.items()
