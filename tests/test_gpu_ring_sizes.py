"""Find batches whose sizes fall around the find-only descent's 64-position chunks (1, 2, 63,
64, 65, 127, 129, 200, 4097 keys) at find tiles and depths that take positions from the
wave's ring (depth <= tile) and that do not, against the oracle; mixed batches of the same
sizes (their read-only pass steps through positions), and one of 600,001 requests, above the
size from which the general read-only pass takes positions from the ring."""
import numpy as np
import pytest

import scenarios as S
from indexingpu_amd import workloads as W

pytestmark = pytest.mark.gpu

SIZES = [1, 2, 63, 64, 65, 127, 129, 200, 4097]


@pytest.fixture(scope="module")
def loaded():
    rng = np.random.default_rng(41)
    mk, mv = 6, 3
    keys, lens = W.unique_random_keys(rng, 12000, 1, mk)
    keys = np.where(np.arange(mk)[None, :] < lens[:, None], keys, 0).astype(np.uint32)
    vals = rng.integers(0, 1 << 32, size=(len(lens), mv), dtype=np.uint64).astype(np.uint32)
    g, o = S.GpuBackend(16, pool_bytes=128 << 20), S.OracleBackend()
    half = len(lens) // 2  # the second half stays absent
    assert np.array_equal(g.insert(keys[:half], lens[:half], vals[:half], None, mk, mv),
                          o.insert(keys[:half], lens[:half], vals[:half], None, mk, mv))
    yield g, o, keys, lens, vals, mk, mv
    g.t.close()


@pytest.mark.parametrize("tile,depth", [(1, 1), (2, 1), (2, 2), (4, 4), (16, 1), (16, 4),
                                        (1, 2), (2, 4)])
def test_find_batch_sizes(loaded, tile, depth):
    g, o, keys, lens, _, mk, mv = loaded
    g.t.set_find_passes(2)
    g.t.set_find_tile(tile)
    g.t.set_find_depth(depth)
    g.t.set_record_tile(0)
    rng = np.random.default_rng(tile * 10 + depth)
    for n in SIZES:
        pick = rng.integers(0, len(lens), size=n)
        gv, gl = g.find(keys[pick], lens[pick], mk, mv)
        ov, ol = o.find(keys[pick], lens[pick], mk, mv)
        assert np.array_equal(gl, ol), n
        assert np.array_equal(gv, ov), n


def test_mixed_batch_sizes(loaded):
    g, o, keys, lens, vals, mk, mv = loaded
    g.t.set_find_passes(2)
    g.t.set_find_tile(2)
    g.t.set_find_depth(2)
    rng = np.random.default_rng(43)
    for n in SIZES:
        pick = rng.choice(len(lens), size=n, replace=False)
        types = rng.choice(np.array([0, 1, 2, 2, 2, 2], np.uint8), size=n)
        v = vals[pick].copy()
        vl = np.full(n, mv, np.uint32)
        for a, b in zip(g.mixed(types, keys[pick], lens[pick], v, vl, mk, mv),
                        o.mixed(types, keys[pick], lens[pick], v.copy(), vl, mk, mv)):
            assert np.array_equal(a, b), n
    ok, *_, msg = g.validate()
    assert ok, msg


def test_large_mixed_batch(loaded):
    """About 600,000 requests, no key twice: finds, updates and erases of every loaded key,
    inserts and finds (misses) of fresh keys.  The general read-only pass takes its
    positions from the ring at this size."""
    g, o, keys, lens, vals, mk, mv = loaded
    g.t.set_find_passes(2)
    g.t.set_find_tile(2)
    g.t.set_find_depth(1)
    rng = np.random.default_rng(47)
    fresh, flens = W.unique_random_keys(rng, 600_000, 1, mk)
    fresh = np.where(np.arange(mk)[None, :] < flens[:, None], fresh, 0).astype(np.uint32)
    # drop fresh keys equal to loaded ones (requests of one batch must not share a key)
    loaded_h = set(W.key_hash64(keys, lens).tolist())
    keep = np.array([h not in loaded_h for h in W.key_hash64(fresh, flens).tolist()])
    fresh, flens = fresh[keep], flens[keep]
    k = np.concatenate([keys, fresh])
    kl = np.concatenate([lens, flens]).astype(np.uint32)
    types = np.concatenate([rng.choice(np.array([1, 2, 2, 3], np.uint8), size=len(lens)),
                            rng.choice(np.array([0, 2], np.uint8), size=len(flens))])
    p = rng.permutation(len(kl))
    k, kl, types = k[p], kl[p], types[p]
    n = len(kl)
    assert n > (1 << 19)
    v = rng.integers(0, 1 << 32, size=(n, mv), dtype=np.uint64).astype(np.uint32)
    vl = np.full(n, mv, np.uint32)
    for a, b in zip(g.mixed(types, k, kl, v, vl, mk, mv), o.mixed(types, k, kl, v.copy(), vl, mk, mv)):
        assert np.array_equal(a, b)
    ok, *_, msg = g.validate()
    assert ok, msg
