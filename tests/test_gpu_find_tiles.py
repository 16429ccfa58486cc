"""Read-only finds at every find tile width (1, 2, 4, 8, 16 lanes, and 0 = the index's tile),
every in-flight depth and both pass counts against the oracle: keys of 1..40 slices sharing prefixes
(trie layers, suffix records whose tails span slabs), values of 0..70 slices (inline
values, records of several slabs), hits and misses, truncation by max_value_length."""
import numpy as np
import pytest

import scenarios as S
from indexingpu_amd import workloads as W

pytestmark = pytest.mark.gpu


def _keyset(seed):
    rng = np.random.default_rng(seed)
    mk = 40
    base, blen = W.unique_random_keys(rng, 3000, 1, mk)
    # shared prefixes: copy the first slices of other keys so layers form
    keys = base.copy()
    lens = blen.copy()
    donor = rng.integers(0, len(lens), size=len(lens))
    share = rng.integers(0, mk, size=len(lens))
    for i in range(0, len(lens), 3):
        p = min(share[i], lens[i] - 1, lens[donor[i]])
        keys[i, :p] = base[donor[i], :p]
    h = W.key_hash64(keys, lens)
    _, first = np.unique(h, return_index=True)
    first = np.sort(first)
    return keys[first], lens[first], mk


@pytest.fixture(scope="module")
def loaded():
    keys, lens, mk = _keyset(21)
    n = len(lens)
    rng = np.random.default_rng(22)
    mv = 70
    vl = rng.choice([0, 1, 1, 2, 5, 30, 31, 62, 70], size=n).astype(np.uint32)
    vals = rng.integers(0, 1 << 32, size=(n, mv), dtype=np.uint64).astype(np.uint32)
    g, o = S.GpuBackend(16, pool_bytes=128 << 20), S.OracleBackend()
    half = n // 2  # second half stays absent: misses that share prefixes with hits
    assert np.array_equal(g.insert(keys[:half], lens[:half], vals[:half], vl[:half], mk, mv),
                          o.insert(keys[:half], lens[:half], vals[:half], vl[:half], mk, mv))
    yield g, o, keys, lens, mk
    g.t.close()


@pytest.mark.parametrize("find_tile", [0, 1, 2, 4, 8, 16])
@pytest.mark.parametrize("depth", [1, 2, 4, 8])
@pytest.mark.parametrize("max_value", [70, 3])
@pytest.mark.parametrize("passes", [1, 2])
def test_find_tiles_match_oracle(loaded, find_tile, depth, max_value, passes):
    g, o, keys, lens, mk = loaded
    g.t.set_find_tile(find_tile)
    g.t.set_find_depth(depth)
    g.t.set_find_passes(passes)
    got_v, got_l = g.find(keys, lens, mk, max_value)
    exp_v, exp_l = o.find(keys, lens, mk, max_value)
    assert np.array_equal(got_l, exp_l)
    assert np.array_equal(got_v, exp_v)


@pytest.mark.parametrize("record_tile", [0, 1, 2, 4, 8, 16])
@pytest.mark.parametrize("find_tile", [2, 4])
@pytest.mark.parametrize("max_value", [70, 3])
def test_record_tiles_match_oracle(loaded, record_tile, find_tile, max_value):
    """Two-pass finds with the record pass at every tile width (0 = the find tile)."""
    g, o, keys, lens, mk = loaded
    g.t.set_find_passes(2)
    g.t.set_find_tile(find_tile)
    g.t.set_find_depth(1)
    g.t.set_record_tile(record_tile)
    got_v, got_l = g.find(keys, lens, mk, max_value)
    exp_v, exp_l = o.find(keys, lens, mk, max_value)
    assert np.array_equal(got_l, exp_l)
    assert np.array_equal(got_v, exp_v)
