"""Dataset path on the GPU: text keys (URL-like lines) converted to slices, inserted,
found, and scanned back in byte order; values compared with the oracle."""
import numpy as np
import pytest

import scenarios as S
from indexingpu_amd import workloads as W
from test_dataset_keys import _urls

pytestmark = pytest.mark.gpu


def test_text_keys_insert_find_scan():
    lines = sorted(set(_urls(5000, seed=9)))
    keys, lens, _, mk = W.dataset_keys(lines)
    n = len(lens)
    mv = 2
    vals = W.make_values(keys, lens, mv)
    g, o = S.GpuBackend(16, pool_bytes=64 << 20), S.OracleBackend()
    perm = np.random.default_rng(1).permutation(n)
    assert np.array_equal(g.insert(keys[perm], lens[perm], vals[perm], None, mk, mv),
                          o.insert(keys[perm], lens[perm], vals[perm], None, mk, mv))
    for a, b in zip(g.find(keys, lens, mk, mv), o.find(keys, lens, mk, mv)):
        assert np.array_equal(a, b)
    # one scan from the smallest key returns the first 200 lines in byte order
    c, v, vl, k, kl = g.scan(keys[:1], lens[:1], mk, 200, maxV=mv)
    assert int(c[0]) == min(200, n)
    for i in range(int(c[0])):
        got = tuple(k[i * mk:i * mk + int(kl[i])].tolist())
        assert got == tuple(keys[i, :lens[i]].tolist())
    assert g.items() == o.items()
