"""gmt_find_noalloc: the read-only find whose scratch is a fixed array in the library
instead of a buffer the index allocates.  Called through the C ABI and checked against
gmt_find (read-only form) on the same index, bit-exact, on two calls in a row -- the
second smaller, after the first left its scratch words behind -- at every find
configuration that uses the scratch, and on a layered key set; a batch above the
maximum is rejected.  Alone in a fresh process it also runs before any allocating find
has grown the index's own scratch."""
import numpy as np
import pytest
import torch

import scenarios as S
from indexingpu_amd import MasstreeError, workloads as W

pytestmark = pytest.mark.gpu


def _find(t, fn, keys, lens, mk, mv, n):
    v = torch.zeros(max(n * mv, 1), dtype=torch.int32, device="cuda")
    l = torch.zeros(max(n, 1), dtype=torch.int32, device="cuda")
    fn(keys, mk, lens, v, mv, l, n)
    torch.cuda.synchronize()
    t.check_device_error()
    return v.cpu().numpy().view(np.uint32)[:n * mv], l.cpu().numpy().view(np.uint32)[:n]


@pytest.mark.parametrize("config", [(2, 2, 1, 0), (2, 4, 2, 2), (2, 1, 1, 0), (2, 16, 1, 8),
                                    (1, 2, 1, 0)])
def test_find_noalloc_matches_find(config):
    passes, ft, depth, rt = config
    d = W.plain_batch_workload(300_000, seed=17)
    mk, mv = d["max_key_length"], d["max_value_length"]
    g = S.GpuBackend(16, pool_bytes=256 << 20)
    t = g.t
    t.set_find_passes(passes)
    t.set_find_tile(ft)
    t.set_find_depth(depth)
    t.set_record_tile(rt)
    n = len(d["lens"])
    t.insert(g._d(d["keys"]), mk, g._d(d["lens"]), g._d(d["values"]), mv, None, n)
    fk, fl = g._d(d["find_keys"]), g._d(d["find_lens"])
    nf = len(d["find_lens"])
    # the no-alloc form first (the index's own scratch has not been allocated yet)
    a1 = _find(t, t.find_noalloc, fk, fl, mk, mv, nf)
    b1 = _find(t, t.find, fk, fl, mk, mv, nf)
    assert np.array_equal(a1[1], b1[1]) and np.array_equal(a1[0], b1[0])
    # a second, smaller call: scratch words of the first call are left over
    m = nf // 3
    fk2 = fk.view(nf, mk)[nf - m:].reshape(-1).contiguous()
    fl2 = fl[nf - m:].contiguous()
    a2 = _find(t, t.find_noalloc, fk2, fl2, mk, mv, m)
    b2 = _find(t, t.find, fk2, fl2, mk, mv, m)
    assert np.array_equal(a2[1], b2[1]) and np.array_equal(a2[0], b2[0])
    assert (a1[1] > 0).sum() == nf // 2  # the generator's 50% hits
    t.close()


def test_find_noalloc_layered_keys_and_limit():
    """Keys under shared prefixes (records behind layers), values longer than the output
    width (full lengths reported), then a batch above the maximum is refused."""
    rng = np.random.default_rng(23)
    keys, lens = W.unique_random_keys(rng, 50_000, 2, 9)
    keys[:, 0] %= 16
    keys[:, 1] %= 4
    _, first = np.unique(keys, axis=0, return_index=True)
    keys, lens = keys[np.sort(first)], lens[np.sort(first)]
    mk, mv = keys.shape[1], 5
    vl = rng.integers(0, 9, size=len(lens)).astype(np.uint32)
    vals = rng.integers(0, 1 << 32, size=(len(lens), 8), dtype=np.uint64).astype(np.uint32)
    g = S.GpuBackend(32, pool_bytes=256 << 20)
    t = g.t
    n = len(lens)
    t.insert(g._d(keys), mk, g._d(lens), g._d(vals), 8, g._d(vl), n)
    dk, dl = g._d(keys), g._d(lens)
    for n2 in (n, n // 5):
        a = _find(t, t.find_noalloc, dk, dl, mk, mv, n2)
        b = _find(t, t.find, dk, dl, mk, mv, n2)
        assert np.array_equal(a[1], b[1]) and np.array_equal(a[0], b[0])
    assert np.array_equal(a[1], vl[: n // 5])
    big = t.NOALLOC_MAX_KEYS + 1
    with pytest.raises(MasstreeError):
        t.find_noalloc(torch.zeros(big, dtype=torch.int32, device="cuda"), 1, None,
                       torch.zeros(big, dtype=torch.int32, device="cuda"), 1,
                       torch.zeros(big, dtype=torch.int32, device="cuda"), big)
    t.close()
