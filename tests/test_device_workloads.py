"""The torch generators against the numpy ones (same formulas), and the batched-form
workload's expected outputs against the oracle replaying its batches (CPU device)."""
import numpy as np
import torch

from indexingpu_amd import device_workloads as D
from indexingpu_amd import workloads as W
from oracle.oracle import Oracle


def test_mix64_matches_numpy():
    rng = np.random.default_rng(1)
    x = rng.integers(0, 1 << 64, size=4096, dtype=np.uint64)
    x[:4] = [0, 1, (1 << 63), (1 << 64) - 1]
    got = D.mix64(torch.from_numpy(x.view(np.int64))).numpy().view(np.uint64)
    assert np.array_equal(got, W._mix64(x))


def test_feistel_matches_numpy_and_is_a_bijection():
    for n in (1000, 4096, 77777):
        x = np.arange(n, dtype=np.uint64)
        want = W.feistel_permute(x, n, seed=0)
        got = D.feistel_permute(torch.arange(n, dtype=torch.int64), n, seed=0).numpy()
        assert np.array_equal(got.astype(np.uint64), want)
        assert np.array_equal(np.sort(got), np.arange(n))


def test_zipf_ranks_match_numpy_transform():
    z = W.Zipfian(100000, 0.99)
    rng = np.random.default_rng(3)
    u = rng.random(20000)
    uz = u * z.zetan
    want = ((z.n - 1) * np.power(z.eta * u - z.eta + 1, z.alpha)).astype(np.int64)
    want = np.where(uz < 1.0 + 0.5 ** z.theta, 1, want)
    want = np.where(uz < 1.0, 0, want)
    want = np.clip(want, 0, z.n - 1)
    got = D.zipf_ranks(z, torch.from_numpy(u)).numpy()
    assert np.array_equal(got, want)
    assert (got == 0).mean() > 0.05  # rank 0 is the hottest


def _u32(t):
    return np.ascontiguousarray(t.numpy().view(np.uint32))


def test_prefix_workload_expectations_match_oracle_replay():
    wl = D.PrefixBatchWorkload(5000, n_prefixes=8, device="cpu", seed=11)
    keys, lens, vals = wl.preload()
    mk, mv = wl.key_len, wl.mv
    assert keys.shape == (5000, 16) and (lens == 16).all()
    o = Oracle()
    r = o.insert(_u32(keys).reshape(-1), mk, _u32(lens), _u32(vals).reshape(-1), mv, None, 5000)
    assert r.all()  # preload keys are distinct
    sizes = D.log_uniform_sizes(np.random.default_rng(2), 12, 50, 3000)
    batches = [wl.batch(int(s)) for s in sizes]
    cat = D.concat_batches(batches)
    assert cat["offsets"][-1].item() == sum(cat["sizes"]) == len(cat["types"])
    for b in batches:
        n = len(b["types"])
        t = b["types"].numpy()
        # no key is written twice in a batch and no written key is also read
        k = _u32(b["keys"])
        h = W.key_hash64(k, _u32(b["lens"]))
        w = t != 2
        assert len(np.unique(h[w])) == w.sum()
        assert not np.isin(h[w], h[~w]).any()
        v, vl, res = o.mixed(t, k.reshape(-1).copy(), mk, _u32(b["lens"]),
                             _u32(b["values"]).reshape(-1).copy(), mv,
                             np.full(n, mv, np.uint32), n)
        assert np.array_equal(vl, _u32(b["expect_len"]))
        assert np.array_equal(v.reshape(n, mv), _u32(b["expect_values"]))
        assert np.array_equal(res, b["expect_results"].numpy())
    # keys share 12-slice prefixes
    k = _u32(batches[0]["keys"])
    assert len({tuple(r[:12]) for r in k}) <= 8
