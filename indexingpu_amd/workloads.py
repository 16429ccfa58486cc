"""Synthetic workloads for the Masstree benchmarks and parity tests (numpy, seeded).

Keys are rows of u32 slices ([n, max_key_length], unused tail slots zero) with a
length per key; values are [n, max_value_length].  Every generator is deterministic
for its seed.  Conventions follow the reference's workload generator
(evaluation/generate_workload.hpp in hyoungjook/IndexinGPU):
  * values are FNV-1a-32 over the key bytes (benchmarks/masstree_bench.cu:325-334),
    remapped away from 0xFFFFFFFF, the CPU adapter's "absent" marker
    (evaluation/cpu_masstree_adapter.hpp:42);
  * Zipfian ranks use the YCSB-style generator of generate_workload.hpp:82-120 and a
    Feistel permutation of the rank space (generate_workload.hpp:54-80);
  * within one batch no key is written twice and no written key is also read, so
    outcomes do not depend on execution order.
"""
import numpy as np

U32 = np.uint32
U64 = np.uint64
MASK64 = (1 << 64) - 1


def _mix64(v):
    """murmur3 fmix64 on a uint64 array (wrapping arithmetic)."""
    v = v.astype(U64, copy=True)
    v ^= v >> U64(33)
    v *= U64(0xFF51AFD7ED558CCD)
    v ^= v >> U64(33)
    v *= U64(0xC4CEB9FE1A85EC53)
    v ^= v >> U64(33)
    return v


def key_hash64(keys, lengths):
    """64-bit hash of each key (slices beyond the length ignored)."""
    n, maxk = keys.shape
    h = _mix64(lengths.astype(U64) + U64(0x9E3779B97F4A7C15))
    for s in range(maxk):
        sl = np.where(lengths > s, keys[:, s], 0).astype(U64)
        h = _mix64(h ^ (sl + U64(s * 0x100000001B3 + 1)))
    return h


def fnv1a32_values(keys, lengths):
    """FNV-1a-32 over the little-endian bytes of each key, never 0xFFFFFFFF."""
    n, maxk = keys.shape
    h = np.full(n, 2166136261, dtype=U32)
    prime = U32(16777619)
    with np.errstate(over="ignore"):
        for s in range(maxk):
            live = lengths > s
            sl = keys[:, s]
            for b in range(4):
                byte = (sl >> U32(8 * b)) & U32(0xFF)
                hn = (h ^ byte) * prime
                h = np.where(live, hn, h)
    h = np.where(h == U32(0xFFFFFFFF), U32(0xFFFFFFFE), h).astype(U32)
    return h


def unique_random_keys(rng, n, min_len, max_len, exclude_hashes=None):
    """n distinct keys, length uniform in [min_len, max_len], uniform random slices."""
    out_k = np.zeros((0, max_len), U32)
    out_l = np.zeros(0, U32)
    seen = np.zeros(0, U64)
    need = n
    while need > 0:
        m = int(need * 1.02) + 16
        lens = rng.integers(min_len, max_len + 1, size=m, dtype=np.int64).astype(U32)
        ks = rng.integers(0, 1 << 32, size=(m, max_len), dtype=np.uint64).astype(U32)
        ks[np.arange(max_len)[None, :] >= lens[:, None]] = 0
        h = key_hash64(ks, lens)
        _, first = np.unique(h, return_index=True)
        first = np.sort(first)
        ks, lens, h = ks[first], lens[first], h[first]
        keep = ~np.isin(h, seen)
        if exclude_hashes is not None:
            keep &= ~np.isin(h, exclude_hashes)
        ks, lens, h = ks[keep][:need], lens[keep][:need], h[keep][:need]
        out_k = np.concatenate([out_k, ks])
        out_l = np.concatenate([out_l, lens])
        seen = np.concatenate([seen, h])
        need = n - len(out_l)
    return np.ascontiguousarray(out_k), out_l


def make_values(keys, lengths, max_value_length):
    """Value slice 0 = FNV-1a-32 of the key; further slices = hash rotated by the slot."""
    h = fnv1a32_values(keys, lengths)
    vals = np.zeros((len(lengths), max_value_length), U32)
    for j in range(max_value_length):
        vals[:, j] = ((h << U32(j)) | (h >> U32((32 - j) % 32))) if j else h
    return vals


def feistel_permute(x, n, seed=0):
    """Bijection on [0, n) (generate_workload.hpp:54-80), vectorized with cycle walking."""
    x = np.asarray(x, dtype=U64).copy()
    if n <= 1:
        return np.zeros_like(x)
    half = 0
    while (1 << (2 * half)) < n:
        half += 1
    mask = U64((1 << half) - 1)
    todo = np.ones(x.shape, bool)
    out = np.zeros_like(x)
    cur = x
    while todo.any():
        L = cur >> U64(half)
        R = cur & mask
        for r in range(4):
            t = L
            L = R
            R = t ^ (_mix64(R + U64(seed + r)) & mask)
        cur = (L << U64(half)) | R
        done = todo & (cur < U64(n))
        out[done] = cur[done]
        todo &= ~done
    return out


class Zipfian:
    """YCSB-style Zipfian ranks in [0, n) (generate_workload.hpp:82-120)."""

    def __init__(self, n, theta):
        self.n = n
        self.theta = theta
        idx = np.arange(1, n + 1, dtype=np.float64)
        self.zetan = float(np.sum(1.0 / np.power(idx, theta)))
        zeta2 = 1.0 + 1.0 / (2.0 ** theta)
        self.alpha = 1.0 / (1.0 - theta)
        self.eta = (1 - (2.0 / n) ** (1 - theta)) / (1 - zeta2 / self.zetan)

    def sample(self, rng, size):
        u = rng.random(size)
        uz = u * self.zetan
        base = ((self.n - 1) * np.power(self.eta * u - self.eta + 1, self.alpha)).astype(np.int64)
        base = np.where(uz < 1.0 + 0.5 ** self.theta, 1, base)
        base = np.where(uz < 1.0, 0, base)
        return np.clip(base, 0, self.n - 1).astype(U64)


# ----------------------------------------------------------------------------- configs
def plain_batch_workload(n_keys, seed=2, min_len=2, max_len=8, max_value_length=2,
                         n_find=None, n_erase=None):
    """Plain batch path: insert n_keys unique var-len keys, one find batch (50% hits,
    50% fresh keys) and one erase batch of distinct inserted keys."""
    rng = np.random.default_rng(seed)
    n_find = n_keys if n_find is None else n_find
    n_erase = n_keys // 2 if n_erase is None else n_erase
    keys, lens = unique_random_keys(rng, n_keys, min_len, max_len)
    vals = make_values(keys, lens, max_value_length)
    hit = n_find // 2
    miss = n_find - hit
    hit_idx = rng.integers(0, n_keys, size=hit)
    fresh_k, fresh_l = unique_random_keys(rng, miss, min_len, max_len,
                                          exclude_hashes=key_hash64(keys, lens))
    fk = np.concatenate([keys[hit_idx], fresh_k])
    fl = np.concatenate([lens[hit_idx], fresh_l])
    perm = rng.permutation(n_find)
    fk, fl = np.ascontiguousarray(fk[perm]), fl[perm].copy()
    er = rng.permutation(n_keys)[:n_erase]
    return dict(keys=keys, lens=lens, values=vals, find_keys=fk, find_lens=fl,
                erase_keys=np.ascontiguousarray(keys[er]), erase_lens=lens[er].copy(),
                max_key_length=max_len, max_value_length=max_value_length)


def cpu_path_workload(n_keys, seed=1):
    """CPU-runnable case: n_keys unique 2-slice keys from 64-bit draws, fixed length,
    1-slice values; n_keys finds, half existing and half fresh."""
    rng = np.random.default_rng(seed)
    keys, lens = unique_random_keys(rng, n_keys, 2, 2)
    vals = make_values(keys, lens, 1)
    hit = n_keys // 2
    fresh_k, fresh_l = unique_random_keys(rng, n_keys - hit, 2, 2,
                                          exclude_hashes=key_hash64(keys, lens))
    fk = np.concatenate([keys[rng.integers(0, n_keys, size=hit)], fresh_k])
    perm = rng.permutation(n_keys)
    return dict(keys=keys, values=vals, find_keys=np.ascontiguousarray(fk[perm]),
                find_hit=perm < hit, max_key_length=2, max_value_length=1)


class MixedWorkload:
    """Mixed batches over a preloaded key set: Zipf(theta) finds, fresh inserts, erases of
    live keys not read in the same batch; plus range scans (lower bound uniform random,
    per-query limit uniform in [1, max_scan])."""

    def __init__(self, n_preload, key_len=4, max_value_length=2, theta=0.99, seed=3,
                 find_frac=0.90, insert_frac=0.05):
        self.rng = np.random.default_rng(seed)
        self.key_len = key_len
        self.max_value_length = max_value_length
        self.keys, self.lens = unique_random_keys(self.rng, n_preload, key_len, key_len)
        self.values = make_values(self.keys, self.lens, max_value_length)
        self.n_preload = n_preload
        self.zipf = Zipfian(n_preload, theta)
        self.find_frac = find_frac
        self.insert_frac = insert_frac
        self.live = np.ones(n_preload, bool)
        self.hashes = key_hash64(self.keys, self.lens)

    def batch(self, n_ops, check_fresh=True):
        """One batch: types, keys, lens, values (input for inserts), plus the expected
        find result of every request (`expect_len`, 0 = absent) from the live-set model.
        check_fresh=False skips the membership test of fresh keys against every key ever
        generated (for keys of >= 2 random slices a collision has probability < 2^-60)."""
        rng = self.rng
        nf = int(n_ops * self.find_frac)
        ni = int(n_ops * self.insert_frac)
        ne = n_ops - nf - ni
        ranks = self.zipf.sample(rng, nf)
        fidx = feistel_permute(ranks, self.n_preload, seed=0).astype(np.int64)
        cand = rng.integers(0, self.n_preload, size=4 * ne + 64)
        cand = cand[self.live[cand] & ~np.isin(cand, fidx)]
        _, first = np.unique(cand, return_index=True)
        eidx = cand[np.sort(first)][:ne]
        ne = len(eidx)
        find_live = self.live[fidx].copy()
        self.live[eidx] = False
        ik, il = unique_random_keys(rng, ni, self.key_len, self.key_len,
                                    exclude_hashes=self.hashes if check_fresh else None)
        if check_fresh:
            self.hashes = np.concatenate([self.hashes, key_hash64(ik, il)])
        iv = make_values(ik, il, self.max_value_length)
        types = np.concatenate([np.full(nf, 2), np.full(ni, 0), np.full(ne, 1)]).astype(np.uint8)
        keys = np.concatenate([self.keys[fidx], ik, self.keys[eidx]])
        lens = np.concatenate([self.lens[fidx], il, self.lens[eidx]])
        vals = np.concatenate([np.zeros((nf, self.max_value_length), U32), iv,
                               np.zeros((ne, self.max_value_length), U32)])
        # expected outcome: finds hit iff the key is still live; inserts of fresh keys and
        # erases of live keys succeed
        expect_len = np.concatenate([np.where(find_live, self.max_value_length, 0),
                                     np.zeros(ni + ne, np.int64)]).astype(U32)
        expect_val = np.concatenate([np.where(find_live[:, None], self.values[fidx], 0),
                                     iv, np.zeros((ne, self.max_value_length), U32)]).astype(U32)
        perm = rng.permutation(len(types))
        return dict(types=types[perm].copy(), keys=np.ascontiguousarray(keys[perm]),
                    lens=lens[perm].copy(), values=np.ascontiguousarray(vals[perm]),
                    expect_len=expect_len[perm].copy(),
                    expect_values=np.ascontiguousarray(expect_val[perm]))

    def scans(self, n_queries, max_scan=100):
        rng = self.rng
        lo = rng.integers(0, 1 << 32, size=(n_queries, self.key_len), dtype=np.uint64).astype(U32)
        lim = rng.integers(1, max_scan + 1, size=n_queries).astype(U32)
        return dict(lower=np.ascontiguousarray(lo), limits=lim, max_count=max_scan)


# ----------------------------------------------------------------------------- datasets
def dataset_keys(lines, big_endian=False, append_zero_at_end=False):
    """Text keys (one per line) as u32 slice rows, as the reference's dataset loader
    builds them (evaluation/generate_workload.hpp:131-190): a trailing '\\r' is dropped,
    empty lines are skipped, each key is cut into 4-byte slices whose numeric value puts
    the first byte highest (zero padded), so numeric slice order is byte order;
    big_endian=True byte-swaps every slice; append_zero_at_end adds one zero slice to a
    key whose length is a multiple of 4 and shorter than the longest key.
    Returns (keys [n, max_len] u32, lengths u32, min_len, max_len)."""
    raw = []
    for ln in lines:
        b = ln.encode() if isinstance(ln, str) else bytes(ln)
        if b.endswith(b"\r"):
            b = b[:-1]
        if b:
            raw.append(b)
    if not raw:
        raise ValueError("dataset has no keys")
    nbytes = np.array([len(b) for b in raw], np.int64)
    lens = ((nbytes + 3) // 4).astype(U32)
    max_len = int(lens.max())
    buf = np.zeros((len(raw), max_len * 4), np.uint8)
    for i, b in enumerate(raw):
        buf[i, :len(b)] = np.frombuffer(b, np.uint8)
    keys = buf.view(">u4").astype(U32)
    if big_endian:
        keys = keys.byteswap()
    min_len = int(lens.min())
    if append_zero_at_end:
        grow = (lens < max_len) & (nbytes == lens.astype(np.int64) * 4)
        lens = lens + grow.astype(U32)
    return np.ascontiguousarray(keys), lens, min_len, max_len


def load_dataset_file(path, big_endian=False, append_zero_at_end=False):
    with open(path, "rb") as f:
        return dataset_keys(f.read().split(b"\n"), big_endian, append_zero_at_end)
