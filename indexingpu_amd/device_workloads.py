"""Generators that build large workloads directly in device memory with torch.

Same formulas as the numpy generators in workloads.py (which the parity tests use at
small sizes): murmur3 fmix64, the 4-round Feistel bijection of the rank space and the
YCSB Zipfian rank transform (evaluation/generate_workload.hpp:54-120 in
hyoungjook/IndexinGPU).  u64 values live in int64 tensors; multiplication wraps and
right shifts are made logical by masking.

PrefixBatchWorkload builds the batched-form workload: keys of 16 u32 slices whose
first 12 slices are one of 256 random prefixes and whose last 4 are random, so every
lookup walks 12 single-entry trie layers before the per-prefix B-link tree; uneven
mixed batches (90% Zipf finds / 5% fresh inserts / 5% erases of live keys not read in
the batch); the expected result of every request from a live-set model.
"""
import math

import numpy as np
import torch

from . import workloads as W


def _s64(c):
    """u64 constant as the int64 with the same bits."""
    c &= (1 << 64) - 1
    return c - (1 << 64) if c >= (1 << 63) else c


_M1 = _s64(0xFF51AFD7ED558CCD)
_M2 = _s64(0xC4CEB9FE1A85EC53)


def srl(v, s):
    """Logical right shift of the u64 bits held in an int64 tensor."""
    return (v >> s) & ((1 << (64 - s)) - 1)


def mix64(v):
    """murmur3 fmix64 (same as workloads._mix64)."""
    v = v ^ srl(v, 33)
    v = v * _M1
    v = v ^ srl(v, 33)
    v = v * _M2
    return v ^ srl(v, 33)


def feistel_permute(x, n, seed=0):
    """Bijection on [0, n) (same as workloads.feistel_permute), int64 in and out."""
    if n <= 1:
        return torch.zeros_like(x)
    half = 0
    while (1 << (2 * half)) < n:
        half += 1
    mask = (1 << half) - 1
    out = torch.empty_like(x)
    todo = torch.ones_like(x, dtype=torch.bool)
    cur = x.clone()
    while bool(todo.any()):
        L = cur >> half
        R = cur & mask
        for r in range(4):
            L, R = R, L ^ (mix64(R + (seed + r)) & mask)
        cur = (L << half) | R
        done = todo & (cur < n)
        out[done] = cur[done]
        todo &= ~done
    return out


def zipf_ranks(z, u):
    """Zipfian ranks from uniforms u in [0, 1) (same transform as workloads.Zipfian)."""
    base = ((z.n - 1) * torch.pow(z.eta * u - z.eta + 1, z.alpha)).to(torch.int64)
    uz = u * z.zetan
    base = torch.where(uz < 1.0 + 0.5 ** z.theta, torch.ones_like(base), base)
    base = torch.where(uz < 1.0, torch.zeros_like(base), base)
    return base.clamp_(0, z.n - 1)


def log_uniform_sizes(rng, count, lo, hi):
    """floor(exp(U(ln lo, ln hi))) batch sizes (numpy Generator)."""
    return np.floor(np.exp(rng.uniform(math.log(lo), math.log(hi), size=count))).astype(np.int64)


def _u32_as_i32(v64):
    return (v64 & 0xFFFFFFFF).to(torch.int32)  # wraps to the same 32 bits


class PrefixBatchWorkload:
    """Keys of prefix_len + tail_len slices; prefix = one of n_prefixes random rows."""

    def __init__(self, n_preload, n_prefixes=256, prefix_len=12, tail_len=4,
                 max_value_length=2, theta=0.99, seed=4, device="cuda",
                 find_frac=0.90, insert_frac=0.05):
        self.device = device
        self.n = n_preload
        self.prefix_len, self.tail_len = prefix_len, tail_len
        self.key_len = prefix_len + tail_len
        self.mv = max_value_length
        self.find_frac, self.insert_frac = find_frac, insert_frac
        self.gen = torch.Generator(device=device)
        self.gen.manual_seed(seed)
        rng = np.random.default_rng(seed)
        pre = rng.integers(0, 1 << 32, size=(n_prefixes, prefix_len), dtype=np.uint64)
        self.prefixes = torch.from_numpy(pre.astype(np.uint32).view(np.int32)).to(device)
        self.pid, self.tails = self._draw(n_preload)
        self.values = self._values(self.pid, self.tails)
        self.live = torch.ones(n_preload, dtype=torch.bool, device=device)
        self.mark = torch.full((n_preload,), -1, dtype=torch.int32, device=device)
        self.batch_id = 0
        self.zipf = W.Zipfian(n_preload, theta)

    def _draw(self, n):
        pid = torch.randint(0, self.prefixes.shape[0], (n,), generator=self.gen,
                            device=self.device)
        tails = torch.randint(-(1 << 31), 1 << 31, (n, self.tail_len), generator=self.gen,
                              device=self.device, dtype=torch.int64).to(torch.int32)
        return pid, tails

    def _values(self, pid, tails):
        """Value slices: fmix64 chain over (prefix id, tail slices); a function of the key."""
        h = mix64(pid.to(torch.int64) + 0x5F3759DF)
        for s in range(self.tail_len):
            h = mix64(h ^ (tails[:, s].to(torch.int64) & 0xFFFFFFFF))
        cols = []
        for j in range(self.mv):
            cols.append(_u32_as_i32(srl(h, 32) if j % 2 else h))
            if j % 2:
                h = mix64(h + j)
        return torch.stack(cols, dim=1)

    def keys_of(self, pid, tails):
        """[n, key_len] int32 key rows (u32 bits)."""
        return torch.cat([self.prefixes[pid], tails], dim=1).contiguous()

    def preload(self):
        keys = self.keys_of(self.pid, self.tails)
        lens = torch.full((self.n,), self.key_len, dtype=torch.int32, device=self.device)
        return keys, lens, self.values

    def batch(self, n_ops):
        """One batch of n_ops requests as device tensors, with expected outputs."""
        dev, gen = self.device, self.gen
        nf = int(n_ops * self.find_frac)
        ni = int(n_ops * self.insert_frac)
        ne = n_ops - nf - ni
        u = torch.rand(nf, generator=gen, device=dev, dtype=torch.float64)
        fidx = feistel_permute(zipf_ranks(self.zipf, u), self.n, seed=0)
        bid = self.batch_id
        self.batch_id += 1
        self.mark[fidx] = bid
        cand = torch.randint(0, self.n, (4 * ne + 64,), generator=gen, device=dev)
        cand = cand[self.live[cand] & (self.mark[cand] != bid)]
        cand = torch.unique(cand)
        cand = cand[torch.randperm(len(cand), generator=gen, device=dev)]
        eidx = cand[:ne]
        ne = len(eidx)
        find_live = self.live[fidx]
        self.live[eidx] = False
        ipid, itails = self._draw(ni)
        ivals = self._values(ipid, itails)
        n = nf + ni + ne
        types = torch.cat([torch.full((nf,), 2, dtype=torch.uint8, device=dev),
                           torch.zeros(ni, dtype=torch.uint8, device=dev),
                           torch.ones(ne, dtype=torch.uint8, device=dev)])
        keys = torch.cat([self.keys_of(self.pid[fidx], self.tails[fidx]),
                          self.keys_of(ipid, itails),
                          self.keys_of(self.pid[eidx], self.tails[eidx])])
        zeros_v = torch.zeros((nf, self.mv), dtype=torch.int32, device=dev)
        vals = torch.cat([zeros_v, ivals, torch.zeros((ne, self.mv), dtype=torch.int32,
                                                        device=dev)])
        # value_lengths is in/out: finds write 0 or the stored length, inserts and
        # erases leave the input (max_value_length) in place
        exp_len = torch.cat([find_live.to(torch.int32) * self.mv,
                             torch.full((ni + ne,), self.mv, dtype=torch.int32, device=dev)])
        exp_val = torch.cat([torch.where(find_live[:, None], self.values[fidx], zeros_v),
                             ivals, torch.zeros((ne, self.mv), dtype=torch.int32, device=dev)])
        exp_res = torch.cat([torch.zeros(nf, dtype=torch.uint8, device=dev),
                             torch.ones(ni + ne, dtype=torch.uint8, device=dev)])
        perm = torch.randperm(n, generator=gen, device=dev)
        return dict(types=types[perm], keys=keys[perm].contiguous(),
                    lens=torch.full((n,), self.key_len, dtype=torch.int32, device=dev),
                    values=vals[perm].contiguous(), expect_len=exp_len[perm],
                    expect_values=exp_val[perm].contiguous(), expect_results=exp_res[perm])


def concat_batches(batches):
    """Back-to-back request arrays plus the u64 offsets of each batch (device)."""
    dev = batches[0]["types"].device
    sizes = [len(b["types"]) for b in batches]
    offs = torch.tensor(np.concatenate([[0], np.cumsum(sizes)]), dtype=torch.int64, device=dev)
    out = {k: torch.cat([b[k] for b in batches]) for k in batches[0]}
    out["offsets"] = offs
    out["sizes"] = sizes
    return out


def key_values(keys, lens, mv, chunk=1 << 25):
    """Value slices of var-length keys ([n, L] int32 rows, zero past the length): an
    fmix64 chain over (length, slices); a function of the key alone."""
    out = torch.empty((keys.shape[0], mv), dtype=torch.int32, device=keys.device)
    for a in range(0, keys.shape[0], chunk):
        b = min(keys.shape[0], a + chunk)
        h = mix64(lens[a:b].to(torch.int64) + 0x2545F491)
        for s in range(keys.shape[1]):
            h = mix64(h ^ ((keys[a:b, s].to(torch.int64) & 0xFFFFFFFF) + (s << 40)))
        for j in range(mv):
            out[a:b, j] = _u32_as_i32(srl(h, 32) if j % 2 else h)
            if j % 2:
                h = mix64(h + j)
    return out


def random_keys(gen, n, min_len, max_len, region, device, chunk=1 << 25):
    """n keys of min_len..max_len random slices; region 0 puts the first slice below
    2^31, region 1 at or above it.  Rows are zero past the key length.  Generated in
    chunks so the int64 temporaries stay small."""
    lens = torch.randint(min_len, max_len + 1, (n,), generator=gen, device=device,
                         dtype=torch.int32)
    keys = torch.empty((n, max_len), dtype=torch.int32, device=device)
    col = torch.arange(max_len, device=device, dtype=torch.int32)[None, :]
    for a in range(0, n, chunk):
        b = min(n, a + chunk)
        k = torch.randint(-(1 << 31), 1 << 31, (b - a, max_len), generator=gen, device=device,
                          dtype=torch.int64).to(torch.int32)
        first = k[:, 0] & 0x7FFFFFFF
        k[:, 0] = first if region == 0 else first | torch.tensor(-(1 << 31), dtype=torch.int32,
                                                                  device=device)
        keys[a:b] = torch.where(col < lens[a:b, None], k, torch.zeros_like(k))
    return keys, lens


class ServeWorkload:
    """Served-form workload over a resident index of n keys (min_len..max_len random
    slices, max_value_length-slice values = key_values of the key).  The key space is
    split by the top bit of the first slice: the read region (finds and scans) is never
    written; the write region holds the erase targets (each erased once) and receives
    the fresh inserts.  Hence every request's outcome is independent of the others and
    of their timing, and known in advance: finds hit with the key's value, inserts and
    erases succeed, a scan returns the same rows whenever it runs.

    trace(): requests of floor(exp(U(0, ln max_ops))) operations with Poisson arrivals at
    `rate` requests per second; each operation is a find / insert / erase / scan with
    probabilities mix; scans start at a random read-region key and stop at the end of
    the read region or after scan_limit results."""

    def __init__(self, n_preload, seed=5, device="cuda", min_len=2, max_len=8,
                 max_value_length=2):
        self.device = device
        self.mk, self.mv = max_len, max_value_length
        self.min_len = min_len
        self.gen = torch.Generator(device=device)
        self.gen.manual_seed(seed)
        self.rng = np.random.default_rng(seed)
        nr = n_preload // 2
        rk, rl = random_keys(self.gen, nr, min_len, max_len, 0, device)
        wk, wl = random_keys(self.gen, n_preload - nr, min_len, max_len, 1, device)
        self.keys = torch.cat([rk, wk]).contiguous()
        self.lens = torch.cat([rl, wl]).contiguous()
        self.values = key_values(self.keys, self.lens, self.mv)
        self.n_read = nr
        self.erase_order = nr + torch.randperm(n_preload - nr, generator=self.gen, device=device)
        self.erase_cursor = 0
        hi = torch.full((self.mk,), -1, dtype=torch.int32, device=device)
        hi[0] = 0x7FFFFFFF
        self.upper = hi  # largest key of the read region

    def preload(self):
        return self.keys, self.lens, self.values

    def trace(self, n_requests, rate, max_ops=4096, mix=(0.80, 0.10, 0.05, 0.05),
              scan_limit=50):
        rng, dev, mk, mv = self.rng, self.device, self.mk, self.mv
        sizes = np.floor(np.exp(rng.uniform(0.0, np.log(max_ops), size=n_requests))).astype(np.int64)
        sizes = np.clip(sizes, 1, max_ops)
        per = rng.multinomial(sizes, mix)  # [n, 4]: find, insert, erase, scan
        nf, ni, ne, ns = per[:, 0], per[:, 1], per[:, 2], per[:, 3]
        nops = nf + ni + ne
        if self.erase_cursor + int(ne.sum()) > len(self.erase_order):
            raise ValueError("erase targets exhausted: preload more keys")
        # per-op layout: each request holds its finds, then inserts, then erases
        types = np.repeat(np.tile(np.array([2, 0, 1], np.uint8), n_requests),
                          per[:, :3].reshape(-1))
        n_op = len(types)
        t_types = torch.from_numpy(types).to(dev)
        is_f, is_i, is_e = t_types == 2, t_types == 0, t_types == 1
        keys = torch.zeros((n_op, mk), dtype=torch.int32, device=dev)
        lens = torch.zeros(n_op, dtype=torch.int32, device=dev)
        vals = torch.zeros((n_op, mv), dtype=torch.int32, device=dev)
        exp_v = torch.zeros((n_op, mv), dtype=torch.int32, device=dev)
        F = int(nf.sum())
        fidx = torch.randint(0, self.n_read, (F,), generator=self.gen, device=dev)
        keys[is_f] = self.keys[fidx]
        lens[is_f] = self.lens[fidx]
        exp_v[is_f] = self.values[fidx]
        I = int(ni.sum())
        ik, il = random_keys(self.gen, I, self.min_len, mk, 1, dev)
        iv = key_values(ik, il, mv)
        keys[is_i], lens[is_i], vals[is_i], exp_v[is_i] = ik, il, iv, iv
        E = int(ne.sum())
        eidx = self.erase_order[self.erase_cursor:self.erase_cursor + E]
        self.erase_cursor += E
        keys[is_e] = self.keys[eidx]
        lens[is_e] = self.lens[eidx]
        S = int(ns.sum())
        lo, lol = random_keys(self.gen, S, 1, mk, 0, dev)
        req = np.zeros(n_requests, dtype=_request_dtype())
        gaps = rng.exponential(1e6 / rate, size=n_requests)
        req["arrival_us"] = np.cumsum(gaps) - gaps[0]
        req["op_offset"] = np.concatenate([[0], np.cumsum(nops)[:-1]])
        req["scan_offset"] = np.concatenate([[0], np.cumsum(ns)[:-1]])
        req["num_ops"], req["num_scans"], req["num_writes"] = nops, ns, ni
        arena = dict(types=t_types, keys=keys, max_key_length=mk, key_lengths=lens,
                     values=vals, max_value_length=mv,
                     value_lengths=torch.full((n_op,), mv, dtype=torch.int32, device=dev),
                     results=torch.zeros(n_op, dtype=torch.uint8, device=dev),
                     scan_lower=lo, scan_lower_lengths=lol,
                     scan_upper=self.upper.repeat(S, 1).contiguous(),
                     scan_upper_lengths=torch.full((S,), mk, dtype=torch.int32, device=dev),
                     scan_limits=torch.full((S,), scan_limit, dtype=torch.int32, device=dev),
                     scan_max_count=scan_limit,
                     scan_counts=torch.zeros(S, dtype=torch.int32, device=dev))
        expect = dict(value_lengths=torch.full((n_op,), mv, dtype=torch.int32, device=dev),
                      values=exp_v,
                      results=(t_types != 2).to(torch.uint8))
        return arena, req, expect

    @staticmethod
    def with_ring(arena, rows):
        """Adds scan result ring buffers of `rows` rows to an arena."""
        dev = arena["types"].device
        mk, mv = arena["max_key_length"], arena["max_value_length"]
        a = dict(arena)
        a["ring_keys"] = torch.zeros(rows * mk, dtype=torch.int32, device=dev)
        a["ring_key_lengths"] = torch.zeros(rows, dtype=torch.int32, device=dev)
        a["ring_values"] = torch.zeros(rows * mv, dtype=torch.int32, device=dev)
        a["ring_value_lengths"] = torch.zeros(rows, dtype=torch.int32, device=dev)
        return a

    @staticmethod
    def fresh_outputs(arena):
        """A copy of an arena whose in/out and output tensors are new (inputs shared)."""
        a = dict(arena)
        for k in ("values", "value_lengths", "results", "scan_counts"):
            a[k] = arena[k].clone()
        return a


def _request_dtype():
    from .serve import REQUEST_DTYPE
    return REQUEST_DTYPE
