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
