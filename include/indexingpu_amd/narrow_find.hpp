// Read-only find with narrow tiles (the find<concurrent=false> batch path).
//
// A tile of T lanes (1, 2, 4, 8 or 16) reads a 128-byte node as T pieces of 128/T bytes:
// lane j holds slots S*j .. S*j + S - 1 (S = 16 / T), fetched with one dwordx2 (T = 16)
// or 16/(2T) dwordx4 loads (T <= 8).  A 64-lane wave therefore advances
// 64/T lookups per node step, and each lookup keeps K of its own in flight, so a wave
// has 64K/T independent node loads outstanding.  Per node step a tile needs:
//   * the header and sibling words (slot 15) from its last lane (ds_bpermute);
//   * interior nodes: the child slot is the number of bounds below the searched slice
//     (bounds are sorted), counted with one ballot per slot column and a popcount;
//   * border nodes: the entry equal to (slice, continues?) found with one ballot per
//     slot column, its keystate from slot 14.
// Records (key tail + value) are read the same way, 32 words per slab spread over the
// tile.  Semantics are those of op_step's find path (layers, suffix records, values of
// any length); loads are plain cached loads, since no writer runs during the launch.
#pragma once
#include "kernels.hpp"

namespace indexingpu_amd {

template <u32 T>
struct narrow_tile {
  static_assert(T == 1 || T == 2 || T == 4 || T == 8 || T == 16,
                "narrow tiles are 1, 2, 4, 8 or 16 lanes");
  static constexpr u32 S = kSlots / T;  // slots per lane
  static constexpr u32 U = 1;           // key slices cached per lane (T per tile)
  u32 j;       // lane within the tile
  u32 base;    // first lane of the tile within the wave
  u64 tmask;   // the tile's bits of a wave ballot

  __device__ __forceinline__ narrow_tile() {
    const u32 l = threadIdx.x & 63u;
    j = l & (T - 1);
    base = l & ~(T - 1);
    tmask = ((1ull << T) - 1) << base;
  }
  // value v of lane src of this tile (a one-lane tile holds everything itself)
  __device__ __forceinline__ u32 from(u32 src, u32 v) const {
    if constexpr (T == 1) {
      (void)src;
      return v;
    } else {
      return (u32)__builtin_amdgcn_ds_bpermute((int)((base + src) << 2), (int)v);
    }
  }
  __device__ __forceinline__ u64 bits(bool p) const {
    if constexpr (T == 1) return p ? tmask : 0ull;
    else return __ballot(p) & tmask;
  }
  __device__ __forceinline__ u32 count(bool p) const {
    if constexpr (T == 1) return p ? 1u : 0u;
    else return (u32)__popcll(bits(p));
  }
  // lane index (within the tile) of the first lane with p, or T
  __device__ __forceinline__ u32 first(bool p) const {
    if constexpr (T == 1) {
      return p ? 0u : 1u;
    } else {
      const u64 m = bits(p);
      return m ? (u32)(__ffsll((unsigned long long)m) - 1) - base : T;
    }
  }
};

template <u32 S>
struct node_piece {
  u64 s[S];
};

template <u32 T>
__device__ __forceinline__ node_piece<kSlots / T> load_piece(const u64* pool, u32 idx, u32 j) {
  constexpr u32 S = kSlots / T;
  node_piece<S> r;
  const u64* p = pool + (u64)idx * kSlots + (u64)j * S;
  if constexpr (S == 1) {
    r.s[0] = p[0];
  } else {
#pragma unroll
    for (u32 t = 0; t < S; t += 2) {
      const ulonglong2 v = *reinterpret_cast<const ulonglong2*>(p + t);
      r.s[t] = v.x;
      r.s[t + 1] = v.y;
    }
  }
  return r;
}

// element e (uniform over the tile) of a lane's piece
template <u32 S>
__device__ __forceinline__ u64 piece_at(const node_piece<S>& n, u32 e) {
  u64 v = n.s[0];
#pragma unroll
  for (u32 t = 1; t < S; t++) v = (e == t) ? n.s[t] : v;
  return v;
}

template <u32 T, u32 K>
__global__ void __launch_bounds__(kBlockThreads) narrow_find_kernel(tree_view tv, batch_args a) {
  using tile = narrow_tile<T>;
  constexpr u32 S = tile::S, U = tile::U;
  const tile c;
  const u64* pool = tv.pool.pool;
  const u32 nslabs = tv.pool.num_slabs;
  const u32 tiles_per_block = kBlockThreads / T;
  const u64 stride = (u64)gridDim.x * tiles_per_block * K;
  const u64 first = ((u64)blockIdx.x * tiles_per_block + threadIdx.x / T) * K;
  const u32 maxK = a.max_key_length, maxV = a.max_value_length;
  u64 idx[K];
  u32 node[K], layer[K], phase[K], klen[K], q[K], rkind[K];
  u32 kc[K][U];

  auto set_error = [&](u32 code) {
    if (c.j == 0) atomicCAS(tv.err, 0u, code);
  };
  // slice l of lookup k: from the tile's cache for l < T, else from the key array
  auto slice = [&](u32 k, u32 l) -> u32 {
    if (l >= U * T) return a.keys[idx[k] * (u64)maxK + l];
    u32 v = kc[k][0];
#pragma unroll
    for (u32 u = 1; u < U; u++) v = (l / T == u) ? kc[k][u] : v;
    return c.from(l % T, v);
  };
  auto start = [&](u32 k) {
    phase[k] = kFindIdle;
    while (idx[k] < a.n) {
      const u64 i = idx[k];
      const u32 L = a.key_lengths ? a.key_lengths[i] : maxK;
      if (L == 0 || L > maxK) {
        set_error(kErrBadLength);
        if (c.j == 0) a.out_value_lengths[i] = 0;
        idx[k] += stride;
        continue;
      }
      klen[k] = L;
      const u32* key = a.keys + i * (u64)maxK;
#pragma unroll
      for (u32 u = 0; u < U; u++) {
        const u32 s = c.j + u * T;
        kc[k][u] = s < L ? key[s] : 0u;
      }
      node[k] = tv.root;
      layer[k] = 0;
      phase[k] = kFindDescend;
      q[k] = slice(k, 0);
      return;
    }
  };
  auto finish = [&](u32 k, u32 len) {
    if (c.j == 0) a.out_value_lengths[idx[k]] = len;
    idx[k] += stride;
    start(k);
  };

#pragma unroll
  for (u32 k = 0; k < K; k++) {
    idx[k] = first + k;
    start(k);
  }
  for (u32 guard = 0; guard < (1u << 24); guard++) {
    bool any = false;
#pragma unroll
    for (u32 k = 0; k < K; k++) any |= phase[k] != kFindIdle;
    if (!any) break;
    node_piece<S> n[K];
#pragma unroll
    for (u32 k = 0; k < K; k++) {
      if (phase[k] != kFindIdle && node[k] < nslabs) n[k] = load_piece<T>(pool, node[k], c.j);
    }
#pragma unroll
    for (u32 k = 0; k < K; k++) {
      if (phase[k] == kFindIdle) continue;
      if (node[k] >= nslabs) {
        set_error(kErrBadIndex);
        finish(k, 0);
        continue;
      }
      const u64 i = idx[k];
      const u32 l = layer[k];
      if (phase[k] == kFindRecord) {
        // record: data word d of slab b is word d % 31 of the b-th slab; lane j holds
        // words 2Sj .. 2Sj + 2S - 1 of the slab it read
        const u32 lw = c.from(0, slot_lo(n[k].s[0]));
        const u32 tl = rec_tail_len(lw), vl = rec_value_len(lw);
        const u32 total = 1 + tl + vl;
        const u32 vcopy = vl < maxV ? vl : maxV;
        bool hit = true;
        const u32* key = a.keys + i * (u64)maxK;
        u32* out = a.out_values + i * (u64)maxV;
        if (rkind[k] == kSuffix && tl != klen[k] - l - 1) hit = false;
        node_piece<S> cur = n[k];
        for (u32 d0 = 0; hit;) {
          // tail words precede value words, so by the time a slab's value words are
          // copied every tail word up to them has been compared
          bool mis = false;
          if (rkind[k] == kSuffix) {
#pragma unroll
            for (u32 t = 0; t < S; t++) {
#pragma unroll
              for (u32 h = 0; h < 2; h++) {
                const u32 w = 2 * (S * c.j + t) + h;  // word within the slab
                const u32 d = d0 + w;
                const u32 v = h ? slot_hi(cur.s[t]) : slot_lo(cur.s[t]);
                if (w < kRecordWordsPerSlab && d >= 1 && d <= tl) mis |= v != key[l + d];
              }
            }
          }
          if (c.bits(mis)) {
            hit = false;
            break;
          }
#pragma unroll
          for (u32 t = 0; t < S; t++) {
#pragma unroll
            for (u32 h = 0; h < 2; h++) {
              const u32 w = 2 * (S * c.j + t) + h;
              const u32 d = d0 + w;
              const u32 v = h ? slot_hi(cur.s[t]) : slot_lo(cur.s[t]);
              if (w < kRecordWordsPerSlab && d >= 1 + tl && d < 1 + tl + vcopy) out[d - 1 - tl] = v;
            }
          }
          d0 += kRecordWordsPerSlab;
          if (d0 >= total) break;
          const u32 next = c.from(T - 1, slot_hi(cur.s[S - 1]));  // word 31
          if (next >= nslabs) {
            set_error(kErrBadIndex);
            hit = false;
            break;
          }
          cur = load_piece<T>(pool, next, c.j);
        }
        finish(k, hit ? vl : 0u);
        continue;
      }
      const u32 h = c.from(T - 1, slot_lo(n[k].s[S - 1]));  // header (slot 15)
      const u32 nk = hdr_nkeys(h);
      if (!hdr_border(h)) {
        // child slot = number of bounds below q (bounds are sorted; the last is the
        // node's high key): nk means q lies to the right of this node
        u32 below = 0;
#pragma unroll
        for (u32 t = 0; t < S; t++) {
          const u32 slot = S * c.j + t;
          below += c.count(slot < nk && q[k] > slot_lo(n[k].s[t]));
        }
        u32 next;
        if (below < nk)
          next = c.from(below / S, slot_hi(piece_at<S>(n[k], below % S))) & kChildMask;
        else
          next = c.from(T - 1, slot_hi(n[k].s[S - 1]));  // right sibling
        node[k] = next;
        if (next == kInvalid) {
          set_error(kErrBadIndex);
          finish(k, 0);
        }
        continue;
      }
      // border: move right past the high key (slot 14), as a B-link reader does
      const u32 high = c.from(kHighSlot / S, slot_lo(n[k].s[kHighSlot % S]));
      if (q[k] > high) {
        const u32 sib = c.from(T - 1, slot_hi(n[k].s[S - 1]));
        if (sib != kInvalid) {
          node[k] = sib;
          continue;
        }
      }
      const u32 ksb = c.from(kHighSlot / S, slot_hi(n[k].s[kHighSlot % S]));
      const bool more = l + 1 < klen[k];
      u32 pos = kInvalid;
#pragma unroll
      for (u32 t = 0; t < S; t++) {
        const u32 slot = S * c.j + t;
        const bool m = slot < nk && slot_lo(n[k].s[t]) == q[k] && ks_more(ks_get(ksb, slot)) == more;
        const u32 f = c.first(m);
        if (f < T) pos = S * f + t;
      }
      if (pos == kInvalid) {
        finish(k, 0);
        continue;
      }
      const u32 ks = ks_get(ksb, pos);
      const u32 val = c.from(pos / S, slot_hi(piece_at<S>(n[k], pos % S)));
      if (ks == kLink) {
        layer[k] = l + 1;
        node[k] = val;
        q[k] = slice(k, l + 1);
      } else if (ks == kValue) {
        if (c.j == 0 && maxV >= 1) a.out_values[i * (u64)maxV] = val;
        finish(k, 1);
      } else {
        rkind[k] = ks;
        node[k] = val;
        phase[k] = kFindRecord;
      }
    }
  }
}

// ---------------------------------------------------------------------------------
// Two-pass read-only find.  Within a wave, the tiles' lookups are at different depths,
// so one combined kernel executes the interior, border and record code paths on nearly
// every step.  Here pass 1 (narrow_descend_kernel) runs one uniform step for interior
// and border nodes -- both the "bounds below the slice" count and the border match are
// computed, one of them is used -- through the trie layers, and parks each lookup that
// ends at a record: out_value_lengths[i] = kPendingRecord | record index, and
// park[i] = keystate << 16 | layer.  Pass 2 (narrow_record_kernel) reads every parked
// record in one uniform sweep.  Results equal the one-pass kernel's; values of absent
// keys are not written by either pass.
constexpr u32 kPendingRecord = 0x80000000u;  // record indices stay below 2^31

template <u32 T, u32 K>
__global__ void __launch_bounds__(kBlockThreads) narrow_descend_kernel(tree_view tv, batch_args a,
                                                                        u32* park) {
  using tile = narrow_tile<T>;
  constexpr u32 S = tile::S;
  const tile c;
  const u64* pool = tv.pool.pool;
  const u32 nslabs = tv.pool.num_slabs;
  const u32 tiles_per_block = kBlockThreads / T;
  const u64 stride = (u64)gridDim.x * tiles_per_block * K;
  const u64 first = ((u64)blockIdx.x * tiles_per_block + threadIdx.x / T) * K;
  const u32 maxK = a.max_key_length, maxV = a.max_value_length;
  u64 idx[K];
  u32 node[K], layer[K], klen[K], q[K], kc[K];
  bool active[K];

  auto set_error = [&](u32 code) {
    if (c.j == 0) atomicCAS(tv.err, 0u, code);
  };
  auto slice = [&](u32 k, u32 l) -> u32 {
    if (l >= T) return a.keys[idx[k] * (u64)maxK + l];
    return c.from(l, kc[k]);
  };
  auto start = [&](u32 k) {
    active[k] = false;
    while (idx[k] < a.n) {
      const u64 i = idx[k];
      const u32 L = a.key_lengths ? a.key_lengths[i] : maxK;
      if (L == 0 || L > maxK) {
        set_error(kErrBadLength);
        if (c.j == 0) a.out_value_lengths[i] = 0;
        idx[k] += stride;
        continue;
      }
      klen[k] = L;
      kc[k] = c.j < L ? a.keys[i * (u64)maxK + c.j] : 0u;
      node[k] = tv.root;
      layer[k] = 0;
      active[k] = true;
      q[k] = slice(k, 0);
      return;
    }
  };
  auto finish = [&](u32 k, u32 len) {
    if (c.j == 0) a.out_value_lengths[idx[k]] = len;
    idx[k] += stride;
    start(k);
  };

#pragma unroll
  for (u32 k = 0; k < K; k++) {
    idx[k] = first + k;
    start(k);
  }
  for (u32 guard = 0; guard < (1u << 24); guard++) {
    bool any = false;
#pragma unroll
    for (u32 k = 0; k < K; k++) any |= active[k];
    if (!any) break;
    node_piece<S> n[K];
#pragma unroll
    for (u32 k = 0; k < K; k++)
      if (active[k] && node[k] < nslabs) n[k] = load_piece<T>(pool, node[k], c.j);
#pragma unroll
    for (u32 k = 0; k < K; k++) {
      if (!active[k]) continue;
      if (node[k] >= nslabs) {
        set_error(kErrBadIndex);
        finish(k, 0);
        continue;
      }
      const u32 l = layer[k];
      const u32 h = c.from(T - 1, slot_lo(n[k].s[S - 1]));    // header (slot 15)
      const u32 sib = c.from(T - 1, slot_hi(n[k].s[S - 1]));  // right sibling
      const u32 high = c.from(kHighSlot / S, slot_lo(n[k].s[kHighSlot % S]));
      const u32 ksb = c.from(kHighSlot / S, slot_hi(n[k].s[kHighSlot % S]));
      const u32 nk = hdr_nkeys(h);
      const bool border = hdr_border(h);
      const bool more = l + 1 < klen[k];
      // interior: child slot = bounds below q; border: the entry equal to (q, more)
      u32 below = 0, pos = kInvalid;
#pragma unroll
      for (u32 t = 0; t < S; t++) {
        const u32 slot = S * c.j + t;
        const u32 key = slot_lo(n[k].s[t]);
        const bool valid = slot < nk;
        below += c.count(valid && q[k] > key);
        const u32 f = c.first(valid && key == q[k] && ks_more(ks_get(ksb, slot)) == more);
        if (f < T) pos = S * f + t;
      }
      const bool right = border ? (q[k] > high && sib != kInvalid) : below >= nk;
      const u32 sel = border ? (pos < kSlots ? pos : 0u) : (below < kSlots ? below : 0u);
      const u32 v = c.from(sel / S, slot_hi(piece_at<S>(n[k], sel % S)));
      if (right) {
        node[k] = sib;
        if (sib == kInvalid) {
          set_error(kErrBadIndex);
          finish(k, 0);
        }
        continue;
      }
      if (!border) {
        node[k] = v & kChildMask;
        continue;
      }
      if (pos == kInvalid) {
        finish(k, 0);
        continue;
      }
      const u32 ks = ks_get(ksb, pos);
      if (ks == kLink) {
        layer[k] = l + 1;
        node[k] = v;
        q[k] = slice(k, l + 1);
      } else if (ks == kValue) {
        if (c.j == 0 && maxV >= 1) a.out_values[idx[k] * (u64)maxV] = v;
        finish(k, 1);
      } else if (v >= nslabs) {
        set_error(kErrBadIndex);
        finish(k, 0);
      } else {
        if (c.j == 0) park[idx[k]] = (ks << 16) | l;
        finish(k, kPendingRecord | v);
      }
    }
  }
}

template <u32 T>
__global__ void __launch_bounds__(kBlockThreads) narrow_record_kernel(tree_view tv, batch_args a,
                                                                       const u32* park) {
  using tile = narrow_tile<T>;
  constexpr u32 S = tile::S;
  const tile c;
  const u64* pool = tv.pool.pool;
  const u32 nslabs = tv.pool.num_slabs;
  const u32 tiles_per_block = kBlockThreads / T;
  const u64 ntiles = (u64)gridDim.x * tiles_per_block;
  const u32 maxK = a.max_key_length, maxV = a.max_value_length;
  for (u64 i = (u64)blockIdx.x * tiles_per_block + threadIdx.x / T; i < a.n; i += ntiles) {
    const u32 w = a.out_value_lengths[i];
    if (!(w & kPendingRecord)) continue;
    const u32 rec = w & ~kPendingRecord;
    const u32 pk = park[i];
    const u32 rkind = pk >> 16, l = pk & 0xFFFFu;
    const u32 klen = a.key_lengths ? a.key_lengths[i] : maxK;
    node_piece<S> cur = load_piece<T>(pool, rec, c.j);
    const u32 lw = c.from(0, slot_lo(cur.s[0]));
    const u32 tl = rec_tail_len(lw), vl = rec_value_len(lw);
    const u32 total = 1 + tl + vl;
    const u32 vcopy = vl < maxV ? vl : maxV;
    const u32* key = a.keys + i * (u64)maxK;
    u32* out = a.out_values + i * (u64)maxV;
    bool hit = !(rkind == kSuffix && tl != klen - l - 1);
    for (u32 d0 = 0; hit;) {
      // tail words precede value words: every tail word up to a slab's value words has
      // been compared when they are copied
      bool mis = false;
      if (rkind == kSuffix) {
#pragma unroll
        for (u32 t = 0; t < S; t++) {
#pragma unroll
          for (u32 hh = 0; hh < 2; hh++) {
            const u32 wd = 2 * (S * c.j + t) + hh;
            const u32 d = d0 + wd;
            const u32 v = hh ? slot_hi(cur.s[t]) : slot_lo(cur.s[t]);
            if (wd < kRecordWordsPerSlab && d >= 1 && d <= tl) mis |= v != key[l + d];
          }
        }
      }
      if (c.bits(mis)) {
        hit = false;
        break;
      }
#pragma unroll
      for (u32 t = 0; t < S; t++) {
#pragma unroll
        for (u32 hh = 0; hh < 2; hh++) {
          const u32 wd = 2 * (S * c.j + t) + hh;
          const u32 d = d0 + wd;
          const u32 v = hh ? slot_hi(cur.s[t]) : slot_lo(cur.s[t]);
          if (wd < kRecordWordsPerSlab && d >= 1 + tl && d < 1 + tl + vcopy) out[d - 1 - tl] = v;
        }
      }
      d0 += kRecordWordsPerSlab;
      if (d0 >= total) break;
      const u32 next = c.from(T - 1, slot_hi(cur.s[S - 1]));  // word 31
      if (next >= nslabs) {
        if (c.j == 0) atomicCAS(tv.err, 0u, (u32)kErrBadIndex);
        hit = false;
        break;
      }
      cur = load_piece<T>(pool, next, c.j);
    }
    if (c.j == 0) a.out_value_lengths[i] = hit ? vl : 0u;
  }
}

}  // namespace indexingpu_amd
