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
  const u32* start_table = find_start_table(tv);

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
      layer[k] = 0;
      phase[k] = kFindDescend;
      q[k] = slice(k, 0);
      node[k] = start_table ? start_table[q[k] >> tv.shortcut_shift] : tv.root;
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
// every step.  Here pass 1 (narrow_descend_kernel) walks interior and border nodes
// through the trie layers and parks each lookup that ends at a record:
// out_value_lengths[i] = kPendingRecord | record index, and park[i - lo] = keystate << 30
// | layer << 15 | tail length (key length - layer - 1), lo being the first request of the
// launch (of its range, in the batched form); so pass 2 reads no key lengths.  Keys of
// 2^15 slices or more take keystate << 16 | layer, and pass 2 reads their lengths.
// Pass 2 (narrow_record_kernel) reads every parked record in one uniform sweep.  Results
// equal the one-pass kernel's; values of absent keys are not written by either pass.
// In a mixed batch both passes answer the finds only.
constexpr u32 kPendingRecord = 0x80000000u;  // record indices stay below 2^31
constexpr u32 kParkPacked = 1u << 15;        // max_key_length below this: packed park words
constexpr u32 kNoOutput = 0xFFFFFFFEu;       // a request the find passes do not answer
constexpr u32 kHintDone = 0xFFFFFFFDu;       // a write request: its start hint is written

// Exchanges inside a tile for the descent.  Tiles of two or four lanes use DPP quad
// permutations (an operand modifier of a VALU instruction: no LDS traffic, no wait);
// wider tiles use ds_bpermute.  Every lane of a tile takes the same branches, so the
// lanes a permutation reads are active.
template <u32 T>
struct tile_xchg {
  u32 j;     // lane within the tile
  u32 base;  // first lane of the tile within the wave
  __device__ __forceinline__ tile_xchg() {
    const u32 l = threadIdx.x & 63u;
    j = l & (T - 1);
    base = l & ~(T - 1);
  }
  // quad_perm control: lane i of each quad reads lane sel(i) of the same quad
  template <u32 L0, u32 L1, u32 L2, u32 L3>
  static __device__ __forceinline__ u32 quad(u32 v) {
    constexpr int ctrl = (int)(L0 | (L1 << 2) | (L2 << 4) | (L3 << 6));
    return (u32)__builtin_amdgcn_update_dpp(0, (int)v, ctrl, 0xF, 0xF, false);
  }
  // v of tile lane Src
  template <u32 Src>
  __device__ __forceinline__ u32 lane(u32 v) const {
    static_assert(Src < T, "source lane outside the tile");
    if constexpr (T == 1) {
      return v;
    } else if constexpr (T == 2) {
      return quad<Src, Src, 2 + Src, 2 + Src>(v);
    } else if constexpr (T == 4) {
      return quad<Src, Src, Src, Src>(v);
    } else {
      return (u32)__builtin_amdgcn_ds_bpermute((int)((base + Src) << 2), (int)v);
    }
  }
  // v of the lane whose index differs from this one's in bit M
  template <u32 M>
  __device__ __forceinline__ u32 partner(u32 v) const {
    if constexpr (M == 1) {
      return quad<1, 0, 3, 2>(v);
    } else if constexpr (M == 2) {
      return quad<2, 3, 0, 1>(v);
    } else {
      const u32 src = (threadIdx.x & 63u) ^ M;
      return (u32)__builtin_amdgcn_ds_bpermute((int)(src << 2), (int)v);
    }
  }
  __device__ __forceinline__ u32 min(u32 v) const {
    if constexpr (T >= 2) v = umin(v, partner<1>(v));
    if constexpr (T >= 4) v = umin(v, partner<2>(v));
    if constexpr (T >= 8) v = umin(v, partner<4>(v));
    if constexpr (T >= 16) v = umin(v, partner<8>(v));
    return v;
  }
  __device__ __forceinline__ u32 bor(u32 v) const {
    if constexpr (T >= 2) v |= partner<1>(v);
    if constexpr (T >= 4) v |= partner<2>(v);
    if constexpr (T >= 8) v |= partner<4>(v);
    if constexpr (T >= 16) v |= partner<8>(v);
    return v;
  }
  static __device__ __forceinline__ u32 umin(u32 a, u32 b) { return a < b ? a : b; }
};

// One descent step per node: every lane finds the lowest slot of its piece that
// qualifies (interior: the first bound >= the slice; border: the entry equal to
// (slice, continues?)) with two selects per slot, then the tile takes the minimum slot
// and the owner's child/value with two exchanges.  Slots at or past the node's key count
// hold no entry, but every such slot lies above every real entry, so the tile's lowest
// qualifying slot is a real entry exactly when it is below the key count.
//
// General = false: a read-only find batch (every request a find, in array order).
// General = true adds what the passes over mixed and write batches need: request kinds
// (only finds are answered), the device-read range of the batched form, a scattered
// request order (request_index), and the start hints of write requests.  The two are
// separate instantiations, so find batches carry none of it.
template <u32 T, u32 K, bool General = false, bool UseRing = true>
__global__ void __launch_bounds__(kBlockThreads) narrow_descend_kernel(tree_view tv, batch_args a,
                                                                        u32* park) {
  constexpr u32 S = kSlots / T;  // slots per lane
  constexpr u32 kNone = 32;      // no qualifying slot in this lane's piece
  // positions from the wave's ring (below) unless the launch asks for tiles stepping
  // through their own positions (UseRing = false: the host's choice for the batched form's
  // device-read ranges, scattered orders and general batches too small to fill the grid
  // twice, where the ring reads no fewer key lines or runs in more rounds)
  constexpr bool Ring = UseRing && K <= T;
  const tile_xchg<T> x;
  const u64* pool = tv.pool.pool;
  const u32 nslabs = tv.pool.num_slabs;
  const u32 tiles_per_block = kBlockThreads / T;
  const u64 stride = (u64)gridDim.x * tiles_per_block * K;
  const u64 first = ((u64)blockIdx.x * tiles_per_block + threadIdx.x / T) * K;
  const u32 maxK = a.max_key_length, maxV = a.max_value_length;
  const u32 slot0 = S * x.j;  // this lane's first slot
  // launch positions [lo, hi): the batch, or the range read on the device (batched form)
  u64 lo = 0, hi = a.n;
  if constexpr (General) {
    if (a.range) {
      hi = min(a.range[1], a.n);
      lo = min(a.range[0], hi);
    }
  }
  if (lo >= hi) return;
  u64 idx[K];                  // launch position of lookup k
  u32 node[K], layer[K], klen[K], q[K];
  u64 req[K], next_req[K];     // (General) request index of lookup k, and of the next
  u32 next_len[K], next_q[K];  // length and first slice of lookup idx[k], loaded ahead
  u32 next_kind[K], kind[K];   // (General) request kinds: only finds are answered
  u32 lroot[K], par[K];        // (General) layer root and level-1 node passed (hints)
  bool active[K], q_stale[K];
  // finds may start at border level (find_start_table); the write requests of the general
  // passes (their start hints want the level-1 node passed) at level 1
  const u32* find_table = find_start_table(tv);
  const u32* write_table = tv.shortcut;
  // the node a lookup of first slice q0 and request kind kd starts at
  auto start_node = [&](u32 q0, u32 kd) -> u32 {
    bool is_find = true;
    if constexpr (General) is_find = a.types ? kd == op_find : a.kind == op_find;
    const u32* t = is_find ? find_table : write_table;
    return t ? t[q0 >> tv.shortcut_shift] : tv.root;
  };

  auto request = [&](u32 k) -> u64 {
    if constexpr (General) return req[k];
    else return idx[k];
  };
  // loads of the tile's following lookup (the position clamped into the batch, so the
  // loads are unconditional and land in the registers the next start reads)
  auto fetch = [&](u32 k, u64 i) {
    u64 c = i < hi ? i : hi - 1;
    if constexpr (General) {
      c = request_index(a, c);
      next_req[k] = c;
      if (a.types) next_kind[k] = a.types[c];
    }
    next_len[k] = a.key_lengths ? a.key_lengths[c] : maxK;
    next_q[k] = a.keys[c * (u64)maxK];
  };
  // begins lookup idx[k] (its length and first slice are in next_len/next_q, loaded one
  // lookup ahead) and loads those of the tile's following lookup, which arrive while
  // this one descends.  A length outside 1..max_key_length is reported at the lookup's
  // first step.
  auto start = [&](u32 k) {
    // copy the values into their own registers here, before the following loads are
    // issued into next_len/next_q (a copy placed after those loads would wait for them)
    u32 L, q0;
    asm volatile("v_mov_b32 %0, %1" : "=v"(L) : "v"(next_len[k]));
    asm volatile("v_mov_b32 %0, %1" : "=v"(q0) : "v"(next_q[k]));
    klen[k] = L;
    q[k] = q0;
    u32 kd = op_find;
    if constexpr (General) {
      req[k] = next_req[k];
      if (a.types) {
        asm volatile("v_mov_b32 %0, %1" : "=v"(kd) : "v"(next_kind[k]));
        kind[k] = kd;
      }
      lroot[k] = tv.root;
      par[k] = kInvalid;
    }
    node[k] = start_node(q0, kd);
    layer[k] = 0;
    q_stale[k] = false;
    active[k] = idx[k] < hi;
    fetch(k, idx[k] + stride);
  };

  // (Ring) The wave takes the launch's positions in chunks of 64: chunk c is positions
  // lo + 64c .. lo + 64c + 63, and the wave's chunks are its wave number plus multiples of
  // the grid's wave count.  Lane l holds the first slice, length and (General) kind of
  // the request at position l of the current chunk and of the next one (loaded one chunk
  // ahead, one coalesced load each in array order), and a tile whose lookup ends takes
  // the next position of the ring from the lane that holds it.  Each key row and length
  // line is then read once per chunk, where tiles that step through their own positions
  // (the other instantiations) read the same lines again at different times.  A round of
  // takes asks for at most 64 positions (K <= T), so it stays within the two chunks held.
  const u32 lane = threadIdx.x & 63u;
  const u64 waves = (u64)gridDim.x * (kBlockThreads / 64);
  u64 cur = (u64)blockIdx.x * (kBlockThreads / 64) + threadIdx.x / 64, nxt = cur + waves;
  u32 ring_q = 0, ring_len = 0, ring_kd = 0, next_ring_q = 0, next_ring_len = 0,
      next_ring_kd = 0;
  u32 head = 0;          // positions of the current chunk already taken
  bool want[K];          // lookup k has ended and takes a new position
  auto ring_load = [&](u64 chunk, u32& rq, u32& rl, u32& rk) {
    const u64 p = lo + chunk * 64 + lane;
    if (p < hi) {
      const u64 c = General ? request_index(a, p) : p;
      rl = a.key_lengths ? a.key_lengths[c] : maxK;
      rq = a.keys[c * (u64)maxK];
      if constexpr (General) {
        if (a.types) rk = a.types[c];
      }
    }
  };
  auto take = [&]() {
    u32 tq[K], tl[K], tk[K];
    u64 tpos[K];
    // every position and ring value first, then the loads (a ring value read after a
    // load was issued would wait for it)
#pragma unroll
    for (u32 k = 0; k < K; k++) {
      const u64 b = __ballot(want[k] && x.j == 0);  // one bit per taking tile
      const u32 slot = head + (u32)__popcll(b & ((1ull << x.base) - 1ull));
      const int src = (int)((slot & 63u) << 2);
      const u32 qc = (u32)__builtin_amdgcn_ds_bpermute(src, (int)ring_q);
      const u32 qn = (u32)__builtin_amdgcn_ds_bpermute(src, (int)next_ring_q);
      const u32 lc = (u32)__builtin_amdgcn_ds_bpermute(src, (int)ring_len);
      const u32 ln = (u32)__builtin_amdgcn_ds_bpermute(src, (int)next_ring_len);
      const bool in_cur = slot < 64;
      tq[k] = in_cur ? qc : qn;
      tl[k] = in_cur ? lc : ln;
      tk[k] = op_find;
      if constexpr (General) {
        if (a.types) {
          const u32 kc = (u32)__builtin_amdgcn_ds_bpermute(src, (int)ring_kd);
          const u32 kn = (u32)__builtin_amdgcn_ds_bpermute(src, (int)next_ring_kd);
          tk[k] = in_cur ? kc : kn;
        }
      }
      tpos[k] = lo + (in_cur ? cur : nxt) * 64 + (slot & 63u);
      head += (u32)__popcll(b);
    }
    if (head >= 64) {  // the current chunk is used up: the next one becomes current
      head -= 64;
      asm volatile("v_mov_b32 %0, %1" : "=v"(ring_q) : "v"(next_ring_q));
      asm volatile("v_mov_b32 %0, %1" : "=v"(ring_len) : "v"(next_ring_len));
      if constexpr (General) {
        if (a.types) asm volatile("v_mov_b32 %0, %1" : "=v"(ring_kd) : "v"(next_ring_kd));
      }
      cur = nxt;
      nxt += waves;
      ring_load(nxt, next_ring_q, next_ring_len, next_ring_kd);
    }
#pragma unroll
    for (u32 k = 0; k < K; k++) {
      if (!want[k]) continue;
      want[k] = false;
      idx[k] = tpos[k];
      active[k] = tpos[k] < hi;
      klen[k] = tl[k];
      q[k] = tq[k];
      if constexpr (General) {
        req[k] = request_index(a, tpos[k]);
        kind[k] = tk[k];
        lroot[k] = tv.root;
        par[k] = kInvalid;
      }
      node[k] = start_node(tq[k], tk[k]);
      layer[k] = 0;
      q_stale[k] = false;
    }
  };

  if constexpr (Ring) {
    ring_load(cur, ring_q, ring_len, ring_kd);
    ring_load(nxt, next_ring_q, next_ring_len, next_ring_kd);
#pragma unroll
    for (u32 k = 0; k < K; k++) want[k] = true;
    take();
  } else {
#pragma unroll
    for (u32 k = 0; k < K; k++) {
      idx[k] = lo + first + k;
      fetch(k, idx[k]);
      start(k);
    }
  }
  for (u32 guard = 0; guard < (1u << 24); guard++) {
    bool any = false;
#pragma unroll
    for (u32 k = 0; k < K; k++) any |= active[k];
    if constexpr (Ring) {
      if (__ballot(any) == 0) break;  // the ring's lanes stay in the loop together
    } else {
      if (!any) break;
    }
    // node[k] is a pool index here (the root after start; a step that leaves an index
    // outside the pool finishes the lookup), so every tile loads unconditionally.  A
    // lookup that entered a new trie layer loads that layer's slice beside its node.
    node_piece<S> n[K];
#pragma unroll
    for (u32 k = 0; k < K; k++) {
      if constexpr (Ring) {
        if (!active[k]) continue;  // an idle tile of the ring loads nothing
      }
      if (layer[k] != 0 && q_stale[k]) {
        q[k] = a.keys[request(k) * (u64)maxK + layer[k]];
        q_stale[k] = false;
      }
      n[k] = load_piece<T>(pool, node[k], x.j);
    }
#pragma unroll
    for (u32 k = 0; k < K; k++) {
      if (!active[k]) continue;
      const u32 h = x.template lane<T - 1>(slot_lo(n[k].s[S - 1]));    // header (slot 15)
      const u32 sib = x.template lane<T - 1>(slot_hi(n[k].s[S - 1]));  // right sibling
      const u32 nk = hdr_nkeys(h);
      const bool border = hdr_border(h);
      const u32 l = layer[k];
      u32 lt = kNone, cand = 0, high = 0, ksb = 0;
      if (!border) {
#pragma unroll
        for (int t = S - 1; t >= 0; t--) {
          const bool c = slot_lo(n[k].s[t]) >= q[k];
          lt = c ? (u32)t : lt;
          cand = c ? slot_hi(n[k].s[t]) : cand;
        }
      } else {
        high = x.template lane<kHighSlot / S>(slot_lo(n[k].s[kHighSlot % S]));
        ksb = x.template lane<kHighSlot / S>(slot_hi(n[k].s[kHighSlot % S]));
        // the continues-bit of slot s is bit 2s + 1 of the keystates; after the shift,
        // bit 2t is that of this lane's slot t, inverted when the key ends here
        u32 P = ksb >> (2 * slot0 + 1);
        if (!(l + 1 < klen[k])) P = ~P;
#pragma unroll
        for (int t = S - 1; t >= 0; t--) {
          const bool c = slot_lo(n[k].s[t]) == q[k] && ((P >> (2 * t)) & 1u);
          lt = c ? (u32)t : lt;
          cand = c ? slot_hi(n[k].s[t]) : cand;
        }
      }
      const u32 mine = slot0 + lt;
      const u32 gid = x.min(mine);
      const u32 v = x.bor(mine == gid ? cand : 0u);
      const bool found = gid < nk;
      // (General) a write request is walked to its border node when start hints are
      // wanted, else left alone: the pass that runs the writes starts there
      bool write = false;
      if constexpr (General) write = a.types ? kind[k] != op_find : a.kind != op_find;
      u32 len = kInvalid;   // kInvalid: the lookup goes on
      u32 hint = kInvalid;  // write: the border node reached
      if (write && !a.hints) {
        len = kNoOutput;
      } else if (klen[k] == 0 || klen[k] > maxK) {
        if (write) {
          len = kHintDone;  // the write pass reports the length
        } else {
          if (x.j == 0) atomicCAS(tv.err, 0u, (u32)kErrBadLength);
          len = 0;
        }
      } else if (!border) {
        // no bound >= q: q lies right of this node (B-link move right)
        if constexpr (General) {
          if (found && hdr_level(h) == 1) par[k] = node[k];
        }
        node[k] = found ? (v & kChildMask) : sib;
      } else if (q[k] > high && sib != kInvalid) {
        node[k] = sib;
      } else if (found && ks_get(ksb, gid) == kLink) {
        layer[k] = l + 1;  // < key length: the entry continues
        node[k] = v;
        if constexpr (General) {
          lroot[k] = v;
          par[k] = kInvalid;
        }
        q_stale[k] = true;
      } else if (write) {
        hint = node[k];
        len = kHintDone;
      } else if (!found) {
        len = 0;
      } else {
        const u32 ks = ks_get(ksb, gid);
        if (ks == kValue) {
          if (x.j == 0 && maxV >= 1) a.out_values[request(k) * (u64)maxV] = v;
          len = 1;
        } else if (v < nslabs) {
          if (x.j == 0)
            park[idx[k] - lo] = maxK < kParkPacked ? (ks << 30) | (l << 15) | (klen[k] - l - 1)
                                                   : (ks << 16) | l;
          len = kPendingRecord | v;
        } else {
          node[k] = v;  // reported below
        }
      }
      if (len == kInvalid && node[k] >= nslabs) {
        if (x.j == 0) atomicCAS(tv.err, 0u, (u32)kErrBadIndex);
        len = write ? kHintDone : 0u;
      }
      if (len != kInvalid) {
        if (x.j == 0) {
          if constexpr (General) {
            if (len == kHintDone) {
              u32* hp = a.hints + (u64)kHintWords * (idx[k] - lo);
              *reinterpret_cast<uint4*>(hp) = make_uint4(hint, lroot[k], par[k], layer[k]);
            } else if (len != kNoOutput) {
              a.out_value_lengths[request(k)] = len;
            }
          } else {
            a.out_value_lengths[idx[k]] = len;
          }
        }
        if constexpr (Ring) {
          active[k] = false;
          want[k] = true;
        } else {
          idx[k] += stride;
          start(k);
        }
      }
    }
    if constexpr (Ring) take();
  }
}

// Pass 2: every parked lookup reads its record.  A tile takes entries i, i + tiles, ...;
// the output word, parked state and key length of its next entry are loaded one entry
// ahead, and the key-tail words a match compares in the record's first slab are loaded
// beside the record, so a lookup waits for one memory round trip.
template <u32 T>
__global__ void __launch_bounds__(kBlockThreads) narrow_record_kernel(tree_view tv, batch_args a,
                                                                       const u32* park) {
  constexpr u32 S = kSlots / T;  // slots per lane: words 2*S*j .. 2*S*j + 2*S - 1 of a slab
  constexpr u32 W = 2 * S;
  const tile_xchg<T> x;
  const u64* pool = tv.pool.pool;
  const u32 nslabs = tv.pool.num_slabs;
  const u32 tiles_per_block = kBlockThreads / T;
  const u64 ntiles = (u64)gridDim.x * tiles_per_block;
  const u32 maxK = a.max_key_length, maxV = a.max_value_length;
  const u32 w0 = W * x.j;  // this lane's first word of a slab
  u64 lo = 0, hi = a.n;
  if (a.range) {
    hi = min(a.range[1], a.n);
    lo = min(a.range[0], hi);
  }
  if (lo >= hi) return;
  u64 i = lo + (u64)blockIdx.x * tiles_per_block + threadIdx.x / T;
  const bool packed = maxK < kParkPacked;  // park words carry the tail length
  u32 next_w, next_pk, next_len, next_kind = op_find;
  auto fetch = [&](u64 e) {
    const u64 c = e < hi ? e : hi - 1;  // clamped: the loads are unconditional
    next_w = a.out_value_lengths[c];
    next_pk = park[c - lo];
    next_len = (!packed && a.key_lengths) ? a.key_lengths[c] : maxK;
    if (a.types) next_kind = a.types[c];
  };
  fetch(i);
  for (; i < hi; i += ntiles) {
    // copies made before the next entry's loads are issued into the same registers
    u32 w, pk, klen, kd = op_find;
    asm volatile("v_mov_b32 %0, %1" : "=v"(w) : "v"(next_w));
    asm volatile("v_mov_b32 %0, %1" : "=v"(pk) : "v"(next_pk));
    asm volatile("v_mov_b32 %0, %1" : "=v"(klen) : "v"(next_len));
    if (a.types) asm volatile("v_mov_b32 %0, %1" : "=v"(kd) : "v"(next_kind));
    fetch(i + ntiles);
    // (in a mixed batch the length word of a write request is its input value length)
    if (!(w & kPendingRecord) || kd != op_find) continue;
    const u32 rec = w & ~kPendingRecord;  // below the pool size: checked by pass 1
    const u32 rkind = packed ? pk >> 30 : pk >> 16;
    const u32 l = packed ? (pk >> 15) & 0x7FFFu : pk & 0xFFFFu;
    const u32* key = a.keys + i * (u64)maxK;
    u32* out = a.out_values + i * (u64)maxV;
    node_piece<S> cur = load_piece<T>(pool, rec, x.j);
    // a suffix record matches only with a tail of exactly klen - l - 1 slices; the key's
    // tail words for the first slab's data words 1 .. 30 load beside the record
    const u32 want = rkind == kSuffix ? (packed ? pk & 0x7FFFu : klen - l - 1) : 0u;
    u32 kw[W];
#pragma unroll
    for (u32 e = 0; e < W; e++) {
      const u32 d = w0 + e;
      kw[e] = (d >= 1 && d <= want && d < kRecordWordsPerSlab) ? key[l + d] : 0u;
    }
    const u32 lw = x.template lane<0>(slot_lo(cur.s[0]));
    const u32 tl = rec_tail_len(lw), vl = rec_value_len(lw);
    const u32 total = 1 + tl + vl;
    const u32 vcopy = vl < maxV ? vl : maxV;
    bool hit = !(rkind == kSuffix && tl != want);
    // first slab: tail words (compared) precede value words (copied)
    if (hit) {
      bool mis = false;
#pragma unroll
      for (u32 e = 0; e < W; e++) {
        const u32 d = w0 + e;
        const u32 v = (e & 1) ? slot_hi(cur.s[e / 2]) : slot_lo(cur.s[e / 2]);
        if (d >= 1 && d <= tl && d < kRecordWordsPerSlab) mis |= v != kw[e];
      }
      if (x.bor(mis ? 1u : 0u)) hit = false;
    }
    if (hit) {
#pragma unroll
      for (u32 e = 0; e < W; e++) {
        const u32 d = w0 + e;
        const u32 v = (e & 1) ? slot_hi(cur.s[e / 2]) : slot_lo(cur.s[e / 2]);
        if (d < kRecordWordsPerSlab && d >= 1 + tl && d < 1 + tl + vcopy) out[d - 1 - tl] = v;
      }
    }
    // further slabs (records of more than 31 words): key words loaded as compared
    for (u32 d0 = kRecordWordsPerSlab; hit && d0 < total; d0 += kRecordWordsPerSlab) {
      const u32 nxt = x.template lane<T - 1>(slot_hi(cur.s[S - 1]));  // word 31
      if (nxt >= nslabs) {
        if (x.j == 0) atomicCAS(tv.err, 0u, (u32)kErrBadIndex);
        hit = false;
        break;
      }
      cur = load_piece<T>(pool, nxt, x.j);
      bool mis = false;
      if (rkind == kSuffix) {
#pragma unroll
        for (u32 e = 0; e < W; e++) {
          const u32 wd = w0 + e, d = d0 + wd;
          const u32 v = (e & 1) ? slot_hi(cur.s[e / 2]) : slot_lo(cur.s[e / 2]);
          if (wd < kRecordWordsPerSlab && d <= tl) mis |= v != key[l + d];
        }
      }
      if (x.bor(mis ? 1u : 0u)) {
        hit = false;
        break;
      }
#pragma unroll
      for (u32 e = 0; e < W; e++) {
        const u32 wd = w0 + e, d = d0 + wd;
        const u32 v = (e & 1) ? slot_hi(cur.s[e / 2]) : slot_lo(cur.s[e / 2]);
        if (wd < kRecordWordsPerSlab && d >= 1 + tl && d < 1 + tl + vcopy) out[d - 1 - tl] = v;
      }
    }
    if (x.j == 0) a.out_value_lengths[i] = hit ? vl : 0u;
  }
}

}  // namespace indexingpu_amd
