// gfx950 tile policy: the node, record and slab primitives of the Masstree state
// machine (tree_ops.hpp), written for wave64.
//
// A tile is T = 16, 32 or 64 consecutive lanes of a wave.  Lane l holds node slot
// (l & 15); with T > 16 the extra lanes mirror the first 16 (their loads coalesce
// into the same 128-byte line, their stores are skipped).  Searches are one 64-bit
// wave ballot masked to the tile's 16 slot bits plus one ds_bpermute broadcast;
// inserts, removals and splits are lane-parallel shuffles (ds_bpermute) followed by
// one coalesced store of the whole node.
//
// Memory ordering (concurrent batches).  Every access to shared tree memory is a
// relaxed agent-scope atomic (sc1 loads and stores: served coherently across the eight
// XCDs), and ordering between them comes from completion: dev::order() waits for every
// outstanding memory operation of the wave (s_waitcnt vmcnt(0)) behind a compiler
// barrier.  No C++ acquire/release fences are used on these paths: on gfx950 they are
// whole-L2 write-back/invalidate operations.
//   * Validated read: the whole line (header included) in one relaxed atomic load; a
//     locked header fails; order(); the header is loaded again and must be unchanged.
//   * Lock: compare-and-swap of the unlocked header seen by a validated read (relaxed),
//     then order() before any store to the node.
//   * Write + unlock: relaxed atomic stores of the node, order() (the stores are
//     complete), then a relaxed fetch_add that clears the lock bit and advances the
//     version (unlock_bump), or a fetch_and that only clears it (unlock_same).
//   * Interior nodes may come from a plain cached load when unlocked and tagged with
//     the header's version in both halves of the line (layout.hpp).
// Read-only batches (find<concurrent=false>, scan) use plain cached loads, as the
// reference's weak memory order does (include/gpu_masstree.hpp:206); kernel
// boundaries order them against writes.
#pragma once
#include "hip_util.hpp"
#include "layout.hpp"
#include "reclaim.hpp"
#include "slab_alloc.hpp"
#include "tree_ops.hpp"

namespace indexingpu_amd {

struct gpu_node {
  u64 s;  // this lane's slot
};

// Coherent = true when writers may run during the launch: node and record reads then
// use agent-scope atomic (sc1) loads; read-only launches use plain cached loads.
template <u32 T, bool Coherent = true>
struct wave_tile {
  static_assert(T == 16 || T == 32 || T == 64, "tile width must be 16, 32 or 64 lanes");
  using node_t = gpu_node;

  slab_pool_view pool;
  retire_view rv;
  u32* err;
  u32 lane;       // lane within the tile
  u32 slot;       // node slot owned by this lane
  u32 shift;      // position of the tile inside the wave ballot
  u32 key_cache;  // key[lane] of the current request
  u64 rec_slot;   // this lane's 8 bytes of the loaded record head slab
  u32 rec_idx;
  bool fatal;
  tile_slab_cache ac;
  tile_retire_cache rc;

  u32 tree_root;  // root of the top layer (a write restarts there after a removed layer)

  __device__ __forceinline__ wave_tile(const slab_pool_view& p, const retire_view& r, u32* e,
                                       u32 root = kInvalid)
      : pool(p), rv(r), err(e), tree_root(root) {
    const u32 l64 = __lane_id();
    lane = l64 & (T - 1);
    slot = lane & 15u;
    shift = l64 & ~(T - 1);
    key_cache = 0;
    rec_slot = 0;
    rec_idx = kInvalid;
    fatal = false;
  }

  // ---------------- tile collectives ----------------
  __device__ __forceinline__ u32 ballot(bool p) const {
    return (u32)(__ballot(p) >> shift) & 0xFFFFu;
  }
  __device__ __forceinline__ u32 bcast(u32 v, u32 src) const { return __shfl(v, (int)src, (int)T); }
  __device__ __forceinline__ u64 bcast64(u64 v, u32 src) const {
    return __shfl(v, (int)src, (int)T);
  }
  // lane of slot s in this lane's mirror group
  __device__ __forceinline__ u32 src_lane(u32 s) const { return (lane & ~15u) + s; }
  __device__ __forceinline__ u64 slot_from(u64 v, u32 s) const {
    return __shfl(v, (int)src_lane(s), (int)T);
  }

  __device__ __forceinline__ void set_error(u32 code) {
    if (lane == 0) dev::set_error(err, code);
  }
  // set_error that reports (to every lane) whether this call's code was the first
  __device__ __forceinline__ bool set_error_first(u32 code) {
    u32 won = 0;
    if (lane == 0) {
      u32 expected = 0;
      won = __hip_atomic_compare_exchange_strong(err, &expected, code, __ATOMIC_RELAXED,
                                                 __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    return bcast(won, 0) != 0;
  }
  __device__ __forceinline__ u32 max_op_time() const { return kMaxOpTime; }
  __device__ __forceinline__ void count_merge() {
    if (lane == 0) dev::atomic_add(err + kMergeEpochWord, 1u);
  }
  // Border-level shortcut table of the launch (tree_view::shortcut_border; null: none) and
  // its shift.  A merge that removes layer-0 border node `right` (exclusive low bound
  // `low`) rewrites the run of entries naming it, which starts at the first prefix above
  // `low`, to `target` (tree_ops.hpp, ph_merge_lock).  shortcut_prev: the entry of the
  // prefix holding `low` (kInvalid: no table).  shortcut_patch: false when there is no
  // table or the run is longer than kShortcutPatchMax.
  // A table already retired (a merge counted since its build) is not rewritten: its
  // entries may name nodes freed since, so shortcut_prev gives kInvalid for it too.
  u32* sc_border = nullptr;
  const u32* sc_epoch = nullptr;  // the merge count the table was built at
  u32 sc_shift = 0;
  __device__ __forceinline__ u32 slab_count() const { return pool.num_slabs; }
  __device__ __forceinline__ u32 shortcut_prev(u32 low) const {
    if (!sc_border) return kInvalid;
    u32 v = kInvalid;
    if (lane == 0 && dev::coherent_load(err + kMergeEpochWord) == dev::coherent_load(sc_epoch))
      v = dev::coherent_load(sc_border + (low >> sc_shift));
    return bcast(v, 0);
  }
  __device__ __forceinline__ bool shortcut_patch(u32 right, u32 target, u32 low) {
    if (!sc_border) return false;
    u32 ok = 0;
    if (lane == 0) {
      const u32 entries = 1u << (32u - sc_shift);
      u32 i = (low >> sc_shift) + 1u;  // entries at or below low name `left` or nodes left of it
      u32 n = 0;
      while (i < entries && n < kShortcutPatchMax && dev::coherent_load(sc_border + i) == right) {
        dev::coherent_store(sc_border + i, target);
        i++;
        n++;
      }
      ok = (i >= entries || dev::coherent_load(sc_border + i) != right) ? 1u : 0u;
    }
    return bcast(ok, 0) != 0;
  }
  // the guard's trip record, kept after the error word (layout.hpp)
  __device__ __forceinline__ void store_trip(const u32* w) {
    if (lane == 0)
      for (u32 k = 0; k < kTripWords; k++) dev::coherent_store(err + kTripOffset + k, w[k]);
  }
  __device__ __forceinline__ u32 peek_header(u32 idx) const {
    return idx < pool.num_slabs ? bcast(dev::coherent_load(hdr_ptr(idx)), 0) : kInvalid;
  }
  __device__ __forceinline__ bool failed() const { return fatal; }
  // coarse device time for the per-operation time guard: the 100 MHz wall clock in
  // units of 1024 ticks (about 10 us)
  __device__ __forceinline__ u32 coarse_clock() const { return (u32)(wall_clock64() >> 10); }
  __device__ __forceinline__ bool check(u32 idx) {
    if (idx < pool.num_slabs) return true;
    set_error(kErrBadIndex);
    fatal = true;
    return false;
  }
  __device__ __forceinline__ u64* slab(u32 idx) const { return pool.pool + (u64)idx * kSlots; }
  __device__ __forceinline__ u64 ld(const u64* p) const {
    if (Coherent) return dev::coherent_load(p);
    return *p;
  }
  __device__ __forceinline__ u32* hdr_ptr(u32 idx) const {
    return reinterpret_cast<u32*>(slab(idx) + kMetaSlot);
  }

  // ---------------- node reads / locking ----------------
  // cached_ok: the node is expected to be interior; it is first read with a plain cached
  // load and accepted if unlocked with a matching version tag (see layout.hpp), which
  // keeps the hot upper levels of the tree in L1/L2 instead of one memory channel.
  template <bool Validate>
  __device__ __forceinline__ bool read_node(u32 idx, node_t& n, bool cached_ok = false) {
    if (!check(idx)) return false;
    const u64* p = slab(idx) + slot;
    if (!Validate) {
      n.s = *p;
      return true;
    }
    if (cached_ok) {
      n.s = *p;
      const u32 h = bcast(slot_lo(n.s), kMetaSlot);
      const u32 tag = ballot(slot < kTagSlots && (slot_hi(n.s) >> 31) != 0) & 0xFFu;
      if (!hdr_border(h) && !hdr_locked(h) && tag == hdr_tag(h)) return true;
    }
    // seqlock read: the whole line (header included) in one load, then the header again
    n.s = dev::coherent_load(p);
    const u32 h = bcast(slot_lo(n.s), kMetaSlot);
    if (hdr_locked(h)) return false;
    dev::order();
    const u32 h2 = dev::coherent_load(hdr_ptr(idx));
    return h2 == h;
  }
  __device__ __forceinline__ void read_node_locked(u32 idx, node_t& n) {
    n.s = dev::coherent_load(slab(idx) + slot);
  }
  __device__ __forceinline__ bool try_lock(u32 idx, u32& old) {
    if (!check(idx)) return false;
    // test-and-test-and-set: a locked header is seen with a load, so waiting tiles do
    // not queue atomics in front of the holder's unlock
    u32 o = kLockBit;
    if (lane == 0) {
      o = dev::coherent_load(hdr_ptr(idx));
      if (!hdr_locked(o)) o = dev::atomic_or(hdr_ptr(idx), kLockBit);
    }
    old = bcast(o, 0);
    if (hdr_locked(old)) return false;
    dev::order();
    return true;
  }
  // Locks node idx only if its header still equals `expect` (the unlocked header of a
  // validated read): kLockOk, kLockBusy (locked by another tile) or kLockStale (the
  // node changed: re-read it).  On a retry the compare-and-swap is attempted only when
  // a plain coherent load shows the expected header, so waiters do not queue atomics;
  // `direct` (the header was validated by the step just before) skips that load.
  __device__ __forceinline__ u32 try_lock_snapshot(u32 idx, u32 expect, bool direct = false) {
    if (!check(idx)) return kLockStale;
    u32 r = kLockStale;
    if (lane == 0) {
      u32* hp = hdr_ptr(idx);
      u32 h = direct ? expect : dev::coherent_load(hp);
      if (h == expect) {
        u32 e = expect;
        if (__hip_atomic_compare_exchange_strong(hp, &e, expect | kLockBit, __ATOMIC_RELAXED,
                                                 __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT))
          r = kLockOk;
        else
          h = e;
      }
      if (r != kLockOk) r = hdr_locked(h) ? kLockBusy : kLockStale;
    }
    r = bcast(r, 0);
    if (r == kLockOk) dev::order();
    return r;
  }
  __device__ __forceinline__ void unlock_same(u32 idx) {
    dev::order();
    if (lane == 0) dev::atomic_and(hdr_ptr(idx), ~kLockBit);
  }
  __device__ __forceinline__ void unlock_bump(u32 idx) {
    dev::order();  // the node's stores complete before the unlock publishes them
    if (lane == 0) dev::atomic_add(hdr_ptr(idx), kVersionOne - kLockBit);
  }
  __device__ __forceinline__ void write_node(u32 idx, const node_t& n) {
    const u32 h = bcast(slot_lo(n.s), kMetaSlot);
    u64 s = n.s;
    if (!hdr_border(h) && slot < kTagSlots) {
      const u64 bit = (u64)((hdr_publish_tag(h) >> slot) & 1u) << 63;
      s = (s & ~(1ull << 63)) | bit;
    }
    if (lane < 16) dev::coherent_store(slab(idx) + slot, s);
  }
  __device__ __forceinline__ void fence_release() { dev::order(); }

  // ---------------- node queries ----------------
  __device__ __forceinline__ u32 header(const node_t& n) const {
    return bcast(slot_lo(n.s), kMetaSlot);
  }
  __device__ __forceinline__ u32 nkeys_of(const node_t& n) const { return hdr_nkeys(header(n)); }
  __device__ __forceinline__ u32 sibling(const node_t& n) const {
    return bcast(slot_hi(n.s), kMetaSlot);
  }
  __device__ __forceinline__ u32 high(const node_t& n) const { return bcast(slot_lo(n.s), kHighSlot); }
  __device__ __forceinline__ u32 ksbits(const node_t& n) const {
    return bcast(slot_hi(n.s), kHighSlot);
  }
  __device__ __forceinline__ void set_header(node_t& n, u32 h) const {
    if (slot == kMetaSlot) n.s = slot_make(h, slot_hi(n.s));
  }
  __device__ __forceinline__ void set_sibling(node_t& n, u32 s) const {
    if (slot == kMetaSlot) n.s = slot_make(slot_lo(n.s), s);
  }
  __device__ __forceinline__ void set_high(node_t& n, u32 hk) const {
    if (slot == kHighSlot) n.s = slot_make(hk, slot_hi(n.s));
  }

  __device__ __forceinline__ u32 find_child(const node_t& n, u32 q) const {
    const u32 nk = nkeys_of(n);
    const u32 m = ballot(slot < nk && q <= slot_lo(n.s));
    const u32 child = bcast(slot_hi(n.s), m ? (u32)__builtin_ctz(m) : 0u) & kChildMask;
    return m ? child : kInvalid;
  }
  // value / object index of border entry pos
  __device__ __forceinline__ u32 entry_value(const node_t& n, u32 pos) const {
    return bcast(slot_hi(n.s), pos);
  }
  // slot j of the first bound >= q and its child; false when q is above every bound
  __device__ __forceinline__ bool find_cover(const node_t& n, u32 q, u32& j, u32& child) const {
    const u32 nk = nkeys_of(n);
    const u32 m = ballot(slot < nk && q <= slot_lo(n.s));
    j = m ? (u32)__builtin_ctz(m) : 0u;
    child = bcast(slot_hi(n.s), j) & kChildMask;
    return m != 0;
  }
  __device__ __forceinline__ bool find_child_slot(const node_t& n, u32 child, u32& j) const {
    const u32 nk = nkeys_of(n);
    const u32 m = ballot(slot < nk && child_of(n.s) == child);
    j = m ? (u32)__builtin_ctz(m) : 0u;
    return m != 0;
  }
  __device__ __forceinline__ int lookup(const node_t& n, u32 q, bool more, u32& ks,
                                        u32& val) const {
    const u32 nk = nkeys_of(n);
    const u32 b = ksbits(n);
    const u32 myks = ks_get(b, slot);
    const u32 m = ballot(slot < nk && slot_lo(n.s) == q && ks_more(myks) == more);
    const u32 pos = m ? (u32)__builtin_ctz(m) : 0u;
    const u32 v = bcast(slot_hi(n.s), pos);
    if (!m) return -1;
    ks = ks_get(b, pos);
    val = v;
    return (int)pos;
  }
  __device__ __forceinline__ u32 insert_position(const node_t& n, u32 q, bool more) const {
    const u32 nk = nkeys_of(n);
    const u32 myks = ks_get(ksbits(n), slot);
    return (u32)__builtin_popcount(
        ballot(slot < nk && entry_less(slot_lo(n.s), ks_more(myks), q, more)));
  }

  // ---------------- node edits ----------------
  __device__ __forceinline__ void border_insert(node_t& n, u32 pos, u32 q, u32 ks, u32 val) const {
    const u32 h = header(n), b = ksbits(n), nk = hdr_nkeys(h);
    const u64 prev = slot_from(n.s, slot > 0 ? slot - 1 : 0);
    const u32 lowmask = pos == 0 ? 0u : ((1u << (2 * pos)) - 1u);
    const u32 nb = (b & lowmask) | ((b & ~lowmask) << 2) | (ks << (2 * pos));
    if (slot < kBorderCap) {
      if (slot == pos) n.s = slot_make(q, val);
      else if (slot > pos && slot <= nk) n.s = prev;
    } else if (slot == kHighSlot) {
      n.s = slot_make(slot_lo(n.s), nb);
    } else {
      n.s = slot_make(hdr_with_nkeys(h, nk + 1), slot_hi(n.s));
    }
  }
  __device__ __forceinline__ void border_remove(node_t& n, u32 pos) const {
    const u32 h = header(n), b = ksbits(n), nk = hdr_nkeys(h);
    const u64 next = slot_from(n.s, slot < 15 ? slot + 1 : 15);
    const u32 lowmask = pos == 0 ? 0u : ((1u << (2 * pos)) - 1u);
    const u32 highpart = (b >> (2 * (pos + 1))) << (2 * pos);
    if (slot < kBorderCap) {
      if (slot >= pos) {
        if (slot + 1 < nk) n.s = next;
        else if (slot + 1 == nk) n.s = 0;
      }
    } else if (slot == kHighSlot) {
      n.s = slot_make(slot_lo(n.s), (b & lowmask) | highpart);
    } else {
      n.s = slot_make(hdr_with_nkeys(h, nk - 1), slot_hi(n.s));
    }
  }
  __device__ __forceinline__ void border_set(node_t& n, u32 pos, u32 ks, u32 val) const {
    const u32 b = ksbits(n);
    if (slot == pos) n.s = slot_make(slot_lo(n.s), val);
    if (slot == kHighSlot)
      n.s = slot_make(slot_lo(n.s), (b & ~(3u << (2 * pos))) | (ks << (2 * pos)));
  }
  // Splits n's entries plus the pending (q, ks, val) at ipos into two nodes; returns
  // the separator (largest slice of the left node).  Entries with one slice stay
  // together.  Sets only nkeys in the result headers.
  __device__ __forceinline__ u32 border_split(const node_t& n, u32 ipos, u32 q, u32 ks, u32 val,
                                              node_t& left, node_t& right) const {
    const u32 b = ksbits(n), nk = nkeys_of(n), total = nk + 1;
    u32 m = (total + 1) / 2;
    {
      const u32 i0 = m - 1, i1 = m;
      const u32 s0 = bcast(slot_lo(n.s), i0 < ipos ? i0 : (i0 > ipos ? i0 - 1 : 0));
      const u32 s1 = bcast(slot_lo(n.s), i1 < ipos ? i1 : (i1 > ipos ? i1 - 1 : 0));
      const u32 v0 = (i0 == ipos) ? q : s0;
      const u32 v1 = (i1 == ipos) ? q : s1;
      if (v0 == v1) m -= 1;
    }
    const u32 li = slot, ri = m + slot;
    const u32 lsrc = li < ipos ? li : (li > ipos ? li - 1 : 0);
    u32 rsrc = ri < ipos ? ri : (ri > ipos ? ri - 1 : 0);
    rsrc = rsrc > 15 ? 15 : rsrc;
    u64 lv = slot_from(n.s, lsrc);
    u64 rv = slot_from(n.s, rsrc);
    if (li == ipos) lv = slot_make(q, val);
    if (ri == ipos) rv = slot_make(q, val);
    u32 lb = 0, rb = 0;
    for (u32 i = 0; i < total; i++) {
      const u32 k = i < ipos ? ks_get(b, i) : (i == ipos ? ks : ks_get(b, i - 1));
      if (i < m) lb |= k << (2 * i);
      else rb |= k << (2 * (i - m));
    }
    left.s = (slot < m) ? lv : 0;
    right.s = (slot < total - m) ? rv : 0;
    if (slot == kHighSlot) {
      left.s = slot_make(0, lb);
      right.s = slot_make(0, rb);
    } else if (slot == kMetaSlot) {
      left.s = slot_make(hdr_with_nkeys(0, m), kInvalid);
      right.s = slot_make(hdr_with_nkeys(0, total - m), kInvalid);
    }
    const u32 i0 = m - 1;
    const u32 s0 = bcast(slot_lo(n.s), i0 < ipos ? i0 : (i0 > ipos ? i0 - 1 : 0));
    return (i0 == ipos) ? q : s0;
  }
  __device__ __forceinline__ void make_interior_root(node_t& n, u32 sep, u32 l, u32 r, u32 level,
                                                     u32 h) const {
    const u32 keep = h & (kLockBit | ~((1u << kVersionShift) - 1u));
    u64 v = 0;
    if (slot == 0) v = slot_make(sep, l);
    if (slot == 1) v = slot_make(kMaxSlice, r);
    if (slot == kMetaSlot) v = slot_make(keep | hdr_make(2, false, true, level), kInvalid);
    n.s = v;
  }
  __device__ __forceinline__ void make_single_root(node_t& x, u32 t0, u32 ks, u32 val) const {
    u64 v = 0;
    if (slot == 0) v = slot_make(t0, val);
    if (slot == kHighSlot) v = slot_make(kMaxSlice, ks);
    if (slot == kMetaSlot) v = slot_make(hdr_make(1, true, true, 0), kInvalid);
    x.s = v;
  }
  __device__ __forceinline__ void interior_insert(node_t& n, u32 j, u32 sep, u32 lc, u32 rc) const {
    const u32 h = header(n), nk = hdr_nkeys(h);
    const u32 old_bound = bcast(slot_lo(n.s), j);
    const u64 prev = slot_from(n.s, slot > 0 ? slot - 1 : 0);
    if (slot < kInteriorCap) {
      if (slot == j) n.s = slot_make(sep, lc);
      else if (slot == j + 1) n.s = slot_make(old_bound, rc);
      else if (slot > j + 1 && slot <= nk) n.s = prev;
    } else {
      n.s = slot_make(hdr_with_nkeys(h, nk + 1), slot_hi(n.s));
    }
  }
  // removes child slot j (j > 0): slot j - 1 keeps its child and takes slot j's bound
  __device__ __forceinline__ void interior_remove(node_t& n, u32 j) const {
    const u32 h = header(n), nk = hdr_nkeys(h);
    const u32 bj = bcast(slot_lo(n.s), j);
    const u64 next = slot_from(n.s, slot + 1 < kSlots ? slot + 1 : kSlots - 1);
    if (slot < kInteriorCap) {
      if (slot == j - 1) n.s = slot_make(bj, slot_hi(n.s));
      else if (slot >= j && slot + 1 < nk) n.s = next;
      else if (slot >= j && slot < nk) n.s = 0;
    } else {
      n.s = slot_make(hdr_with_nkeys(h, nk - 1), slot_hi(n.s));
    }
  }
  __device__ __forceinline__ u32 entry_child(const node_t& n, u32 j) const {
    return bcast(slot_hi(n.s), j) & kChildMask;
  }
  __device__ __forceinline__ u32 interior_split(const node_t& n, u32 j, u32 sep, u32 lc, u32 rc,
                                                node_t& left, node_t& right) const {
    const u32 nk = nkeys_of(n), total = nk + 1, m = total / 2;
    const u32 bj = bcast(slot_lo(n.s), j);
    const u32 li = slot, ri = m + slot;
    auto src = [&](u32 i) -> u32 {
      u32 s = i < j ? i : (i > j + 1 ? i - 1 : 0);
      return s > 15 ? 15 : s;
    };
    u64 lv = slot_from(n.s, src(li));
    u64 rv = slot_from(n.s, src(ri));
    if (li == j) lv = slot_make(sep, lc);
    if (li == j + 1) lv = slot_make(bj, rc);
    if (ri == j) rv = slot_make(sep, lc);
    if (ri == j + 1) rv = slot_make(bj, rc);
    left.s = (slot < m) ? lv : 0;
    right.s = (slot < total - m) ? rv : 0;
    if (slot == kMetaSlot) {
      left.s = slot_make(hdr_with_nkeys(0, m), kInvalid);
      right.s = slot_make(hdr_with_nkeys(0, total - m), kInvalid);
    }
    const u32 i0 = m - 1;
    const u32 s0 = bcast(slot_lo(n.s), i0 < j ? i0 : (i0 > j + 1 ? i0 - 1 : 0));
    return i0 == j ? sep : (i0 == j + 1 ? bj : s0);
  }

  // ---------------- keys and values ----------------
  __device__ __forceinline__ void load_key_cache(const request& rq) {
    key_cache = (lane < rq.key_length) ? rq.key[lane] : 0u;
  }
  __device__ __forceinline__ u32 key_slice(const request& rq, u32 l) const {
    return l < T ? bcast(key_cache, l) : rq.key[l];
  }
  // per-lane gather of key slice k (k < key length); every lane must call it
  __device__ __forceinline__ u32 key_gather(const request& rq, u32 k) const {
    const u32 c = __shfl(key_cache, (int)(k & (T - 1)), (int)T);
    return k < T ? c : rq.key[k];
  }
  __device__ __forceinline__ u32 value_slice(const request& rq, u32 i) const { return rq.value[i]; }
  __device__ __forceinline__ void write_inline_value(const request& rq, u32 v) const {
    if (lane == 0 && rq.max_value_length >= 1) rq.out_value[0] = v;
  }

  // ---------------- slabs ----------------
  __device__ __forceinline__ u32 alloc_slab() {
    if (ac.bits == 0) {
      u32 w, b;
      const bool ok = dev::claim_word(pool, lane == 0, w, b);
      w = bcast(w, 0);
      b = bcast(b, 0);
      const u32 okv = bcast((u32)ok, 0);
      if (!okv) {
        set_error(kErrOutOfMemory);
        return kInvalid;
      }
      ac.word = w;
      ac.bits = b;
    }
    const u32 j = (u32)__builtin_ctz(ac.bits);
    ac.bits &= ac.bits - 1;
    return ac.word * 32u + j;
  }
  __device__ __forceinline__ void flush() {
    if (ac.bits != 0 && lane == 0) dev::atomic_and(&pool.bitmap[ac.word], ~ac.bits);
    ac.bits = 0;
    if (rc.base != kInvalid) {
      const u32 pad = kRetireChunk - rc.used;
      if (lane < pad) rv.list[rc.base + rc.used + lane] = kInvalid;
      rc.base = kInvalid;
    }
  }
  __device__ __forceinline__ void retire(u32 idx) {
    if (!rv.enabled) return;
    if (rv.bag) {  // device API inside a critical section: this block's current bag
      u32 ok = 0;
      if (lane == 0) {
        const u32 cur = rv.bag[kBagCurrent];
        const u32 k = dev::atomic_add(&rv.bag[cur], 1u);
        if (k < rv.bag_cap) {
          rv.bag[kBagHeader + cur * rv.bag_cap + k] = idx;
          ok = 1;
        }
      }
      if (bcast(ok, 0)) return;  // a full bag falls back to the launch list
    }
    if (rc.base == kInvalid || rc.used == kRetireChunk) {
      u32 b = 0;
      if (lane == 0) b = dev::atomic_add(rv.count, kRetireChunk);
      b = bcast(b, 0);
      if (b + kRetireChunk > rv.capacity) {
        if (lane == 0) dev::atomic_add(rv.leaked, 1u);
        rc.base = kInvalid;
        return;
      }
      rc.base = b;
      rc.used = 0;
    }
    if (lane == 0) rv.list[rc.base + rc.used] = idx;
    rc.used += 1;
  }

  // ---------------- records ----------------
  __device__ __forceinline__ bool rec_load(u32 idx, u32& len_word) {
    if (!check(idx)) return false;
    rec_idx = idx;
    rec_slot = ld(slab(idx) + slot);
    len_word = bcast(slot_lo(rec_slot), 0);
    return true;
  }
  __device__ __forceinline__ bool rec_next(u64& cur) {
    const u32 nxt = bcast(slot_hi(cur), kMetaSlot);
    if (!check(nxt)) return false;
    cur = ld(slab(nxt) + slot);
    return true;
  }
  // compares the record tail with key[off, off + len); the record head is loaded
  __device__ __forceinline__ bool rec_tail_equal(const u32* key, u32 off, u32 len, u32 len_word) {
    if (rec_tail_len(len_word) != len) return false;
    const u32 need = 1 + len;
    request kr;
    kr.key = key;
    u64 cur = rec_slot;
    for (u32 base = 0;; base += kRecordWordsPerSlab) {
      const u32 w0 = 2 * slot, w1 = 2 * slot + 1;
      const u32 d0 = base + w0, d1 = base + w1;
      const u32 k0 = off + d0 - 1, k1 = off + d1 - 1;
      const bool c0 = d0 >= 1 && d0 < need;
      const bool c1 = w1 < kRecordWordsPerSlab && d1 >= 1 && d1 < need;
      const u32 g0 = key_gather(kr, c0 ? k0 : 0u);
      const u32 g1 = key_gather(kr, c1 ? k1 : 0u);
      const bool mis = (c0 && g0 != slot_lo(cur)) || (c1 && g1 != slot_hi(cur));
      if (ballot(mis)) return false;
      if (base + kRecordWordsPerSlab >= need) return true;
      if (!rec_next(cur)) return false;
    }
  }
  __device__ __forceinline__ u32 rec_copy_value(const request& rq, u32 len_word) {
    const u32 tl = rec_tail_len(len_word), vl = rec_value_len(len_word);
    const u32 n = vl < rq.max_value_length ? vl : rq.max_value_length;
    const u32 first = 1 + tl, end = first + n;
    u64 cur = rec_slot;
    for (u32 base = 0; base < end; base += kRecordWordsPerSlab) {
      if (base + kRecordWordsPerSlab > first && lane < 16) {
        const u32 d0 = base + 2 * slot, d1 = d0 + 1;
        if (d0 >= first && d0 < end) rq.out_value[d0 - first] = slot_lo(cur);
        if (2 * slot + 1 < kRecordWordsPerSlab && d1 >= first && d1 < end)
          rq.out_value[d1 - first] = slot_hi(cur);
      }
      if (base + kRecordWordsPerSlab >= end) break;
      if (!rec_next(cur)) break;
    }
    return vl;
  }
  // Frees the slabs of a record chain whose allocation failed part way: the slabs from
  // `head` up to (not including) `cur` are written and linked, `cur` is allocated but
  // not written.  Nothing else has seen them.
  __device__ __forceinline__ void rec_free_partial(u32 head, u32 cur) {
    dev::order();  // this tile's stores of the chain are complete before it is read back
    u32 idx = head;
    while (idx != cur) {
      const u32 nxt = bcast(slot_hi(dev::coherent_load(slab(idx) + slot)), kMetaSlot);
      if (lane == 0) dev::free_slab(pool, idx);
      idx = nxt;
    }
    if (lane == 0) dev::free_slab(pool, cur);
  }
  // writes a new record whose data words come from word(d) (lane-parallel)
  template <class WordFn>
  __device__ __forceinline__ u32 rec_write(u32 total, WordFn word) {
    const u32 nslabs = rec_num_slabs(total);
    u32 cur = alloc_slab();
    if (cur == kInvalid) return kInvalid;
    const u32 head = cur;
    for (u32 s = 0; s < nslabs; s++) {
      u32 nxt = kInvalid;
      if (s + 1 < nslabs) {
        nxt = alloc_slab();
        if (nxt == kInvalid) {
          rec_free_partial(head, cur);
          return kInvalid;
        }
      }
      const u32 d0 = s * kRecordWordsPerSlab + 2 * slot, d1 = d0 + 1;
      const u32 v0 = word(d0, s, 0);
      const u32 v1 = (2 * slot + 1 == kRecordNextWord) ? nxt : word(d1, s, 1);
      if (lane < 16) dev::coherent_store(slab(cur) + slot, slot_make(v0, v1));
      cur = nxt;
    }
    return head;
  }
  __device__ __forceinline__ u32 rec_create(const u32* key, u32 off, u32 tail_len, const u32* value,
                                            u32 vlen) {
    if (tail_len > kMaxLength || vlen > kMaxLength) {
      set_error(kErrBadLength);
      return kInvalid;
    }
    const u32 total = 1 + tail_len + vlen;
    const u32 lw = rec_len_word(tail_len, vlen);
    return rec_write(total, [&](u32 d, u32, u32) -> u32 {
      if (d == 0) return lw;
      if (d <= tail_len) return key[off + d - 1];
      if (d < total) return value[d - 1 - tail_len];
      return 0u;
    });
  }
  // new record = old record's data words shifted down by one (drops tail[0]);
  // the old record head is loaded
  __device__ __forceinline__ u32 rec_clone_shift(u32 new_len_word, u32 count, u32 old_total) {
    u64 oa = rec_slot;
    u64 ob = 0;
    bool have_b = old_total > kRecordWordsPerSlab;
    if (have_b) {
      ob = oa;
      if (!rec_next(ob)) return kInvalid;
    }
    const u32 total = 1 + count;
    const u32 nslabs = rec_num_slabs(total);
    u32 cur = alloc_slab();
    if (cur == kInvalid) return kInvalid;
    const u32 head = cur;
    for (u32 s = 0; s < nslabs; s++) {
      u32 nxt = kInvalid;
      if (s + 1 < nslabs) {
        nxt = alloc_slab();
        if (nxt == kInvalid) {
          rec_free_partial(head, cur);
          return kInvalid;
        }
      }
      // new word d' = 31 s + w takes old word d' + 1
      const u64 right = slot_from(oa, slot < 15 ? slot + 1 : 15);
      const u64 b0 = slot_from(ob, 0);
      const u32 d0 = s * kRecordWordsPerSlab + 2 * slot, d1 = d0 + 1;
      u32 v0 = (slot < 15) ? slot_hi(oa) : slot_lo(b0);
      u32 v1 = (slot < 15) ? slot_lo(right) : nxt;
      if (d0 == 0) v0 = new_len_word;
      else if (d0 >= total) v0 = 0;
      if (slot < 15 && d1 >= total) v1 = 0;
      if (lane < 16) dev::coherent_store(slab(cur) + slot, slot_make(v0, v1));
      cur = nxt;
      // advance the old window by one slab
      oa = ob;
      const bool more_old = (s + 2) * kRecordWordsPerSlab < old_total;
      if (more_old) {
        if (!rec_next(ob)) return kInvalid;
      } else {
        ob = 0;
      }
    }
    return head;
  }
  __device__ __forceinline__ bool rec_push_down(u32 rec, u32 len_word, u32& t0, u32& eks, u32& eval) {
    (void)rec;
    const u32 tl = rec_tail_len(len_word), vl = rec_value_len(len_word);
    t0 = bcast(slot_hi(rec_slot), 0);  // data word 1
    const u32 old_total = 1 + tl + vl;
    if (tl == 1 && vl == 1) {
      eks = kValue;
      eval = bcast(slot_lo(rec_slot), 1);  // data word 2
      return true;
    }
    if (tl == 1) {
      eks = kLongVal;
      eval = rec_clone_shift(rec_len_word(0, vl), vl, old_total);
    } else {
      eks = kSuffix;
      eval = rec_clone_shift(rec_len_word(tl - 1, vl), tl - 1 + vl, old_total);
    }
    return eval != kInvalid;
  }
};

}  // namespace indexingpu_amd
