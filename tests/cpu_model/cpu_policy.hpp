// CPU tile policy for the Masstree state machine (test infrastructure only).
//
// Implements the node/record/allocation primitives that tree_ops.hpp needs with plain
// loops over the 16 slots of a node, on the same 128-byte slab layout the gfx950
// kernels use.  Racy accesses use GCC atomics so that several host threads (each one
// acting as a tile) can run operations concurrently against one pool.
#pragma once
#include <atomic>
#include <chrono>
#include <cstdint>
#include <cstring>
#include <mutex>
#include <vector>

#include "indexingpu_amd/layout.hpp"
#include "indexingpu_amd/tree_ops.hpp"

namespace cpu_model {
using namespace indexingpu_amd;

struct node_img {
  u64 s[kSlots];
};

struct shared_pool {
  std::vector<u64> slots;  // num_slabs * 16
  u32 num_slabs = 0;
  std::atomic<u32> next{0};      // bump allocation (the model never reuses slabs)
  std::atomic<u32> error{0};
  // injected allocation failures (0 = off): every k-th allocation fails, or every k-th
  // allocation made while a split is inserting into its parent
  u32 fail_any_every = 0, fail_parent_every = 0;
  std::atomic<u32> any_allocs{0}, parent_allocs{0};
  std::mutex retire_mu;
  std::vector<u32> retired;
  u32 max_op_time = kMaxOpTime;  // the livelock guard's time limit (coarse ticks)
  std::atomic<u32> merges{0};     // layer-0 border nodes merged away without a table rewrite
  u32* sc_border = nullptr;       // border-level shortcut table (null: none) and its shift
  u32 sc_shift = 0;
  u32 sc_built = 0;               // `merges` when the table was built
  u32 trip[kTripWords] = {};     // the guard's record of the operation it abandoned
  explicit shared_pool(u32 n) : slots((size_t)n * kSlots, 0), num_slabs(n) {}
  u64* slab(u32 idx) { return &slots[(size_t)idx * kSlots]; }
};

struct cpu_tile {
  using node_t = node_img;
  shared_pool& P;
  // record head cache (first slab of the record loaded by rec_load)
  u32 rec_idx = kInvalid;

  u32 tree_root = kInvalid;  // root of the top layer (writes restart there)
  explicit cpu_tile(shared_pool& p, u32 root = kInvalid) : P(p), tree_root(root) {}

  void set_error(u32 code) {
    u32 z = 0;
    P.error.compare_exchange_strong(z, code);
  }
  bool set_error_first(u32 code) {
    u32 z = 0;
    return P.error.compare_exchange_strong(z, code);
  }
  u32 max_op_time() const { return P.max_op_time; }
  void count_merge() { P.merges.fetch_add(1); }
  // the device policy's shortcut_prev / shortcut_patch on the model's table
  u32 slab_count() const { return P.num_slabs; }
  u32 shortcut_prev(u32 low) const {
    if (!P.sc_border || P.merges.load() != P.sc_built) return kInvalid;
    return __atomic_load_n(P.sc_border + (low >> P.sc_shift), __ATOMIC_RELAXED);
  }
  bool shortcut_patch(u32 right, u32 target, u32 low) {
    if (!P.sc_border) return false;
    const u32 entries = 1u << (32u - P.sc_shift);
    u32 i = (low >> P.sc_shift) + 1u;
    u32 n = 0;
    while (i < entries && n < kShortcutPatchMax &&
           __atomic_load_n(P.sc_border + i, __ATOMIC_RELAXED) == right) {
      __atomic_store_n(P.sc_border + i, target, __ATOMIC_RELAXED);
      i++;
      n++;
    }
    return i >= entries || __atomic_load_n(P.sc_border + i, __ATOMIC_RELAXED) != right;
  }
  void store_trip(const u32* w) {
    for (u32 k = 0; k < kTripWords; k++) P.trip[k] = w[k];
  }
  u32 peek_header(u32 idx) {
    return idx < P.num_slabs ? __atomic_load_n(hdr_ptr(idx), __ATOMIC_RELAXED) : kInvalid;
  }
  bool fatal = false;
  bool failed() const { return fatal; }
  // same unit as the device policy: 10.24 us
  u32 coarse_clock() const {
    const auto ns = std::chrono::duration_cast<std::chrono::nanoseconds>(
                        std::chrono::steady_clock::now().time_since_epoch()).count();
    return (u32)((uint64_t)ns / 10240u);
  }
  bool check(u32 idx) {
    if (idx >= P.num_slabs) {
      set_error(kErrBadIndex);
      fatal = true;
      return false;
    }
    return true;
  }
  static u64 ld(const u64* p) { return __atomic_load_n(p, __ATOMIC_RELAXED); }
  static void st(u64* p, u64 v) { __atomic_store_n(p, v, __ATOMIC_RELAXED); }
  u32* hdr_ptr(u32 idx) { return reinterpret_cast<u32*>(P.slab(idx) + kMetaSlot); }

  // ---------------- node reads / locking ----------------
  template <bool Validate>
  bool read_node(u32 idx, node_t& n, bool cached_ok = false) {
    (void)cached_ok;  // host caches are coherent: every read is the validated one
    if (!check(idx)) return false;
    const u64* p = P.slab(idx);
    if (!Validate) {
      for (u32 i = 0; i < kSlots; i++) n.s[i] = ld(p + i);
      return true;
    }
    // seqlock read: header, then the slots, then the header again
    n.s[kMetaSlot] = __atomic_load_n(p + kMetaSlot, __ATOMIC_ACQUIRE);
    u32 h = slot_lo(n.s[kMetaSlot]);
    if (hdr_locked(h)) return false;
    for (u32 i = 0; i < kMetaSlot; i++) n.s[i] = ld(p + i);
    __atomic_thread_fence(__ATOMIC_ACQUIRE);
    u32 h2 = __atomic_load_n(hdr_ptr(idx), __ATOMIC_RELAXED);
    return h2 == h;
  }
  void read_node_locked(u32 idx, node_t& n) {
    const u64* p = P.slab(idx);
    for (u32 i = 0; i < kSlots; i++) n.s[i] = ld(p + i);
  }
  bool try_lock(u32 idx, u32& old) {
    if (!check(idx)) return false;
    old = __atomic_fetch_or(hdr_ptr(idx), kLockBit, __ATOMIC_ACQUIRE);
    if (hdr_locked(old)) return false;
    __atomic_thread_fence(__ATOMIC_RELEASE);
    return true;
  }
  u32 try_lock_snapshot(u32 idx, u32 expect, bool direct = false) {
    if (!check(idx)) return kLockStale;
    u32 h = direct ? expect : __atomic_load_n(hdr_ptr(idx), __ATOMIC_ACQUIRE);
    if (h == expect) {
      u32 e = expect;
      if (__atomic_compare_exchange_n(hdr_ptr(idx), &e, expect | kLockBit, false,
                                      __ATOMIC_ACQUIRE, __ATOMIC_RELAXED)) {
        __atomic_thread_fence(__ATOMIC_RELEASE);
        return kLockOk;
      }
      h = e;
    }
    return hdr_locked(h) ? kLockBusy : kLockStale;
  }
  void unlock_same(u32 idx) { __atomic_fetch_and(hdr_ptr(idx), ~kLockBit, __ATOMIC_RELEASE); }
  void unlock_bump(u32 idx) {
    __atomic_fetch_add(hdr_ptr(idx), kVersionOne - kLockBit, __ATOMIC_RELEASE);
  }
  void write_node(u32 idx, const node_t& n) {
    u64* p = P.slab(idx);
    const u32 h = header(n);
    for (u32 i = 0; i < kSlots; i++) {
      u64 s = n.s[i];
      if (!hdr_border(h) && i < kTagSlots)
        s = (s & ~(1ull << 63)) | ((u64)((hdr_publish_tag(h) >> i) & 1u) << 63);
      st(p + i, s);
    }
  }
  void fence_release() { __atomic_thread_fence(__ATOMIC_RELEASE); }

  // ---------------- node queries ----------------
  u32 header(const node_t& n) const { return slot_lo(n.s[kMetaSlot]); }
  u32 nkeys_of(const node_t& n) const { return hdr_nkeys(header(n)); }
  u32 sibling(const node_t& n) const { return slot_hi(n.s[kMetaSlot]); }
  u32 high(const node_t& n) const { return slot_lo(n.s[kHighSlot]); }
  u32 ksbits(const node_t& n) const { return slot_hi(n.s[kHighSlot]); }
  void set_header(node_t& n, u32 h) { n.s[kMetaSlot] = slot_make(h, sibling(n)); }
  void set_sibling(node_t& n, u32 s) { n.s[kMetaSlot] = slot_make(header(n), s); }
  void set_high(node_t& n, u32 hk) { n.s[kHighSlot] = slot_make(hk, ksbits(n)); }
  void set_ksbits(node_t& n, u32 b) { n.s[kHighSlot] = slot_make(high(n), b); }

  u32 find_child(const node_t& n, u32 q) const {
    u32 nk = nkeys_of(n);
    for (u32 i = 0; i < nk; i++)
      if (q <= slot_lo(n.s[i])) return child_of(n.s[i]);
    return kInvalid;
  }
  u32 entry_value(const node_t& n, u32 pos) const { return slot_hi(n.s[pos]); }
  bool find_cover(const node_t& n, u32 q, u32& j, u32& child) const {
    u32 nk = nkeys_of(n);
    for (u32 i = 0; i < nk; i++)
      if (q <= slot_lo(n.s[i])) {
        j = i;
        child = child_of(n.s[i]);
        return true;
      }
    return false;
  }
  bool find_child_slot(const node_t& n, u32 child, u32& j) const {
    u32 nk = nkeys_of(n);
    for (u32 i = 0; i < nk; i++)
      if (child_of(n.s[i]) == child) {
        j = i;
        return true;
      }
    return false;
  }
  int lookup(const node_t& n, u32 q, bool more, u32& ks, u32& val) const {
    u32 nk = nkeys_of(n), b = ksbits(n);
    for (u32 i = 0; i < nk; i++) {
      u32 k = ks_get(b, i);
      if (slot_lo(n.s[i]) == q && ks_more(k) == more) {
        ks = k;
        val = slot_hi(n.s[i]);
        return (int)i;
      }
    }
    return -1;
  }
  u32 insert_position(const node_t& n, u32 q, bool more) const {
    u32 nk = nkeys_of(n), b = ksbits(n), c = 0;
    for (u32 i = 0; i < nk; i++)
      if (entry_less(slot_lo(n.s[i]), ks_more(ks_get(b, i)), q, more)) c++;
    return c;
  }

  // ---------------- node edits ----------------
  void border_insert(node_t& n, u32 pos, u32 q, u32 ks, u32 val) {
    u32 nk = nkeys_of(n), b = ksbits(n);
    for (u32 i = nk; i > pos; i--) n.s[i] = n.s[i - 1];
    n.s[pos] = slot_make(q, val);
    u32 lowmask = (pos == 0) ? 0u : ((1u << (2 * pos)) - 1u);
    u32 nb = (b & lowmask) | ((b & ~lowmask) << 2) | (ks << (2 * pos));
    set_ksbits(n, nb);
    set_header(n, hdr_with_nkeys(header(n), nk + 1));
  }
  void border_remove(node_t& n, u32 pos) {
    u32 nk = nkeys_of(n), b = ksbits(n);
    for (u32 i = pos; i + 1 < nk; i++) n.s[i] = n.s[i + 1];
    n.s[nk - 1] = 0;
    u32 lowmask = (pos == 0) ? 0u : ((1u << (2 * pos)) - 1u);
    u32 highpart = (pos + 1 >= 16) ? 0u : ((b >> (2 * (pos + 1))) << (2 * pos));
    set_ksbits(n, (b & lowmask) | highpart);
    set_header(n, hdr_with_nkeys(header(n), nk - 1));
  }
  void border_set(node_t& n, u32 pos, u32 ks, u32 val) {
    n.s[pos] = slot_make(slot_lo(n.s[pos]), val);
    u32 b = ksbits(n);
    b = (b & ~(3u << (2 * pos))) | (ks << (2 * pos));
    set_ksbits(n, b);
  }
  // splits the 15 entries (n's 14 plus the pending one at ipos) into left/right
  u32 border_split(const node_t& n, u32 ipos, u32 q, u32 ks, u32 val, node_t& left,
                   node_t& right) {
    u32 nk = nkeys_of(n), b = ksbits(n);
    u32 total = nk + 1;
    u64 e[kSlots];
    u32 k[kSlots];
    for (u32 i = 0; i < total; i++) {
      if (i < ipos) {
        e[i] = n.s[i];
        k[i] = ks_get(b, i);
      } else if (i == ipos) {
        e[i] = slot_make(q, val);
        k[i] = ks;
      } else {
        e[i] = n.s[i - 1];
        k[i] = ks_get(b, i - 1);
      }
    }
    u32 m = (total + 1) / 2;
    if (slot_lo(e[m - 1]) == slot_lo(e[m])) m -= 1;
    std::memset(&left, 0, sizeof(left));
    std::memset(&right, 0, sizeof(right));
    u32 lb = 0, rb = 0;
    for (u32 i = 0; i < m; i++) {
      left.s[i] = e[i];
      lb |= k[i] << (2 * i);
    }
    for (u32 i = m; i < total; i++) {
      right.s[i - m] = e[i];
      rb |= k[i] << (2 * (i - m));
    }
    left.s[kHighSlot] = slot_make(0, lb);
    right.s[kHighSlot] = slot_make(0, rb);
    left.s[kMetaSlot] = slot_make(hdr_with_nkeys(0, m), kInvalid);
    right.s[kMetaSlot] = slot_make(hdr_with_nkeys(0, total - m), kInvalid);
    return slot_lo(e[m - 1]);
  }
  void make_interior_root(node_t& n, u32 sep, u32 l, u32 r, u32 level, u32 h) {
    u32 keep = h & (kLockBit | ~((1u << kVersionShift) - 1u));
    std::memset(&n, 0, sizeof(n));
    n.s[0] = slot_make(sep, l);
    n.s[1] = slot_make(kMaxSlice, r);
    n.s[kMetaSlot] = slot_make(keep | hdr_make(2, false, true, level), kInvalid);
  }
  void make_single_root(node_t& x, u32 t0, u32 ks, u32 val) {
    std::memset(&x, 0, sizeof(x));
    x.s[0] = slot_make(t0, val);
    x.s[kHighSlot] = slot_make(kMaxSlice, ks);
    x.s[kMetaSlot] = slot_make(hdr_make(1, true, true, 0), kInvalid);
  }
  void interior_insert(node_t& n, u32 j, u32 sep, u32 left, u32 right) {
    u32 nk = nkeys_of(n);
    u32 old_bound = slot_lo(n.s[j]);
    for (u32 i = nk; i > j + 1; i--) n.s[i] = n.s[i - 1];
    n.s[j] = slot_make(sep, left);
    n.s[j + 1] = slot_make(old_bound, right);
    set_header(n, hdr_with_nkeys(header(n), nk + 1));
  }
  void interior_remove(node_t& n, u32 j) {
    u32 nk = nkeys_of(n);
    n.s[j - 1] = slot_make(slot_lo(n.s[j]), slot_hi(n.s[j - 1]));
    for (u32 i = j; i + 1 < nk; i++) n.s[i] = n.s[i + 1];
    n.s[nk - 1] = 0;
    set_header(n, hdr_with_nkeys(header(n), nk - 1));
  }
  u32 entry_child(const node_t& n, u32 j) const { return child_of(n.s[j]); }
  u32 interior_split(const node_t& n, u32 j, u32 sep, u32 lchild, u32 rchild, node_t& left,
                     node_t& right) {
    u32 nk = nkeys_of(n);
    u32 total = nk + 1;
    u64 e[kSlots + 1];
    for (u32 i = 0, o = 0; i < nk; i++) {
      if (i == j) {
        e[o++] = slot_make(sep, lchild);
        e[o++] = slot_make(slot_lo(n.s[i]), rchild);
      } else {
        e[o++] = n.s[i];
      }
    }
    u32 m = total / 2;
    std::memset(&left, 0, sizeof(left));
    std::memset(&right, 0, sizeof(right));
    for (u32 i = 0; i < m; i++) left.s[i] = e[i];
    for (u32 i = m; i < total; i++) right.s[i - m] = e[i];
    left.s[kMetaSlot] = slot_make(hdr_with_nkeys(0, m), kInvalid);
    right.s[kMetaSlot] = slot_make(hdr_with_nkeys(0, total - m), kInvalid);
    return slot_lo(e[m - 1]);
  }

  // ---------------- keys and values ----------------
  u32 key_slice(const request& rq, u32 l) const { return rq.key[l]; }
  u32 value_slice(const request& rq, u32 i) const { return rq.value[i]; }
  void write_inline_value(const request& rq, u32 v) {
    if (rq.max_value_length >= 1) rq.out_value[0] = v;
  }

  // ---------------- slabs ----------------
  bool in_parent = false;  // set by the batch loop while the operation is in ph_parent_lock
  u32 alloc_slab() {
    if ((P.fail_any_every && P.any_allocs.fetch_add(1) % P.fail_any_every == 0) ||
        (P.fail_parent_every && in_parent &&
         P.parent_allocs.fetch_add(1) % P.fail_parent_every == 0)) {
      set_error(kErrOutOfMemory);
      return kInvalid;
    }
    u32 i = P.next.fetch_add(1);
    if (i >= P.num_slabs) {
      set_error(kErrOutOfMemory);
      return kInvalid;
    }
    return i;
  }
  void retire(u32 idx) {
    std::lock_guard<std::mutex> g(P.retire_mu);
    P.retired.push_back(idx);
  }

  // ---------------- records ----------------
  u32 rec_word(u32 head, u32 d) {
    u32 cur = head;
    for (u32 s = 0; s < d / kRecordWordsPerSlab; s++)
      cur = reinterpret_cast<u32*>(P.slab(cur))[kRecordNextWord];
    return reinterpret_cast<u32*>(P.slab(cur))[d % kRecordWordsPerSlab];
  }
  bool rec_load(u32 idx, u32& len_word) {
    if (!check(idx)) return false;
    rec_idx = idx;
    len_word = reinterpret_cast<u32*>(P.slab(idx))[0];
    return true;
  }
  bool rec_tail_equal(const u32* key, u32 off, u32 len, u32 len_word) {
    if (rec_tail_len(len_word) != len) return false;
    for (u32 i = 0; i < len; i++)
      if (rec_word(rec_idx, 1 + i) != key[off + i]) return false;
    return true;
  }
  u32 rec_copy_value(const request& rq, u32 len_word) {
    u32 tl = rec_tail_len(len_word), vl = rec_value_len(len_word);
    u32 n = vl < rq.max_value_length ? vl : rq.max_value_length;
    for (u32 i = 0; i < n; i++) rq.out_value[i] = rec_word(rec_idx, 1 + tl + i);
    return vl;
  }
  u32 rec_write(const std::vector<u32>& words) {
    u32 nslabs = rec_num_slabs((u32)words.size());
    std::vector<u32> ids(nslabs);
    for (u32 s = 0; s < nslabs; s++) {
      ids[s] = alloc_slab();
      if (ids[s] == kInvalid) return kInvalid;
    }
    for (u32 s = 0; s < nslabs; s++) {
      u32* w = reinterpret_cast<u32*>(P.slab(ids[s]));
      for (u32 i = 0; i < kRecordWordsPerSlab; i++) {
        size_t d = (size_t)s * kRecordWordsPerSlab + i;
        w[i] = d < words.size() ? words[d] : 0u;
      }
      w[kRecordNextWord] = (s + 1 < nslabs) ? ids[s + 1] : kInvalid;
    }
    return ids[0];
  }
  u32 rec_create(const u32* key, u32 off, u32 tail_len, const u32* value, u32 vlen) {
    if (tail_len > kMaxLength || vlen > kMaxLength) {
      set_error(kErrBadLength);
      return kInvalid;
    }
    std::vector<u32> w;
    w.push_back(rec_len_word(tail_len, vlen));
    for (u32 i = 0; i < tail_len; i++) w.push_back(key[off + i]);
    for (u32 i = 0; i < vlen; i++) w.push_back(value[i]);
    return rec_write(w);
  }
  bool rec_push_down(u32 rec, u32 len_word, u32& t0, u32& eks, u32& eval) {
    u32 tl = rec_tail_len(len_word), vl = rec_value_len(len_word);
    t0 = rec_word(rec, 1);
    std::vector<u32> value(vl);
    for (u32 i = 0; i < vl; i++) value[i] = rec_word(rec, 1 + tl + i);
    if (tl == 1) {
      if (vl == 1) {
        eks = kValue;
        eval = value[0];
        return true;
      }
      eks = kLongVal;
      std::vector<u32> w{rec_len_word(0, vl)};
      w.insert(w.end(), value.begin(), value.end());
      eval = rec_write(w);
      return eval != kInvalid;
    }
    eks = kSuffix;
    std::vector<u32> w{rec_len_word(tl - 1, vl)};
    for (u32 i = 1; i < tl; i++) w.push_back(rec_word(rec, 1 + i));
    w.insert(w.end(), value.begin(), value.end());
    eval = rec_write(w);
    return eval != kInvalid;
  }
};

}  // namespace cpu_model
