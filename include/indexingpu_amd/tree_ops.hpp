// Per-tile operation state machine of the MI355X Masstree.
//
// One tile (16, 32 or 64 lanes of a wave64) owns one operation at a time.  Each call
// of `step()` visits at most one node (one coalesced 128-byte load) and then either
// advances the operation or leaves it where it is ("retry": the node was locked, an
// optimistic read was torn, or a try-lock failed).  Nothing inside a step waits on
// another tile.  This matters on CDNA: the tiles of one wave share a program counter,
// so a tile spinning in an inner loop on a lock held by a sibling tile of the same
// wave would never let that sibling run.  A batch kernel calls `step()` in one flat
// loop, so every tile of a wave advances once per iteration and lock holders always
// make progress.
//
// The machine is generic over a tile policy `C` (node loads and stores, lane-parallel
// node edits, record access, slab allocation, retirement).  The product instantiates
// the gfx950 wave policy (tile_policy.hpp); the tests also instantiate a sequential
// CPU policy over the same memory layout to check interleavings and concurrency on
// the host.
//
// Algorithm (behaviour follows the reference's cooperative_{find,insert,update,erase},
// include/gpu_masstree.hpp:194-248, :397-627, :629-749, :751-971; the concurrency
// scheme is this project's own):
//   * B-link tree per layer with high keys and right-sibling links; readers move right
//     when the slice is above a node's high key.
//   * Writers lock only border nodes on the way down; a full border node is split
//     bottom-up: the new right node is written locked, the left node stays locked
//     until the parent has been updated (lock order child -> parent, left -> right,
//     so no cycles), and a full parent splits the same way.  Layer roots split in
//     place, so a layer's root index never changes.
//   * A key tail that collides with an existing suffix record is pushed one layer
//     down (a new layer root holding the old entry) and the insert continues there.
//   * Erase removes the entry without merging nodes.  A layer root below the top layer
//     that an erase leaves empty is removed: marked removed while locked, its LINK taken
//     out of the layer above, then unlocked and retired (cascading up while layer roots
//     empty).  Writers that meet a removed node restart from the tree root; readers see
//     an empty node.  Records and removed nodes are retired to the reclaimer and freed
//     once no reader of the launch can hold them.
#pragma once
#include "layout.hpp"

namespace indexingpu_amd {

// Operation kinds; the first four values are the reference's request_type
// (include/macros.hpp:51-58).
enum op_kind : u32 {
  op_insert = 0,
  op_erase = 1,
  op_find = 2,
  op_update = 3,
  op_upsert = 4,  // insert<update_if_exists=true>
};

enum op_phase : u32 {
  ph_descend = 0,         // optimistic descent in the current layer
  ph_record = 1,          // find: read the record `node`
  ph_lock = 2,            // lock border `node` and apply the write
  ph_parent_descend = 3,  // split: find the level that must receive (sep -> left, right)
  ph_parent_lock = 4,     // split: lock that parent and insert the separator
  ph_done = 5,
  ph_unlink_descend = 6,  // erase emptied layer root a0: walk from the tree root to the
                          // border node of layer a1 that holds its LINK
  ph_unlink_lock = 7,     // lock that node and remove the LINK
  ph_merge_descend = 8,   // erase emptied border node a0 (held locked): find its parent
  ph_merge_lock = 9,      // lock the parent and the left sibling, fold a0 into the sibling
};

// One request, uniform over the tile.
struct request {
  u32 kind;
  u32 key_length;
  const u32* key;
  const u32* value;     // insert / update / upsert input
  u32 value_length;
  u32* out_value;       // find output (max_value_length slices)
  u32 max_value_length;
  u32 flags;  // kReqNoSuffix | kReqKeepEmptyLayers | kReqNoMerge | fanout (bits 8-15)
};
// insert / upsert with enable_suffix = false: a new key's slices past the current layer
// become one single-entry layer each instead of a suffix record
constexpr u32 kReqNoSuffix = 1u;
// erase<do_remove_empty_root = false>: a layer emptied by an erase stays in place
constexpr u32 kReqKeepEmptyLayers = 2u;
// erase<..., do_merge = false>: a border node emptied by an erase stays in place
constexpr u32 kReqNoMerge = 4u;
// Node fanout (an extension; the reference's nodes are fixed at 16 slots): bits 8-11 give
// the entries at which a border node splits (2..kBorderCap) and bits 12-15 the children
// at which an interior node splits (3..kInteriorCap); 0 means the layout's maximum.  Any
// setting reads every tree, since readers use each node's own key count.
constexpr u32 kReqBorderCapShift = 8;
constexpr u32 kReqInteriorCapShift = 12;
IGPU_HD u32 border_cap(const request& rq) {
  const u32 c = (rq.flags >> kReqBorderCapShift) & 15u;
  return c ? c : kBorderCap;
}
IGPU_HD u32 interior_cap(const request& rq) {
  const u32 c = (rq.flags >> kReqInteriorCapShift) & 15u;
  return c ? c : kInteriorCap;
}

template <class NodeImage>
struct op_state {
  u32 phase;
  u32 layer;       // index of the key slice handled by the current layer
  u32 node;        // node (or record) visited by the next step
  u32 layer_root;  // root of the current layer
  u32 snap_hdr;    // header of `img` when it was validated (kInvalid = must reload)
  u32 a0, a1, a2, a3;  // split propagation: left, right, separator, parent level
  u32 result;      // find: value length; writes: 1 = success
  u32 steps;       // steps and waiting iterations so far
  u32 t0;          // coarse clock when the time guard started (0 = not yet)
  u32 hint[4];     // last interior node visited at levels 1..4 of the current layer
  u32 lvl;         // expected level of `node` (kInvalid: unknown)
  u32 fresh;       // bit 0: `snap_hdr` was validated by the step just before this lock
                   // attempt; bit 1: read interior nodes coherently (after a restart)
  NodeImage img;   // last validated image of `node`
};

// Livelock guard.  An operation may legitimately wait a long time behind a hot node
// (thousands of tiles queue on it), so the guard is time based: after kClockCheckSteps
// steps-or-waiting-iterations the operation starts a clock, and kMaxOpTime coarse ticks
// (about 21 s; the policy's max_op_time()) later it is abandoned with kErrStepLimit,
// leaving a record of its state (record_trip).  kMaxStepsPerOp bounds the
// count itself.
constexpr u32 kClockCheckSteps = 1u << 16;
constexpr u32 kMaxOpTime = 1u << 21;
constexpr u32 kMaxStepsPerOp = 1u << 31;
constexpr u32 kHintLevels = 4;

// true while the operation holds node locks (a split waiting for its parent, an emptied
// layer root waiting to be unlinked): such a tile must retry at once rather than back off
template <class S>
IGPU_HD bool op_holds_locks(const S& st) {
  return st.phase == ph_parent_descend || st.phase == ph_parent_lock ||
         st.phase == ph_unlink_descend || st.phase == ph_unlink_lock ||
         st.phase == ph_merge_descend || st.phase == ph_merge_lock;
}
template <class S>
IGPU_HD bool op_in_merge(const S& st) {
  return st.phase == ph_merge_descend || st.phase == ph_merge_lock;
}
template <class S>
IGPU_HD bool op_in_split(const S& st) {
  return st.phase == ph_parent_descend || st.phase == ph_parent_lock;
}

// Restarts a write from the tree root (it met a node that an erase removed).  Interior
// nodes are read coherently from then on: a stale cached interior line could route it to
// the removed node again.
template <class S>
IGPU_HD void op_restart(S& st, u32 root) {
  st.layer = 0;
  st.node = root;
  st.layer_root = root;
  st.snap_hdr = kInvalid;
  st.fresh |= 2u;
  clear_hints(st);
  st.phase = ph_descend;
}

// Starts unlinking the emptied layer root st.node (locked, marked removed, written):
// its LINK lives in layer st.layer - 1; the walk starts at the tree root.
template <class S>
IGPU_HD void begin_unlink(S& st, u32 root) {
  st.a0 = st.node;
  st.a1 = st.layer - 1;
  op_restart(st, root);
  st.phase = ph_unlink_descend;
}

// Gives a removed layer root back its live header (an unlink that cannot complete):
// an empty layer stays in place, as when removal is off.
template <class C>
IGPU_HD void revive_layer_root(C& c, u32 idx) {
  typename C::node_t m;
  c.read_node_locked(idx, m);
  c.set_header(m, hdr_revive(c.header(m)));
  c.write_node(idx, m);
  c.unlock_bump(idx);
}

template <class S>
IGPU_HD void clear_hints(S& st) {
  for (u32 k = 0; k < kHintLevels; k++) st.hint[k] = kInvalid;
  st.lvl = kInvalid;
}

// Continue a split at parent level `st.a3`: try the remembered node of that level
// first (the B-link move-right in ph_parent_lock finds the parent if it split), else
// descend from the layer root.
template <class S>
IGPU_HD void goto_parent(S& st) {
  const u32 h = (st.a3 - 1 < kHintLevels) ? st.hint[st.a3 - 1] : kInvalid;
  st.snap_hdr = kInvalid;
  st.node = (h != kInvalid) ? h : st.layer_root;
  st.phase = ph_parent_descend;
}

// results of try_lock_snapshot()
constexpr u32 kLockOk = 0, kLockBusy = 1, kLockStale = 2;

template <class C, class S>
IGPU_HD void op_begin(C& c, S& st, u32 root) {
  st.phase = ph_done;
  st.layer = 0;
  st.node = root;
  st.layer_root = root;
  st.snap_hdr = kInvalid;
  st.a0 = st.a1 = st.a2 = st.a3 = 0;
  st.result = 0;
  st.steps = 0;
  st.t0 = 0;
  st.fresh = 0;
  clear_hints(st);
  st.phase = ph_descend;
  (void)c;
}

// Starts an operation (already begun at the root) at border node `node` of layer
// `layer`, whose root is `layer_root` and whose level-1 interior node on the way down
// was `parent` (kInvalid: unknown).  Any node of the layer's border level left of the
// key's is a valid start in a B-link tree: the descent moves right from it.
template <class S>
IGPU_HD void op_begin_at(S& st, u32 node, u32 layer_root, u32 parent, u32 layer) {
  st.layer = layer;
  st.node = node;
  st.layer_root = layer_root;
  st.hint[0] = parent;
  st.lvl = 0;
}

template <class S>
IGPU_HD void op_finish(S& st, u32 result) {
  st.result = result;
  st.phase = ph_done;
}

template <class S>
IGPU_HD void op_next_layer(S& st, u32 root) {
  st.layer += 1;
  st.node = root;
  st.layer_root = root;
  st.snap_hdr = kInvalid;
  clear_hints(st);
  st.phase = ph_descend;
}

// Builds the entry that stores `rq`'s value at a key that ends at this slice.
template <class C>
IGPU_HD bool make_value_entry(C& c, const request& rq, u32& ks, u32& val) {
  if (rq.value_length == 1) {
    ks = kValue;
    val = c.value_slice(rq, 0);
    return true;
  }
  ks = kLongVal;
  val = c.rec_create(rq.key, 0, 0, rq.value, rq.value_length);
  return val != kInvalid;
}

// enable_suffix = false (the reference's insert<..., enable_suffix = false>,
// include/gpu_masstree.hpp:591-603): the key's slices key[from..len) become a chain of
// new single-entry layer roots, one per slice, the last holding the value; returns the
// first root (kInvalid when out of slabs: what was built is retired, nothing is
// published).  The chain is private until the caller links it.
template <class C>
IGPU_HD u32 make_layer_chain(C& c, const request& rq, u32 from) {
  using node_t = typename C::node_t;
  u32 ks, val;
  if (!make_value_entry(c, rq, ks, val)) return kInvalid;
  for (u32 l = rq.key_length - 1;; l--) {
    const u32 x = c.alloc_slab();
    if (x == kInvalid) {
      // retire the part built so far (layers l+1.. and the value record)
      c.fence_release();
      while (ks == kLink) {
        node_t m;
        c.read_node_locked(val, m);
        const u32 nks = ks_get(c.ksbits(m), 0), nval = c.entry_value(m, 0);
        c.retire(val | kRetiredNode);
        ks = nks;
        val = nval;
      }
      if (ks == kLongVal) c.retire(val);
      return kInvalid;
    }
    node_t n;
    c.make_single_root(n, c.key_slice(rq, l), ks, val);
    c.write_node(x, n);
    ks = kLink;
    val = x;
    if (l == from) break;
  }
  c.fence_release();
  return val;
}

// true when the border node whose locked header is h is a layer root below the top
// layer that the removal of one entry leaves empty, and empty layers are removed
template <class S>
IGPU_HD bool layer_emptied(const request& rq, const S& st, u32 h) {
  return hdr_root(h) && hdr_border(h) && hdr_nkeys(h) == 1 && st.layer > 0 &&
         !(rq.flags & kReqKeepEmptyLayers);
}

// An entry was removed from a non-root border node (locked header h before the removal)
// that is empty now.
IGPU_HD bool border_emptied(const request& rq, u32 h) {
  return !hdr_root(h) && hdr_border(h) && hdr_nkeys(h) == 1 && !(rq.flags & kReqNoMerge);
}

// Starts folding the emptied non-root border node st.node (locked, image n written)
// into its left sibling (ph_merge_*); it stays locked meanwhile.  The reference merges
// and borrows between underflowing nodes (include/gpu_masstree.hpp:1151-1330); here a
// node is merged once it is empty.
template <class C, class S>
IGPU_HD void begin_merge(C& c, S& st, const typename C::node_t& n) {
  c.fence_release();
  st.a0 = st.node;
  st.a1 = c.sibling(n);
  st.a2 = c.high(n);
  st.a3 = 1;
  goto_parent(st);
  st.phase = ph_merge_descend;
}

// An erase removed an entry from the locked border node st.node (image n, locked header
// h before the removal): publish it.  A layer root left empty (below the top layer) is
// marked removed, stays locked, and its LINK is removed from the layer above
// (ph_unlink_*) before it is unlocked and retired, as the reference's
// do_remove_empty_root does (include/gpu_masstree.hpp:916-966).
template <class C, class S>
IGPU_HD void erase_publish(C& c, const request& rq, S& st, typename C::node_t& n, u32 h) {
  if (layer_emptied(rq, st, h)) {
    c.set_header(n, hdr_remove(hdr_with_nkeys(h, 0)));
    c.write_node(st.node, n);
    c.fence_release();
    st.result = 1;
    begin_unlink(st, c.tree_root);
    return;
  }
  c.write_node(st.node, n);
  st.result = 1;
  if (border_emptied(rq, h)) {
    begin_merge(c, st, n);
    return;
  }
  c.unlock_bump(st.node);
  op_finish(st, 1);
}

// The merge of an emptied border node gives up (contention, or its parent slot is not
// where a merge can use it): the node stays, empty and linked, which every reader and
// writer handles.
template <class C, class S>
IGPU_HD void merge_give_up(C& c, S& st) {
  c.unlock_bump(st.a0);
  op_finish(st, st.result);
}

// The guard's record of the operation it abandons (layout at kTripOffset), taken before
// it releases anything; the node's header is read as it is now.
template <class C, class S>
IGPU_HD void record_trip(C& c, const request& rq, const S& st) {
  u32 w[kTripWords];
  w[0] = 1;
  w[1] = st.phase;
  w[2] = st.node;
  w[3] = st.layer;
  w[4] = st.layer_root;
  w[5] = st.a0;
  w[6] = st.a1;
  w[7] = st.a2;
  w[8] = st.a3;
  w[9] = st.snap_hdr;
  w[10] = c.peek_header(st.node);
  w[11] = st.steps;
  w[12] = rq.kind;
  w[13] = rq.key_length;
  w[14] = st.layer < rq.key_length ? c.key_slice(rq, st.layer) : 0u;
  w[15] = (st.lvl & 0xFFFFu) | (st.fresh << 16);
  c.store_trip(w);
}

// Returns true when the step made progress (false: retry later).
template <bool Validate, class C, class S>
IGPU_HD bool op_step(C& c, const request& rq, S& st) {
  using node_t = typename C::node_t;
  st.steps += 1;
  if (c.failed()) {  // a corrupt object index was seen: abandon the operation
    op_finish(st, 0);
    return true;
  }
  if (st.steps > kClockCheckSteps) {
    const u32 now = c.coarse_clock() | 1u;
    if (st.t0 == 0) st.t0 = now;
    if (now - st.t0 > c.max_op_time() || st.steps > kMaxStepsPerOp) {
      if (c.set_error_first(kErrStepLimit)) record_trip(c, rq, st);
      if (op_in_split(st)) {
        // a split waiting for its parent holds both halves (written, not yet
        // published); publish them: the right half is reachable through the left
        // one's sibling link, and a later split of it links it from its left neighbour
        c.unlock_bump(st.a0);
        c.unlock_bump(st.a1);
      } else if (op_in_merge(st)) {
        c.unlock_bump(st.a0);  // the emptied node stays, empty and linked
      } else if (op_holds_locks(st)) {
        revive_layer_root(c, st.a0);  // the emptied layer stays, live
      }
      op_finish(st, op_holds_locks(st) ? st.result : 0);
      return true;
    }
  }
  const u32 q = c.key_slice(rq, st.layer);
  const bool more = st.layer + 1 < rq.key_length;
  switch (st.phase) {
    case ph_descend: {
      node_t& n = st.img;
      // interior nodes (expected level > 0, or unknown at a layer root) may come from
      // the cache; border nodes are always read coherently, since writers lock them by
      // the version they read
      if (!c.template read_node<Validate>(st.node, n, st.lvl != 0 && !(st.fresh & 2u))) return false;
      const u32 h = c.header(n);
      if (!hdr_border(h)) {
        if (hdr_level(h) - 1 < kHintLevels) st.hint[hdr_level(h) - 1] = st.node;
        u32 child = c.find_child(n, q);
        u32 next_level = hdr_level(h) - 1;
        if (child == kInvalid) {
          child = c.sibling(n);
          next_level = hdr_level(h);
        }
        if (child == kInvalid) {
          c.set_error(kErrBadIndex);
          op_finish(st, 0);
          return true;
        }
        st.node = child;
        st.lvl = next_level;
        return true;
      }
      st.lvl = 0;
      if (hdr_removed(h)) {
        // a removed node, reached through a pointer read before its removal: an
        // emptied layer root (the key's place is in the layer above now) or a border
        // node merged into its left sibling (whose range now covers the key).  Start
        // again from the tree root, reading interior nodes coherently, so the walk
        // sees the parent slot the merge rewrote.
        op_restart(st, c.tree_root);
        return true;
      }
      if (q > c.high(n) && c.sibling(n) != kInvalid) {
        st.node = c.sibling(n);
        return true;
      }
      u32 ks = 0, val = 0;
      const int pos = c.lookup(n, q, more, ks, val);
      if (rq.kind == op_find) {
        if (pos < 0) {
          op_finish(st, 0);
        } else if (ks == kLink) {
          op_next_layer(st, val);
        } else if (ks == kValue) {
          c.write_inline_value(rq, val);
          op_finish(st, 1);
        } else {
          st.a0 = ks;
          st.node = val;
          st.phase = ph_record;
        }
        return true;
      }
      if (pos >= 0 && ks == kLink) {
        op_next_layer(st, val);
        return true;
      }
      if (rq.kind == op_insert && pos >= 0 && !more) {  // key ends here and exists
        op_finish(st, 0);
        return true;
      }
      if ((rq.kind == op_erase || rq.kind == op_update) && pos < 0) {  // absent
        op_finish(st, 0);
        return true;
      }
      st.snap_hdr = h;
      st.fresh |= 1u;
      st.phase = ph_lock;
      return true;
    }

    case ph_record: {
      // records are immutable once published: no validation needed
      u32 len_word;
      if (!c.rec_load(st.node, len_word)) {
        op_finish(st, 0);
        return true;
      }
      if (st.a0 == kSuffix) {
        const u32 tail = rq.key_length - st.layer - 1;
        if (!c.rec_tail_equal(rq.key, st.layer + 1, tail, len_word)) {
          op_finish(st, 0);
          return true;
        }
      }
      op_finish(st, c.rec_copy_value(rq, len_word));
      return true;
    }

    case ph_lock: {
      // Lock only the version we validated: a node that changed since is re-read
      // optimistically (it may have split or become interior) instead of being locked
      // just to find that out, which would serialize every waiter on a hot node.  A
      // changed node means another writer got there first, so the re-read waits out a
      // backoff like a busy lock: retrying at once would have every waiter re-read and
      // race for the lock after each release of a hot node (a batch whose keys all land
      // past the current right edge of the tree ran 5-10x slower that way).
      // the first attempt right after the validating read goes straight to the
      // compare-and-swap; a retry first checks the header with a plain load
      const u32 r = c.try_lock_snapshot(st.node, st.snap_hdr, (st.fresh & 1u) != 0);
      st.fresh &= ~1u;
      if (r == kLockBusy) return false;
      if (r == kLockStale) {
        st.phase = ph_descend;
        return false;
      }
      node_t& n = st.img;
      // the in-memory header carries the lock bit while we hold the lock; keep the image
      // identical so that write_node() never publishes an unlocked header
      const u32 h = st.snap_hdr | kLockBit;
      c.set_header(n, h);
      u32 ks = 0, val = 0;
      const int pos = c.lookup(n, q, more, ks, val);
      if (pos >= 0 && ks == kLink) {
        c.unlock_same(st.node);
        op_next_layer(st, val);
        return true;
      }
      if (pos < 0) {
        if (rq.kind == op_erase || rq.kind == op_update) {
          c.unlock_same(st.node);
          op_finish(st, 0);
          return true;
        }
        // insert a new entry
        u32 nks, nval;
        bool ok;
        if (more && (rq.flags & kReqNoSuffix)) {
          nks = kLink;
          nval = make_layer_chain(c, rq, st.layer + 1);
          ok = nval != kInvalid;
        } else if (more) {
          nks = kSuffix;
          nval = c.rec_create(rq.key, st.layer + 1, rq.key_length - st.layer - 1, rq.value,
                              rq.value_length);
          ok = nval != kInvalid;
        } else {
          ok = make_value_entry(c, rq, nks, nval);
        }
        if (!ok) {
          c.unlock_same(st.node);
          op_finish(st, 0);
          return true;
        }
        const u32 ipos = c.insert_position(n, q, more);
        if (hdr_nkeys(h) < border_cap(rq)) {
          c.border_insert(n, ipos, q, nks, nval);
          c.write_node(st.node, n);
          c.unlock_bump(st.node);
          op_finish(st, 1);
          return true;
        }
        // split the full border node
        node_t left, right;
        if (hdr_root(h)) {
          const u32 li = c.alloc_slab();
          const u32 ri = c.alloc_slab();
          if (li == kInvalid || ri == kInvalid) {
            c.unlock_same(st.node);
            op_finish(st, 0);
            return true;
          }
          const u32 sep = c.border_split(n, ipos, q, nks, nval, left, right);
          c.set_header(left, hdr_make(c.nkeys_of(left), true, false, 0));
          c.set_sibling(left, ri);
          c.set_high(left, sep);
          c.set_header(right, hdr_make(c.nkeys_of(right), true, false, 0));
          c.set_sibling(right, kInvalid);
          c.set_high(right, c.high(n));
          c.write_node(li, left);
          c.write_node(ri, right);
          c.fence_release();
          c.make_interior_root(n, sep, li, ri, 1, h);
          c.write_node(st.node, n);
          c.unlock_bump(st.node);
          op_finish(st, 1);
          return true;
        }
        const u32 ri = c.alloc_slab();
        if (ri == kInvalid) {
          c.unlock_same(st.node);
          op_finish(st, 0);
          return true;
        }
        const u32 sep = c.border_split(n, ipos, q, nks, nval, left, right);
        c.set_header(right, hdr_make(c.nkeys_of(right), true, false, 0) | kLockBit);
        c.set_sibling(right, c.sibling(n));
        c.set_high(right, c.high(n));
        c.set_header(left, hdr_with_nkeys(h, c.nkeys_of(left)));
        c.set_sibling(left, ri);
        c.set_high(left, sep);
        c.write_node(ri, right);
        c.write_node(st.node, left);
        c.fence_release();
        st.a0 = st.node;
        st.a1 = ri;
        st.a2 = sep;
        st.a3 = 1;
        st.result = 1;
        goto_parent(st);
        return true;
      }
      // the entry (q, more) exists: VALUE / LONGVAL (key ends here) or SUFFIX
      if (ks == kSuffix) {
        u32 len_word;
        if (!c.rec_load(val, len_word)) {
          c.unlock_same(st.node);
          op_finish(st, 0);
          return true;
        }
        const u32 tail = rq.key_length - st.layer - 1;
        const bool equal = c.rec_tail_equal(rq.key, st.layer + 1, tail, len_word);
        if (equal) {
          if (rq.kind == op_insert) {
            c.unlock_same(st.node);
            op_finish(st, 0);
            return true;
          }
          if (rq.kind == op_erase) {
            c.border_remove(n, pos);
            c.retire(val);
            erase_publish(c, rq, st, n, h);
            return true;
          }
          // update / upsert: same tail, new value
          const u32 nrec = c.rec_create(rq.key, st.layer + 1, tail, rq.value, rq.value_length);
          if (nrec == kInvalid) {
            c.unlock_same(st.node);
            op_finish(st, 0);
            return true;
          }
          c.border_set(n, pos, kSuffix, nrec);
          c.write_node(st.node, n);
          c.unlock_bump(st.node);
          c.retire(val);
          op_finish(st, 1);
          return true;
        }
        if (rq.kind == op_erase || rq.kind == op_update) {
          c.unlock_same(st.node);
          op_finish(st, 0);
          return true;
        }
        // insert / upsert with a different tail: push the existing entry one layer down
        const u32 xi = c.alloc_slab();
        if (xi == kInvalid) {
          c.unlock_same(st.node);
          op_finish(st, 0);
          return true;
        }
        u32 eks, eval;
        u32 t0;
        if (!c.rec_push_down(val, len_word, t0, eks, eval)) {
          c.unlock_same(st.node);
          op_finish(st, 0);
          return true;
        }
        node_t x;
        c.make_single_root(x, t0, eks, eval);
        c.write_node(xi, x);
        c.fence_release();
        c.border_set(n, pos, kLink, xi);
        c.write_node(st.node, n);
        c.unlock_bump(st.node);
        c.retire(val);
        op_next_layer(st, xi);
        return true;
      }
      // VALUE or LONGVAL
      if (rq.kind == op_insert) {
        c.unlock_same(st.node);
        op_finish(st, 0);
        return true;
      }
      if (rq.kind == op_erase) {
        c.border_remove(n, pos);
        if (ks == kLongVal) c.retire(val);
        erase_publish(c, rq, st, n, h);
        return true;
      }
      {
        u32 nks, nval;
        if (!make_value_entry(c, rq, nks, nval)) {
          c.unlock_same(st.node);
          op_finish(st, 0);
          return true;
        }
        c.border_set(n, pos, nks, nval);
        c.write_node(st.node, n);
        c.unlock_bump(st.node);
        if (ks == kLongVal) c.retire(val);
        op_finish(st, 1);
        return true;
      }
    }

    case ph_parent_descend: {
      node_t& n = st.img;
      if (!c.template read_node<true>(st.node, n)) return false;
      const u32 h = c.header(n);
      const u32 lvl = hdr_level(h);
      if (lvl < st.a3) {
        c.set_error(kErrBadIndex);
        op_finish(st, 0);
        return true;
      }
      if (lvl == st.a3) {
        st.snap_hdr = h;
        st.fresh |= 1u;
        st.phase = ph_parent_lock;
        return true;
      }
      if (lvl - 1 < kHintLevels) st.hint[lvl - 1] = st.node;
      u32 child = c.find_child(n, st.a2);
      if (child == kInvalid) child = c.sibling(n);
      st.node = child;
      return true;
    }

    case ph_parent_lock: {
      const u32 r = c.try_lock_snapshot(st.node, st.snap_hdr, (st.fresh & 1u) != 0);
      st.fresh &= ~1u;
      if (r == kLockBusy) return false;
      if (r == kLockStale) {  // re-read the candidate (it may have split or grown)
        st.phase = ph_parent_descend;
        return false;
      }
      node_t& n = st.img;
      const u32 h = st.snap_hdr | kLockBit;
      c.set_header(n, h);
      u32 j;
      u32 lc = st.a0;  // child left of the new separator
      if (!c.find_child_slot(n, st.a0, j)) {
        if (!c.find_cover(n, st.a2, j, lc)) {
          // the separator lies right of this node (the parent split meanwhile, or is
          // still being linked): continue with the right sibling, else from the root
          const u32 sib = c.sibling(n);
          c.unlock_same(st.node);
          st.node = (sib != kInvalid) ? sib : st.layer_root;
          st.phase = ph_parent_descend;
          return true;
        }
        // This node covers the separator but has no slot for a0.  While a0 is locked
        // by this operation no other split can be linking it, so a0 is a node whose
        // own parent entry was never made (an earlier split ran out of slabs for its
        // parent, or was abandoned by the time guard).  The slot covering the
        // separator points at a node left of a0 on the same level, from which B-link
        // move-right reaches a0.  Linking the new right half after that slot keeps
        // every route valid: keys up to the separator still go through the left node
        // and move right to a0; keys above it go straight to a1.
      }
      if (hdr_nkeys(h) < interior_cap(rq)) {
        c.interior_insert(n, j, st.a2, lc, st.a1);
        c.write_node(st.node, n);
        c.unlock_bump(st.node);
        c.unlock_bump(st.a0);
        c.unlock_bump(st.a1);
        op_finish(st, st.result);
        return true;
      }
      node_t left, right;
      if (hdr_root(h)) {
        const u32 li = c.alloc_slab();
        const u32 ri = c.alloc_slab();
        if (li == kInvalid || ri == kInvalid) {
          // out of slabs: the separator cannot be placed.  The right half stays
          // reachable through the sibling link, so every route is still correct, and a
          // later split of it links it from its left neighbour's slot (above).
          c.unlock_same(st.node);
          c.unlock_bump(st.a0);
          c.unlock_bump(st.a1);
          op_finish(st, st.result);
          return true;
        }
        const u32 sep2 = c.interior_split(n, j, st.a2, lc, st.a1, left, right);
        c.set_header(left, hdr_make(c.nkeys_of(left), false, false, st.a3));
        c.set_sibling(left, ri);
        c.set_header(right, hdr_make(c.nkeys_of(right), false, false, st.a3));
        c.set_sibling(right, kInvalid);
        c.write_node(li, left);
        c.write_node(ri, right);
        c.fence_release();
        c.make_interior_root(n, sep2, li, ri, st.a3 + 1, h);
        c.write_node(st.node, n);
        c.unlock_bump(st.node);
        c.unlock_bump(st.a0);
        c.unlock_bump(st.a1);
        op_finish(st, st.result);
        return true;
      }
      const u32 ri = c.alloc_slab();
      if (ri == kInvalid) {
        c.unlock_same(st.node);
        c.unlock_bump(st.a0);
        c.unlock_bump(st.a1);
        op_finish(st, st.result);
        return true;
      }
      const u32 sep2 = c.interior_split(n, j, st.a2, lc, st.a1, left, right);
      c.set_header(right, hdr_make(c.nkeys_of(right), false, false, st.a3) | kLockBit);
      c.set_sibling(right, c.sibling(n));
      c.set_header(left, hdr_with_nkeys(h, c.nkeys_of(left)));
      c.set_sibling(left, ri);
      c.write_node(ri, right);
      c.write_node(st.node, left);
      c.fence_release();
      c.unlock_bump(st.a0);
      c.unlock_bump(st.a1);
      st.a0 = st.node;
      st.a1 = ri;
      st.a2 = sep2;
      st.a3 = st.a3 + 1;
      goto_parent(st);
      return true;
    }

    case ph_unlink_descend: {
      // the same B-link descent as ph_descend, through LINK entries down to layer a1
      node_t& n = st.img;
      if (!c.template read_node<Validate>(st.node, n, st.lvl != 0 && !(st.fresh & 2u))) return false;
      const u32 h = c.header(n);
      if (!hdr_border(h)) {
        u32 child = c.find_child(n, q);
        u32 next_level = hdr_level(h) - 1;
        if (child == kInvalid) {
          child = c.sibling(n);
          next_level = hdr_level(h);
        }
        if (child == kInvalid) {
          c.set_error(kErrBadIndex);
          revive_layer_root(c, st.a0);
          op_finish(st, st.result);
          return true;
        }
        st.node = child;
        st.lvl = next_level;
        return true;
      }
      st.lvl = 0;
      if (hdr_removed(h)) {
        // a border node merged away (its entries live in its neighbour now), reached
        // through a pointer read before the merge: walk again from the tree root
        const u32 target = st.a1;
        op_restart(st, c.tree_root);
        st.a1 = target;
        st.phase = ph_unlink_descend;
        return true;
      }
      if (q > c.high(n) && c.sibling(n) != kInvalid) {
        st.node = c.sibling(n);
        return true;
      }
      if (st.layer == st.a1) {
        st.snap_hdr = h;
        st.fresh |= 1u;
        st.phase = ph_unlink_lock;
        return true;
      }
      u32 ks = 0, val = 0;
      if (c.lookup(n, q, true, ks, val) < 0 || ks != kLink) {
        // the path to the emptied layer is gone (not expected: only the holder of a
        // layer root's lock removes its LINK, and a layer with a LINK is not empty)
        revive_layer_root(c, st.a0);
        op_finish(st, st.result);
        return true;
      }
      st.layer += 1;
      st.node = val;
      st.layer_root = val;
      st.snap_hdr = kInvalid;
      clear_hints(st);
      return true;
    }

    case ph_unlink_lock: {
      const u32 r = c.try_lock_snapshot(st.node, st.snap_hdr, (st.fresh & 1u) != 0);
      st.fresh &= ~1u;
      if (r == kLockBusy) return false;
      if (r == kLockStale) {
        st.phase = ph_unlink_descend;
        return false;
      }
      node_t& n = st.img;
      const u32 h = st.snap_hdr | kLockBit;
      c.set_header(n, h);
      u32 ks = 0, val = 0;
      const int pos = c.lookup(n, q, true, ks, val);
      if (pos < 0 || ks != kLink || val != st.a0) {
        const u32 sib = c.sibling(n);
        c.unlock_same(st.node);
        if (q > c.high(n) && sib != kInvalid) {  // the node split meanwhile: move right
          st.node = sib;
          st.phase = ph_unlink_descend;
          return true;
        }
        revive_layer_root(c, st.a0);
        op_finish(st, st.result);
        return true;
      }
      c.border_remove(n, pos);
      const u32 removed = st.a0;
      if (layer_emptied(rq, st, h)) {
        // this layer's root is empty now too: remove it the same way (cascade)
        c.set_header(n, hdr_remove(hdr_with_nkeys(h, 0)));
        c.write_node(st.node, n);
        c.fence_release();
        c.unlock_bump(removed);
        c.retire(removed | kRetiredNode);
        begin_unlink(st, c.tree_root);
        return true;
      }
      c.write_node(st.node, n);
      c.unlock_bump(removed);  // stays marked removed: later visitors restart
      c.retire(removed | kRetiredNode);
      if (border_emptied(rq, h)) {
        begin_merge(c, st, n);
        return true;
      }
      c.unlock_bump(st.node);
      op_finish(st, st.result);
      return true;
    }

    case ph_merge_descend: {
      // B-link descent to the level-1 node covering the emptied node's high key (a2),
      // from the level-1 hint of the erase, else the layer root
      node_t& n = st.img;
      if (!c.template read_node<true>(st.node, n)) return false;
      const u32 h = c.header(n);
      const u32 lvl = hdr_level(h);
      if (hdr_border(h) || lvl < st.a3 || hdr_removed(h)) {
        merge_give_up(c, st);  // the layer has one level (no parent), or a stale hint
        return true;
      }
      if (lvl == st.a3) {
        st.snap_hdr = h;
        st.fresh |= 1u;
        st.phase = ph_merge_lock;
        return true;
      }
      if (lvl - 1 < kHintLevels) st.hint[lvl - 1] = st.node;
      u32 child = c.find_child(n, st.a2);
      if (child == kInvalid) child = c.sibling(n);
      if (child == kInvalid) {
        merge_give_up(c, st);
        return true;
      }
      st.node = child;
      return true;
    }

    case ph_merge_lock: {
      const u32 r = c.try_lock_snapshot(st.node, st.snap_hdr, (st.fresh & 1u) != 0);
      st.fresh &= ~1u;
      if (r == kLockBusy) return false;
      if (r == kLockStale) {
        st.phase = ph_merge_descend;
        return false;
      }
      node_t& n = st.img;
      const u32 h = st.snap_hdr | kLockBit;
      c.set_header(n, h);
      u32 j;
      if (!c.find_child_slot(n, st.a0, j)) {
        u32 jc, cc;
        if (!c.find_cover(n, st.a2, jc, cc) && c.sibling(n) != kInvalid) {
          // the slot moved right with a split of this node
          const u32 sib = c.sibling(n);
          c.unlock_same(st.node);
          st.node = sib;
          st.phase = ph_merge_descend;
          return true;
        }
        c.unlock_same(st.node);  // no slot for it (linked from its left neighbour)
        merge_give_up(c, st);
        return true;
      }
      // Merge with the neighbour under the same parent: the left one (slot j - 1), or,
      // for the parent's first child, the right one (slot 1).  Of the two nodes, the
      // left one stays: it takes the right one's range, sibling link and (when the
      // emptied node is the left one) entries; the parent drops the right one's slot,
      // whose bound goes to the left one's; the right one is marked removed and retired.
      if (j == 0 && hdr_nkeys(h) < 2) {
        c.unlock_same(st.node);  // the parent's only child: leave it
        merge_give_up(c, st);
        return true;
      }
      // The neighbour gets one lock attempt: splits lock it before this parent, so
      // waiting for it here could deadlock; a busy neighbour ends the merge.
      const u32 other = c.entry_child(n, j ? j - 1 : 1);
      u32 oold;
      if (!c.try_lock(other, oold)) {
        c.unlock_same(st.node);
        merge_give_up(c, st);
        return true;
      }
      node_t on;
      c.read_node_locked(other, on);
      if (!hdr_border(oold) || hdr_removed(oold) ||
          (j ? c.sibling(on) != st.a0 : other != st.a1)) {
        c.unlock_same(other);  // not the adjacent border node (a split is linking one)
        c.unlock_same(st.node);
        merge_give_up(c, st);
        return true;
      }
      node_t en;
      c.read_node_locked(st.a0, en);
      const u32 eh = c.header(en);
      // the exclusive low bound of the node that goes: the left one's high key
      const u32 low = j ? c.high(on) : c.high(en);
      u32 left = other, right = st.a0;
      if (j) {
        c.set_header(on, oold | kLockBit);
        c.set_sibling(on, st.a1);
        c.set_high(on, st.a2);
        c.write_node(left, on);
        c.set_header(en, hdr_remove(eh));
        c.write_node(right, en);
      } else {
        // the emptied node takes the right neighbour's image (entries, high key,
        // keystates, sibling link) under its own header
        left = st.a0;
        right = other;
        c.set_header(on, hdr_with_nkeys(eh, hdr_nkeys(oold)));
        c.write_node(left, on);
        c.set_header(on, hdr_remove(oold | kLockBit));
        c.write_node(right, on);
      }
      c.interior_remove(n, j ? j : 1);
      c.write_node(st.node, n);
      // A border-level shortcut table (entries are layer-0 border nodes; every node some
      // entry names is named by one run of entries that starts at the first prefix above
      // the node's low bound) is rewritten where it named the right node: that run starts
      // right after the entry of the prefix holding `low`, and takes that entry's node
      // (`prev`: the left node or one left of it, a valid start for the whole run), which
      // keeps both properties.  `prev` is held locked meanwhile (a merge that would remove
      // it waits, then sees its longer run).  When the table cannot be rewritten (none
      // given, `prev` busy, too long a run), the merge is counted, which retires the table.
      bool patched = false;
      u32 held = kInvalid;
      if (st.layer == 0) {
        const u32 prev = c.shortcut_prev(low);
        if (prev == left) {
          patched = c.shortcut_patch(right, prev, low);
        } else if (prev < c.slab_count()) {
          u32 pold;
          if (c.try_lock(prev, pold)) {
            held = prev;
            patched = c.shortcut_patch(right, prev, low);
          }
        }
      }
      c.fence_release();
      if (held != kInvalid) c.unlock_same(held);
      c.unlock_bump(left);
      c.unlock_bump(st.node);
      c.unlock_bump(right);
      c.retire(right | kRetiredNode);
      if (st.layer == 0 && !patched) c.count_merge();
      op_finish(st, st.result);
      return true;
    }

    default:
      return true;
  }
}

}  // namespace indexingpu_amd
