// Kernels of the MI355X Masstree.
//
//  * batch_kernel: persistent grid; every tile owns one request at a time and calls
//    op_step() in one flat loop.  A tile that finishes its request immediately loads
//    its next one (static stride over the batch), so tiles of a wave stay busy and a
//    wave keeps up to 64/T independent node loads in flight per iteration.  Replaces
//    the reference's batch_kernel + device funcs (include/kernels.hpp:31-79,116-456).
//  * scan_kernel: one range query per tile at a time (gpu_masstree.hpp:250-395).
//  * drain_kernel / reset_kernel: free the records retired by the previous launch.
//  * init_root_kernel: allocates and writes the empty layer-0 root.
#pragma once
#include "hip_util.hpp"
#include "layout.hpp"
#include "reclaim.hpp"
#include "slab_alloc.hpp"
#include "tile_policy.hpp"
#include "tree_ops.hpp"

namespace indexingpu_amd {

constexpr u32 kBlockThreads = 256;
constexpr u32 kMaxBackoffShift = 8;
// A tile that holds locks (a split waiting for its parent) backs off far less: its
// locks keep other tiles waiting, but retrying at once would hammer the parent's line.
constexpr u32 kHolderBackoffShift = 0;

struct tree_view {
  slab_pool_view pool;
  retire_view retire;
  u32* err;
  u32 root;
  // Optional layer-0 shortcut table for read-only descents (null: none): entry i is an
  // interior node of level 1 (or the tree root) whose key range starts at or below the
  // first slices of prefix i = slice >> shortcut_shift.  Interior nodes of layer 0 are
  // never freed, never change level (only the root grows, in place) and their low bounds
  // never rise, so an entry stays a valid B-link starting point after any later write.
  const u32* shortcut;
  u32 shortcut_shift;
  // Optional border-level twin of the table (finds): entry i is the border node of layer 0
  // reached by the smallest first slice of prefix i.  Border nodes can be merged away and
  // freed: a merge rewrites the entries naming the node it removes (shortcut_patch), and
  // one that cannot is counted (err word kMergeEpochWord); the table is used only while
  // that count equals the count it was built at (*shortcut_epoch).
  u32* shortcut_border;  // (merges of the launch's writes keep it current: shortcut_patch)
  const u32* shortcut_epoch;
};

// The table a find starts from: the border-level one while every merge since it was built
// has rewritten it, else the level-1 one (or none).
__device__ __forceinline__ const u32* find_start_table(const tree_view& tv) {
  if (tv.shortcut_border &&
      dev::coherent_load(tv.err + kMergeEpochWord) == dev::coherent_load(tv.shortcut_epoch))
    return tv.shortcut_border;
  return tv.shortcut;
}

// Arguments of one batch.  `types` selects per-request kinds (mixed batches); when it
// is null every request has `kind`.  Find requests write min(len, maxV) value slices
// at out_values + i * maxV and their length to out_value_lengths[i]; write requests
// report success in results[i] when results is not null.
struct batch_args {
  const uint8_t* types;
  u32 kind;
  const u32* keys;
  u32 max_key_length;
  const u32* key_lengths;
  const u32* values;
  u32 max_value_length;
  const u32* value_lengths;
  u32* out_values;
  u32* out_value_lengths;
  bool* results;
  u64 n;
  // Optional processing order (batch_kernel): request index i of the launch is request
  // (perm_lo + i) * perm_mul % perm_n of the arrays; perm_mul = 0 means i itself.  With
  // perm_mul coprime to perm_n this is a bijection, so consecutive launches over
  // [0, perm_n) in stages visit every request once, each stage a spread-out sample.
  u64 perm_lo;
  u64 perm_mul;
  u64 perm_n;
  // Optional request range read on the device (batched form, one launch per batch):
  // when set, the launch covers requests [range[0], range[1]) clamped to n, and the
  // grid need not match the batch size.
  const u64* range;
  // Mixed batches run as a read-only pass over their finds followed by a pass over the
  // rest: with skip_finds set (and types given), batch_kernel leaves find requests alone.
  u32 skip_finds;
  // Optional start hints of write requests, kHintWords words per launch position
  // (counted from the launch's first request): border node, layer root, level-1 parent
  // and layer, as a read-only pass over the same positions found them (node kInvalid:
  // none).  The write launch then starts each write at its border node instead of the
  // root.
  u32* hints;
  // kReqNoSuffix (insert / upsert with enable_suffix = false), kReqKeepEmptyLayers
  // (erase<do_remove_empty_root = false>), kReqNoMerge (erase<..., do_merge = false>)
  u32 flags;
};
constexpr u32 kHintWords = 4;

__device__ __forceinline__ u64 request_index(const batch_args& a, u64 i) {
  return a.perm_mul ? (a.perm_lo + i) * a.perm_mul % a.perm_n : i;
}

template <u32 T>
__device__ __forceinline__ void load_request(const batch_args& a, u64 i, request& rq) {
  rq.kind = a.types ? (u32)a.types[i] : a.kind;
  rq.key_length = a.key_lengths ? a.key_lengths[i] : a.max_key_length;
  rq.key = a.keys + i * (u64)a.max_key_length;
  rq.value = a.values ? a.values + i * (u64)a.max_value_length : nullptr;
  rq.value_length = a.value_lengths ? a.value_lengths[i] : a.max_value_length;
  rq.out_value = a.out_values ? a.out_values + i * (u64)a.max_value_length : nullptr;
  rq.max_value_length = a.out_values ? a.max_value_length : 0u;
  rq.flags = a.flags;
}

// After a corruption or livelock guard fired, the remaining requests of the launch are
// skipped so that the kernel drains quickly.
__device__ __forceinline__ bool abort_pending(u32* err) {
  const u32 e = dev::coherent_load(err);
  return e == kErrStepLimit || e == kErrBadIndex;
}

// Five waves per SIMD (at most 96 VGPRs).  The kernel needs ~114 VGPRs to run without
// scratch spills (4 waves); the peak is in the border-node write step (record push-down,
// tail compare, split).  Measured on 16M-key batches: 5 waves (36 B of spills per lane)
// run erases and mixed batches ~5% faster than 6 waves (80 VGPRs, 96 B of spills) at the
// same insert rate; 4 waves (no spills) lose 10% on inserts, 8 waves (64 VGPRs, 336 B of
// spills) lose 45%.
template <u32 T, bool Validate>
__global__ void __launch_bounds__(kBlockThreads) __attribute__((amdgpu_waves_per_eu(5)))
batch_kernel(tree_view tv, batch_args a) {
  wave_tile<T, Validate> c(tv.pool, tv.retire, tv.err, tv.root);
  c.sc_border = tv.shortcut_border;
  c.sc_epoch = tv.shortcut_epoch;
  c.sc_shift = tv.shortcut_shift;
  const u32 tiles_per_block = kBlockThreads / T;
  const u64 ntiles = (u64)gridDim.x * tiles_per_block;
  u64 i = (u64)blockIdx.x * tiles_per_block + threadIdx.x / T;
  u64 lo = 0, hi = a.n;
  if (a.range) {
    hi = min(a.range[1], a.n);
    lo = min(a.range[0], hi);
    i += lo;
  }
  u32 ri = 0;  // request_index(a, i); batch launches hold fewer than 2^32 requests
  op_state<gpu_node> st;
  request rq;
  // with skip_finds, the tile's first 16 lanes look at the next 16 requests the tile
  // would take and it moves to the first that is not a find
  auto seek = [&]() {
    if (!a.skip_finds || !a.types) return;
    while (i < hi) {
      const u64 e = i + (u64)c.slot * ntiles;
      const u32 m = c.ballot(e < hi && a.types[request_index(a, e)] != op_find);
      if (m) {
        i += (u64)(__ffs(m) - 1) * ntiles;
        return;
      }
      i += 16 * ntiles;
    }
  };
  seek();
  bool active = i < hi;
  auto start = [&]() {
    ri = (u32)request_index(a, i);
    load_request<T>(a, ri, rq);
    const bool writes = rq.kind == op_insert || rq.kind == op_update || rq.kind == op_upsert;
    const bool bad_len = rq.key_length == 0 || rq.key_length > a.max_key_length ||
                         (writes && rq.value_length > a.max_value_length);
    if (bad_len) rq.key_length = 0;  // nothing of the request is read
    c.load_key_cache(rq);
    op_begin(c, st, tv.root);
    if (a.hints && rq.kind != op_find && !bad_len) {
      // start at the border node the read-only pass reached (B-link: a node that split
      // since is left by moving right; one that became interior is descended)
      const u32* hp = a.hints + (u64)kHintWords * (i - lo);
      const u32 hn = hp[0], hr = hp[1], hpar = hp[2], hl = hp[3];
      const u32 ns = tv.pool.num_slabs;
      if (hn < ns && hr < ns && hl < rq.key_length)
        op_begin_at(st, hn, hr, hpar < ns ? hpar : kInvalid, hl);
    }
    if (bad_len) {
      c.set_error(kErrBadLength);
      op_finish(st, 0);
    } else if (abort_pending(tv.err)) {
      op_finish(st, 0);
    }
  };
  if (active) start();
  // A tile whose step could not progress (lock busy, torn read) sits out a randomized,
  // exponentially growing number of iterations before retrying, so hot nodes are not
  // hammered by thousands of tiles; waiting iterations count toward the step budget.
  u32 backoff = 0, wait = 0;
  while (active) {
    bool progressed = true;
    if (wait > 0) {
      wait -= 1;
      st.steps += 1;
      progressed = false;
    } else if (st.phase != ph_done) {
      progressed = op_step<Validate>(c, rq, st);
      if (progressed) {
        backoff = 0;
      } else {
        const u32 cap = op_holds_locks(st) ? kHolderBackoffShift : kMaxBackoffShift;
        backoff = backoff < cap ? backoff + 1 : cap;
        const u32 span = 1u << backoff;
        const u32 jitter = (u32)(i * 0x9E3779B1u) ^ (st.steps * 0x85EBCA6Bu);
        wait = span - 1 + ((jitter >> 7) & (span - 1));
      }
    }
    if (st.phase == ph_done) {
      backoff = 0;
      wait = 0;
      if (c.lane == 0) {
        if (rq.kind == op_find) {
          if (a.out_value_lengths) a.out_value_lengths[ri] = st.result;
        } else if (a.results) {
          a.results[ri] = st.result != 0;
        }
      }
      i += ntiles;
      seek();
      active = i < hi;
      if (active) start();
    }
    // every tile of this wave is waiting: let the other waves of the SIMD run
    if (__ballot(!progressed) == __ballot(true)) __builtin_amdgcn_s_sleep(1);
  }
  c.flush();
}

// ---------------------------------------------------------------------------------
// Batched form: one persistent launch applies a sequence of mixed batches in order
// (batch b + 1 sees every effect of batch b).  Requests of all batches are concatenated;
// offsets[b] .. offsets[b + 1] delimit batch b.  Tiles take the requests of the current
// batch by static stride.  Between batches runs a two-level grid barrier:
//   * a tile with no request left adds 1 to its group's arrival counter (the tiles of
//     block g of the grid are in group g % groups, each counter on its own 128-B line,
//     so arrivals spread over `groups` lines instead of queueing on one);
//   * the tile that completes its group adds 1 to the top counter; the one that
//     completes the top counter stores b + 1 into every group's release line;
//   * waiting tiles poll their group's release line and move to the next batch.
// No LDS is used: the kernel keeps the batch kernel's register and LDS budget.
// Counters only grow during a launch, so no reset is needed between batches.  The
// barrier steps run inside the same flat loop as the operations, so waiting tiles never
// stall working sibling tiles of their wave.  Records retired by any batch are drained
// once, after the launch.
constexpr u32 kMaxBarrierGroups = 64;
constexpr u32 kLineWords = 32;
// words of the barrier area: arrival lines, top line, release lines
constexpr u32 kBarrierWords = (2 * kMaxBarrierGroups + 1) * kLineWords;

struct multi_args {
  batch_args a;         // a.n = total requests of all batches
  const u64* offsets;   // num_batches + 1 request offsets
  u32 num_batches;
  u32 groups;           // barrier groups, 1 .. kMaxBarrierGroups, <= gridDim.x
  u32* barrier;         // kBarrierWords words, zero at launch
};

constexpr u32 kPollEvery = 8;                         // barrier polls: one iteration in 8
constexpr u64 kBarrierTicks = 20ull * 100000000ull;   // 20 s at the 100 MHz wall clock

// The batch index and its bounds are uniform over a wave (its tiles advance together),
// so they are kept in scalar registers.
template <u32 T>
__global__ void __launch_bounds__(kBlockThreads) __attribute__((amdgpu_waves_per_eu(5)))
multi_batch_kernel(tree_view tv, multi_args m) {
  wave_tile<T, true> c(tv.pool, tv.retire, tv.err, tv.root);
  c.sc_border = tv.shortcut_border;
  c.sc_epoch = tv.shortcut_epoch;
  c.sc_shift = tv.shortcut_shift;
  const batch_args& a = m.a;
  const u32 tiles_per_block = kBlockThreads / T;
  const u64 ntiles = (u64)gridDim.x * tiles_per_block;
  const u32 tile_in_block = threadIdx.x / T;
  const u64 gtile = (u64)blockIdx.x * tiles_per_block + tile_in_block;
  op_state<gpu_node> st;
  request rq;
  u32 b = 0;  // current batch (wave-uniform)
  // offsets are clamped to a.n so a malformed offsets array cannot index past the
  // request arrays
  u64 end = min(m.offsets[1], a.n);
  u64 i = min(m.offsets[0], a.n) + gtile;
  bool have = i < end;
  bool marked = false;
  u32 backoff = 0, wait = 0;  // backoff wait while working; barrier iterations while not
  u64 wait_t0 = 0;            // wall clock when a long barrier wait started
  auto start = [&]() {
    load_request<T>(a, i, rq);
    const bool writes = rq.kind == op_insert || rq.kind == op_update || rq.kind == op_upsert;
    const bool bad_len = rq.key_length == 0 || rq.key_length > a.max_key_length ||
                         (writes && rq.value_length > a.max_value_length);
    if (bad_len) rq.key_length = 0;
    c.load_key_cache(rq);
    op_begin(c, st, tv.root);
    if (bad_len) {
      c.set_error(kErrBadLength);
      op_finish(st, 0);
    } else if (abort_pending(tv.err)) {
      op_finish(st, 0);
    }
  };
  if (have) start();
  while (b < m.num_batches) {
    bool progressed = true;
    if (have) {
      if (wait > 0) {
        wait -= 1;
        st.steps += 1;
        progressed = false;
      } else if (st.phase != ph_done) {
        progressed = op_step<true>(c, rq, st);
        if (progressed) {
          backoff = 0;
        } else {
          const u32 cap = op_holds_locks(st) ? kHolderBackoffShift : kMaxBackoffShift;
          backoff = backoff < cap ? backoff + 1 : cap;
          const u32 span = 1u << backoff;
          const u32 jitter = (u32)(i * 0x9E3779B1u) ^ (st.steps * 0x85EBCA6Bu);
          wait = span - 1 + ((jitter >> 7) & (span - 1));
        }
      }
      if (st.phase == ph_done) {
        backoff = 0;
        wait = 0;
        if (c.lane == 0) {
          if (rq.kind == op_find) {
            if (a.out_value_lengths) a.out_value_lengths[i] = st.result;
          } else if (a.results) {
            a.results[i] = st.result != 0;
          }
        }
        i += ntiles;
        have = i < end;
        if (have) start();
      }
    } else {
      progressed = false;
      const u32 g = blockIdx.x % m.groups;
      if (!marked) {
        // arrive: one count per tile on the group's line; the tile that completes its
        // group counts on the top line, and the one that completes the top line
        // releases every group
        marked = true;
        wait = 0;
        if (c.lane == 0) {
          dev::order();  // this tile's tree writes are complete before it is counted
          const u32 members =
              (gridDim.x / m.groups + (g < gridDim.x % m.groups ? 1u : 0u)) * tiles_per_block;
          const u32 arrived = dev::atomic_add(m.barrier + g * kLineWords, 1u) + 1u;
          if (arrived == (b + 1) * members) {
            u32* top = m.barrier + kMaxBarrierGroups * kLineWords;
            const u32 done_groups = dev::atomic_add(top, 1u) + 1u;
            if (done_groups == (b + 1) * m.groups) {
              dev::order();
              u32* rel = top + kLineWords;
              for (u32 k = 0; k < m.groups; k++) dev::coherent_store(rel + k * kLineWords, b + 1);
            }
          }
        }
      }
      // Poll the group's release line on one iteration in kPollEvery: the tiles of a
      // wave share the line (one request per wave), so they read one value and advance
      // together, and waiting waves do not crowd the memory channels that the last
      // working tiles need.  The wait is bounded in time (kBarrierTicks of the 100 MHz
      // wall clock), not in iterations, whose length depends on that contention.
      wait += 1;
      u32 rel = b;
      if ((wait & (kPollEvery - 1)) == 1) {
        u32 r = 0;
        if (c.lane == 0) r = dev::coherent_load(m.barrier + (kMaxBarrierGroups + 1 + g) * kLineWords);
        rel = (u32)__builtin_amdgcn_readfirstlane(c.bcast(r, 0));
      }
      if (rel > b) {
        b = (u32)__builtin_amdgcn_readfirstlane(b + 1);
        marked = false;
        wait = 0;
        if (b < m.num_batches) {
          end = min(m.offsets[b + 1], a.n);
          i = min(m.offsets[b], a.n) + gtile;
          have = i < end;
          if (have) start();
        }
        progressed = true;
      } else if ((wait & 0xFFFu) == 0) {
        const u64 now = (u64)wall_clock64();
        if (wait == 0x1000u) wait_t0 = now;
        else if (now - wait_t0 > kBarrierTicks) {  // a tile never arrived: give up
          c.set_error(kErrStepLimit);
          break;
        }
      }
    }
    if (__ballot(!progressed) == __ballot(true)) __builtin_amdgcn_s_sleep(1);
  }
  c.flush();
}

// Builds the layer-0 shortcut table (tree_view::shortcut): entry i is the level-1 node
// (the root when the tree has fewer levels) reached from the root by the smallest first
// slice of prefix i.  Runs with no writer in the launch; the walk moves right past bounds
// below the slice as any descent does.
__global__ void shortcut_build_kernel(tree_view tv, u32* table, u32* border_table,
                                      u32* built_epoch, u32 entries, u32 shift) {
  const u32 i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i == 0) *built_epoch = dev::coherent_load(tv.err + kMergeEpochWord);
  if (i >= entries) return;
  const u32 q = i << shift;
  u32 node = tv.root, level1 = tv.root;
  for (u32 guard = 0; guard < 4096; guard++) {
    const u64* p = tv.pool.pool + (u64)node * kSlots;
    const u32 h = slot_lo(p[kMetaSlot]);
    if (hdr_border(h)) break;
    if (hdr_level(h) <= 1) level1 = node;
    const u32 nk = hdr_nkeys(h);
    u32 next = kInvalid;
    for (u32 j = 0; j < nk && j < kInteriorCap; j++) {
      const u64 e = p[j];
      if (q <= slot_lo(e)) {
        next = child_of(e);
        break;
      }
    }
    if (next == kInvalid) next = slot_hi(p[kMetaSlot]);  // right of every bound: sibling
    if (next >= tv.pool.num_slabs) {
      node = level1 = tv.root;  // not expected in a quiescent tree: fall back to the root
      break;
    }
    node = next;
  }
  table[i] = level1;
  border_table[i] = node;
}

__global__ void zero_kernel(u32* p, u32 n) {
  for (u32 k = threadIdx.x; k < n; k += blockDim.x) p[k] = 0;
}

// ---------------------------------------------------------------------------------
// Range scan (read-only).  Outputs per query q at slot q * max_count + c: values
// (max_value_length slices), value lengths, keys (max_key_length slices), key
// lengths; counts[q].  Semantics follow the reference's cooperative_scan: ascending
// keys k with lower <= k (inclusive) and, with upper keys, k <= upper (inclusive).
// Per tile, `stack` holds 3 words per trie layer: (slice, border node, upper-bound
// flag) of every LINK entry on the current path; the key prefix is its slices.
struct scan_args {
  const u32* lower_keys;
  const u32* lower_key_lengths;
  u32 max_key_length;
  u32 max_count;
  u64 num_queries;
  const u32* upper_keys;
  const u32* upper_key_lengths;
  u32* counts;
  u32* values;
  u32 max_value_length;
  u32* value_lengths;
  u32* out_keys;
  u32* out_key_lengths;
  const u32* limits;  // optional per-query limit (<= max_count)
  u32* stack;         // per-tile layer stack, stack_layers * 3 words per tile
  u32 stack_layers;
};

template <u32 T>
struct scan_query {
  const u32* lo;
  u32 nlo;
  const u32* hi;
  u32 nhi;
  bool use_hi;
  u32 limit;
  u32* vals;
  u32* vlens;
  u32* keys;
  u32* klens;
  u32 maxV;
  u32 maxK;
};

// Layer stack of a scan: 3 words per trie layer entered (slice of the LINK entry,
// border node holding it, upper-bound flag); the key prefix of an emitted entry is the
// stacked slices.  Every lane of the tile calls these (uniform arguments); lane 0
// writes.
// Flat form: the tile's region of a scratch buffer (scan_kernel).
struct flat_scan_stack {
  u32* base;
  u32 layers;
  template <class C>
  __device__ __forceinline__ bool reserve(C&, u32 l) { return l + 1 < layers; }
  __device__ __forceinline__ u32* at(u32 l) const { return base + 3u * l; }
  template <class C>
  __device__ __forceinline__ void release(C&) {}
};
// Slab form (device API): 10 layers per 128-byte slab taken from the tile's allocator,
// word 30 of a slab links the next; slabs are taken as the scan goes deeper and freed
// when it ends.  Nothing else ever sees them.
struct slab_scan_stack {
  static constexpr u32 kPer = 10, kNext = 30;
  u64* pool = nullptr;
  u32 head = kInvalid;
  u32 slabs = 0;
  __device__ __forceinline__ u32* slab_words(u32 idx) const {
    return reinterpret_cast<u32*>(pool + (u64)idx * kSlots);
  }
  __device__ __forceinline__ u32 nth(u32 k) const {
    u32 idx = head;
    for (u32 i = 0; i < k; i++) idx = slab_words(idx)[kNext];
    return idx;
  }
  template <class C>
  __device__ __forceinline__ bool reserve(C& c, u32 l) {
    pool = c.pool.pool;
    while (l / kPer >= slabs) {
      const u32 s = c.alloc_slab();
      if (s == kInvalid) return false;
      if (c.lane == 0) {
        if (slabs == 0) head = s;
        else slab_words(nth(slabs - 1))[kNext] = s;
      }
      head = c.bcast(head, 0);
      slabs += 1;
      __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");
    }
    return true;
  }
  __device__ __forceinline__ u32* at(u32 l) const { return slab_words(nth(l / kPer)) + 3u * (l % kPer); }
  template <class C>
  __device__ __forceinline__ void release(C& c) {
    u32 idx = head;
    for (u32 k = 0; k < slabs; k++) {
      const u32 nxt = slab_words(idx)[kNext];
      if (c.lane == 0) dev::free_slab(c.pool, idx);
      idx = nxt;
    }
    slabs = 0;
    head = kInvalid;
  }
};

// number of entries e with (e.slice, e.more) <= (s, m)
template <u32 T, bool C>
__device__ __forceinline__ u32 entries_le(const wave_tile<T, C>& c, const gpu_node& n, u32 s, bool m) {
  const u32 nk = c.nkeys_of(n);
  const u32 myks = ks_get(c.ksbits(n), c.slot);
  return (u32)__builtin_popcount(
      c.ballot(c.slot < nk && !entry_less(s, m, slot_lo(n.s), ks_more(myks))));
}

// compares record tail (record head loaded in c) with key[off, off+len): -1, 0, +1
template <u32 T, bool C>
__device__ __forceinline__ int rec_tail_cmp(wave_tile<T, C>& c, const u32* key, u32 off, u32 len,
                                            u32 len_word) {
  const u32 tl = rec_tail_len(len_word);
  const u32 common = tl < len ? tl : len;
  const u32 need = 1 + common;
  u64 cur = c.rec_slot;
  for (u32 base = 0;; base += kRecordWordsPerSlab) {
    if (base >= need) break;
    const u32 d0 = base + 2 * c.slot, d1 = d0 + 1;
    const bool v0 = d0 >= 1 && d0 < need;
    const bool v1 = (2 * c.slot + 1 < kRecordWordsPerSlab) && d1 >= 1 && d1 < need;
    const u32 k0 = v0 ? key[off + d0 - 1] : 0u;
    const u32 k1 = v1 ? key[off + d1 - 1] : 0u;
    const bool ne0 = v0 && k0 != slot_lo(cur);
    const bool ne1 = v1 && k1 != slot_hi(cur);
    // first differing word index: word order is lane0.lo, lane0.hi, lane1.lo, ...
    const u32 m0 = c.ballot(ne0), m1 = c.ballot(ne1);
    if (m0 | m1) {
      const u32 l0 = m0 ? (u32)__builtin_ctz(m0) : 99u;
      const u32 l1 = m1 ? (u32)__builtin_ctz(m1) : 99u;
      const bool use_lo = l0 <= l1;
      const u32 ln = use_lo ? l0 : l1;
      const u32 rw = use_lo ? c.bcast(slot_lo(cur), ln) : c.bcast(slot_hi(cur), ln);
      const u32 kw = use_lo ? c.bcast(k0, ln) : c.bcast(k1, ln);
      return rw < kw ? -1 : 1;
    }
    if (base + kRecordWordsPerSlab >= need) break;
    if (!c.rec_next(cur)) return 0;
  }
  return tl < len ? -1 : (tl > len ? 1 : 0);
}

template <u32 T, bool Validate, bool C>
__device__ __forceinline__ void scan_read(wave_tile<T, C>& c, u32 idx, gpu_node& n) {
  for (u32 spin = 0; spin < (1u << 20); spin++) {
    if (c.template read_node<Validate>(idx, n)) return;
    if (c.failed()) return;
    __builtin_amdgcn_s_sleep(1);
  }
  c.set_error(kErrStepLimit);
  c.fatal = true;
}

// Returns the number of results.
template <u32 T, bool Validate, bool C, class Stack>
__device__ u32 tile_scan(wave_tile<T, C>& c, u32 root, const scan_query<T>& q, Stack& stk) {
  if (q.limit == 0) return 0;
  u32 count = 0;
  u32 l = 0;
  bool lo = true;       // lower bound constrains this layer
  bool hi = q.use_hi;   // upper bound constrains this layer
  bool resume = false;  // continue after (rs, more) in the stacked border node
  u32 rs = 0;
  u32 node = root;
  gpu_node n;
  for (u32 guard = 0; guard < (1u << 20); guard++) {
    if (c.failed()) return count;
    // ---- position on the border node that holds the start of this layer's range
    const u32 target = resume ? rs : (lo ? q.lo[l] : 0u);
    while (true) {
      scan_read<T, Validate>(c, node, n);
      if (c.failed()) return count;
      const u32 h = c.header(n);
      if (!hdr_border(h)) {
        u32 child = c.find_child(n, target);
        node = (child != kInvalid) ? child : c.sibling(n);
        continue;
      }
      if (target > c.high(n) && c.sibling(n) != kInvalid) {
        node = c.sibling(n);
        continue;
      }
      break;
    }
    const bool lmore = l + 1 < q.nlo;
    const bool umore = l + 1 < q.nhi;
    const u32 uslice = hi ? q.hi[l] : 0u;
    u32 p;
    if (resume) p = entries_le(c, n, rs, true);
    else if (lo) p = c.insert_position(n, q.lo[l], lmore);
    else p = 0;
    bool first_node = true;
    bool descended = false;
    while (true) {
      const u32 nk = c.nkeys_of(n);
      const u32 ksb = c.ksbits(n);
      const u32 end = hi ? min(nk, entries_le(c, n, uslice, umore)) : nk;
      const bool hi_stop = hi && (end < nk || c.high(n) >= uslice);
      // lower-bound check of a suffix entry equal to (lo[l], more)
      if (lo && !resume && first_node && p < end && lmore) {
        const u32 s0 = c.bcast(slot_lo(n.s), p);
        const u32 k0 = ks_get(ksb, p);
        if (s0 == q.lo[l] && k0 == kSuffix) {
          u32 lw;
          if (!c.rec_load(c.bcast(slot_hi(n.s), p), lw)) return count;
          if (rec_tail_cmp(c, q.lo, l + 1, q.nlo - l - 1, lw) < 0) p += 1;
        }
      }
      // first link entry in [p, end)
      const u32 myks = ks_get(ksb, c.slot);
      const u32 links = c.ballot(c.slot >= p && c.slot < end && myks == kLink);
      const u32 lk = links ? (u32)__builtin_ctz(links) : end;
      u32 out_end = lk;
      bool stop_after = false;
      // upper-bound check of a suffix entry equal to (hi[l], true)
      if (hi && umore && out_end > p) {
        const u32 e = out_end - 1;
        const u32 se = c.bcast(slot_lo(n.s), e);
        if (se == uslice && ks_get(ksb, e) == kSuffix) {
          u32 lw;
          if (!c.rec_load(c.bcast(slot_hi(n.s), e), lw)) return count;
          if (rec_tail_cmp(c, q.hi, l + 1, q.nhi - l - 1, lw) > 0) {
            out_end = e;
            stop_after = true;
          }
        }
      }
      // emit entries [p, out_end), at most limit - count.  Lane e of the tile (e < 16)
      // emits entry e itself from the first 64 bytes of its record, so the records of a
      // node's entries are read at once; a record of more than 16 words is emitted by the
      // whole tile afterwards, one at a time.
      u32 take = out_end > p ? out_end - p : 0u;
      if (take > q.limit - count) take = q.limit - count;
      {
        const u32 e = c.slot;
        const bool mine = c.lane < 16 && e >= p && e < p + take;
        const u32 slice = slot_lo(n.s), val = slot_hi(n.s);
        const u32 ks = ks_get(ksb, e);
        if (c.ballot(mine && ks != kValue && val >= c.pool.num_slabs)) {
          c.set_error(kErrBadIndex);
          c.fatal = true;
          return count;
        }
        bool longrec = false;
        if (mine) {
          const u32 o = count + (e - p);
          if (q.keys) {
            u32* dst = q.keys + (u64)o * q.maxK;
            for (u32 j = 0; j <= l && j < q.maxK; j++) dst[j] = (j == l) ? slice : stk.at(j)[0];
          }
          if (ks == kValue) {
            if (q.vlens) q.vlens[o] = 1;
            if (q.klens) q.klens[o] = l + 1;
            if (q.vals && q.maxV >= 1) q.vals[(u64)o * q.maxV] = val;
          } else {
            u64 w[8];
#pragma unroll
            for (u32 i = 0; i < 8; i++) w[i] = c.ld(c.slab(val) + i);
            const u32 lw = slot_lo(w[0]);
            const u32 tl = rec_tail_len(lw), vl = rec_value_len(lw);
            const u32 total = 1 + tl + vl;
            if (q.vlens) q.vlens[o] = vl;
            if (q.klens) q.klens[o] = l + 1 + tl;
            if (total <= 16) {
#pragma unroll
              for (u32 d = 1; d < 16; d++) {
                const u32 word = (d & 1) ? slot_hi(w[d / 2]) : slot_lo(w[d / 2]);
                if (d >= total) continue;
                if (d <= tl) {
                  if (q.keys && l + d < q.maxK) q.keys[(u64)o * q.maxK + l + d] = word;
                } else if (q.vals && d - 1 - tl < q.maxV) {
                  q.vals[(u64)o * q.maxV + d - 1 - tl] = word;
                }
              }
            } else {
              longrec = true;
            }
          }
        }
        for (u32 lm = c.ballot(longrec); lm; lm &= lm - 1) {
          const u32 e2 = (u32)__builtin_ctz(lm);
          const u32 o = count + (e2 - p);
          u32 lw;
          if (!c.rec_load(c.bcast(slot_hi(n.s), e2), lw)) return count;
          const u32 tl = rec_tail_len(lw), vl = rec_value_len(lw);
          // record words: tail -> key slices l+1.., value -> value slots
          const u32 total = 1 + tl + vl;
          u64 cur = c.rec_slot;
          for (u32 base = 0; base < total; base += kRecordWordsPerSlab) {
            if (c.lane < 16) {
              for (u32 hh = 0; hh < 2; hh++) {
                const u32 w = 2 * c.slot + hh;
                if (w >= kRecordWordsPerSlab) continue;
                const u32 d = base + w;
                if (d == 0 || d >= total) continue;
                const u32 word = hh ? slot_hi(cur) : slot_lo(cur);
                if (d <= tl) {
                  const u32 kk = l + d;
                  if (q.keys && kk < q.maxK) q.keys[(u64)o * q.maxK + kk] = word;
                } else {
                  const u32 vv = d - 1 - tl;
                  if (q.vals && vv < q.maxV) q.vals[(u64)o * q.maxV + vv] = word;
                }
              }
            }
            if (base + kRecordWordsPerSlab >= total) break;
            if (!c.rec_next(cur)) return count;
          }
        }
      }
      count += take;
      if (count >= q.limit || stop_after) return count;
      if (lk < end) {
        // descend into the next layer
        if (!stk.reserve(c, l)) {
          c.set_error(kErrScanScratch);
          return count;
        }
        const u32 slice = c.bcast(slot_lo(n.s), lk);
        const u32 child = c.bcast(slot_hi(n.s), lk);
        u32* top = stk.at(l);
        if (c.lane == 0) {
          top[0] = slice;
          top[1] = node;
          top[2] = hi ? 1u : 0u;
        }
        // other lanes of the tile read the stack later (key prefixes, pops)
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");
        const bool at_lo = lo && !resume && first_node && lk == p && slice == q.lo[l] && lmore;
        lo = at_lo;
        hi = hi && slice == uslice && umore;
        l += 1;
        node = child;
        resume = false;
        descended = true;
        break;
      }
      if (hi_stop) return count;
      const u32 sib = c.sibling(n);
      if (sib == kInvalid) break;  // layer exhausted
      node = sib;
      scan_read<T, Validate>(c, node, n);
      if (c.failed()) return count;
      p = 0;
      first_node = false;
      lo = false;
      resume = false;
    }
    if (descended) continue;
    // pop to the previous layer
    if (l == 0) return count;
    l -= 1;
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
    const u32* top = stk.at(l);
    rs = top[0];
    node = top[1];
    hi = top[2] != 0;
    lo = false;
    resume = true;
  }
  c.set_error(kErrStepLimit);
  return count;
}

template <u32 T, bool Validate>
__global__ void __launch_bounds__(kBlockThreads) scan_kernel(tree_view tv, scan_args a) {
  wave_tile<T, Validate> c(tv.pool, tv.retire, tv.err);
  const u32 tiles_per_block = kBlockThreads / T;
  const u64 ntiles = (u64)gridDim.x * tiles_per_block;
  const u64 gtile = (u64)blockIdx.x * tiles_per_block + threadIdx.x / T;
  for (u64 i = gtile; i < a.num_queries; i += ntiles) {
    scan_query<T> q;
    q.nlo = a.lower_key_lengths ? a.lower_key_lengths[i] : a.max_key_length;
    q.lo = a.lower_keys + i * (u64)a.max_key_length;
    q.use_hi = a.upper_keys != nullptr;
    q.hi = q.use_hi ? a.upper_keys + i * (u64)a.max_key_length : nullptr;
    q.nhi = q.use_hi ? (a.upper_key_lengths ? a.upper_key_lengths[i] : a.max_key_length) : 0u;
    u32 lim = a.max_count;
    if (a.limits && a.limits[i] < lim) lim = a.limits[i];
    q.limit = lim;
    const u64 slot0 = i * (u64)a.max_count;
    q.vals = a.values ? a.values + slot0 * a.max_value_length : nullptr;
    q.vlens = a.value_lengths ? a.value_lengths + slot0 : nullptr;
    q.keys = a.out_keys ? a.out_keys + slot0 * a.max_key_length : nullptr;
    q.klens = a.out_key_lengths ? a.out_key_lengths + slot0 : nullptr;
    q.maxV = a.max_value_length;
    q.maxK = a.max_key_length;
    flat_scan_stack stk{a.stack + gtile * (u64)a.stack_layers * 3u, a.stack_layers};
    u32 cnt = 0;
    if (q.nlo == 0 || q.nlo > a.max_key_length || (q.use_hi && (q.nhi == 0 || q.nhi > a.max_key_length))) {
      c.set_error(kErrBadLength);
    } else {
      cnt = tile_scan<T, Validate>(c, tv.root, q, stk);
    }
    if (c.lane == 0 && a.counts) a.counts[i] = cnt;
  }
}

// ---------------------------------------------------------------------------------
__global__ void drain_kernel(slab_pool_view pool, retire_view rv) {
  const u32 raw = *rv.count;
  const u32 count = raw < rv.capacity ? raw : rv.capacity;
  const u32 stride = gridDim.x * blockDim.x;
  for (u32 i = blockIdx.x * blockDim.x + threadIdx.x; i < count; i += stride)
    dev::free_retired(pool, rv.list[i]);
}

__global__ void set_error_kernel(u32* err, u32 code) { dev::set_error(err, code); }

__global__ void reset_kernel(retire_view rv) {
  if (threadIdx.x == 0 && blockIdx.x == 0) *rv.count = 0;
}

__global__ void init_root_kernel(slab_pool_view pool, u32* err, u32* out_root) {
  if (threadIdx.x >= 16) return;
  wave_tile<16, true> c(pool, retire_view{}, err);
  const u32 r = c.alloc_slab();
  if (r != kInvalid) {
    gpu_node n;
    n.s = 0;
    if (c.slot == kHighSlot) n.s = slot_make(kMaxSlice, 0);
    if (c.slot == kMetaSlot) n.s = slot_make(hdr_make(0, true, true, 0), kInvalid);
    c.write_node(r, n);
  }
  c.flush();
  if (c.lane == 0) *out_root = r;
}

// ---------------------------------------------------------------------------------
// Read-only find (the find<concurrent=false> batch path).  Each tile keeps K lookups in
// flight: every iteration issues the K node (or record) loads back to back, then
// advances each lookup by one node.  A wave of 64/T tiles therefore has 64K/T
// independent 128-byte loads outstanding per iteration, which is what the latency-bound
// descent needs.  Semantics are those of op_step's find path (layers, suffix records,
// long values), with plain cached loads as no writer runs during the launch.
constexpr u32 kFindDescend = 0, kFindRecord = 1, kFindIdle = 2;

template <u32 T, u32 K>
__global__ void __launch_bounds__(kBlockThreads) find_kernel(tree_view tv, batch_args a) {
  wave_tile<T, false> c(tv.pool, tv.retire, tv.err);
  const u32 tiles_per_block = kBlockThreads / T;
  const u64 stride = (u64)gridDim.x * tiles_per_block * K;
  const u64 first = ((u64)blockIdx.x * tiles_per_block + threadIdx.x / T) * K;
  u64 idx[K];
  u32 node[K], layer[K], phase[K], klen[K], kc[K], rkind[K];
  const u32 maxK = a.max_key_length, maxV = a.max_value_length;

  auto start = [&](u32 k) {
    phase[k] = kFindIdle;
    while (idx[k] < a.n) {
      const u64 i = idx[k];
      const u32 L = a.key_lengths ? a.key_lengths[i] : maxK;
      if (L == 0 || L > maxK) {
        c.set_error(kErrBadLength);
        if (c.lane == 0) a.out_value_lengths[i] = 0;
        idx[k] += stride;
        continue;
      }
      klen[k] = L;
      const u32* key = a.keys + i * (u64)maxK;
      kc[k] = (c.lane < L) ? key[c.lane] : 0u;
      node[k] = tv.root;
      layer[k] = 0;
      phase[k] = kFindDescend;
      return;
    }
  };
  auto finish = [&](u32 k, u32 len) {
    if (c.lane == 0) a.out_value_lengths[idx[k]] = len;
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
    // issue the K loads together
    gpu_node n[K];
#pragma unroll
    for (u32 k = 0; k < K; k++) {
      n[k].s = 0;
      if (phase[k] != kFindIdle && node[k] < tv.pool.num_slabs)
        n[k].s = c.slab(node[k])[c.slot];
    }
#pragma unroll
    for (u32 k = 0; k < K; k++) {
      if (phase[k] == kFindIdle) continue;
      if (node[k] >= tv.pool.num_slabs) {
        c.set_error(kErrBadIndex);
        finish(k, 0);
        continue;
      }
      const u64 i = idx[k];
      const u32 l = layer[k];
      const u32* key = a.keys + i * (u64)maxK;
      if (phase[k] == kFindRecord) {
        c.rec_slot = n[k].s;
        c.rec_idx = node[k];
        const u32 lw = c.bcast(slot_lo(n[k].s), 0);
        request rq;
        rq.key = key;
        rq.key_length = klen[k];
        rq.out_value = a.out_values + i * (u64)maxV;
        rq.max_value_length = maxV;
        const u32 saved = c.key_cache;
        c.key_cache = kc[k];
        bool hit = true;
        if (rkind[k] == kSuffix) hit = c.rec_tail_equal(key, l + 1, klen[k] - l - 1, lw);
        const u32 len = hit ? c.rec_copy_value(rq, lw) : 0u;
        c.key_cache = saved;
        finish(k, len);
        continue;
      }
      const u32 q = l < T ? c.bcast(kc[k], l) : key[l];
      const u32 h = c.header(n[k]);
      if (!hdr_border(h)) {
        u32 child = c.find_child(n[k], q);
        node[k] = (child != kInvalid) ? child : c.sibling(n[k]);
        continue;
      }
      if (q > c.high(n[k]) && c.sibling(n[k]) != kInvalid) {
        node[k] = c.sibling(n[k]);
        continue;
      }
      const bool more = l + 1 < klen[k];
      u32 ks = 0, val = 0;
      const int pos = c.lookup(n[k], q, more, ks, val);
      if (pos < 0) {
        finish(k, 0);
      } else if (ks == kLink) {
        layer[k] = l + 1;
        node[k] = val;
      } else if (ks == kValue) {
        if (c.lane == 0 && maxV >= 1) a.out_values[i * (u64)maxV] = val;
        finish(k, 1);
      } else {
        rkind[k] = ks;
        node[k] = val;
        phase[k] = kFindRecord;
      }
    }
  }
}

}  // namespace indexingpu_amd
