// CPU model of the Masstree batch engine (test infrastructure only).
//
// Runs the same per-tile state machine as the gfx950 kernels (tree_ops.hpp) with the
// CPU tile policy, either
//   * interleaved: many simulated tiles, one step at a time, tiles picked by a seeded
//     random scheduler (reproducible interleavings at step granularity), or
//   * threaded: one host thread per tile, real concurrency through GCC atomics.
// The host walker then validates the structure and dumps the entries in key order.
#include <cstdint>
#include <cstring>
#include <random>
#include <thread>
#include <vector>

#include "cpu_policy.hpp"
#include "indexingpu_amd/host_walk.hpp"

using namespace indexingpu_amd;
using cpu_model::cpu_tile;
using cpu_model::shared_pool;

namespace {
struct model {
  shared_pool pool;
  u32 root;
  u32 hinted = 0;  // writes start at hints from a read-only walk (as the GPU write passes)
  u32 no_suffix = 0;  // inserts build layers instead of suffix records (enable_suffix=false)
  u32 erase_flags = 0;  // kReqNoMerge | kReqKeepEmptyLayers | fanout bits for every request
  std::vector<u32> shortcut;  // border-level shortcut table (cm_build_shortcut)
  explicit model(u32 n) : pool(n) {
    cpu_tile t(pool);
    root = t.alloc_slab();
    cpu_model::node_img r;
    std::memset(&r, 0, sizeof(r));
    r.s[kHighSlot] = slot_make(kMaxSlice, 0);
    r.s[kMetaSlot] = slot_make(hdr_make(0, true, true, 0), kInvalid);
    t.write_node(root, r);
  }
};

struct batch_args {
  const uint8_t* kinds;
  uint32_t default_kind;
  const uint32_t* keys;
  uint32_t maxK;
  const uint32_t* klens;
  const uint32_t* values;
  uint32_t maxV;
  const uint32_t* vlens;
  uint32_t* out_values;
  uint32_t* out_vlens;
  uint8_t* results;
  uint32_t no_suffix;
  uint32_t erase_flags;
};

request make_request(const batch_args& a, uint64_t i, uint32_t* scratch_out) {
  request rq;
  rq.kind = a.kinds ? a.kinds[i] : a.default_kind;
  rq.key_length = a.klens ? a.klens[i] : a.maxK;
  rq.key = a.keys + i * (uint64_t)a.maxK;
  rq.value = a.values ? a.values + i * (uint64_t)a.maxV : nullptr;
  rq.value_length = a.vlens ? a.vlens[i] : a.maxV;
  rq.flags = (a.no_suffix ? kReqNoSuffix : 0u) | a.erase_flags;
  rq.out_value = a.out_values ? a.out_values + i * (uint64_t)a.maxV : scratch_out;
  rq.max_value_length = a.out_values ? a.maxV : 0;
  // a find in a mixed batch writes its value over the request's value slots
  if (rq.kind == op_find && a.values && !a.out_values) {
    rq.out_value = const_cast<uint32_t*>(rq.value);
    rq.max_value_length = a.maxV;
  }
  return rq;
}

void store_result(const batch_args& a, const request& rq, uint64_t i, uint32_t result) {
  if (rq.kind == op_find) {
    if (a.out_vlens) a.out_vlens[i] = result;
  } else if (a.results) {
    a.results[i] = (uint8_t)result;
  }
}

using state_t = op_state<cpu_model::node_img>;

// The start hint of a write, as the GPU's read-only pass (narrow_descend_kernel<T, K, true>)
// leaves it before the write pass: the border node the key reaches with no writer running,
// the root of that layer, the level-1 node passed on the way (kInvalid: none) and the layer.
struct start_hint {
  u32 node = kInvalid, root = kInvalid, par = kInvalid, layer = 0;
};

start_hint read_only_hint(cpu_tile& c, u32 root, const request& rq) {
  start_hint h;
  if (rq.key_length == 0) return h;
  u32 node = root, lroot = root, par = kInvalid, layer = 0;
  for (u32 guard = 0; guard < (1u << 24); guard++) {
    cpu_model::node_img n;
    if (!c.read_node<false>(node, n)) return h;
    const u32 hd = c.header(n);
    const u32 q = rq.key[layer];
    if (!hdr_border(hd)) {
      const u32 child = c.find_child(n, q);
      if (child != kInvalid && hdr_level(hd) == 1) par = node;
      node = child != kInvalid ? child : c.sibling(n);
      if (node == kInvalid) return h;
      continue;
    }
    if (q > c.high(n) && c.sibling(n) != kInvalid) {
      node = c.sibling(n);
      continue;
    }
    u32 ks = 0, val = 0;
    const int pos = c.lookup(n, q, layer + 1 < rq.key_length, ks, val);
    if (pos >= 0 && ks == kLink) {
      layer += 1;
      if (layer >= rq.key_length) return h;
      node = lroot = val;
      par = kInvalid;
      continue;
    }
    h.node = node;
    h.root = lroot;
    h.par = par;
    h.layer = layer;
    return h;
  }
  return h;
}

// hints of every write request of the batch (computed before any write runs)
std::vector<start_hint> batch_hints(model& M, const batch_args& a, uint64_t n) {
  std::vector<start_hint> hs(M.hinted ? n : 0);
  cpu_tile c(M.pool);
  uint32_t scratch = 0;
  for (uint64_t i = 0; i < hs.size(); i++) {
    request r = make_request(a, i, &scratch);
    if (r.kind != op_find) hs[i] = read_only_hint(c, M.root, r);
  }
  return hs;
}

// op_begin, then the hinted start of a write when there is a hint
void begin_request(model& M, cpu_tile& c, state_t& s, const request& r,
                   const std::vector<start_hint>& hs, uint64_t i) {
  op_begin(c, s, M.root);
  if (r.key_length == 0) {
    op_finish(s, 0);
    return;
  }
  if (i < hs.size() && r.kind != op_find && hs[i].node != kInvalid &&
      hs[i].layer < r.key_length)
    op_begin_at(s, hs[i].node, hs[i].root, hs[i].par, hs[i].layer);
}
}  // namespace

extern "C" {

void* cm_create(uint32_t num_slabs) { return new model(num_slabs); }
void cm_destroy(void* h) { delete static_cast<model*>(h); }
uint32_t cm_error(void* h) { return static_cast<model*>(h)->pool.error.load(); }
uint32_t cm_slabs_used(void* h) { return static_cast<model*>(h)->pool.next.load(); }
uint64_t cm_retired(void* h) { return static_cast<model*>(h)->pool.retired.size(); }
void cm_set_hinted(void* h, uint32_t on) { static_cast<model*>(h)->hinted = on; }
void cm_set_no_suffix(void* h, uint32_t on) { static_cast<model*>(h)->no_suffix = on; }
void cm_set_erase_flags(void* h, uint32_t f) { static_cast<model*>(h)->erase_flags = f; }
// The livelock guard's time limit in coarse ticks (10.24 us each; default about 21 s).
void cm_set_max_op_time(void* h, uint32_t ticks) { static_cast<model*>(h)->pool.max_op_time = ticks; }
// The guard's record of the operation it abandoned (kTripWords words, all 0 if none).
void cm_trip(void* h, uint32_t* out) {
  for (u32 k = 0; k < kTripWords; k++) out[k] = static_cast<model*>(h)->pool.trip[k];
}
uint32_t cm_root(void* h) { return static_cast<model*>(h)->root; }
// Sets (on != 0) or clears the lock bit of node idx, as a holder that never finishes would
// leave it; returns the header before.
uint32_t cm_set_lock(void* h, uint32_t idx, uint32_t on) {
  model& M = *static_cast<model*>(h);
  u32* hp = reinterpret_cast<u32*>(M.pool.slab(idx) + kMetaSlot);
  return on ? __atomic_fetch_or(hp, kLockBit, __ATOMIC_RELAXED)
            : __atomic_fetch_and(hp, ~kLockBit, __ATOMIC_RELAXED);
}
// Allocation failure injection (0 = off): every k-th allocation, and every k-th
// allocation of a split inserting into its parent.  Also clears the error word.
void cm_set_fail(void* h, uint32_t any_every, uint32_t parent_every) {
  model& M = *static_cast<model*>(h);
  M.pool.fail_any_every = any_every;
  M.pool.fail_parent_every = parent_every;
  M.pool.error.store(0);
}
// Trie layers reachable from the root.
uint64_t cm_layers(void* h) {
  model& M = *static_cast<model*>(h);
  host_walker w([&](uint32_t idx) -> const uint64_t* {
    return idx < M.pool.num_slabs ? M.pool.slab(idx) : nullptr;
  }, nullptr);
  return w.run(M.root).layers;
}
// Non-root nodes that no parent slot names (reachable through sibling links only).
uint64_t cm_unlinked(void* h) {
  model& M = *static_cast<model*>(h);
  host_walker w([&](uint32_t idx) -> const uint64_t* {
    return idx < M.pool.num_slabs ? M.pool.slab(idx) : nullptr;
  }, nullptr);
  return w.run(M.root).unlinked;
}
// Reachable nodes (which = 0), border nodes (1), empty non-root border nodes (2), the
// largest border entry count (3) and interior child count (4); merges of layer-0 border nodes so far (5).
uint64_t cm_nodes(void* h, uint32_t which) {
  model& M = *static_cast<model*>(h);
  if (which == 5) return M.pool.merges.load();
  host_walker w([&](uint32_t idx) -> const uint64_t* {
    return idx < M.pool.num_slabs ? M.pool.slab(idx) : nullptr;
  }, nullptr);
  const walk_stats s = w.run(M.root);
  switch (which) {
    case 0: return s.nodes;
    case 1: return s.border_nodes;
    case 2: return s.empty_borders;
    case 3: return s.max_border_keys;
    default: return s.max_interior_children;
  }
}

// The layer-0 border node reached from the root by first slice q (the walk of the
// device's shortcut_build_kernel: the first bound >= q, else the sibling).
static u32 border_of(model& M, u32 q) {
  u32 node = M.root;
  for (u32 guard = 0; guard < 4096; guard++) {
    const u64* p = M.pool.slab(node);
    const u32 h = slot_lo(p[kMetaSlot]);
    if (hdr_border(h)) break;
    u32 next = kInvalid;
    for (u32 j = 0; j < hdr_nkeys(h) && j < kInteriorCap; j++) {
      if (q <= slot_lo(p[j])) {
        next = child_of(p[j]);
        break;
      }
    }
    if (next == kInvalid) next = slot_hi(p[kMetaSlot]);
    if (next >= M.pool.num_slabs) return kInvalid;
    node = next;
  }
  return node;
}
// Builds a border-level shortcut table of 2^bits entries (entry i: the border node of
// first slice i << (32 - bits)); merges of later batches rewrite it (shortcut_patch).
void cm_build_shortcut(void* h, uint32_t bits) {
  model& M = *static_cast<model*>(h);
  const u32 entries = 1u << bits, shift = 32u - bits;
  M.shortcut.assign(entries, kInvalid);
  for (u32 i = 0; i < entries; i++) M.shortcut[i] = border_of(M, i << shift);
  M.pool.sc_border = M.shortcut.data();
  M.pool.sc_shift = shift;
  M.pool.sc_built = M.pool.merges.load();
}
// Entries of the table that are no valid start for their prefix: not a live layer-0
// border node (on the chain of sibling links from the leftmost one), or right of the
// node that holds the prefix's smallest slice.
uint64_t cm_check_shortcut(void* h) {
  model& M = *static_cast<model*>(h);
  std::vector<u32> chain, highs;
  std::vector<int64_t> pos(M.pool.num_slabs, -1);
  for (u32 n = border_of(M, 0); n < M.pool.num_slabs && chain.size() < M.pool.num_slabs;
       n = slot_hi(M.pool.slab(n)[kMetaSlot])) {
    pos[n] = (int64_t)chain.size();
    chain.push_back(n);
    highs.push_back(slot_lo(M.pool.slab(n)[kHighSlot]));
  }
  uint64_t bad = 0;
  const u32 shift = M.pool.sc_shift;
  size_t t = 0;  // the chain node holding q: the first whose high is >= q (q ascends)
  for (size_t i = 0; i < M.shortcut.size(); i++) {
    const u32 q = (u32)i << shift, x = M.shortcut[i];
    while (t + 1 < chain.size() && highs[t] < q) t++;
    if (x >= M.pool.num_slabs || pos[x] < 0 || (size_t)pos[x] > t) bad++;
  }
  return bad;
}

// Runs one batch.  nthreads == 0: interleaved simulation with ntiles tiles and the
// given scheduler seed; nthreads > 0: that many host threads (static partition).
void cm_batch(void* h, const uint8_t* kinds, uint32_t default_kind, const uint32_t* keys,
              uint32_t maxK, const uint32_t* klens, const uint32_t* values, uint32_t maxV,
              const uint32_t* vlens, uint32_t* out_values, uint32_t* out_vlens,
              uint8_t* results, uint64_t n, uint32_t ntiles, uint32_t nthreads,
              uint64_t seed) {
  model& M = *static_cast<model*>(h);
  batch_args a{kinds, default_kind, keys, maxK, klens, values, maxV, vlens,
               out_values, out_vlens, results, M.no_suffix, M.erase_flags};
  const std::vector<start_hint> hs = batch_hints(M, a, n);
  if (nthreads == 0) {
    std::mt19937_64 rng(seed);
    std::vector<state_t> st(ntiles);
    std::vector<uint64_t> cur(ntiles);
    std::vector<request> rq(ntiles);
    std::vector<cpu_tile> ctx;
    ctx.reserve(ntiles);
    std::vector<uint32_t> scratch(ntiles, 0);
    std::vector<uint32_t> active;
    for (uint32_t t = 0; t < ntiles; t++) {
      ctx.emplace_back(M.pool, M.root);
      cur[t] = t;
      if (t < n) {
        rq[t] = make_request(a, t, &scratch[t]);
        begin_request(M, ctx[t], st[t], rq[t], hs, t);
        active.push_back(t);
      }
    }
    while (!active.empty()) {
      size_t pick = rng() % active.size();
      uint32_t t = active[pick];
      if (st[t].phase != ph_done) {
        ctx[t].in_parent = st[t].phase == ph_parent_lock;
        op_step<true>(ctx[t], rq[t], st[t]);
      }
      if (st[t].phase == ph_done) {
        store_result(a, rq[t], cur[t], st[t].result);
        cur[t] += ntiles;
        if (cur[t] < n) {
          rq[t] = make_request(a, cur[t], &scratch[t]);
          begin_request(M, ctx[t], st[t], rq[t], hs, cur[t]);
        } else {
          active[pick] = active.back();
          active.pop_back();
        }
      }
    }
    return;
  }
  std::vector<std::thread> th;
  for (uint32_t t = 0; t < nthreads; t++) {
    th.emplace_back([&, t]() {
      cpu_tile c(M.pool, M.root);
      state_t s;
      uint32_t scratch = 0;
      for (uint64_t i = t; i < n; i += nthreads) {
        request r = make_request(a, i, &scratch);
        begin_request(M, c, s, r, hs, i);
        while (s.phase != ph_done) {
          c.in_parent = s.phase == ph_parent_lock;
          if (!op_step<true>(c, r, s)) std::this_thread::yield();
        }
        store_result(a, r, i, s.result);
      }
    });
  }
  for (auto& x : th) x.join();
}

// Validates the structure and dumps entries in key order (same buffer protocol as
// orc_dump).  Returns the entry count, or ~0 on a structural error (message in errbuf).
uint64_t cm_dump(void* h, uint32_t* key_offsets, uint32_t* key_data, uint32_t* value_offsets,
                 uint32_t* value_data, uint64_t* key_slices, uint64_t* value_slices,
                 char* errbuf, uint32_t errlen, uint64_t* nodes) {
  model& M = *static_cast<model*>(h);
  uint64_t ks = 0, vs = 0, i = 0;
  host_walker w(
      [&](uint32_t idx) -> const uint64_t* {
        return idx < M.pool.num_slabs ? M.pool.slab(idx) : nullptr;
      },
      [&](const std::vector<uint32_t>& k, const std::vector<uint32_t>& v) {
        if (key_offsets) key_offsets[i] = (uint32_t)ks;
        if (value_offsets) value_offsets[i] = (uint32_t)vs;
        if (key_data) std::memcpy(key_data + ks, k.data(), k.size() * 4);
        if (value_data) std::memcpy(value_data + vs, v.data(), v.size() * 4);
        ks += k.size();
        vs += v.size();
        i++;
      });
  walk_stats s = w.run(M.root);
  if (key_offsets) key_offsets[i] = (uint32_t)ks;
  if (value_offsets) value_offsets[i] = (uint32_t)vs;
  if (key_slices) *key_slices = ks;
  if (value_slices) *value_slices = vs;
  if (nodes) *nodes = s.nodes;
  if (!s.error.empty()) {
    if (errbuf && errlen) {
      std::strncpy(errbuf, s.error.c_str(), errlen - 1);
      errbuf[errlen - 1] = 0;
    }
    return ~0ull;
  }
  return i;
}

}  // extern "C"
