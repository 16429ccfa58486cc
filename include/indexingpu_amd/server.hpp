// Served form: a request loop over one resident index (continuous batching).
//
// Requests arrive over time; each carries point operations (find / insert / erase /
// update, the mixed-batch layout) and range scans.  Their data sits in device memory
// ("arena"), back to back in arrival order, so a run of consecutive requests is one
// contiguous slice of it.  The loop
//   * receives the requests whose arrival time has come (trace replay in real time);
//   * admits them in arrival order while their budgets fit: point ops and scans per
//     wave, free slabs (every write reserves slabs until the next pool census) and
//     scan output rows in a ring of fixed size;
//   * launches an admitted run of requests as one wave -- one mixed-batch launch and
//     one scan launch -- on any of `streams` HIP streams that is idle, so waves overlap
//     on the device and requests that arrive while others run join the next wave;
//   * completes a wave's requests when its event fires (latency = completion time -
//     arrival time) and frees its scan rows.
// With coalesce = false it runs the plain path instead: the index's own mixed_batch /
// scan calls on one stream, in arrival order, one call pair at a time -- one request
// per pair, or (plain_batch) every request that arrived while the previous pair ran
// (interval batching, the natural way to serve with the plain calls).
//
// Ordered waves (ordered = 1).  The point ops of the waves in flight are
// queued on one HIP stream instead of running side by side, so no wave's writes run
// while another wave's read-only pass reads.  A wave's point ops then run as two
// passes: the read-only descent (finds answered, with the record pass; every write
// walked to its border node, where it leaves a start hint) and the writes starting at
// those hints.  Its scans run beside them on the wave slot's own stream with the
// validated scan kernel (scans only read, and may meet any wave's writes), so a long
// scan does not hold up the next wave's point ops; the wave completes when both are
// done.  A trace whose reads never touch keys that other requests write (the served
// workload) gets the same outcomes either way.  ordered = 0 (the default, which
// measured as fast or faster at every load of the served workload: its time goes to
// the scans) keeps the waves on separate streams, each one validated point-op launch
// plus one scan launch.
//
// Reclamation across overlapping waves.  Waves of one index run concurrently, so the
// launch-epoch drain of the plain path is not safe here.  The server rotates three
// retire lists.  Waves retire into the active list; a flip makes the next free list
// active.  A closed list is drained once (1) every wave that used it has completed and
// then (2) every wave that was running at that moment has completed too: a record it
// holds was retired before (1), so only waves started before (1) can still reach it.
// Waves are numbered in launch order; "all waves below n completed" is tracked as the
// lowest incomplete number.
#pragma once
#include <algorithm>
#include <chrono>
#include <cstdint>
#include <deque>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "gpu_masstree.hpp"
#include "scan_ring.hpp"

namespace indexingpu_amd {

struct serve_config {
  uint32_t streams = 4;              // waves in flight
  uint64_t max_wave_ops = 1u << 20;  // point ops per wave
  uint64_t max_wave_scans = 1u << 16;
  uint64_t ring_rows = 1u << 22;     // scan output ring (rows of max_count results)
  uint32_t slabs_per_write = 4;      // slabs an insert/update may allocate (admission)
  uint32_t epoch_waves = 32;         // waves between retire-list flips
  uint32_t coalesce = 1;             // 0: plain path (one request per launch pair)
  uint32_t ordered = 0;              // 1: point ops of all waves on one stream, read-only passes
  uint32_t plain_batch = 0;          // plain path: 1 = one call pair over all arrived requests
};

// Device arena of a trace.  Point ops: mixed-batch layout (types, keys, key_lengths,
// values in/out, value_lengths in/out, results).  Scans: lower keys, optional upper
// keys, per-query limits (<= scan_max_count), counts out; result rows go to the ring:
// request r's scans occupy ring rows [ring_row[r], ring_row[r] + num_scans * max_count).
struct serve_arena {
  const uint8_t* types;
  const uint32_t* keys;
  uint32_t max_key_length;
  const uint32_t* key_lengths;
  uint32_t* values;
  uint32_t max_value_length;
  uint32_t* value_lengths;
  uint8_t* results;    // one byte per op (0 / 1)
  const uint32_t* scan_lower;
  const uint32_t* scan_lower_lengths;
  const uint32_t* scan_upper;          // nullable
  const uint32_t* scan_upper_lengths;  // nullable
  const uint32_t* scan_limits;         // nullable
  uint32_t scan_max_count;
  uint32_t* scan_counts;
  uint32_t* ring_keys;          // ring_rows * max_key_length
  uint32_t* ring_key_lengths;   // ring_rows
  uint32_t* ring_values;        // ring_rows * max_value_length
  uint32_t* ring_value_lengths; // ring_rows
};

struct serve_request {
  double arrival_us;    // arrival time from the start of the run
  uint64_t op_offset;   // first point op in the arena
  uint64_t scan_offset; // first scan in the arena
  uint32_t num_ops;
  uint32_t num_scans;
  uint32_t num_writes;  // inserts + updates among the ops (admission estimate)
};

struct serve_outcome {
  double admit_us;      // when its wave was launched
  double done_us;       // when its wave completed (as seen by the loop)
  uint64_t ring_row;    // first ring row of its scans
  uint32_t wave;        // wave number
};

struct serve_stats {
  uint64_t waves = 0;
  uint64_t ops = 0;
  uint64_t scans = 0;
  uint64_t flips = 0;
  uint64_t drains = 0;
  uint64_t census = 0;
  uint64_t admission_waits = 0;  // loop turns where the queue head did not fit
  double elapsed_us = 0;
  uint32_t max_inflight = 0;
  uint32_t device_error = 0;
};

__global__ void popcount_kernel(const u32* words, u64 n, unsigned long long* out) {
  u64 local = 0;
  for (u64 i = (u64)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (u64)gridDim.x * blockDim.x)
    local += (u64)__popc(words[i]);
  unsigned long long v = local;
  for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off);
  if ((threadIdx.x & 63) == 0 && v) atomicAdd(out, v);
}

template <class Tree>
class masstree_server {
 public:
  masstree_server(Tree& tree, const serve_config& cfg) : tree_(tree), cfg_(cfg) {
    cfg_.streams = std::max<uint32_t>(1, std::min<uint32_t>(cfg_.streams, 16));
    // one wave is one batch launch, whose request indices are 32-bit
    cfg_.max_wave_ops = std::min<uint64_t>(cfg_.max_wave_ops, 1ull << 31);
    const uint32_t nstreams = cfg_.coalesce ? cfg_.streams : 1u;
    ordered_ = cfg_.coalesce && cfg_.ordered;
    for (uint32_t s = 0; s < nstreams; s++) {
      // every slot has its own stream (ordered waves: for their scans; the point ops of
      // all slots queue on ops_stream_)
      hipStream_t st;
      IGPU_HIP_TRY(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
      streams_.push_back(st);
      hipEvent_t ev;
      IGPU_HIP_TRY(hipEventCreateWithFlags(&ev, hipEventDisableTiming));
      events_.push_back(ev);
      IGPU_HIP_TRY(hipEventCreateWithFlags(&ev, hipEventDisableTiming));
      ops_events_.push_back(ev);
    }
    if (ordered_) IGPU_HIP_TRY(hipStreamCreateWithFlags(&ops_stream_, hipStreamNonBlocking));
    IGPU_HIP_TRY(hipStreamCreateWithFlags(&aux_, hipStreamNonBlocking));
    IGPU_HIP_TRY(hipEventCreateWithFlags(&aux_ev_, hipEventDisableTiming));
    if (cfg_.coalesce) {
      for (int l = 0; l < 3; l++) lists_[l].rec.reset(new simple_debra_reclaimer<>());
    }
    IGPU_HIP_TRY(hipMalloc(&census_dev_, 8));
    IGPU_HIP_TRY(hipHostMalloc(&census_host_, 8, hipHostMallocDefault));
    // read-only passes of ordered waves start at the tree's layer-0 shortcut table (built
    // here if due; it stays correct while the waves write), their finds at its
    // border-level twin, which their write passes rewrite when they merge nodes (no write
    // runs beside those read-only passes)
    if (ordered_) {
      IGPU_HIP_TRY(hipDeviceSynchronize());
      shortcut_ = tree_.find_shortcut_table(ops_stream_);
      shortcut_border_ = tree_.find_shortcut_border();
      IGPU_HIP_TRY(hipStreamSynchronize(ops_stream_));
    }
  }
  ~masstree_server() {
    (void)hipDeviceSynchronize();
    for (auto st : streams_) (void)hipStreamDestroy(st);
    for (auto e : ops_events_) (void)hipEventDestroy(e);
    if (ops_stream_) (void)hipStreamDestroy(ops_stream_);
    if (park_) (void)hipFree(park_);
    if (hints_) (void)hipFree(hints_);
    for (auto e : events_) (void)hipEventDestroy(e);
    (void)hipStreamDestroy(aux_);
    (void)hipEventDestroy(aux_ev_);
    for (auto p : stacks_) (void)hipFree(p);
    (void)hipFree(census_dev_);
    (void)hipHostFree(census_host_);
  }

  // Replays `n` requests (sorted by arrival) against the index in real time.
  serve_stats run(const serve_arena& a, const serve_request* req, uint64_t n, serve_outcome* out) {
    serve_stats st;
    if (n == 0) return st;
    for (uint64_t r = 1; r < n; r++)
      if (req[r].arrival_us < req[r - 1].arrival_us ||
          req[r].op_offset != req[r - 1].op_offset + req[r - 1].num_ops ||
          req[r].scan_offset != req[r - 1].scan_offset + req[r - 1].num_scans)
        throw std::invalid_argument("serve: requests must be in arrival order, back to back in the arena");
    const uint64_t ring_need = (uint64_t)cfg_.max_wave_scans * a.scan_max_count;
    if (cfg_.coalesce && ring_need > cfg_.ring_rows)
      throw std::invalid_argument("serve: ring_rows must hold max_wave_scans * scan_max_count rows");
    for (uint64_t r = 0; r < n; r++) {
      if (req[r].num_ops > cfg_.max_wave_ops || req[r].num_scans > cfg_.max_wave_scans)
        throw std::invalid_argument("serve: a request exceeds the per-wave limits");
    }
    if (cfg_.coalesce) prepare_stacks(a);
    if (ordered_ && cfg_.max_wave_ops > scratch_ops_) {
      // find scratch (4 B) and write start hints (16 B) for the largest wave
      IGPU_HIP_TRY(hipDeviceSynchronize());
      if (park_) IGPU_HIP_TRY(hipFree(park_));
      if (hints_) IGPU_HIP_TRY(hipFree(hints_));
      park_ = nullptr;
      hints_ = nullptr;
      IGPU_HIP_TRY(hipMalloc(&park_, cfg_.max_wave_ops * 4));
      IGPU_HIP_TRY(hipMalloc(&hints_, cfg_.max_wave_ops * 4 * kHintWords));
      scratch_ops_ = cfg_.max_wave_ops;
    }
    pool_slabs_ = tree_.allocator_.num_slabs;
    used_slabs_ = tree_.allocator_.num_allocated_slabs();
    reserved_ = 0;
    census_pending_ = false;
    ring_.reset(cfg_.ring_rows);
    census_due_ = census_fresh_ = aux_busy_ = false;
    draining_ = -1;
    waves_.clear();
    next_wave_ = 0;
    lowest_incomplete_ = 0;
    active_ = 0;
    for (int l = 0; l < 3; l++) lists_[l].state = l == 0 ? list_state::active : list_state::free;
    waves_since_flip_ = 0;

    const auto t0 = std::chrono::steady_clock::now();
    auto now_us = [&]() {
      return std::chrono::duration<double, std::micro>(std::chrono::steady_clock::now() - t0).count();
    };
    uint64_t arrived = 0;   // requests whose arrival time has come
    uint64_t admitted = 0;  // requests launched
    uint64_t completed = 0;
    std::vector<int> busy(streams_.size(), -1);  // wave index running on each stream
    while (completed < n) {
      const double now = now_us();
      while (arrived < n && req[arrived].arrival_us <= now) arrived++;
      // completions
      for (size_t s = 0; s < streams_.size(); s++) {
        if (busy[s] < 0) continue;
        if (hipEventQuery(events_[s]) != hipSuccess) continue;
        if (ordered_ && hipEventQuery(ops_events_[s]) != hipSuccess) continue;
        wave& w = waves_[(size_t)busy[s]];
        const double t = now_us();
        for (uint64_t r = w.first; r < w.last; r++) out[r].done_us = t;
        completed += w.last - w.first;
        w.done = true;
        if (w.ring_rows) ring_.release(w.ring_id);
        reserved_ -= w.reserved;
        used_slabs_ += w.reserved;  // pessimistic until the next census
        if (census_pending_) completed_in_census_ += w.reserved;
        census_fresh_ = false;
        busy[s] = -1;
      }
      while (lowest_incomplete_ < waves_.size() && waves_[lowest_incomplete_].done) lowest_incomplete_++;
      if (cfg_.coalesce) {
        advance_reclamation(st);
        census_step(st);
      }
      // launch on idle streams
      for (size_t s = 0; s < streams_.size() && admitted < arrived; s++) {
        if (busy[s] >= 0) continue;
        uint64_t last = admitted, ops = 0, scans = 0, writes = 0;
        while (last < arrived) {
          const serve_request& q = req[last];
          if (last > admitted &&
              ((!cfg_.coalesce && !cfg_.plain_batch) || ops + q.num_ops > cfg_.max_wave_ops ||
               scans + q.num_scans > cfg_.max_wave_scans))
            break;
          // the plain path has no admission control (it is the index's own calls)
          const uint64_t need = (writes + q.num_writes) * (uint64_t)cfg_.slabs_per_write;
          if (cfg_.coalesce && used_slabs_ + reserved_ + need > pool_slabs_) break;
          ops += q.num_ops;
          scans += q.num_scans;
          writes += q.num_writes;
          last++;
        }
        uint64_t row0 = 0, ring_id = 0;
        const uint64_t rows = scans * (uint64_t)a.scan_max_count;
        if (cfg_.coalesce) {
          if (last > admitted && rows && !ring_.reserve(rows, row0, ring_id)) last = admitted;
        } else {
          row0 = req[admitted].scan_offset * (uint64_t)a.scan_max_count;
        }
        if (last == admitted) {
          st.admission_waits++;
          if (inflight(busy) == 0) blocked_idle(st);
          break;
        }
        wave w;
        w.first = admitted;
        w.last = last;
        w.reserved = writes * (uint64_t)cfg_.slabs_per_write;
        w.ring_row = row0;
        w.ring_rows = cfg_.coalesce ? rows : 0;
        w.ring_id = ring_id;
        w.list = active_;
        census_fresh_ = false;
        reserved_ += w.reserved;
        const double t = now_us();
        launch(a, req, w, s, st);
        for (uint64_t r = w.first; r < w.last; r++) {
          out[r].admit_us = t;
          out[r].ring_row = row0 + (req[r].scan_offset - req[w.first].scan_offset) * a.scan_max_count;
          out[r].wave = (uint32_t)waves_.size();
        }
        busy[s] = (int)waves_.size();
        waves_.push_back(w);
        next_wave_ = waves_.size();
        admitted = last;
        st.waves++;
        st.ops += ops;
        st.scans += scans;
        waves_since_flip_++;
        st.max_inflight = std::max<uint32_t>(st.max_inflight, inflight(busy));
      }
    }
    // drain every list once the loop is idle
    IGPU_HIP_TRY(hipDeviceSynchronize());
    if (cfg_.coalesce) {
      for (int l = 0; l < 3; l++) {
        if (lists_[l].state != list_state::free || l == active_) drain_list(l, st);
        lists_[l].state = list_state::free;
      }
      lists_[active_].state = list_state::active;
      IGPU_HIP_TRY(hipStreamSynchronize(aux_));
    }
    st.elapsed_us = now_us();
    uint32_t e = 0;
    IGPU_HIP_TRY(hipMemcpy(&e, tree_.err_, 4, hipMemcpyDeviceToHost));
    st.device_error = e;
    return st;
  }

 private:
  enum class list_state { free, active, closing, grace, draining };
  struct retire_list {
    std::unique_ptr<simple_debra_reclaimer<>> rec;
    list_state state = list_state::free;
    uint64_t wait_below = 0;  // waves below this number must complete first
  };
  struct wave {
    uint64_t first = 0, last = 0;  // requests [first, last)
    uint64_t reserved = 0;         // slabs reserved
    uint64_t ring_row = 0, ring_rows = 0, ring_id = 0;
    int list = 0;
    bool done = false;
  };

  uint32_t inflight(const std::vector<int>& busy) const {
    uint32_t k = 0;
    for (int b : busy) k += b >= 0;
    return k;
  }

  void prepare_stacks(const serve_arena& a) {
    stack_layers_ = std::max<uint32_t>(a.max_key_length, 64u) + 1;
    scan_blocks_ = resident_blocks(scan_kernel<Tree::tile_size_, true>);
    const uint64_t words =
        (uint64_t)scan_blocks_ * (kBlockThreads / Tree::tile_size_) * stack_layers_ * 3u;
    if (stacks_.empty() || words > stack_words_) {
      IGPU_HIP_TRY(hipDeviceSynchronize());
      for (auto p : stacks_) (void)hipFree(p);
      stacks_.assign(streams_.size(), nullptr);
      for (auto& p : stacks_) IGPU_HIP_TRY(hipMalloc(&p, words * 4));
      stack_words_ = words;
    }
  }

  tree_view view_with(int list) const {
    return tree_view{tree_.allocator_, lists_[list].rec->get_device_instance(), tree_.err_,
                     tree_.root_index_, shortcut_.first, shortcut_.second,
                     shortcut_border_.first, shortcut_border_.second};
  }

  void launch(const serve_arena& a, const serve_request* req, const wave& w, size_t s,
              serve_stats& st) {
    hipStream_t stream = streams_[s];
    const serve_request& f = req[w.first];
    const serve_request& l = req[w.last - 1];
    const uint64_t o0 = f.op_offset, o1 = l.op_offset + l.num_ops;
    const uint64_t s0 = f.scan_offset, s1 = l.scan_offset + l.num_scans;
    const uint32_t mk = a.max_key_length, mv = a.max_value_length;
    if (!cfg_.coalesce) {
      // the plain path: the index's own calls (each mutating launch drains after itself)
      if (o1 > o0)
        tree_.mixed_batch(reinterpret_cast<const kernels::request_type*>(a.types + o0),
                          a.keys + o0 * mk, mk, a.key_lengths ? a.key_lengths + o0 : nullptr,
                          a.values + o0 * mv, mv, a.value_lengths + o0, reinterpret_cast<bool*>(a.results + o0),
                          (uint32_t)(o1 - o0), stream);
      if (s1 > s0) {
        const uint64_t r0 = s0 * a.scan_max_count;
        tree_.template scan<true, true>(
            a.scan_lower + s0 * mk, a.scan_lower_lengths ? a.scan_lower_lengths + s0 : nullptr, mk,
            a.scan_max_count, (uint32_t)(s1 - s0), a.scan_upper ? a.scan_upper + s0 * mk : nullptr,
            a.scan_upper_lengths ? a.scan_upper_lengths + s0 : nullptr, a.scan_counts + s0,
            a.ring_values + r0 * mv, mv, a.ring_value_lengths + r0, a.ring_keys + r0 * mk,
            a.ring_key_lengths + r0, stream, a.scan_limits ? a.scan_limits + s0 : nullptr);
      }
      IGPU_HIP_TRY(hipEventRecord(events_[s], stream));
      return;
    }
    const tree_view v = view_with(w.list);
    hipStream_t ostream = stream;  // point ops (ordered waves: the shared ops stream)
    if (o1 > o0) {
      batch_args b{};
      b.types = a.types + o0;
      b.kind = op_find;
      b.keys = a.keys + o0 * mk;
      b.max_key_length = mk;
      b.key_lengths = a.key_lengths ? a.key_lengths + o0 : nullptr;
      b.values = a.values + o0 * mv;
      b.max_value_length = mv;
      b.value_lengths = a.value_lengths + o0;
      b.out_values = a.values + o0 * mv;
      b.out_value_lengths = a.value_lengths + o0;
      b.results = reinterpret_cast<bool*>(a.results + o0);
      b.n = o1 - o0;
      b.flags = tree_.fanout_flags();
      if (ordered_) {
        // read-only pass: finds answered, writes walked to their border nodes
        ostream = ops_stream_;
        b.hints = hints_;
        auto desc = narrow_descend_kernel<kOrderedTile, 1, true>;
        auto rec = narrow_record_kernel<kOrderedTile>;
        constexpr uint32_t ntiles = kBlockThreads / kOrderedTile;
        const uint64_t want = (b.n + ntiles - 1) / ntiles;
        hipLaunchKernelGGL(desc, dim3((uint32_t)std::min<uint64_t>(want, resident_blocks(desc))),
                           dim3(kBlockThreads), 0, ostream, v, b, park_);
        hipLaunchKernelGGL(rec, dim3((uint32_t)std::min<uint64_t>(want, resident_blocks(rec))),
                           dim3(kBlockThreads), 0, ostream, v, b, static_cast<const u32*>(park_));
        b.skip_finds = 1;
      }
      auto kernel = batch_kernel<Tree::tile_size_, true>;
      const uint32_t tiles = kBlockThreads / Tree::tile_size_;
      const uint32_t blocks =
          (uint32_t)std::min<uint64_t>((b.n + tiles - 1) / tiles, resident_blocks(kernel));
      hipLaunchKernelGGL(kernel, dim3(blocks), dim3(kBlockThreads), 0, ostream, v, b);
    }
    if (ordered_) IGPU_HIP_TRY(hipEventRecord(ops_events_[s], ops_stream_));
    if (s1 > s0) {
      scan_args sa{};
      sa.lower_keys = a.scan_lower + s0 * mk;
      sa.lower_key_lengths = a.scan_lower_lengths ? a.scan_lower_lengths + s0 : nullptr;
      sa.max_key_length = mk;
      sa.max_count = a.scan_max_count;
      sa.num_queries = s1 - s0;
      sa.upper_keys = a.scan_upper ? a.scan_upper + s0 * mk : nullptr;
      sa.upper_key_lengths = a.scan_upper_lengths ? a.scan_upper_lengths + s0 : nullptr;
      sa.counts = a.scan_counts + s0;
      sa.values = a.ring_values + w.ring_row * mv;
      sa.max_value_length = mv;
      sa.value_lengths = a.ring_value_lengths + w.ring_row;
      sa.out_keys = a.ring_keys + w.ring_row * mk;
      sa.out_key_lengths = a.ring_key_lengths + w.ring_row;
      sa.limits = a.scan_limits ? a.scan_limits + s0 : nullptr;
      sa.stack = stacks_[s];
      sa.stack_layers = stack_layers_;
      const uint32_t tiles = kBlockThreads / Tree::tile_size_;
      const uint32_t blocks =
          (uint32_t)std::min<uint64_t>((sa.num_queries + tiles - 1) / tiles, scan_blocks_);
      hipLaunchKernelGGL((scan_kernel<Tree::tile_size_, true>), dim3(blocks), dim3(kBlockThreads),
                         0, stream, v, sa);
    }
    IGPU_HIP_TRY(hipEventRecord(events_[s], stream));
    (void)st;
  }

  void drain_list(int l, serve_stats& st) {
    const retire_view rv = lists_[l].rec->get_device_instance();
    hipLaunchKernelGGL(drain_kernel, dim3(256), dim3(256), 0, aux_, tree_.allocator_, rv);
    hipLaunchKernelGGL(reset_kernel, dim3(1), dim3(64), 0, aux_, rv);
    IGPU_HIP_TRY(hipEventRecord(aux_ev_, aux_));
    st.drains++;
  }

  void advance_reclamation(serve_stats& st) {
    for (int l = 0; l < 3; l++) {
      retire_list& L = lists_[l];
      if (L.state == list_state::closing && lowest_incomplete_ >= L.wait_below) {
        L.state = list_state::grace;
        L.wait_below = waves_.size();
      }
      if (L.state == list_state::grace && lowest_incomplete_ >= L.wait_below && !aux_busy_) {
        drain_list(l, st);
        L.state = list_state::draining;
        aux_busy_ = true;
        draining_ = l;
      }
    }
    if (aux_busy_ && draining_ >= 0 && hipEventQuery(aux_ev_) == hipSuccess) {
      aux_busy_ = false;
      lists_[draining_].state = list_state::free;
      draining_ = -1;
      // freed slabs come back at the next census
      census_due_ = true;
    }
    if (waves_since_flip_ >= cfg_.epoch_waves) {
      const int nxt = (active_ + 1) % 3;
      if (lists_[nxt].state == list_state::free) {
        lists_[active_].state = list_state::closing;
        lists_[active_].wait_below = waves_.size();
        lists_[nxt].state = list_state::active;
        active_ = nxt;
        waves_since_flip_ = 0;
        st.flips++;
      }
    }
  }

  // Nothing runs and the queue head does not fit the slab budget: retire what can be
  // retired, take a census of the pool, and give up only if the head still does not fit.
  void blocked_idle(serve_stats& st) {
    if (!cfg_.coalesce) throw std::runtime_error("serve: slab pool exhausted");
    if (waves_since_flip_ > 0) waves_since_flip_ = cfg_.epoch_waves;  // force a flip
    bool all_free = true;
    for (int l = 0; l < 3; l++)
      if (l != active_ && lists_[l].state != list_state::free) all_free = false;
    if (census_fresh_ && all_free && !aux_busy_ && waves_since_flip_ == 0)
      throw std::runtime_error("serve: slab pool exhausted");
    if (all_free && !aux_busy_ && !census_pending_) start_census();
    (void)st;
  }

  void start_census() {
    if (census_pending_ || aux_busy_) return;
    IGPU_HIP_TRY(hipMemsetAsync(census_dev_, 0, 8, aux_));
    const auto& p = tree_.allocator_;
    hipLaunchKernelGGL(popcount_kernel, dim3(1024), dim3(256), 0, aux_, p.bitmap, (u64)p.num_words,
                       census_dev_);
    IGPU_HIP_TRY(hipMemcpyAsync(census_host_, census_dev_, 8, hipMemcpyDeviceToHost, aux_));
    IGPU_HIP_TRY(hipEventRecord(aux_ev_, aux_));
    census_pending_ = true;
    aux_busy_ = true;
    draining_ = -1;
    completed_in_census_ = 0;
  }
  void census_step(serve_stats& st) {
    if (census_pending_ && hipEventQuery(aux_ev_) == hipSuccess) {
      // slabs in use: the census, plus the reservations of waves that completed while it
      // ran (their allocations may postdate the bitmap words it read); waves still
      // running keep their reservations in reserved_
      used_slabs_ = *census_host_ + completed_in_census_;
      census_pending_ = false;
      census_fresh_ = true;
      aux_busy_ = false;
      st.census++;
      return;
    }
    if (census_due_ && !aux_busy_) {
      census_due_ = false;
      start_census();
    }
  }

  Tree& tree_;
  std::pair<const uint32_t*, uint32_t> shortcut_{nullptr, 0u};  // layer-0 shortcut table
  // its border-level twin and built-at merge count (ordered waves' read-only passes)
  std::pair<uint32_t*, const uint32_t*> shortcut_border_{nullptr, nullptr};
  serve_config cfg_;
  static constexpr uint32_t kOrderedTile = 2;  // lanes per lookup of the read-only pass
  bool ordered_ = false;
  hipStream_t ops_stream_ = nullptr;     // ordered waves: every wave's point ops
  std::vector<hipEvent_t> ops_events_;  // ordered waves: a slot's point ops done
  u32* park_ = nullptr;
  u32* hints_ = nullptr;
  uint64_t scratch_ops_ = 0;
  std::vector<hipStream_t> streams_;
  std::vector<hipEvent_t> events_;
  hipStream_t aux_ = nullptr;
  hipEvent_t aux_ev_ = nullptr;
  bool aux_busy_ = false;
  int draining_ = -1;
  std::vector<u32*> stacks_;
  uint64_t stack_words_ = 0;
  uint32_t stack_layers_ = 0;
  uint32_t scan_blocks_ = 0;
  retire_list lists_[3];
  int active_ = 0;
  uint64_t waves_since_flip_ = 0;
  std::vector<wave> waves_;
  uint64_t next_wave_ = 0;
  uint64_t lowest_incomplete_ = 0;
  uint64_t pool_slabs_ = 0, used_slabs_ = 0, reserved_ = 0, completed_in_census_ = 0;
  bool census_pending_ = false, census_due_ = false, census_fresh_ = false;
  unsigned long long* census_dev_ = nullptr;
  unsigned long long* census_host_ = nullptr;
  scan_ring ring_;  // rows of the scan result ring (scan_ring.hpp)
};

}  // namespace indexingpu_amd
