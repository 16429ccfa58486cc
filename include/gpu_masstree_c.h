/* C ABI of the MI355X Masstree (libgpu_masstree.so).
 *
 * Plain pointers and sizes only.  Key/value/output pointers are device pointers (or
 * pinned host pointers, as the reference's tests use).  `stream` is a hipStream_t
 * (NULL = the default stream).  Batch calls are asynchronous on `stream`; every call
 * returns 0 on success or a GMT_E* code for an invalid argument or a HIP failure
 * (message: gmt_last_error_message()).  Errors found inside a kernel (pool exhausted,
 * invalid key length, ...) are reported by gmt_device_error().
 *
 * The reference exposes this path as the C++ template API
 * GpuMasstree::gpu_masstree<Allocator, Reclaimer, tile_size> (include/gpu_masstree.hpp
 * in hyoungjook/IndexinGPU); each entry point below names the member it stands for.
 * The same API is available here as C++ in <indexingpu_amd/gpu_masstree.hpp> (and under
 * the reference's include names, <gpu_masstree.hpp> etc.).
 *
 * Concurrency: every call on one index, read-only ones included, must be ordered on one
 * stream.  Mutating calls share the index's reclaimer, and finds, mixed batches and
 * hinted writes share its scratch (per-request words indexed from each batch's first
 * request); two finds on different streams would overwrite each other's scratch.
 */
#ifndef GPU_MASSTREE_C_H
#define GPU_MASSTREE_C_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct gmt_index gmt_index;

enum gmt_status {
  GMT_OK = 0,
  GMT_EINVAL = 1,   /* bad argument (null handle, unsupported tile size, ...) */
  GMT_EHIP = 2,     /* a HIP runtime call failed */
};

/* Request types of gmt_mixed_batch: kernels::request_type (include/macros.hpp:51-58). */
enum gmt_request_type {
  GMT_REQUEST_INSERT = 0,
  GMT_REQUEST_ERASE = 1,
  GMT_REQUEST_FIND = 2,
  GMT_REQUEST_UPDATE = 3,
};

/* Device-side error codes (gmt_device_error). */
enum gmt_device_error_code {
  GMT_DEV_OK = 0,
  GMT_DEV_OUT_OF_MEMORY = 1,
  GMT_DEV_STEP_LIMIT = 2,
  GMT_DEV_BAD_INDEX = 3,
  GMT_DEV_BAD_LENGTH = 4,
  GMT_DEV_RETIRE_OVERFLOW = 5,
  GMT_DEV_SCAN_SCRATCH = 6,
};

/* Creates an index with its own slab allocator and reclaimer.
 * Replaces: simple_slab_allocator<128>(pool_ratio) (include/simple_slab_alloc.hpp:56-69),
 *           simple_debra_reclaimer<>() (include/simple_debra_reclaim.hpp:56-67),
 *           gpu_masstree(allocator, reclaimer) (include/gpu_masstree.hpp:67-72).
 * tile_size: 16, 32 or 64.  pool_bytes > 0 sizes the slab pool explicitly; otherwise
 * pool_ratio x total device memory is used (the reference's rule).  retire_capacity
 * = 0 picks the default retire-list size. */
int gmt_create(uint32_t tile_size, uint64_t pool_bytes, float pool_ratio,
               uint32_t retire_capacity, gmt_index** out);
int gmt_destroy(gmt_index* index);

/* gpu_masstree::find<concurrent> (include/gpu_masstree.hpp:86-99). value_lengths[i] = 0
 * means absent; values get min(length, max_value_length) slices. */
int gmt_find(gmt_index* index, const uint32_t* keys, uint32_t max_key_length,
             const uint32_t* key_lengths, uint32_t* values, uint32_t max_value_length,
             uint32_t* value_lengths, uint32_t num_keys, int concurrent, void* stream);

/* No counterpart in the reference: for callers that cannot allocate device memory
 * (code under HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * gmt_find with concurrent = 0, whose per-key scratch of the two-pass find is a fixed
 * device array of GMT_NOALLOC_MAX_KEYS words inside the library (one per GPU, held until
 * the process ends) instead of a buffer the index grows with hipMalloc.  It allocates
 * nothing; num_keys above GMT_NOALLOC_MAX_KEYS is rejected with GMT_EINVAL.  Every call
 * on one GPU shares that array, so calls that overlap on one GPU (other streams, other
 * indexes) race: order them. */
#define GMT_NOALLOC_MAX_KEYS (1u << 22)
int gmt_find_noalloc(gmt_index* index, const uint32_t* keys, uint32_t max_key_length,
                     const uint32_t* key_lengths, uint32_t* values, uint32_t max_value_length,
                     uint32_t* value_lengths, uint32_t num_keys, void* stream);

/* gpu_masstree::insert<update_if_exists> (include/gpu_masstree.hpp:101-115).
 * results (nullable, one byte per request) receives per-op success: an extension. */
int gmt_insert(gmt_index* index, const uint32_t* keys, uint32_t max_key_length,
               const uint32_t* key_lengths, const uint32_t* values, uint32_t max_value_length,
               const uint32_t* value_lengths, uint32_t num_keys, int update_if_exists,
               uint8_t* results, void* stream);

/* gpu_masstree::update (include/gpu_masstree.hpp:117-130). */
int gmt_update(gmt_index* index, const uint32_t* keys, uint32_t max_key_length,
               const uint32_t* key_lengths, const uint32_t* values, uint32_t max_value_length,
               const uint32_t* value_lengths, uint32_t num_keys, uint8_t* results, void* stream);

/* gpu_masstree::erase (include/gpu_masstree.hpp:132-145). */
int gmt_erase(gmt_index* index, const uint32_t* keys, uint32_t max_key_length,
              const uint32_t* key_lengths, uint32_t num_keys, uint8_t* results, void* stream);

/* gpu_masstree::scan<use_upper_key> (include/gpu_masstree.hpp:147-171); use_upper_key is
 * (upper_keys != NULL).  per_query_limits (nullable) caps query q at
 * min(per_query_limits[q], max_count_per_query): an extension. */
int gmt_scan(gmt_index* index, const uint32_t* lower_keys, const uint32_t* lower_key_lengths,
             uint32_t max_key_length, uint32_t max_count_per_query, uint32_t num_queries,
             const uint32_t* upper_keys, const uint32_t* upper_key_lengths, uint32_t* counts,
             uint32_t* values, uint32_t max_value_length, uint32_t* value_lengths,
             uint32_t* out_keys, uint32_t* out_key_lengths, const uint32_t* per_query_limits,
             void* stream);

/* gpu_masstree::mixed_batch (include/gpu_masstree.hpp:173-191). types[i] is a
 * gmt_request_type; finds write values/value_lengths in place, writes set results[i]. */
int gmt_mixed_batch(gmt_index* index, const uint8_t* types, const uint32_t* keys,
                    uint32_t max_key_length, const uint32_t* key_lengths, uint32_t* values,
                    uint32_t max_value_length, uint32_t* value_lengths, uint8_t* results,
                    uint32_t num_requests, void* stream);

/* Batched form (no counterpart in the reference; it extends mixed_batch): applies
 * num_batches mixed batches in order (batch b + 1 sees every effect of batch b) with one
 * call and no host round trip.  Requests of all batches are stored back to back with
 * gmt_mixed_batch's layout; `offsets` is a device array of num_batches + 1 request
 * offsets; total_requests = offsets[num_batches].  See gmt_set_multi_mode. */
int gmt_multi_batch(gmt_index* index, const uint8_t* types, const uint32_t* keys,
                    uint32_t max_key_length, const uint32_t* key_lengths, uint32_t* values,
                    uint32_t max_value_length, uint32_t* value_lengths, uint8_t* results,
                    const uint64_t* offsets, uint32_t num_batches, uint64_t total_requests,
                    void* stream);

/* Tuning knobs of gmt_multi_batch: share (1..100 %) of the resident-block capacity its
 * launches use, and how it runs: mode 0 = one persistent launch with a grid barrier
 * between batches; mode 1 (default) = per batch, a read-only pass over its finds when it
 * holds at least the gmt_set_mixed_split size, one launch for the rest, each reading its
 * range from `offsets` on the device, and a reclamation pass (total_requests must be
 * below 2^32, else mode 0 is used).  Outside graph capture mode 1 reads `offsets` back
 * once (a synchronization) to size each batch's launches.  Results do not depend on
 * either knob. */
int gmt_set_multi_grid_percent(gmt_index* index, uint32_t percent);
int gmt_set_multi_mode(gmt_index* index, uint32_t mode);

/* Waits for the device and returns (then clears, if clear != 0) the first device-side
 * error code since the last clear (gmt_device_error_code). */
uint32_t gmt_device_error(gmt_index* index, int clear);

/* The livelock guard's record of the operation it abandoned (GMT_DEV_STEP_LIMIT), kept by
 * the guard whose error code was the first since the last clear: GMT_TRIP_WORDS words,
 * trip_valid (1), phase, node, layer, layer_root, a0, a1, a2, a3, snap_hdr, the node's
 * header when the guard fired, steps, request kind, key length, key slice of the current
 * layer, lvl | fresh << 16.  Returns 1 and fills out when present, 0 when the guard has
 * not fired, a negative GMT_E* code on failure; gmt_device_error(index, 1) clears it.
 * No counterpart in the reference (it asserts); for diagnosing a stuck operation after
 * the fact, without rerunning it. */
#define GMT_TRIP_WORDS 16
int gmt_guard_trip(gmt_index* index, uint32_t* out);

/* Host walk of the whole tree (debug; the reference's validate(),
 * include/gpu_masstree.hpp:1421-1472).  Returns 1 if every invariant holds, 0 if not
 * (first violation in msg), negative on failure.  Any output pointer may be NULL. */
int gmt_validate(gmt_index* index, uint64_t* entries, uint64_t* nodes, uint64_t* record_slabs,
                 char* msg, uint32_t msg_len);

/* Dumps every entry in key order into host buffers (debug and parity checks).  Call
 * with NULL buffers to get the totals: returns the entry count (or -1), sets
 * *key_slices / *value_slices.  key_offsets / value_offsets hold count + 1 entries. */
int64_t gmt_dump(gmt_index* index, uint32_t* key_offsets, uint32_t* key_data,
                 uint32_t* value_offsets, uint32_t* value_data, uint64_t* key_slices,
                 uint64_t* value_slices);

/* Slabs currently allocated (the reference's num_nodes_used(),
 * include/gpu_masstree.hpp:1415-1418) and the pool size in slabs. */
uint64_t gmt_slabs_allocated(gmt_index* index);
uint64_t gmt_pool_slabs(gmt_index* index);

/* Lookups each tile keeps in flight in read-only find batches (1, 2, 4 or 8; other
 * values select 4; default 1).  A tuning knob with no counterpart in the reference. */
int gmt_set_find_depth(gmt_index* index, uint32_t depth);

/* Lanes per lookup of read-only find batches (gmt_find with concurrent = 0): 1, 2, 4, 8, 16
 * (each lane holds 16 / find_tile slots of a node; default 2), or 0 for the index's
 * tile width.  A tuning knob with no counterpart in the reference; results do not
 * depend on it. */
int gmt_set_find_tile(gmt_index* index, uint32_t find_tile);

/* Read-only finds with a narrow tile as one kernel (passes = 1) or as a descent pass and
 * a record pass (2, the default; the first call at a new largest batch size allocates 4
 * bytes per request of scratch and synchronizes).  A tuning knob with no counterpart in
 * the reference; results do not depend on it. */
int gmt_set_find_passes(gmt_index* index, uint32_t passes);

/* Lanes per lookup of the record pass of two-pass finds (1, 2, 4, 8 or 16), or 0 (the
 * default) to use the find tile.  A tuning knob with no counterpart in the reference; results
 * do not depend on it. */
int gmt_set_record_tile(gmt_index* index, uint32_t record_tile);

/* Mixed batches of at least min_requests requests (default 16384) run their finds as a
 * read-only pass with the narrow two-pass kernels, then the other requests; 0 splits
 * every batch, 0xFFFFFFFF none.  The first split batch at a new largest size allocates
 * 4 bytes per request of scratch and synchronizes.  A tuning knob with no counterpart in
 * the reference; the outcome equals that of one launch, since a mixed batch writes no
 * key twice and reads no key it writes. */
int gmt_set_mixed_split(gmt_index* index, uint32_t min_requests);

/* Insert, update and erase batches of at least min_requests requests (default 1) run a
 * read-only descent first that finds each request's border node, and start the writes
 * there instead of at the root; 0xFFFFFFFF turns this off.  Uses 16 bytes of scratch
 * per request for up to 4M requests at a time (allocated at the first larger batch,
 * which synchronizes).  A tuning knob with no counterpart in the reference; results do
 * not depend on it. */
int gmt_set_write_hints(gmt_index* index, uint32_t min_requests);

/* Allocates now the scratch that two-pass finds and split mixed batches use (4 bytes per
 * request) for batches of up to max_requests requests; without it the first larger
 * batch allocates and synchronizes.  Call it before capturing such calls into a graph.
 * No counterpart in the reference. */
int gmt_reserve_find_scratch(gmt_index* index, uint64_t max_requests);

/* Read-only descents (find, the read-only pass of mixed batches, the start hints of write
 * batches) of a tree of at least 2^18 inserted keys whose root is at level 3 or above
 * start at a layer-0 shortcut table (about one 4-byte entry per key, at most 2^26: the
 * level-1 node holding the smallest key of each prefix of the first slice; finds start one
 * level lower, at the border node, through a twin table that merges rewrite, or that a
 * merge it cannot follow retires until the next rebuild)
 * instead of the root: on (the default) or off.  An extension; results do not depend on it.  The table
 * is rebuilt when the insert count has doubled; one left from before later writes is
 * still correct.  Built for the 16-lane tile width only (the instance that owns the tree);
 * at 32 and 64 lanes descents start at the root. */
int gmt_set_find_shortcut(gmt_index* index, int on);
/* log2 of the entries of the largest shortcut table the index holds (0: none). */
uint32_t gmt_find_shortcut_bits(gmt_index* index);

/* Node fanout, an extension: the reference's nodes are fixed at 16 slots (14 keys per
 * border node, 15 children per interior node; include/masstree_node_subwarp.hpp:35).
 * Later writes split a border node once it holds border_cap entries (2..14) and an
 * interior node once it has interior_cap children (3..15); 0 keeps the maximum.  Trees
 * built under any setting stay readable under every other.  For the autotuner. */
int gmt_set_fanout(gmt_index* index, uint32_t border_cap, uint32_t interior_cap);

/* Tile width (lanes per operation: 16, 32 or 64) that later calls on this index use.
 * The reference fixes it per instance as the tile_size template argument
 * (include/gpu_masstree.hpp:42-46); here all widths share one node layout and one tree
 * state, so it can change between calls (a tuning knob).  gmt_tile_size returns the
 * current width. */
int gmt_set_tile_size(gmt_index* index, uint32_t tile_size);
uint32_t gmt_tile_size(gmt_index* index);

/* ------------------------------------------------------------------ served form
 * No counterpart in the reference (its batch calls are the whole interface; it has
 * no request loop).  A server replays a trace of requests against the index in real
 * time with continuous batching: requests are admitted in arrival order under a
 * per-wave op/scan limit, a free-slab budget and a scan-output ring budget.  With
 * ordered = 0 each wave is one mixed-batch launch plus one scan launch on one of
 * `streams` HIP streams, so waves overlap on the device; with ordered = 1 the point ops
 * of up to `streams` waves queue in order on one stream, each wave a read-only pass and
 * then its writes, and its scans run beside them on the wave's own stream.  coalesce = 0 runs the plain path instead
 * (gmt_mixed_batch / gmt_scan pairs in order on one stream: one request per pair, or with
 * plain_batch = 1 everything that arrived while the previous pair ran) for comparison. */
typedef struct gmt_server gmt_server;

typedef struct {
  uint32_t streams;          /* waves in flight (1..16) */
  uint32_t slabs_per_write;  /* slabs reserved per insert/update at admission */
  uint64_t max_wave_ops;     /* point ops per wave */
  uint64_t max_wave_scans;   /* scans per wave */
  uint64_t ring_rows;        /* scan result ring, rows of scan_max_count results */
  uint32_t epoch_waves;      /* waves between retire-list flips */
  uint32_t coalesce;         /* 1: served form; 0: plain path */
  uint32_t ordered;          /* 1: the point ops of all waves queued in order on one stream,
                                each wave a read-only pass (finds, write start hints) then
                                its writes, its scans beside them on the wave's stream;
                                0 (recommended): waves side by side on `streams` streams */
  uint32_t plain_batch;      /* with coalesce = 0: 1 = each gmt_mixed_batch / gmt_scan pair
                                covers every request that has arrived (up to the per-wave
                                limits), one pair at a time (interval batching of the plain
                                path); 0 = one request per pair */
} gmt_serve_config;

/* Device arena of a trace: point ops in gmt_mixed_batch's layout, scans in gmt_scan's
 * (per-query limits <= scan_max_count), all requests back to back in arrival order.
 * Scan results go to the ring: request r's rows start at outcome[r].ring_row, each scan
 * taking scan_max_count rows.  A completed request's rows stay until the ring wraps past
 * them (a ring that holds every scan of the trace keeps all of them).  With coalesce = 0
 * the ring must hold every scan of the trace (rows are at scan index * scan_max_count). */
typedef struct {
  const uint8_t* types;
  const uint32_t* keys;
  uint32_t max_key_length;
  const uint32_t* key_lengths;
  uint32_t* values;
  uint32_t max_value_length;
  uint32_t* value_lengths;
  uint8_t* results;
  const uint32_t* scan_lower;
  const uint32_t* scan_lower_lengths;
  const uint32_t* scan_upper;          /* nullable: no upper bound */
  const uint32_t* scan_upper_lengths;  /* nullable */
  const uint32_t* scan_limits;         /* nullable */
  uint32_t scan_max_count;
  uint32_t* scan_counts;
  uint32_t* ring_keys;
  uint32_t* ring_key_lengths;
  uint32_t* ring_values;
  uint32_t* ring_value_lengths;
} gmt_serve_arena;

typedef struct {
  double arrival_us;     /* from the start of the run; non-decreasing */
  uint64_t op_offset;    /* = previous op_offset + previous num_ops */
  uint64_t scan_offset;  /* = previous scan_offset + previous num_scans */
  uint32_t num_ops;
  uint32_t num_scans;
  uint32_t num_writes;   /* inserts + updates among the ops */
} gmt_serve_request;

typedef struct {
  double admit_us;   /* launch time of its wave */
  double done_us;    /* completion time of its wave, as seen by the loop */
  uint64_t ring_row;
  uint32_t wave;
} gmt_serve_outcome;

typedef struct {
  uint64_t waves, ops, scans, flips, drains, census, admission_waits;
  double elapsed_us;
  uint32_t max_inflight;
  uint32_t device_error;  /* gmt_device_error_code seen at the end of the run */
} gmt_serve_stats;

/* Creates a server for the index at its current tile width. */
int gmt_server_create(gmt_index* index, const gmt_serve_config* config, gmt_server** out);
int gmt_server_destroy(gmt_server* server);
/* Replays n requests (host array) and fills outcomes (host array of n) and stats. */
int gmt_server_run(gmt_server* server, const gmt_serve_arena* arena,
                   const gmt_serve_request* requests, uint64_t n, gmt_serve_outcome* outcomes,
                   gmt_serve_stats* stats);

/* Message of the last failing call on this thread. */
const char* gmt_last_error_message(void);

#ifdef __cplusplus
}
#endif

#endif /* GPU_MASSTREE_C_H */
