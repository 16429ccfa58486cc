"""Python host side of the MI355X Masstree: a ctypes binding of libgpu_masstree.so.

Mirrors the reference's ``GpuMasstree::gpu_masstree`` host API (include/gpu_masstree.hpp
:84-191 in hyoungjook/IndexinGPU): the same call names, argument order and meaning
(flat u32 slice arrays, optional key/value lengths, ``value_lengths == 0`` means not
found, scan outputs strided by ``max_count_per_query``), over torch tensors that live on
the GPU.  Calls are asynchronous on the current torch stream (or the ``stream``
argument).  There is no CPU fallback: constructing an index without the built library
or without a GPU raises.
"""
import ctypes
import os

import torch

from . import build as _build

REQUEST_INSERT = 0
REQUEST_ERASE = 1
REQUEST_FIND = 2
REQUEST_UPDATE = 3

DEVICE_ERRORS = {
    0: "ok",
    1: "slab pool exhausted",
    2: "operation step limit exceeded",
    3: "object index outside the pool",
    4: "invalid key or value length",
    5: "retire list overflow",
    6: "scan layer stack overflow",
}

_lib = None


class MasstreeError(RuntimeError):
    pass


def load_library():
    """Loads the in-tree libgpu_masstree.so; raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(_build.LIB_PATH):
        raise MasstreeError(
            f"{_build.LIB_PATH} is missing: build it with indexingpu_amd.build.build() "
            "(hipcc --offload-arch=gfx950)")
    lib = ctypes.CDLL(_build.LIB_PATH)
    vp, u32, u64, i32 = ctypes.c_void_p, ctypes.c_uint32, ctypes.c_uint64, ctypes.c_int
    P = ctypes.c_void_p
    lib.gmt_create.argtypes = [u32, u64, ctypes.c_float, u32, ctypes.POINTER(vp)]
    lib.gmt_destroy.argtypes = [vp]
    lib.gmt_find.argtypes = [vp, P, u32, P, P, u32, P, u32, i32, vp]
    lib.gmt_find_noalloc.argtypes = [vp, P, u32, P, P, u32, P, u32, vp]
    lib.gmt_insert.argtypes = [vp, P, u32, P, P, u32, P, u32, i32, P, vp]
    lib.gmt_update.argtypes = [vp, P, u32, P, P, u32, P, u32, P, vp]
    lib.gmt_erase.argtypes = [vp, P, u32, P, u32, P, vp]
    lib.gmt_scan.argtypes = [vp, P, P, u32, u32, u32, P, P, P, P, u32, P, P, P, P, vp]
    lib.gmt_mixed_batch.argtypes = [vp, P, P, u32, P, P, u32, P, P, u32, vp]
    lib.gmt_device_error.argtypes = [vp, i32]
    lib.gmt_device_error.restype = u32
    lib.gmt_set_find_shortcut.argtypes = [vp, i32]
    lib.gmt_set_find_shortcut.restype = i32
    lib.gmt_find_shortcut_bits.argtypes = [vp]
    lib.gmt_find_shortcut_bits.restype = u32
    lib.gmt_set_fanout.argtypes = [vp, u32, u32]
    lib.gmt_set_fanout.restype = i32
    lib.gmt_guard_trip.argtypes = [vp, vp]
    lib.gmt_guard_trip.restype = i32
    lib.gmt_validate.argtypes = [vp, ctypes.POINTER(u64), ctypes.POINTER(u64),
                                 ctypes.POINTER(u64), ctypes.c_char_p, u32]
    lib.gmt_dump.argtypes = [vp, P, P, P, P, ctypes.POINTER(u64), ctypes.POINTER(u64)]
    lib.gmt_dump.restype = ctypes.c_int64
    lib.gmt_slabs_allocated.argtypes = [vp]
    lib.gmt_slabs_allocated.restype = u64
    lib.gmt_pool_slabs.argtypes = [vp]
    lib.gmt_pool_slabs.restype = u64
    lib.gmt_set_find_depth.argtypes = [vp, u32]
    lib.gmt_multi_batch.argtypes = [vp, P, P, u32, P, P, u32, P, P, P, u32, u64, vp]
    lib.gmt_set_multi_grid_percent.argtypes = [vp, u32]
    lib.gmt_set_multi_mode.argtypes = [vp, u32]
    lib.gmt_set_tile_size.argtypes = [vp, u32]
    lib.gmt_set_find_tile.argtypes = [vp, u32]
    lib.gmt_set_find_passes.argtypes = [vp, u32]
    lib.gmt_set_record_tile.argtypes = [vp, u32]
    lib.gmt_set_mixed_split.argtypes = [vp, u32]
    lib.gmt_set_write_hints.argtypes = [vp, u32]
    lib.gmt_reserve_find_scratch.argtypes = [vp, ctypes.c_uint64]
    lib.gmt_tile_size.argtypes = [vp]
    lib.gmt_tile_size.restype = u32
    lib.gmt_last_error_message.restype = ctypes.c_char_p
    _lib = lib
    return lib


def _ptr(t, esize, name, need=0):
    """Device pointer of a contiguous CUDA tensor with elements of `esize` bytes and at
    least `need` elements (checked here, so a kernel never indexes past a buffer)."""
    if t is None:
        return None
    if not isinstance(t, torch.Tensor):
        raise TypeError(f"{name} must be a torch tensor")
    if t.element_size() != esize:
        raise TypeError(f"{name} must have {esize}-byte elements, got {t.dtype}")
    if not t.is_contiguous():
        raise ValueError(f"{name} must be contiguous")
    if not t.is_cuda:
        raise ValueError(f"{name} must be a GPU tensor")
    if t.numel() < need:
        raise ValueError(f"{name} holds {t.numel()} elements, the call needs {need}")
    return ctypes.c_void_p(t.data_ptr())


def _stream(stream):
    s = stream if stream is not None else torch.cuda.current_stream()
    return ctypes.c_void_p(s.cuda_stream)


class GpuMasstree:
    """gpu_masstree<simple_slab_allocator<128>, simple_debra_reclaimer<>, tile_size>.

    ``pool_bytes`` sizes the slab pool explicitly; otherwise ``pool_ratio`` x total GPU
    memory is used, as the reference's ``simple_slab_allocator(0.9f)`` does.
    """

    def __init__(self, tile_size=16, pool_bytes=0, pool_ratio=0.9, retire_capacity=0):
        if not torch.cuda.is_available():
            raise MasstreeError("GpuMasstree needs a GPU (torch.cuda.is_available() is False)")
        self._lib = load_library()
        torch.cuda.init()
        h = ctypes.c_void_p()
        self._check(self._lib.gmt_create(tile_size, int(pool_bytes), float(pool_ratio),
                                         int(retire_capacity), ctypes.byref(h)))
        self._h = h

    def close(self):
        if getattr(self, "_h", None):
            self._lib.gmt_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, rc):
        if rc != 0:
            msg = self._lib.gmt_last_error_message().decode()
            raise MasstreeError(f"libgpu_masstree call failed ({rc}): {msg}")

    # ------------------------------------------------------------------ batch API
    def find(self, keys, max_key_length, key_lengths, values, max_value_length, value_lengths,
             num_keys, concurrent=False, stream=None):
        n = int(num_keys)
        self._check(self._lib.gmt_find(
            self._h, _ptr(keys, 4, "keys", n * max_key_length), max_key_length,
            _ptr(key_lengths, 4, "key_lengths", n),
            _ptr(values, 4, "values", n * max_value_length), max_value_length,
            _ptr(value_lengths, 4, "value_lengths", n), n, int(bool(concurrent)),
            _stream(stream)))

    # Largest batch of find_noalloc (GMT_NOALLOC_MAX_KEYS).
    NOALLOC_MAX_KEYS = 1 << 22

    def find_noalloc(self, keys, max_key_length, key_lengths, values, max_value_length,
                     value_lengths, num_keys, stream=None):
        """find() (read-only form) that allocates nothing: its scratch is a fixed array in
        the library (NOALLOC_MAX_KEYS keys at most; calls overlapping on one GPU race)."""
        n = int(num_keys)
        self._check(self._lib.gmt_find_noalloc(
            self._h, _ptr(keys, 4, "keys", n * max_key_length), max_key_length,
            _ptr(key_lengths, 4, "key_lengths", n),
            _ptr(values, 4, "values", n * max_value_length), max_value_length,
            _ptr(value_lengths, 4, "value_lengths", n), n, _stream(stream)))

    def insert(self, keys, max_key_length, key_lengths, values, max_value_length, value_lengths,
               num_keys, update_if_exists=False, results=None, stream=None):
        n = int(num_keys)
        self._check(self._lib.gmt_insert(
            self._h, _ptr(keys, 4, "keys", n * max_key_length), max_key_length,
            _ptr(key_lengths, 4, "key_lengths", n),
            _ptr(values, 4, "values", n * max_value_length), max_value_length,
            _ptr(value_lengths, 4, "value_lengths", n), n, int(bool(update_if_exists)),
            _ptr(results, 1, "results", n), _stream(stream)))

    def update(self, keys, max_key_length, key_lengths, values, max_value_length, value_lengths,
               num_keys, results=None, stream=None):
        n = int(num_keys)
        self._check(self._lib.gmt_update(
            self._h, _ptr(keys, 4, "keys", n * max_key_length), max_key_length,
            _ptr(key_lengths, 4, "key_lengths", n),
            _ptr(values, 4, "values", n * max_value_length), max_value_length,
            _ptr(value_lengths, 4, "value_lengths", n), n, _ptr(results, 1, "results", n),
            _stream(stream)))

    def erase(self, keys, max_key_length, key_lengths, num_keys, results=None, stream=None):
        n = int(num_keys)
        self._check(self._lib.gmt_erase(
            self._h, _ptr(keys, 4, "keys", n * max_key_length), max_key_length,
            _ptr(key_lengths, 4, "key_lengths", n), n, _ptr(results, 1, "results", n),
            _stream(stream)))

    def scan(self, lower_keys, lower_key_lengths, max_key_length, max_count_per_query,
             num_queries, upper_keys=None, upper_key_lengths=None, counts=None, values=None,
             max_value_length=1, value_lengths=None, out_keys=None, out_key_lengths=None,
             per_query_limits=None, stream=None):
        q, c = int(num_queries), int(max_count_per_query)
        self._check(self._lib.gmt_scan(
            self._h, _ptr(lower_keys, 4, "lower_keys", q * max_key_length),
            _ptr(lower_key_lengths, 4, "lower_key_lengths", q), max_key_length, c, q,
            _ptr(upper_keys, 4, "upper_keys", q * max_key_length),
            _ptr(upper_key_lengths, 4, "upper_key_lengths", q), _ptr(counts, 4, "counts", q),
            _ptr(values, 4, "values", q * c * max_value_length), max_value_length,
            _ptr(value_lengths, 4, "value_lengths", q * c),
            _ptr(out_keys, 4, "out_keys", q * c * max_key_length),
            _ptr(out_key_lengths, 4, "out_key_lengths", q * c),
            _ptr(per_query_limits, 4, "per_query_limits", q), _stream(stream)))

    def mixed_batch(self, request_types, keys, max_key_length, key_lengths, values,
                    max_value_length, value_lengths, results, num_requests, stream=None):
        n = int(num_requests)
        self._check(self._lib.gmt_mixed_batch(
            self._h, _ptr(request_types, 1, "request_types", n),
            _ptr(keys, 4, "keys", n * max_key_length), max_key_length,
            _ptr(key_lengths, 4, "key_lengths", n),
            _ptr(values, 4, "values", n * max_value_length), max_value_length,
            _ptr(value_lengths, 4, "value_lengths", n), _ptr(results, 1, "results", n), n,
            _stream(stream)))

    # ------------------------------------------------------------------ checks
    def device_error(self, clear=True):
        """Synchronizes and returns the first device-side error code (0 = none)."""
        return int(self._lib.gmt_device_error(self._h, int(bool(clear))))

    def set_find_shortcut(self, on):
        """Read-only descents of large trees start at the layer-0 shortcut table (True, the
        default) or at the root (False)."""
        self._check(self._lib.gmt_set_find_shortcut(self._h, 1 if on else 0))

    def find_shortcut_bits(self):
        """log2 of the shortcut table's entries (0: no table)."""
        return int(self._lib.gmt_find_shortcut_bits(self._h))

    def set_fanout(self, border_cap=0, interior_cap=0):
        """Node fanout for later writes: border nodes split at border_cap entries (2..14),
        interior nodes at interior_cap children (3..15); 0 = the layout's maximum."""
        self._check(self._lib.gmt_set_fanout(self._h, border_cap, interior_cap))

    def guard_trip(self):
        """The livelock guard's record of the operation it abandoned, as a dict, or None
        when the guard has not fired since the last clear (device_error clears it)."""
        import numpy as np
        out = np.zeros(16, np.uint32)
        rc = self._lib.gmt_guard_trip(self._h, ctypes.c_void_p(out.ctypes.data))
        if rc < 0:
            self._check(-rc)
        if rc == 0:
            return None
        names = ("valid", "phase", "node", "layer", "layer_root", "a0", "a1", "a2", "a3",
                 "snap_hdr", "header_now", "steps", "kind", "key_length", "key_slice",
                 "lvl_fresh")
        return dict(zip(names, (int(v) for v in out)))

    def check_device_error(self):
        e = self.device_error()
        if e != 0:
            raise MasstreeError(f"device error {e}: {DEVICE_ERRORS.get(e, '?')}")

    def validate(self):
        """Host walk of the tree; returns (ok, entries, nodes, record_slabs, message)."""
        e, n, r = ctypes.c_uint64(), ctypes.c_uint64(), ctypes.c_uint64()
        msg = ctypes.create_string_buffer(512)
        rc = self._lib.gmt_validate(self._h, ctypes.byref(e), ctypes.byref(n), ctypes.byref(r),
                                    msg, 512)
        if rc < 0:
            self._check(-rc)
        return rc == 1, e.value, n.value, r.value, msg.value.decode()

    def dump(self):
        """All entries in key order as numpy arrays (key_offsets, keys, value_offsets, values)."""
        import numpy as np
        ks, vs = ctypes.c_uint64(), ctypes.c_uint64()
        n = self._lib.gmt_dump(self._h, None, None, None, None, ctypes.byref(ks), ctypes.byref(vs))
        if n < 0:
            raise MasstreeError(self._lib.gmt_last_error_message().decode())
        ko = np.zeros(n + 1, np.uint32)
        kd = np.zeros(max(ks.value, 1), np.uint32)
        vo = np.zeros(n + 1, np.uint32)
        vd = np.zeros(max(vs.value, 1), np.uint32)
        p = lambda a: ctypes.c_void_p(a.ctypes.data)
        n2 = self._lib.gmt_dump(self._h, p(ko), p(kd), p(vo), p(vd), None, None)
        if n2 != n:
            raise MasstreeError("tree changed between dump calls")
        return ko, kd[:ks.value], vo, vd[:vs.value]

    def multi_batch(self, request_types, keys, max_key_length, key_lengths, values,
                    max_value_length, value_lengths, results, offsets, num_batches,
                    total_requests, stream=None):
        """Batched form: applies num_batches mixed batches (stored back to back, split by
        the int64 `offsets` tensor of num_batches + 1 entries) in order, in one launch."""
        n = int(total_requests)
        self._check(self._lib.gmt_multi_batch(
            self._h, _ptr(request_types, 1, "request_types", n),
            _ptr(keys, 4, "keys", n * max_key_length), max_key_length,
            _ptr(key_lengths, 4, "key_lengths", n),
            _ptr(values, 4, "values", n * max_value_length), max_value_length,
            _ptr(value_lengths, 4, "value_lengths", n), _ptr(results, 1, "results", n),
            _ptr(offsets, 8, "offsets", int(num_batches) + 1), num_batches, n,
            _stream(stream)))

    def set_multi_mode(self, mode):
        """Batched form: 0 = one persistent launch, 1 = one launch per batch."""
        self._check(self._lib.gmt_set_multi_mode(self._h, int(mode)))

    def set_multi_grid_percent(self, p):
        self._check(self._lib.gmt_set_multi_grid_percent(self._h, int(p)))

    @property
    def tile_size(self):
        """Lanes per operation (16, 32 or 64); settable between calls."""
        return int(self._lib.gmt_tile_size(self._h))

    @tile_size.setter
    def tile_size(self, t):
        self._check(self._lib.gmt_set_tile_size(self._h, int(t)))

    def set_find_tile(self, t):
        """Lanes per lookup of read-only find batches: 1, 2, 4, 8, 16, or 0 (the tile
        width)."""
        self._check(self._lib.gmt_set_find_tile(self._h, int(t)))

    def set_find_passes(self, p):
        """Read-only finds as one kernel (1) or descent + record passes (2)."""
        self._check(self._lib.gmt_set_find_passes(self._h, int(p)))

    def set_record_tile(self, t):
        """Lanes per lookup of the record pass of two-pass finds: 1, 2, 4, 8, 16, or 0
        (the find tile)."""
        self._check(self._lib.gmt_set_record_tile(self._h, int(t)))

    def set_mixed_split(self, min_requests):
        """Mixed batches of at least min_requests requests run their finds as a separate
        read-only pass (0: every batch; 0xFFFFFFFF: none)."""
        self._check(self._lib.gmt_set_mixed_split(self._h, int(min_requests) & 0xFFFFFFFF))

    def set_write_hints(self, min_requests):
        """Insert, update and erase batches of at least min_requests requests start their
        writes at border nodes found by a read-only descent (0xFFFFFFFF: never)."""
        self._check(self._lib.gmt_set_write_hints(self._h, int(min_requests) & 0xFFFFFFFF))

    def reserve_find_scratch(self, max_requests):
        """Allocates the two-pass find scratch for batches of up to max_requests requests
        now (before capturing find or mixed-batch calls into a graph)."""
        self._check(self._lib.gmt_reserve_find_scratch(self._h, int(max_requests)))

    def set_find_depth(self, k):
        """Lookups each tile keeps in flight in read-only find batches (1, 2, 4, 8)."""
        self._check(self._lib.gmt_set_find_depth(self._h, int(k)))

    def slabs_allocated(self):
        return int(self._lib.gmt_slabs_allocated(self._h))

    def pool_slabs(self):
        return int(self._lib.gmt_pool_slabs(self._h))
