"""Served form: host binding of gmt_server_* (a request loop with continuous batching).

A trace is an arena of device tensors (point ops in mixed-batch layout, scans, a ring
for scan results) plus a host array of requests (arrival time, op/scan ranges).  The
server replays it in real time and returns, per request, when its wave was launched and
completed; see include/gpu_masstree_c.h for the contract.
"""
import ctypes

import numpy as np
import torch

from .masstree import MasstreeError, _ptr, load_library

u32, u64 = ctypes.c_uint32, ctypes.c_uint64
vp = ctypes.c_void_p


class ServeConfig(ctypes.Structure):
    _fields_ = [("streams", u32), ("slabs_per_write", u32), ("max_wave_ops", u64),
                ("max_wave_scans", u64), ("ring_rows", u64), ("epoch_waves", u32),
                ("coalesce", u32), ("ordered", u32), ("plain_batch", u32)]


class ServeArena(ctypes.Structure):
    _fields_ = [("types", vp), ("keys", vp), ("max_key_length", u32), ("key_lengths", vp),
                ("values", vp), ("max_value_length", u32), ("value_lengths", vp),
                ("results", vp), ("scan_lower", vp), ("scan_lower_lengths", vp),
                ("scan_upper", vp), ("scan_upper_lengths", vp), ("scan_limits", vp),
                ("scan_max_count", u32), ("scan_counts", vp), ("ring_keys", vp),
                ("ring_key_lengths", vp), ("ring_values", vp), ("ring_value_lengths", vp)]


class ServeStats(ctypes.Structure):
    _fields_ = [("waves", u64), ("ops", u64), ("scans", u64), ("flips", u64), ("drains", u64),
                ("census", u64), ("admission_waits", u64), ("elapsed_us", ctypes.c_double),
                ("max_inflight", u32), ("device_error", u32)]


REQUEST_DTYPE = np.dtype([("arrival_us", "f8"), ("op_offset", "u8"), ("scan_offset", "u8"),
                          ("num_ops", "u4"), ("num_scans", "u4"), ("num_writes", "u4")],
                         align=True)
OUTCOME_DTYPE = np.dtype([("admit_us", "f8"), ("done_us", "f8"), ("ring_row", "u8"),
                          ("wave", "u4")], align=True)


def _bind(lib):
    if getattr(lib, "_serve_bound", False):
        return lib
    P = ctypes.POINTER
    lib.gmt_server_create.argtypes = [vp, P(ServeConfig), P(vp)]
    lib.gmt_server_destroy.argtypes = [vp]
    lib.gmt_server_run.argtypes = [vp, P(ServeArena), vp, u64, vp, P(ServeStats)]
    lib._serve_bound = True
    return lib


class MasstreeServer:
    """Request loop over one GpuMasstree (the index's tile width at creation)."""

    def __init__(self, tree, streams=4, max_wave_ops=1 << 20, max_wave_scans=1 << 16,
                 ring_rows=1 << 22, slabs_per_write=4, epoch_waves=32, coalesce=True,
                 ordered=False, plain_batch=False):
        self._lib = _bind(load_library())
        self.tree = tree
        self.config = ServeConfig(streams, slabs_per_write, max_wave_ops, max_wave_scans,
                                  ring_rows, epoch_waves, int(bool(coalesce)), int(bool(ordered)),
                                  int(bool(plain_batch)))
        h = vp()
        self._check(self._lib.gmt_server_create(tree._h, ctypes.byref(self.config), ctypes.byref(h)))
        self._h = h

    def _check(self, rc):
        if rc != 0:
            raise MasstreeError(f"gmt_server call failed ({rc}): "
                                f"{self._lib.gmt_last_error_message().decode()}")

    def close(self):
        if getattr(self, "_h", None):
            self._lib.gmt_server_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def run(self, arena, requests):
        """arena: dict of device tensors (see gmt_serve_arena); requests: array of
        REQUEST_DTYPE.  Returns (outcomes array of OUTCOME_DTYPE, stats dict)."""
        req = np.ascontiguousarray(requests, dtype=REQUEST_DTYPE)
        n = len(req)
        mk, mv = int(arena["max_key_length"]), int(arena["max_value_length"])
        mc = int(arena.get("scan_max_count", 0))
        n_ops = int(req["op_offset"][-1] + req["num_ops"][-1]) if n else 0
        n_scans = int(req["scan_offset"][-1] + req["num_scans"][-1]) if n else 0
        ring = int(self.config.ring_rows) if self.config.coalesce else n_scans * mc
        g = arena.get
        a = ServeArena(
            _ptr(g("types"), 1, "types", n_ops), _ptr(g("keys"), 4, "keys", n_ops * mk), mk,
            _ptr(g("key_lengths"), 4, "key_lengths", n_ops),
            _ptr(g("values"), 4, "values", n_ops * mv), mv,
            _ptr(g("value_lengths"), 4, "value_lengths", n_ops),
            _ptr(g("results"), 1, "results", n_ops),
            _ptr(g("scan_lower"), 4, "scan_lower", n_scans * mk),
            _ptr(g("scan_lower_lengths"), 4, "scan_lower_lengths", n_scans),
            _ptr(g("scan_upper"), 4, "scan_upper", n_scans * mk),
            _ptr(g("scan_upper_lengths"), 4, "scan_upper_lengths", n_scans),
            _ptr(g("scan_limits"), 4, "scan_limits", n_scans), mc,
            _ptr(g("scan_counts"), 4, "scan_counts", n_scans),
            _ptr(g("ring_keys"), 4, "ring_keys", ring * mk if n_scans else 0),
            _ptr(g("ring_key_lengths"), 4, "ring_key_lengths", ring if n_scans else 0),
            _ptr(g("ring_values"), 4, "ring_values", ring * mv if n_scans else 0),
            _ptr(g("ring_value_lengths"), 4, "ring_value_lengths", ring if n_scans else 0))
        out = np.zeros(n, OUTCOME_DTYPE)
        st = ServeStats()
        torch.cuda.synchronize()
        self._check(self._lib.gmt_server_run(self._h, ctypes.byref(a), vp(req.ctypes.data), n,
                                             vp(out.ctypes.data), ctypes.byref(st)))
        stats = {f: getattr(st, f) for f, _ in ServeStats._fields_}
        return out, stats


def latency_summary(requests, outcomes):
    """p50 / p99 / p99.9 / max latency (us) from arrival to completion."""
    lat = outcomes["done_us"] - requests["arrival_us"]
    q = np.percentile(lat, [50, 99, 99.9]) if len(lat) else [0, 0, 0]
    return dict(p50_us=float(q[0]), p99_us=float(q[1]), p999_us=float(q[2]),
                max_us=float(lat.max()) if len(lat) else 0.0)
