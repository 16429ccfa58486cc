"""Oracle (test infrastructure only): numpy-facing wrapper of liboracle.so.

liboracle.so is a host ``std::map<std::vector<u32>, std::vector<u32>>`` replay of
the reference's ``gpu_masstree`` batch API (semantics and reference citations are
in ``masstree_oracle.cpp``).  Only ``tests/``, ``__graft_entry__.smoke()`` and the
CPU-baseline leg of ``bench.py`` import this module; the product never does.
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")
_libs = {}

u32p = ctypes.POINTER(ctypes.c_uint32)
u8p = ctypes.POINTER(ctypes.c_uint8)


def build():
    """Compile liboracle.so with the committed Makefile."""
    subprocess.check_call(["make", "-s", "-C", _HERE])


def _load(path=None):
    """Loads liboracle.so (or, for the pinning tests, a variant built with a planted
    mistake at ``path``)."""
    path = path or _LIB_PATH
    if path in _libs:
        return _libs[path]
    if path == _LIB_PATH and not os.path.exists(path):
        build()
    lib = ctypes.CDLL(path)
    vp = ctypes.c_void_p
    u32 = ctypes.c_uint32
    u64 = ctypes.c_uint64
    lib.orc_create.restype = vp
    lib.orc_destroy.argtypes = [vp]
    lib.orc_size.argtypes = [vp]
    lib.orc_size.restype = u64
    lib.orc_clear.argtypes = [vp]
    lib.orc_insert.argtypes = [vp, u32p, u32, u32p, u32p, u32, u32p, u64, ctypes.c_int, u8p]
    lib.orc_update.argtypes = [vp, u32p, u32, u32p, u32p, u32, u32p, u64, u8p]
    lib.orc_erase.argtypes = [vp, u32p, u32, u32p, u64, u8p]
    lib.orc_find.argtypes = [vp, u32p, u32, u32p, u32p, u32, u32p, u64]
    lib.orc_scan.argtypes = [vp, u32p, u32p, u32, u32, u64, u32p, u32p, u32p, u32p, u32,
                             u32p, u32p, u32p, u32p]
    lib.orc_mixed.argtypes = [vp, u8p, u32p, u32, u32p, u32p, u32, u32p, u8p, u64]
    lib.orc_dump.argtypes = [vp, u32p, u32p, u32p, u32p, ctypes.POINTER(u64),
                             ctypes.POINTER(u64)]
    lib.orc_dump.restype = u64
    _libs[path] = lib
    return lib


def _p32(a):
    if a is None:
        return None
    assert a.dtype == np.uint32 and a.flags["C_CONTIGUOUS"], (a.dtype, a.flags)
    return a.ctypes.data_as(u32p)


def _p8(a):
    if a is None:
        return None
    assert a.dtype in (np.uint8, np.bool_) and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(u8p)


class Oracle:
    """std::map replay with the gpu_masstree batch signatures (numpy in/out)."""

    def __init__(self, lib_path=None):
        self._lib = _load(lib_path)
        self._h = self._lib.orc_create()

    def __del__(self):
        if getattr(self, "_h", None):
            self._lib.orc_destroy(self._h)
            self._h = None

    def __len__(self):
        return int(self._lib.orc_size(self._h))

    def clear(self):
        self._lib.orc_clear(self._h)

    def insert(self, keys, max_key_length, key_lengths, values, max_value_length,
               value_lengths, num_keys, update_if_exists=False, want_results=True):
        res = np.zeros(num_keys, np.uint8) if want_results else None
        self._lib.orc_insert(self._h, _p32(keys), max_key_length, _p32(key_lengths),
                             _p32(values), max_value_length, _p32(value_lengths), num_keys,
                             int(bool(update_if_exists)), _p8(res))
        return res

    def update(self, keys, max_key_length, key_lengths, values, max_value_length,
               value_lengths, num_keys):
        res = np.zeros(num_keys, np.uint8)
        self._lib.orc_update(self._h, _p32(keys), max_key_length, _p32(key_lengths),
                             _p32(values), max_value_length, _p32(value_lengths), num_keys,
                             _p8(res))
        return res

    def erase(self, keys, max_key_length, key_lengths, num_keys):
        res = np.zeros(num_keys, np.uint8)
        self._lib.orc_erase(self._h, _p32(keys), max_key_length, _p32(key_lengths), num_keys,
                            _p8(res))
        return res

    def find(self, keys, max_key_length, key_lengths, max_value_length, num_keys, fill=0):
        values = np.full(num_keys * max_value_length, fill, np.uint32)
        lens = np.zeros(num_keys, np.uint32)
        self._lib.orc_find(self._h, _p32(keys), max_key_length, _p32(key_lengths), _p32(values),
                           max_value_length, _p32(lens), num_keys)
        return values, lens

    def scan(self, lower_keys, lower_lengths, max_key_length, max_count, num_queries,
             upper_keys=None, upper_lengths=None, max_value_length=1, limits=None, fill=0):
        counts = np.zeros(num_queries, np.uint32)
        values = np.full(num_queries * max_count * max_value_length, fill, np.uint32)
        vlens = np.full(num_queries * max_count, fill, np.uint32)
        okeys = np.full(num_queries * max_count * max_key_length, fill, np.uint32)
        oklens = np.full(num_queries * max_count, fill, np.uint32)
        self._lib.orc_scan(self._h, _p32(lower_keys), _p32(lower_lengths), max_key_length,
                           max_count, num_queries, _p32(upper_keys), _p32(upper_lengths),
                           _p32(counts), _p32(values), max_value_length, _p32(vlens),
                           _p32(okeys), _p32(oklens), _p32(limits))
        return counts, values, vlens, okeys, oklens

    def mixed(self, types, keys, max_key_length, key_lengths, values, max_value_length,
              value_lengths, num_requests):
        """Replays a mixed batch.  ``values``/``value_lengths`` are copied, then finds
        write into the copies, as the GPU call writes in place."""
        values = values.copy()
        value_lengths = value_lengths.copy() if value_lengths is not None else \
            np.full(num_requests, max_value_length, np.uint32)
        res = np.zeros(num_requests, np.uint8)
        self._lib.orc_mixed(self._h, _p8(types), _p32(keys), max_key_length, _p32(key_lengths),
                            _p32(values), max_value_length, _p32(value_lengths), _p8(res),
                            num_requests)
        return values, value_lengths, res

    def dump(self):
        """All entries in key order: (key_offsets, key_data, value_offsets, value_data)."""
        ks = ctypes.c_uint64()
        vs = ctypes.c_uint64()
        n = self._lib.orc_dump(self._h, None, None, None, None, ctypes.byref(ks),
                               ctypes.byref(vs))
        ko = np.zeros(n + 1, np.uint32)
        kd = np.zeros(max(ks.value, 1), np.uint32)
        vo = np.zeros(n + 1, np.uint32)
        vd = np.zeros(max(vs.value, 1), np.uint32)
        self._lib.orc_dump(self._h, _p32(ko), _p32(kd), _p32(vo), _p32(vd), None, None)
        return ko, kd[:ks.value], vo, vd[:vs.value]

    def items(self):
        ko, kd, vo, vd = self.dump()
        return [(tuple(kd[ko[i]:ko[i + 1]].tolist()), tuple(vd[vo[i]:vo[i + 1]].tolist()))
                for i in range(len(ko) - 1)]
