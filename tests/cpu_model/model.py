"""ctypes wrapper of libcpu_model.so (test infrastructure only)."""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = os.path.join(_HERE, "libcpu_model.so")
_lib = None
u32p = ctypes.POINTER(ctypes.c_uint32)
u8p = ctypes.POINTER(ctypes.c_uint8)


def _load():
    global _lib
    if _lib is None:
        subprocess.check_call(["make", "-s", "-C", _HERE])
        lib = ctypes.CDLL(_LIB)
        vp, u32, u64 = ctypes.c_void_p, ctypes.c_uint32, ctypes.c_uint64
        lib.cm_create.argtypes = [u32]
        lib.cm_create.restype = vp
        lib.cm_destroy.argtypes = [vp]
        lib.cm_error.argtypes = [vp]
        lib.cm_error.restype = u32
        lib.cm_slabs_used.argtypes = [vp]
        lib.cm_slabs_used.restype = u32
        lib.cm_retired.argtypes = [vp]
        lib.cm_retired.restype = u64
        lib.cm_set_hinted.argtypes = [vp, u32]
        lib.cm_set_fail.argtypes = [vp, u32, u32]
        lib.cm_set_no_suffix.argtypes = [vp, u32]
        lib.cm_set_erase_flags.argtypes = [vp, u32]
        lib.cm_set_max_op_time.argtypes = [vp, u32]
        lib.cm_trip.argtypes = [vp, u32p]
        lib.cm_root.argtypes = [vp]
        lib.cm_root.restype = u32
        lib.cm_set_lock.argtypes = [vp, u32, u32]
        lib.cm_set_lock.restype = u32
        lib.cm_build_shortcut.argtypes = [vp, u32]
        lib.cm_check_shortcut.argtypes = [vp]
        lib.cm_check_shortcut.restype = u64
        lib.cm_unlinked.argtypes = [vp]
        lib.cm_unlinked.restype = u64
        lib.cm_nodes.argtypes = [vp, u32]
        lib.cm_nodes.restype = u64
        lib.cm_layers.argtypes = [vp]
        lib.cm_layers.restype = u64
        lib.cm_batch.argtypes = [vp, u8p, u32, u32p, u32, u32p, u32p, u32, u32p, u32p, u32p,
                                 u8p, u64, u32, u32, u64]
        lib.cm_dump.argtypes = [vp, u32p, u32p, u32p, u32p, ctypes.POINTER(u64),
                                ctypes.POINTER(u64), ctypes.c_char_p, u32, ctypes.POINTER(u64)]
        lib.cm_dump.restype = u64
        _lib = lib
    return _lib


def _p32(a):
    return None if a is None else a.ctypes.data_as(u32p)


def _p8(a):
    return None if a is None else a.ctypes.data_as(u8p)


class CpuModel:
    def __init__(self, num_slabs=1 << 20):
        self.lib = _load()
        self.h = self.lib.cm_create(num_slabs)

    def __del__(self):
        if getattr(self, "h", None):
            self.lib.cm_destroy(self.h)
            self.h = None

    def set_hinted(self, on):
        """Writes of later batches start at hints from a read-only walk made before the
        batch runs (the GPU write passes' start hints)."""
        self.lib.cm_set_hinted(self.h, 1 if on else 0)

    def set_no_suffix(self, on):
        """Inserts of later batches build one layer per remaining key slice instead of a
        suffix record (insert<..., enable_suffix = false>)."""
        self.lib.cm_set_no_suffix(self.h, 1 if on else 0)

    def set_erase_flags(self, flags):
        """Request flags for every request: 4 = no merging of emptied border nodes
        (do_merge = false), 2 = emptied layers stay (do_remove_empty_root = false)."""
        self.lib.cm_set_erase_flags(self.h, flags)

    def set_fanout(self, border_cap=0, interior_cap=0, erase_flags=0):
        """Node fanout for later writes (request flag bits 8-15; 0 = the maximum), with the
        erase flags kept beside it."""
        self.lib.cm_set_erase_flags(self.h, erase_flags | (border_cap << 8) | (interior_cap << 12))

    def set_max_op_time(self, ticks):
        """The livelock guard's time limit in 10.24-us ticks."""
        self.lib.cm_set_max_op_time(self.h, ticks)

    def trip(self):
        """The guard's record of the operation it abandoned (16 words; zeros if none)."""
        out = np.zeros(16, np.uint32)
        self.lib.cm_trip(self.h, _p32(out))
        return out

    @property
    def root(self):
        return int(self.lib.cm_root(self.h))

    def set_lock(self, idx, on):
        """Sets or clears a node's lock bit; returns the header before."""
        return int(self.lib.cm_set_lock(self.h, idx, 1 if on else 0))

    def set_fail(self, any_every=0, parent_every=0):
        """Injected allocation failures: every k-th allocation fails, and every k-th
        allocation of a split inserting into its parent (0 = off).  Clears the error."""
        self.lib.cm_set_fail(self.h, any_every, parent_every)

    @property
    def layers(self):
        """Trie layers reachable from the root."""
        return int(self.lib.cm_layers(self.h))

    @property
    def retired(self):
        """Objects retired so far (records; nodes carry bit 31)."""
        return int(self.lib.cm_retired(self.h))

    def nodes(self, which=0):
        """Reachable nodes (0), border nodes (1), empty non-root border nodes (2), the
        largest border entry count (3) or interior child count (4); merges of layer-0 border nodes counted (5)."""
        return int(self.lib.cm_nodes(self.h, which))

    def build_shortcut(self, bits):
        """Builds a border-level shortcut table of 2**bits entries that later merges rewrite."""
        self.lib.cm_build_shortcut(self.h, bits)

    def check_shortcut(self):
        """Entries of the table that are no valid start for their prefix (0 when it is sound)."""
        return int(self.lib.cm_check_shortcut(self.h))

    @property
    def unlinked(self):
        """Non-root nodes that no parent slot names (reached through sibling links)."""
        return int(self.lib.cm_unlinked(self.h))

    @property
    def error(self):
        return int(self.lib.cm_error(self.h))

    def batch(self, kind, keys, maxK, klens, values=None, maxV=1, vlens=None, n=None,
              kinds=None, want_find_out=False, ntiles=8, nthreads=0, seed=0, mixed=False):
        n = n if n is not None else (len(klens) if klens is not None else keys.size // maxK)
        out_values = out_vlens = None
        if want_find_out:
            out_values = np.zeros(n * maxV, np.uint32)
            out_vlens = np.zeros(n, np.uint32)
        results = np.zeros(n, np.uint8)
        if mixed:
            values = values.copy()
            out_vlens = (vlens.copy() if vlens is not None
                         else np.full(n, maxV, np.uint32))
            vlens = out_vlens
        self.lib.cm_batch(self.h, _p8(kinds), kind, _p32(keys), maxK, _p32(klens),
                          _p32(values), maxV, _p32(vlens), _p32(out_values), _p32(out_vlens),
                          _p8(results), n, ntiles, nthreads, seed)
        if mixed:
            return values, out_vlens, results
        return out_values, out_vlens, results

    def dump(self):
        ks, vs, nodes = ctypes.c_uint64(), ctypes.c_uint64(), ctypes.c_uint64()
        err = ctypes.create_string_buffer(256)
        n = self.lib.cm_dump(self.h, None, None, None, None, ctypes.byref(ks), ctypes.byref(vs),
                             err, 256, ctypes.byref(nodes))
        if n == 2**64 - 1:
            raise AssertionError("structure invalid: " + err.value.decode())
        ko = np.zeros(n + 1, np.uint32)
        kd = np.zeros(max(ks.value, 1), np.uint32)
        vo = np.zeros(n + 1, np.uint32)
        vd = np.zeros(max(vs.value, 1), np.uint32)
        self.lib.cm_dump(self.h, _p32(ko), _p32(kd), _p32(vo), _p32(vd), None, None, err, 256,
                         None)
        return ko, kd[:ks.value], vo, vd[:vs.value]

    def items(self):
        ko, kd, vo, vd = self.dump()
        return [(tuple(kd[ko[i]:ko[i + 1]].tolist()), tuple(vd[vo[i]:vo[i + 1]].tolist()))
                for i in range(len(ko) - 1)]
