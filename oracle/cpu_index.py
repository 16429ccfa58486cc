"""ctypes binding of libcpu_index.so: the multi-threaded CPU ordered index (sharded
std::map behind the reference CPU adapter's per-key shape) used as the timed CPU
baseline.  Test and benchmark infrastructure only."""
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "libcpu_index.so")
_lib = None


def _load():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            subprocess.check_call(["make", "-s", "-C", HERE, "libcpu_index.so"])
        lib = ctypes.CDLL(LIB)
        vp, u32, u64 = ctypes.c_void_p, ctypes.c_uint32, ctypes.c_uint64
        lib.cpx_create.restype = vp
        lib.cpx_create.argtypes = [u32]
        lib.cpx_destroy.argtypes = [vp]
        lib.cpx_size.argtypes = [vp]
        lib.cpx_size.restype = u64
        lib.cpx_insert.argtypes = [vp, vp, u32, vp, vp, u32, vp, u64, ctypes.c_int, vp, u32]
        lib.cpx_erase.argtypes = [vp, vp, u32, vp, u64, vp, u32]
        lib.cpx_find.argtypes = [vp, vp, u32, vp, vp, u32, vp, u64, u32]
        lib.cpx_scan.argtypes = [vp, vp, u32, vp, u32, u64, vp, vp, u32]
        lib.cpx_put1.argtypes = [vp, vp, u32, vp, vp, u64, u32]
        lib.cpx_get1.argtypes = [vp, vp, u32, vp, vp, u64, u32]
        _lib = lib
    return _lib


def _p(a, dtype=np.uint32):
    if a is None:
        return None
    assert a.dtype == dtype and a.flags.c_contiguous
    return ctypes.c_void_p(a.ctypes.data)


def _u32(a):
    return None if a is None else np.ascontiguousarray(a, dtype=np.uint32)


class CpuIndex:
    def __init__(self, shard_bits=8, threads=None):
        self._lib = _load()
        self._h = self._lib.cpx_create(shard_bits)
        self.threads = threads or max(1, os.cpu_count() or 1)

    def __del__(self):
        if getattr(self, "_h", None):
            self._lib.cpx_destroy(self._h)
            self._h = None

    def __len__(self):
        return int(self._lib.cpx_size(self._h))

    def insert(self, keys, maxK, lens, values, maxV, vlens, n, update_if_exists=False):
        keys, lens, values, vlens = _u32(keys), _u32(lens), _u32(values), _u32(vlens)
        r = np.zeros(n, np.uint8)
        self._lib.cpx_insert(self._h, _p(keys), maxK, _p(lens), _p(values), maxV, _p(vlens), n,
                             int(update_if_exists), _p(r, np.uint8), self.threads)
        return r

    def erase(self, keys, maxK, lens, n):
        keys, lens = _u32(keys), _u32(lens)
        r = np.zeros(n, np.uint8)
        self._lib.cpx_erase(self._h, _p(keys), maxK, _p(lens), n, _p(r, np.uint8), self.threads)
        return r

    def find(self, keys, maxK, lens, maxV, n):
        keys, lens = _u32(keys), _u32(lens)
        v = np.zeros(n * maxV, np.uint32)
        l = np.zeros(n, np.uint32)
        self._lib.cpx_find(self._h, _p(keys), maxK, _p(lens), _p(v), maxV, _p(l), n, self.threads)
        return v, l

    def scan(self, lower, maxK, lens, count, nq):
        lower, lens = _u32(lower), _u32(lens)
        v = np.zeros(nq * count, np.uint32)
        c = np.zeros(nq, np.uint32)
        self._lib.cpx_scan(self._h, _p(lower), maxK, _p(lens), count, nq, _p(v), _p(c),
                           self.threads)
        return c, v

    # ---- the reference CPU adapter's per-key shape (one-slice values, 0xFFFFFFFF absent)
    ABSENT = 0xFFFFFFFF

    def put1(self, keys, maxK, lens, values, n):
        """Insert that overwrites an existing value (cpu_masstree_adapter.hpp:96-104)."""
        keys, lens, values = _u32(keys), _u32(lens), _u32(values)
        self._lib.cpx_put1(self._h, _p(keys), maxK, _p(lens), _p(values), n, self.threads)

    def get1(self, keys, maxK, lens, n):
        """The first value slice of each key, ABSENT if the key is not stored
        (cpu_masstree_adapter.hpp:120-125)."""
        keys, lens = _u32(keys), _u32(lens)
        out = np.zeros(n, np.uint32)
        self._lib.cpx_get1(self._h, _p(keys), maxK, _p(lens), _p(out), n, self.threads)
        return out


def result_line(name, mops_runs):
    """The reference bench's result line, "<op>: <avg> Mop/s (<n>; <run> <run> ... )"
    (evaluation/universal_bench.cu:65-75), which its evaluation script parses."""
    avg = sum(mops_runs) / len(mops_runs)
    return f"{name}: {avg:g} Mop/s ({len(mops_runs)}; " + "".join(f"{m:g} " for m in mops_runs) + ")"
