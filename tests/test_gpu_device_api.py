"""The device-side API (gpu_masstree::cooperative_insert / _update / _find / _erase, the
reference's include/gpu_masstree.hpp:194-971) called from a user kernel, one request per
16-lane tile: every result equals the oracle's, over keys of 1..24 slices sharing
prefixes (trie layers, suffix records) and values of 0..40 slices (inline values and
multi-slab records).  The kernel lives in tests/device_api (built by build())."""
import ctypes
import os

import numpy as np
import pytest
import torch

import scenarios as S
from indexingpu_amd import workloads as W

pytestmark = pytest.mark.gpu

LIB = os.path.join(os.path.dirname(__file__), "device_api", "libdevice_api_probe.so")
INSERT, ERASE, FIND, UPDATE, UPSERT = 0, 1, 2, 3, 4


class DeviceApiIndex:
    def __init__(self, pool_bytes=256 << 20):
        if not os.path.exists(LIB):
            raise RuntimeError(f"{LIB} is missing: run __graft_entry__.build()")
        self.lib = ctypes.CDLL(LIB)
        vp, u32 = ctypes.c_void_p, ctypes.c_uint32
        self.lib.dapi_create.restype = vp
        self.lib.dapi_create.argtypes = [ctypes.c_uint64]
        self.lib.dapi_destroy.argtypes = [vp]
        self.lib.dapi_apply.restype = u32
        self.lib.dapi_apply.argtypes = [vp, u32, vp, u32, vp, vp, u32, vp, vp, vp, vp, u32]
        self.h = self.lib.dapi_create(pool_bytes)
        assert self.h

    @staticmethod
    def _d(a, dtype=np.uint32):
        a = np.ascontiguousarray(a, dtype=dtype)
        view = np.int32 if dtype == np.uint32 else np.uint8
        return torch.from_numpy(a.view(view).reshape(-1).copy()).cuda()

    def apply(self, op, keys, lens, maxK, vals=None, vlens=None, maxV=1):
        n = len(lens)
        dk, dl = self._d(keys), self._d(lens)
        dv = self._d(vals) if vals is not None else None
        dvl = self._d(vlens) if vlens is not None else None
        ov = torch.zeros(max(n * maxV, 1), dtype=torch.int32, device="cuda")
        ol = torch.zeros(max(n, 1), dtype=torch.int32, device="cuda")
        r = torch.zeros(max(n, 1), dtype=torch.uint8, device="cuda")
        p = lambda t: ctypes.c_void_p(t.data_ptr()) if t is not None else None
        torch.cuda.synchronize()
        err = self.lib.dapi_apply(self.h, op, p(dk), maxK, p(dl), p(dv), maxV, p(dvl), p(ov), p(ol),
                                  p(r), n)
        torch.cuda.synchronize()
        assert err == 0, f"device error {err}"
        return (ov.cpu().numpy().view(np.uint32)[:n * maxV].reshape(n, maxV),
                ol.cpu().numpy().view(np.uint32)[:n], r.cpu().numpy()[:n])

    def close(self):
        self.lib.dapi_destroy(self.h)


def _keyset(seed, n, mk):
    rng = np.random.default_rng(seed)
    base, blen = W.unique_random_keys(rng, n, 1, mk)
    keys, lens = base.copy(), blen.copy()
    donor = rng.integers(0, len(lens), size=len(lens))
    share = rng.integers(0, mk, size=len(lens))
    for i in range(0, len(lens), 2):
        p = min(share[i], lens[i] - 1, lens[donor[i]])
        keys[i, :p] = base[donor[i], :p]
    _, first = np.unique(W.key_hash64(keys, lens), return_index=True)
    first = np.sort(first)
    return keys[first], lens[first]


def test_device_api_matches_oracle():
    mk, mv = 24, 40
    keys, lens = _keyset(41, 6000, mk)
    n = len(lens)
    rng = np.random.default_rng(42)
    vals = rng.integers(0, 1 << 32, size=(n, mv), dtype=np.uint64).astype(np.uint32)
    vl = rng.choice([0, 1, 2, 7, 30, 31, 40], size=n).astype(np.uint32)
    g, o = DeviceApiIndex(), S.OracleBackend()
    try:
        half = n // 2
        # insert half, then all (the first half reports "exists")
        _, _, r = g.apply(INSERT, keys[:half], lens[:half], mk, vals[:half], vl[:half], mv)
        assert np.array_equal(r, o.insert(keys[:half], lens[:half], vals[:half], vl[:half], mk, mv))
        _, _, r = g.apply(INSERT, keys, lens, mk, vals, vl, mv)
        assert np.array_equal(r, o.insert(keys, lens, vals, vl, mk, mv))
        # find everything (hits, with values truncated to max_value_length = 5 as well)
        for maxv in (mv, 5):
            gv, gl, _ = g.apply(FIND, keys, lens, mk, maxV=maxv)
            ev, el = o.find(keys, lens, mk, maxv)
            assert np.array_equal(gl, el) and np.array_equal(gv, ev)
        # update a third with new values of other lengths; upsert another third
        u = rng.choice(n, n // 3, replace=False)
        nv = vals[u] ^ np.uint32(0x5A5A5A5A)
        nvl = rng.choice([1, 3, 33], size=len(u)).astype(np.uint32)
        _, _, r = g.apply(UPDATE, keys[u], lens[u], mk, nv, nvl, mv)
        assert np.array_equal(r, o.update(keys[u], lens[u], nv, nvl, mk, mv))
        w = rng.choice(n, n // 3, replace=False)
        _, _, r = g.apply(UPSERT, keys[w], lens[w], mk, vals[w], vl[w], mv)
        assert np.array_equal(r, o.insert(keys[w], lens[w], vals[w], vl[w], mk, mv, upsert=True))
        # erase half (some twice across batches), then find everything again
        e = rng.choice(n, n // 2, replace=False)
        _, _, r = g.apply(ERASE, keys[e], lens[e], mk)
        assert np.array_equal(r, o.erase(keys[e], lens[e], mk))
        e2 = rng.choice(n, n // 4, replace=False)
        _, _, r = g.apply(ERASE, keys[e2], lens[e2], mk)
        assert np.array_equal(r, o.erase(keys[e2], lens[e2], mk))
        gv, gl, _ = g.apply(FIND, keys, lens, mk, maxV=mv)
        ev, el = o.find(keys, lens, mk, mv)
        assert np.array_equal(gl, el) and np.array_equal(gv, ev)
    finally:
        g.close()
