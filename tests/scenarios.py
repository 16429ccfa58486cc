"""Backends (oracle on the host, GPU index through the C ABI) with one numpy interface,
and the reference unit-test scenarios written once against it.

Scenario bodies follow test/unittest_masstree.cu:318-708 (hyoungjook/IndexinGPU).  Every
output buffer starts zero-filled in both backends, so outputs can be compared as whole
arrays: slots a call must not write stay zero in both.
"""
import numpy as np

from oracle.oracle import Oracle


def _u32(a):
    return np.ascontiguousarray(a, dtype=np.uint32).reshape(-1)


class OracleBackend:
    name = "oracle"

    def __init__(self):
        self.o = Oracle()

    def insert(self, keys, lens, vals, vlens, maxK, maxV, upsert=False):
        n = len(lens)
        return self.o.insert(_u32(keys), maxK, _u32(lens) if lens is not None else None,
                             _u32(vals), maxV, _u32(vlens) if vlens is not None else None, n,
                             update_if_exists=upsert)

    def update(self, keys, lens, vals, vlens, maxK, maxV):
        return self.o.update(_u32(keys), maxK, _u32(lens), _u32(vals), maxV,
                             _u32(vlens) if vlens is not None else None, len(lens))

    def erase(self, keys, lens, maxK):
        return self.o.erase(_u32(keys), maxK, _u32(lens), len(lens))

    def find(self, keys, lens, maxK, maxV, n=None):
        n = len(lens) if n is None else n
        v, l = self.o.find(_u32(keys), maxK, _u32(lens) if lens is not None else None, maxV, n)
        return v.reshape(n, maxV), l

    def scan(self, lower, lower_lens, maxK, max_count, upper=None, upper_lens=None, maxV=1,
             limits=None):
        nq = len(lower_lens)
        c, v, vl, k, kl = self.o.scan(_u32(lower), _u32(lower_lens), maxK, max_count, nq,
                                      _u32(upper) if upper is not None else None,
                                      _u32(upper_lens) if upper_lens is not None else None,
                                      maxV, _u32(limits) if limits is not None else None)
        return c, v, vl, k, kl

    def mixed(self, types, keys, lens, vals, vlens, maxK, maxV):
        v, l, r = self.o.mixed(np.ascontiguousarray(types, np.uint8), _u32(keys), maxK, _u32(lens),
                               _u32(vals), maxV, _u32(vlens), len(types))
        return v.reshape(len(types), maxV), l, r

    def items(self):
        return self.o.items()

    def close(self):
        pass


class GpuBackend:
    name = "gpu"

    def __init__(self, tile_size=16, pool_bytes=64 << 20, fanout=(0, 0)):
        import torch
        from indexingpu_amd import GpuMasstree
        self.torch = torch
        self.t = GpuMasstree(tile_size=tile_size, pool_bytes=pool_bytes)
        if fanout != (0, 0):
            self.t.set_fanout(*fanout)

    def _d(self, a, dtype=np.uint32):
        torch = self.torch
        if a is None:
            return None
        a = np.ascontiguousarray(a, dtype=dtype).reshape(-1)
        if dtype == np.uint32:
            return torch.from_numpy(a.view(np.int32).copy()).cuda()
        return torch.from_numpy(a.copy()).cuda()

    def _h(self, t, dtype=np.uint32):
        self.torch.cuda.synchronize()
        a = t.cpu().numpy()
        return a.view(dtype) if dtype == np.uint32 else a.astype(dtype)

    def _zeros(self, n, dtype=np.uint32):
        torch = self.torch
        return torch.zeros(max(n, 1), dtype=torch.int32 if dtype == np.uint32 else torch.uint8,
                           device="cuda")

    def _check(self):
        self.t.check_device_error()

    def insert(self, keys, lens, vals, vlens, maxK, maxV, upsert=False):
        n = len(lens)
        r = self._zeros(n, np.uint8)
        self.t.insert(self._d(keys), maxK, self._d(lens), self._d(vals), maxV, self._d(vlens), n,
                      update_if_exists=upsert, results=r)
        out = self._h(r, np.uint8)[:n]
        self._check()
        return out

    def update(self, keys, lens, vals, vlens, maxK, maxV):
        n = len(lens)
        r = self._zeros(n, np.uint8)
        self.t.update(self._d(keys), maxK, self._d(lens), self._d(vals), maxV, self._d(vlens), n,
                      results=r)
        out = self._h(r, np.uint8)[:n]
        self._check()
        return out

    def erase(self, keys, lens, maxK):
        n = len(lens)
        r = self._zeros(n, np.uint8)
        self.t.erase(self._d(keys), maxK, self._d(lens), n, results=r)
        out = self._h(r, np.uint8)[:n]
        self._check()
        return out

    def find(self, keys, lens, maxK, maxV, n=None, concurrent=False):
        n = len(lens) if n is None else n
        v = self._zeros(n * maxV)
        l = self._zeros(n)
        self.t.find(self._d(keys), maxK, self._d(lens), v, maxV, l, n, concurrent=concurrent)
        vo = self._h(v)[:n * maxV].reshape(n, maxV)
        lo = self._h(l)[:n]
        self._check()
        return vo, lo

    def scan(self, lower, lower_lens, maxK, max_count, upper=None, upper_lens=None, maxV=1,
             limits=None):
        nq = len(lower_lens)
        c = self._zeros(nq)
        v = self._zeros(nq * max_count * maxV)
        vl = self._zeros(nq * max_count)
        k = self._zeros(nq * max_count * maxK)
        kl = self._zeros(nq * max_count)
        self.t.scan(self._d(lower), self._d(lower_lens), maxK, max_count, nq, self._d(upper),
                    self._d(upper_lens), c, v, maxV, vl, k, kl, per_query_limits=self._d(limits))
        out = (self._h(c)[:nq], self._h(v)[:nq * max_count * maxV], self._h(vl)[:nq * max_count],
               self._h(k)[:nq * max_count * maxK], self._h(kl)[:nq * max_count])
        self._check()
        return out

    def mixed(self, types, keys, lens, vals, vlens, maxK, maxV):
        n = len(types)
        v = self._d(vals)
        l = self._d(vlens)
        r = self._zeros(n, np.uint8)
        self.t.mixed_batch(self._d(types, np.uint8), self._d(keys), maxK, self._d(lens), v, maxV, l,
                           r, n)
        out = (self._h(v)[:n * maxV].reshape(n, maxV), self._h(l)[:n], self._h(r, np.uint8)[:n])
        self._check()
        return out

    def items(self):
        ko, kd, vo, vd = self.t.dump()
        return [(tuple(kd[ko[i]:ko[i + 1]].tolist()), tuple(vd[vo[i]:vo[i + 1]].tolist()))
                for i in range(len(ko) - 1)]

    def validate(self):
        return self.t.validate()

    def close(self):
        self.t.close()


# ------------------------------------------------------------------------ scenarios
# Each returns a dict of named outputs (numpy arrays or lists).
def sc_validate(b, inp):
    mk, mv = inp.max_key_length, inp.max_value_length
    r = b.insert(inp.keys, inp.lengths, inp.values, inp.value_lengths, mk, mv)
    return dict(insert=r, items=b.items())


def sc_find_exist(b, inp):
    mk, mv = inp.max_key_length, inp.max_value_length
    r = b.insert(inp.keys, inp.lengths, inp.values, inp.value_lengths, mk, mv)
    v, l = b.find(inp.keys, inp.lengths, mk, mv)
    return dict(insert=r, find_values=v, find_lens=l)


def sc_find_notexist(b, inp):
    mk, mv = inp.max_key_length, inp.max_value_length
    b.insert(inp.keys, inp.lengths, inp.values, inp.value_lengths, mk, mv)
    v, l = b.find(inp.keys_not_exist, inp.lengths, mk, mv)
    return dict(find_values=v, find_lens=l)


def sc_update(b, inp):
    mk, mv = inp.max_key_length, inp.max_value_length
    b.insert(inp.keys, inp.lengths, inp.values, inp.value_lengths, mk, mv)
    r = b.update(inp.keys, inp.lengths, inp.values2, inp.value_lengths, mk, mv)
    v, l = b.find(inp.keys, inp.lengths, mk, mv)
    return dict(update=r, find_values=v, find_lens=l, items=b.items())


def sc_erase_all(b, inp):
    mk, mv = inp.max_key_length, inp.max_value_length
    b.insert(inp.keys, inp.lengths, inp.values, inp.value_lengths, mk, mv)
    r = b.erase(inp.keys, inp.lengths, mk)
    v, l = b.find(inp.keys, inp.lengths, mk, mv)
    return dict(erase=r, find_values=v, find_lens=l, items=b.items())


def sc_erase_none(b, inp):
    mk, mv = inp.max_key_length, inp.max_value_length
    b.insert(inp.keys, inp.lengths, inp.values, inp.value_lengths, mk, mv)
    r = b.erase(inp.keys_not_exist, inp.lengths, mk)
    v, l = b.find(inp.keys, inp.lengths, mk, mv)
    return dict(erase=r, find_values=v, find_lens=l)


def sc_insert_twice_erase_all(b, inp):
    """inp built with duplicate=True: the batch holds every key twice."""
    mk, mv = inp.max_key_length, inp.max_value_length
    ri = b.insert(inp.keys, inp.lengths, inp.values, inp.value_lengths, mk, mv)
    re = b.erase(inp.keys, inp.lengths, mk)
    v, l = b.find(inp.keys, inp.lengths, mk, mv)
    return dict(insert_dup=ri, erase_dup=re, find_values=v, find_lens=l, items=b.items())


def sc_erase_all_insert_all(b, inp):
    mk, mv = inp.max_key_length, inp.max_value_length
    b.insert(inp.keys, inp.lengths, inp.values, inp.value_lengths, mk, mv)
    b.erase(inp.keys, inp.lengths, mk)
    v1, l1 = b.find(inp.keys, inp.lengths, mk, mv)
    r = b.insert(inp.keys, inp.lengths, inp.values, inp.value_lengths, mk, mv)
    v2, l2 = b.find(inp.keys, inp.lengths, mk, mv)
    return dict(find1_values=v1, find1_lens=l1, insert2=r, find_values=v2, find_lens=l2)


def sc_erase_all_twice(b, inp):
    """inp built with duplicate=True: insert the first half, erase every key (twice)."""
    mk, mv = inp.max_key_length, inp.max_value_length
    h = (inp.num_keys + 1) // 2
    b.insert(inp.keys[:h], inp.lengths[:h], inp.values[:h], inp.value_lengths[:h], mk, mv)
    re = b.erase(inp.keys, inp.lengths, mk)
    v, l = b.find(inp.keys[:h], inp.lengths[:h], mk, mv)
    return dict(erase_dup=re, find_values=v, find_lens=l, items=b.items())


def sc_range_query(b, inp, rinp):
    mk, mv = inp.max_key_length, inp.max_value_length
    b.insert(inp.keys, inp.lengths, inp.values, inp.value_lengths, mk, mv)
    c, v, vl, k, kl = b.scan(rinp.lower, rinp.lower_len, mk, rinp.max_count, rinp.upper,
                             rinp.upper_len, mv)
    return dict(counts=c, values=v, value_lengths=vl, keys=k, key_lengths=kl)


def sc_concurrent_mix(b, inp, mix):
    mk, mv = inp.max_key_length, inp.max_value_length
    ks = mix.num_keyset
    b.insert(inp.keys[:2 * ks], inp.lengths[:2 * ks], inp.values[:2 * ks],
             inp.value_lengths[:2 * ks], mk, mv)
    v, l, r = b.mixed(mix.types, mix.keys, mix.lens, mix.values, mix.vlens, mk, mv)
    fv, fl = b.find(inp.keys[:4 * ks], inp.lengths[:4 * ks], mk, mv)
    return dict(mix_values=v, mix_lens=l, mix_results=r, find_values=fv, find_lens=fl,
                items=b.items())


class CpuModelBackend:
    """The shared state machine run on the host (tests/cpu_model); no scans."""
    name = "cpu_model"

    def __init__(self, nthreads=0, ntiles=32, seed=0, hinted=False, allow_error=False,
                 no_suffix=False, fanout=(0, 0)):
        from cpu_model.model import CpuModel
        self.m = CpuModel(1 << 21)
        self.m.set_no_suffix(no_suffix)
        if fanout != (0, 0):
            self.m.set_fanout(*fanout)
        self.allow_error = allow_error
        self.m.set_hinted(hinted)
        self.nthreads, self.ntiles, self.seed = nthreads, ntiles, seed

    def _run(self, kind, keys, lens, maxK, vals=None, vlens=None, maxV=1, find=False, kinds=None,
             mixed=False):
        self.seed += 1
        out = self.m.batch(kind, _u32(keys), maxK, _u32(lens),
                           _u32(vals) if vals is not None else None, maxV,
                           _u32(vlens) if vlens is not None else None, n=len(lens),
                           kinds=kinds, want_find_out=find, ntiles=self.ntiles,
                           nthreads=self.nthreads, seed=self.seed, mixed=mixed)
        assert self.allow_error or self.m.error == 0, self.m.error
        return out

    def insert(self, keys, lens, vals, vlens, maxK, maxV, upsert=False):
        return self._run(4 if upsert else 0, keys, lens, maxK, vals, vlens, maxV)[2]

    def update(self, keys, lens, vals, vlens, maxK, maxV):
        return self._run(3, keys, lens, maxK, vals, vlens, maxV)[2]

    def erase(self, keys, lens, maxK):
        return self._run(1, keys, lens, maxK)[2]

    def find(self, keys, lens, maxK, maxV, n=None):
        v, l, _ = self._run(2, keys, lens, maxK, maxV=maxV, find=True)
        return v.reshape(len(lens), maxV), l

    def mixed(self, types, keys, lens, vals, vlens, maxK, maxV):
        v, l, r = self._run(2, keys, lens, maxK, vals, vlens, maxV,
                            kinds=np.ascontiguousarray(types, np.uint8), mixed=True)
        return v.reshape(len(types), maxV), l, r

    def items(self):
        return self.m.items()

    def close(self):
        pass
