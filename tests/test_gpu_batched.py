"""Batched form (gmt_multi_batch): many uneven mixed batches applied in order by one
persistent launch.  Parity: every per-request output equals the oracle replaying the
batches one after another; batches that touch the same keys in turn check the order."""
import numpy as np
import pytest
import torch

import scenarios as S
from indexingpu_amd import workloads as W

pytestmark = pytest.mark.gpu


def _concat(batches, mk, mv):
    types = np.concatenate([b["types"] for b in batches])
    keys = np.concatenate([b["keys"] for b in batches])
    lens = np.concatenate([b["lens"] for b in batches])
    vals = np.concatenate([b["values"] for b in batches])
    offs = np.cumsum([0] + [len(b["types"]) for b in batches]).astype(np.int64)
    return types, keys, lens, vals, offs


def _run_multi(g, types, keys, lens, vals, offs, mk, mv):
    n = len(types)
    v = g._d(vals)
    vl = g._d(np.full(n, mv, np.uint32))
    r = g._zeros(n, np.uint8)
    off = torch.from_numpy(offs).cuda()
    g.t.multi_batch(g._d(types, np.uint8), g._d(keys), mk, g._d(lens), v, mv, vl, r, off,
                    len(offs) - 1, n)
    out = (g._h(v)[:n * mv].reshape(n, mv), g._h(vl)[:n], g._h(r, np.uint8)[:n])
    g.t.check_device_error()
    return out


@pytest.mark.parametrize("tile,mode,split", [(16, 0, None), (64, 0, None), (16, 1, 0xFFFFFFFF),
                                              (32, 1, 0xFFFFFFFF), (16, 1, 0), (32, 1, 0)])
def test_uneven_batches_match_sequential_oracle(tile, mode, split):
    """split 0: every batch runs its finds as a read-only pass before its other requests."""
    wl = W.MixedWorkload(50_000, key_len=4, seed=31)
    mk, mv = 4, 2
    rng = np.random.default_rng(5)
    sizes = np.exp(rng.uniform(np.log(16), np.log(20_000), size=40)).astype(int)
    batches = [wl.batch(int(s)) for s in sizes]
    g, o = S.GpuBackend(tile, pool_bytes=128 << 20), S.OracleBackend()
    g.t.set_multi_mode(mode)
    if split is not None:
        g.t.set_mixed_split(split)
    g.insert(wl.keys, wl.lens, wl.values, None, mk, mv)
    o.insert(wl.keys, wl.lens, wl.values, None, mk, mv)
    types, keys, lens, vals, offs = _concat(batches, mk, mv)
    got = _run_multi(g, types, keys, lens, vals, offs, mk, mv)
    exp_v, exp_l, exp_r = [], [], []
    for b in batches:
        v, l, r = o.mixed(b["types"], b["keys"], b["lens"], b["values"],
                          np.full(len(b["types"]), mv, np.uint32), mk, mv)
        exp_v.append(v)
        exp_l.append(l)
        exp_r.append(r)
    assert np.array_equal(got[0], np.concatenate(exp_v))
    assert np.array_equal(got[1], np.concatenate(exp_l))
    assert np.array_equal(got[2], np.concatenate(exp_r))
    assert g.items() == o.items()


@pytest.mark.parametrize("mode,split", [(0, None), (1, 0xFFFFFFFF), (1, 0)])
def test_batches_see_each_other_in_order(mode, split):
    """insert K / find K / update K / find K / erase K / find K as six batches."""
    rng = np.random.default_rng(9)
    keys, lens = W.unique_random_keys(rng, 3000, 1, 5)
    mk, mv = 5, 2
    n = len(lens)
    v1 = rng.integers(0, 1 << 32, size=(n, mv), dtype=np.uint64).astype(np.uint32)
    v2 = v1 ^ np.uint32(0xFFFF)
    z = np.zeros((n, mv), np.uint32)
    seq = [(0, v1), (2, z), (3, v2), (2, z), (1, z), (2, z)]
    batches = [dict(types=np.full(n, t, np.uint8), keys=keys, lens=lens, values=v)
               for t, v in seq]
    types, k, l, vals, offs = _concat(batches, mk, mv)
    g = S.GpuBackend(16, pool_bytes=64 << 20)
    g.t.set_multi_mode(mode)
    if split is not None:
        g.t.set_mixed_split(split)
    gv, gl, gr = _run_multi(g, types, k, l, vals, offs, mk, mv)
    gv = gv.reshape(6, n, mv)
    gl = gl.reshape(6, n)
    gr = gr.reshape(6, n)
    assert gr[0].all() and gr[2].all() and gr[4].all()
    assert (gl[1] == mv).all() and np.array_equal(gv[1], v1)
    assert (gl[3] == mv).all() and np.array_equal(gv[3], v2)
    assert (gl[5] == 0).all()
    assert g.items() == []


def test_tile_width_switches_between_calls():
    """All tile widths share one tree: insert at 16 lanes, update at 32, find at 64,
    erase at 32, and every step matches the oracle."""
    rng = np.random.default_rng(12)
    keys, lens = W.unique_random_keys(rng, 20_000, 1, 6)
    mk, mv = 6, 3
    n = len(lens)
    v1 = rng.integers(0, 1 << 32, size=(n, mv), dtype=np.uint64).astype(np.uint32)
    vl = rng.integers(1, mv + 1, size=n).astype(np.uint32)
    g, o = S.GpuBackend(16, pool_bytes=64 << 20), S.OracleBackend()
    assert g.t.tile_size == 16
    assert np.array_equal(g.insert(keys, lens, v1, vl, mk, mv), o.insert(keys, lens, v1, vl, mk, mv))
    g.t.tile_size = 32
    half = slice(0, n // 2)
    v2 = v1[half] ^ np.uint32(0x5A5A5A5A)
    assert np.array_equal(g.update(keys[half], lens[half], v2, vl[half], mk, mv),
                          o.update(keys[half], lens[half], v2, vl[half], mk, mv))
    g.t.tile_size = 64
    assert g.t.tile_size == 64
    for a, b in zip(g.find(keys, lens, mk, mv), o.find(keys, lens, mk, mv)):
        assert np.array_equal(a, b)
    g.t.tile_size = 32
    q = slice(n // 4, 3 * n // 4)
    assert np.array_equal(g.erase(keys[q], lens[q], mk), o.erase(keys[q], lens[q], mk))
    assert g.items() == o.items()
    with pytest.raises(Exception):
        g.t.tile_size = 48
    assert g.t.tile_size == 32


def test_batched_form_captured_into_a_graph():
    """The batched form (one launch per batch, finds as a read-only pass) captured into a
    graph: ranges are read on the device, grids are full and the split follows the
    average batch size.  One replay applies the batches; outputs match the oracle."""
    wl = W.MixedWorkload(50_000, key_len=4, seed=41)
    mk, mv = 4, 2
    rng = np.random.default_rng(6)
    sizes = np.exp(rng.uniform(np.log(16), np.log(20_000), size=30)).astype(int)
    batches = [wl.batch(int(s)) for s in sizes]
    g, o = S.GpuBackend(16, pool_bytes=128 << 20), S.OracleBackend()
    g.t.set_multi_mode(1)
    g.t.set_mixed_split(0)
    g.insert(wl.keys, wl.lens, wl.values, None, mk, mv)
    o.insert(wl.keys, wl.lens, wl.values, None, mk, mv)
    types, keys, lens, vals, offs = _concat(batches, mk, mv)
    n = len(types)
    dt, dk, dl = g._d(types, np.uint8), g._d(keys), g._d(lens)
    v = g._d(vals)
    vl = g._d(np.full(n, mv, np.uint32))
    r = g._zeros(n, np.uint8)
    off = torch.from_numpy(offs).cuda()
    g.t.reserve_find_scratch(n)
    torch.cuda.synchronize()
    graph = torch.cuda.CUDAGraph()
    with torch.cuda.graph(graph):
        g.t.multi_batch(dt, dk, mk, dl, v, mv, vl, r, off, len(offs) - 1, n)
    graph.replay()
    torch.cuda.synchronize()
    g.t.check_device_error()
    got = (g._h(v)[:n * mv].reshape(n, mv), g._h(vl)[:n], g._h(r, np.uint8)[:n])
    exp = [o.mixed(b["types"], b["keys"], b["lens"], b["values"],
                   np.full(len(b["types"]), mv, np.uint32), mk, mv) for b in batches]
    for i in range(3):
        assert np.array_equal(got[i], np.concatenate([e[i] for e in exp]))
    assert g.items() == o.items()
    del graph


def _prefix_batches(n_keys, n_batches, seed):
    from indexingpu_amd import device_workloads as D
    wl = D.PrefixBatchWorkload(n_keys, seed=seed)
    rng = np.random.default_rng(seed)
    sizes = D.log_uniform_sizes(rng, n_batches, 64, 8192)
    batches = [wl.batch(int(s)) for s in sizes]
    return wl, batches, D.concat_batches(batches)


def _h32(t):
    return t.cpu().numpy().view(np.uint32) if t.dtype == torch.int32 else t.cpu().numpy()


@pytest.mark.parametrize("mode,split,captured", [(0, None, False), (1, 0xFFFFFFFF, False),
                                                 (1, 0, False), (1, 0, True),
                                                 (1, 0xFFFFFFFF, True)])
def test_prefixed_16_slice_batches(mode, split, captured):
    """The batched form's own workload shape: 16-slice keys whose first 12 slices are one
    of 256 shared prefixes (deep trie layers, suffix records), ~50K preloaded, uneven
    batches of 90% Zipf finds / 5% fresh inserts / 5% erases, applied by one multi_batch
    call in both modes, eagerly and captured into a graph.  Every output equals the oracle
    replaying the batches one after another and the generator's own expectation."""
    wl, batches, cat = _prefix_batches(50_000, 60, seed=13 + mode)
    mk, mv = wl.key_len, wl.mv
    pk, pl, pv = wl.preload()
    g, o = S.GpuBackend(16, pool_bytes=256 << 20), S.OracleBackend()
    g.t.set_multi_mode(mode)
    if split is not None:
        g.t.set_mixed_split(split)
    keys_np, lens_np, vals_np = _h32(pk), _h32(pl), _h32(pv)
    n0 = len(lens_np)
    r = g._zeros(n0, np.uint8)
    g.t.insert(pk.view(-1), mk, pl, pv.reshape(-1), mv, None, n0, results=r)
    o.insert(keys_np.reshape(n0, mk), lens_np, vals_np.reshape(n0, mv), None, mk, mv)
    total = len(cat["types"])
    v = cat["values"].reshape(-1).clone()
    vl = torch.full((total,), mv, dtype=torch.int32, device="cuda")
    res = torch.zeros(total, dtype=torch.uint8, device="cuda")
    nb = len(batches)

    def call():
        g.t.multi_batch(cat["types"], cat["keys"].reshape(-1), mk, cat["lens"], v, mv, vl, res,
                        cat["offsets"], nb, total)
    if captured:
        g.t.reserve_find_scratch(total)
        torch.cuda.synchronize()
        graph = torch.cuda.CUDAGraph()
        with torch.cuda.graph(graph):
            call()
        graph.replay()
    else:
        call()
    torch.cuda.synchronize()
    g.t.check_device_error()
    gv, gl, gr = _h32(v).reshape(total, mv), _h32(vl), res.cpu().numpy()
    # the generator's expectation
    assert np.array_equal(gl, _h32(cat["expect_len"]))
    assert np.array_equal(gv, _h32(cat["expect_values"]).reshape(total, mv))
    assert np.array_equal(gr, cat["expect_results"].cpu().numpy())
    # the oracle replaying the batches in order
    exp = []
    for b in batches:
        n = len(b["types"])
        exp.append(o.mixed(b["types"].cpu().numpy(), _h32(b["keys"]).reshape(n, mk),
                           _h32(b["lens"]), _h32(b["values"]).reshape(n, mv),
                           np.full(n, mv, np.uint32), mk, mv))
    for i, got in enumerate((gv, gl, gr)):
        assert np.array_equal(got, np.concatenate([e[i] for e in exp]))
    assert g.items() == o.items()
    if captured:
        del graph
