"""Served form: the same trace through the request loop (continuous batching, waves
overlapping on several streams, retire lists rotated every few waves) and through the
plain path (each request alone: its own mixed-batch and scan calls, in order).  Every
request's result must be the same alone and in the mix -- point-op values, value
lengths, success flags, scan counts and scan rows -- and equal to the generator's
expectation; scans also match the oracle; both indexes end with the same entries."""
import numpy as np
import pytest
import torch

import scenarios as S
from indexingpu_amd import device_workloads as D
from indexingpu_amd.serve import MasstreeServer, latency_summary

pytestmark = pytest.mark.gpu

# above 2^18 keys: the index builds its layer-0 shortcut tables, which the ordered waves'
# read-only passes start from (their finds at the border-level one)
N_PRELOAD = 300_000


def _loaded(wl):
    g = S.GpuBackend(16, pool_bytes=256 << 20)
    k, l, v = wl.preload()
    g.t.insert(k, wl.mk, l, v, wl.mv, None, len(l))
    g.t.check_device_error()
    return g


def _rows(arena, first_row, count, mc):
    mk, mv = arena["max_key_length"], arena["max_value_length"]
    r0, r1 = first_row, first_row + count * mc
    return (arena["ring_keys"][r0 * mk:r1 * mk].cpu(), arena["ring_key_lengths"][r0:r1].cpu(),
            arena["ring_values"][r0 * mv:r1 * mv].cpu(), arena["ring_value_lengths"][r0:r1].cpu())


def _scan_rows_equal(a_mix, out_mix, a_alone, req, mc):
    """Compares the used rows of every request's scans."""
    counts = a_mix["scan_counts"].cpu().numpy()
    for r in np.nonzero(req["num_scans"])[0]:
        s0, ns = int(req["scan_offset"][r]), int(req["num_scans"][r])
        m = _rows(a_mix, int(out_mix["ring_row"][r]), ns, mc)
        p = _rows(a_alone, s0 * mc, ns, mc)
        mk, mv = a_mix["max_key_length"], a_mix["max_value_length"]
        for q in range(ns):
            c = int(counts[s0 + q])
            b = q * mc
            assert torch.equal(m[1][b:b + c], p[1][b:b + c])
            assert torch.equal(m[0][b * mk:(b + c) * mk], p[0][b * mk:(b + c) * mk])
            assert torch.equal(m[3][b:b + c], p[3][b:b + c])
            assert torch.equal(m[2][b * mv:(b + c) * mv], p[2][b * mv:(b + c) * mv])


@pytest.mark.parametrize("ordered", [1, 0], ids=["ordered_waves", "side_by_side"])
@pytest.mark.parametrize("streams,wave_ops,epoch", [(4, 4096, 3), (2, 1 << 16, 1)])
def test_same_result_alone_and_in_the_mix(streams, wave_ops, epoch, ordered):
    wl = D.ServeWorkload(N_PRELOAD, seed=11)
    arena, req, expect = wl.trace(1500, rate=3e6)  # arrivals far faster than service
    mc = arena["scan_max_count"]
    n_scans = int(req["scan_offset"][-1] + req["num_scans"][-1])
    # rings large enough that no row is reused, so every request's rows can be compared
    mix = D.ServeWorkload.with_ring(D.ServeWorkload.fresh_outputs(arena), n_scans * mc)
    alone = D.ServeWorkload.with_ring(D.ServeWorkload.fresh_outputs(arena), n_scans * mc)

    ga, gb = _loaded(wl), _loaded(wl)
    sv = MasstreeServer(ga.t, streams=streams, max_wave_ops=wave_ops, max_wave_scans=256,
                        ring_rows=n_scans * mc, epoch_waves=epoch, ordered=ordered)
    if ordered:
        assert ga.t.find_shortcut_bits() > 0  # the ordered waves start from the tables
    out_mix, st = sv.run(mix, req)
    ga.t.check_device_error()
    assert st["device_error"] == 0
    assert st["waves"] > 1 and st["flips"] >= 1 and st["drains"] >= 1
    plain = MasstreeServer(gb.t, coalesce=False)
    out_alone, st_alone = plain.run(alone, req)
    gb.t.check_device_error()
    assert st_alone["waves"] == len(req)
    assert (out_mix["done_us"] >= req["arrival_us"]).all()

    for k in ("values", "value_lengths", "results"):
        assert torch.equal(mix[k], alone[k]), k
        assert torch.equal(mix[k].view(-1), expect[k].reshape(-1).to(mix[k].dtype)), k
    assert torch.equal(mix["scan_counts"], alone["scan_counts"])
    _scan_rows_equal(mix, out_mix, alone, req, mc)
    assert ga.items() == gb.items()
    ok, *_ = ga.t.validate()
    assert ok
    lat = latency_summary(req, out_mix)
    assert lat["p50_us"] > 0


@pytest.mark.parametrize("ordered", [1, 0], ids=["ordered_waves", "side_by_side"])
def test_scans_match_oracle_and_ring_backpressure(ordered):
    """A ring that holds one full wave's scans: admission waits for rows to free,
    results still match the oracle replaying the read region."""
    wl = D.ServeWorkload(60_000, seed=12)
    arena, req, _ = wl.trace(300, rate=1e7, max_ops=2048)
    mc = arena["scan_max_count"]
    g = _loaded(wl)
    o = S.OracleBackend()
    k, l, v = (t.cpu().numpy().view(np.uint32) for t in wl.preload())
    o.insert(k, l, v, None, wl.mk, wl.mv)
    a = D.ServeWorkload.with_ring(D.ServeWorkload.fresh_outputs(arena), 256 * mc)
    sv = MasstreeServer(g.t, streams=4, max_wave_ops=4096, max_wave_scans=256,
                        ring_rows=256 * mc, epoch_waves=2, ordered=ordered)
    out, st = sv.run(a, req)
    g.t.check_device_error()
    assert st["admission_waits"] > 0
    # replay the scans of the last requests (their rows are still in the ring)
    counts = a["scan_counts"].cpu().numpy()
    lo = arena["scan_lower"].cpu().numpy().view(np.uint32)
    lol = arena["scan_lower_lengths"].cpu().numpy().view(np.uint32)
    up = arena["scan_upper"].cpu().numpy().view(np.uint32)
    upl = arena["scan_upper_lengths"].cpu().numpy().view(np.uint32)
    ec, ev, evl, ek, ekl = o.scan(lo, lol, wl.mk, mc, up, upl, wl.mv)
    assert np.array_equal(counts.view(np.uint32), ec)
    last = [r for r in range(len(req)) if req["num_scans"][r]][-1]
    s0, ns = int(req["scan_offset"][last]), int(req["num_scans"][last])
    rows = _rows(a, int(out["ring_row"][last]), ns, mc)
    mk, mv = wl.mk, wl.mv
    # a reused ring holds old data past each row's key and value lengths: compare rows
    # up to their lengths
    gk, gkl, gv, gvl = (t.numpy().view(np.uint32) for t in rows)
    for q in range(ns):
        for i in range(int(counts[s0 + q])):
            row, erow = q * mc + i, (s0 + q) * mc + i
            kl = int(ekl[erow])
            assert int(gkl[row]) == kl
            assert np.array_equal(gk[row * mk:row * mk + kl], ek[erow * mk:erow * mk + kl])
            vl = min(int(evl[erow]), mv)
            assert int(gvl[row]) == int(evl[erow])
            assert np.array_equal(gv[row * mv:row * mv + vl], ev[erow * mv:erow * mv + vl])
