"""Benchmark of the MI355X Masstree batch path.

Workload (plain GPU batch path, one GPU): 16,777,216 unique keys of 2..8 u32 slices
(uniform length, uniform random slices), 2-slice values (FNV-1a of the key), inserted
with one batch; a step is one find batch of 16,777,216 keys (50% inserted keys, 50%
fresh keys) with inputs resident in HBM.  Also measured once: the insert batch, one
erase batch of 8,388,608 inserted keys, and (with --mixed) mixed batches over 64M
preloaded 4-slice keys.  Prints one JSON line.

Every find result of the timed batches is checked on the host against the values the
generator defines (hit: 2 slices = FNV-derived value; miss: length 0), a size-
independent parity check at full size.
"""
import argparse
import json
import os
import platform
import subprocess
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "Mops/s (find/insert/mixed) at 64M varlen keys; p50/p99 served-request latency"
HBM_PEAK_GBS = 8000.0
LINE = 64  # bytes per random HBM line in the roofline's unit


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def up(torch, a, dtype=np.uint32):
    a = np.ascontiguousarray(a, dtype=dtype).reshape(-1)
    if dtype == np.uint32:
        return torch.from_numpy(a.view(np.int32)).cuda()
    return torch.from_numpy(a).cuda()


def find_bytes(find_lens, hits, maxv):
    """Algorithmic bytes of a find batch: streamed I/O (key slices, key length, value
    length, value slices of hits) plus 64-B random lines: one leaf line per find and one
    record line per find whose first slice is present (every hit here)."""
    streamed = 4.0 * find_lens.sum() + 8.0 * len(find_lens) + 4.0 * maxv * hits
    lines = LINE * (len(find_lens) + hits)
    return float(streamed + lines)


def cpu_baseline(n_keys, seed):
    """Oracle (std::map replay, one thread) on a bounded sample of the same workload."""
    from indexingpu_amd import workloads as W
    from oracle.oracle import Oracle
    d = W.plain_batch_workload(n_keys, seed=seed)
    mk, mv = d["max_key_length"], d["max_value_length"]
    o = Oracle()
    t0 = time.perf_counter()
    o.insert(d["keys"].reshape(-1).copy(), mk, d["lens"], d["values"].reshape(-1).copy(), mv,
             None, n_keys, want_results=False)
    t1 = time.perf_counter()
    o.find(d["find_keys"].reshape(-1).copy(), mk, d["find_lens"], mv, n_keys)
    t2 = time.perf_counter()
    return dict(find_mops=n_keys / (t2 - t1) / 1e6, insert_mops=n_keys / (t1 - t0) / 1e6,
                seconds=t2 - t0)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--keys", type=int, default=1 << 24)
    ap.add_argument("--tile", type=int, default=16)
    ap.add_argument("--pool-gb", type=float, default=24.0)
    ap.add_argument("--cpu-sample", type=int, default=2_000_000)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--mixed", action="store_true", help="also run 64M-key mixed batches")
    ap.add_argument("--seed", type=int, default=2)
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only: every box in this pool has one GPU, so "
              f"--gpus {args.gpus} is not supported (the index shards by key range across "
              f"GPUs; see DESIGN.md).", file=sys.stderr)
        sys.exit(2)

    import torch
    from indexingpu_amd import GpuMasstree, workloads as W

    log(f"[bench] generating {args.keys} keys ...")
    t0 = time.time()
    d = W.plain_batch_workload(args.keys, seed=args.seed)
    mk, mv = d["max_key_length"], d["max_value_length"]
    n = args.keys
    n_erase = len(d["erase_lens"])
    log(f"[bench] generated in {time.time() - t0:.1f}s; uploading")
    keys, lens, vals = up(torch, d["keys"]), up(torch, d["lens"]), up(torch, d["values"])
    fkeys, flens = up(torch, d["find_keys"]), up(torch, d["find_lens"])
    ekeys, elens = up(torch, d["erase_keys"]), up(torch, d["erase_lens"])
    out_v = torch.zeros(n * mv, dtype=torch.int32, device="cuda")
    out_l = torch.zeros(n, dtype=torch.int32, device="cuda")
    tree = GpuMasstree(tile_size=args.tile, pool_bytes=int(args.pool_gb * (1 << 30)))
    stream = torch.cuda.current_stream()

    def timed(fn):
        s, e = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        torch.cuda.synchronize()
        s.record(stream)
        fn()
        e.record(stream)
        torch.cuda.synchronize()
        return s.elapsed_time(e)

    log("[bench] insert batch ...")
    t_ins = timed(lambda: tree.insert(keys, mk, lens, vals, mv, None, n))
    tree.check_device_error()
    log(f"[bench] insert {n / t_ins / 1e3:.1f} Mops/s; find warmup")
    for _ in range(args.warmup):
        tree.find(fkeys, mk, flens, out_v, mv, out_l, n)
    torch.cuda.synchronize()
    # timed region: K find batches, one kernel each, HIP events on the launch stream
    ev = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True))
          for _ in range(args.steps)]
    torch.cuda.synchronize()
    wall0 = time.perf_counter()
    for s, e in ev:
        s.record(stream)
        tree.find(fkeys, mk, flens, out_v, mv, out_l, n)
        e.record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - wall0
    per = [s.elapsed_time(e) for s, e in ev]
    tree.check_device_error()
    log(f"[bench] find batches: {['%.2f' % p for p in per]} ms")

    # full-size parity check of the last find batch (values are a function of the key)
    got_l = out_l.cpu().numpy().view(np.uint32)
    got_v = out_v.cpu().numpy().view(np.uint32).reshape(n, mv)
    hit_mask = got_l > 0
    exp_v = W.make_values(d["find_keys"], d["find_lens"], mv)
    n_hits = int(hit_mask.sum())
    verified = (n_hits == n // 2 and bool((got_l[hit_mask] == mv).all())
                and bool((got_v[hit_mask] == exp_v[hit_mask]).all())
                and bool((got_v[~hit_mask] == 0).all()))
    del got_v, exp_v

    log("[bench] erase batch ...")
    t_er = timed(lambda: tree.erase(ekeys, mk, elens, n_erase))
    tree.check_device_error()

    mean_ms = float(np.mean(per))
    value = n * args.steps / (sum(per) / 1e3) / 1e6
    abytes = find_bytes(d["find_lens"], n // 2, mv)
    achieved = abytes / (mean_ms / 1e3) / 1e9

    result = {
        "metric": METRIC,
        "value": round(value, 2),
        "unit": "Mops/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(mean_ms, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u32",
        "data": "synthetic",
        "config": {
            "workload": "plain batch path: 16M unique var-len keys (2..8 u32 slices), "
                        "2-slice values; step = one find batch of all keys' count "
                        "(50% hits) on the loaded index",
            "keys": n, "find_batch": n, "erase_batch": n_erase, "max_key_length": mk,
            "max_value_length": mv, "tile_size": args.tile,
        },
        "insert_mops": round(n / t_ins / 1e3, 2),
        "erase_mops": round(n_erase / t_er / 1e3, 2),
        "find_verified_full_size": verified,
        "roofline": {
            "bound": "hbm",
            "achieved": round(achieved, 2),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": None,
            "kernel": f"batch_kernel<{args.tile}, false> (find)",
            "algorithmic_bytes_per_launch": abytes,
        },
    }
    tree.close()
    del keys, lens, vals, fkeys, flens, ekeys, elens, out_v, out_l
    torch.cuda.empty_cache()

    if not args.no_cpu_baseline:
        log(f"[bench] CPU baseline: oracle on {args.cpu_sample} keys ...")
        cb = cpu_baseline(args.cpu_sample, seed=args.seed)
        result["cpu_baseline"] = {
            "value": round(cb["find_mops"], 3), "unit": "Mops/s", "cores": 1, "kind": "port",
            "sample": f"std::map replay (oracle), 1 thread: {args.cpu_sample} inserts then "
                      f"{args.cpu_sample} finds (50% hits) of the same workload shape; "
                      f"insert {cb['insert_mops']:.3f} Mops/s; host {platform.processor() or platform.machine()}, "
                      f"{os.cpu_count()} logical CPUs",
        }
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
