"""Autotuner for the index's run-time knobs, per operation class and key-length class.

The knobs (all switchable between calls on one index, same tree state):
  * tile width: lanes per operation, 16 / 32 / 64 (the reference's tile_size template
    argument, include/gpu_masstree.hpp:42-46, fixed per instance there);
  * find tile: lanes per lookup of read-only find batches (4/8/16: each lane holds
    16 / width slots of a node; 0 = the tile width);
  * find depth: lookups each tile keeps in flight in read-only find batches (1/2/4/8);
  * find passes: read-only finds as one kernel or as a descent and a record pass;
  * record tile: lanes per lookup of the record pass (1/2/4/8/16; 0 = the find tile);
  * grid share: percent of the resident-block capacity the batched form launches;
  * batched-form mode: one persistent launch with grid barriers, or one launch per batch.

  * node fanout: the entries at which later writes split a border node (2..14) and the
    children at which they split an interior node (3..15); (0, 0) = the layout's maximum.
    Nodes stay 128-byte slabs of 16 slots; a smaller fanout leaves room in each node
    (fewer entries per lock, more nodes and levels).

The slab size class is not a knob: every node and record slab is 128 bytes, the line one
16-lane tile reads in one coalesced access, and the node layout (16 eight-byte slots) is
built on it.  The tuner searches find tile x depth (read-only finds), and tile x grid share
x fanout (mutating work).

Choices are cached per (operation class, key-length class): the key-length class is the
maximum key length rounded up to a power of two (1, 2, 4, 8, 16, 32, 64, ...), since the
length decides how many trie layers a lookup walks and how much key data each tile
streams.

Two tuning modes:
  * tune_find: read-only and repeatable, so every candidate runs the same batch;
  * tune_online: for mutating work every candidate runs fresh units of the same
    distribution (``run_unit(config)`` draws and runs the next unit), so no operation is
    applied twice.
"""
import itertools
import time

TILES = (16, 32, 64)
FIND_TILES = (1, 2, 4, 8, 16)
FIND_DEPTHS = (1, 2, 4, 8)
FIND_PASSES = (1, 2)
RECORD_TILES = (0, 2, 4, 8)
GRID_PERCENTS = (100, 50, 25)
FANOUTS = ((0, 0), (10, 12), (7, 8))


def length_class(max_key_length):
    """Key-length class: max key length rounded up to a power of two."""
    m = max(1, int(max_key_length))
    return 1 << (m - 1).bit_length()


class _EventTimer:
    """Elapsed milliseconds of fn() on torch's current stream (HIP events)."""

    def __init__(self):
        import torch
        self.torch = torch

    def __call__(self, fn):
        torch = self.torch
        stream = torch.cuda.current_stream()
        s = torch.cuda.Event(enable_timing=True)
        e = torch.cuda.Event(enable_timing=True)
        s.record(stream)
        fn()
        e.record(stream)
        e.synchronize()
        return s.elapsed_time(e)


class WallTimer:
    """Elapsed milliseconds by the host clock (for callers without a GPU stream)."""

    def __call__(self, fn):
        t = time.perf_counter()
        fn()
        return (time.perf_counter() - t) * 1e3


class Autotuner:
    """Picks knob settings for an index by measurement and applies them on request.

    ``index`` needs ``tile_size`` (settable), ``set_find_depth`` and
    ``set_multi_grid_percent`` -- the GpuMasstree wrapper has all three.
    ``timer(fn) -> ms`` defaults to HIP events on torch's current stream.
    """

    def __init__(self, index, tiles=TILES, find_tiles=FIND_TILES, find_depths=FIND_DEPTHS,
                 grid_percents=GRID_PERCENTS, trials=3, timer=None, find_passes=FIND_PASSES,
                 record_tiles=RECORD_TILES):
        self.index = index
        self.find_passes = tuple(find_passes)
        self.record_tiles = tuple(record_tiles)
        self.tiles = tuple(tiles)
        self.find_tiles = tuple(find_tiles)
        self.find_depths = tuple(find_depths)
        self.grid_percents = tuple(grid_percents)
        self.trials = max(1, int(trials))
        self.timer = timer if timer is not None else _EventTimer()
        self.table = {}     # (op, length class) -> config dict
        self.history = {}   # (op, length class) -> [(config, ms per unit), ...]

    # ------------------------------------------------------------------ knobs
    def apply_config(self, cfg):
        if "tile" in cfg:
            self.index.tile_size = cfg["tile"]
        if "find_tile" in cfg:
            self.index.set_find_tile(cfg["find_tile"])
        if "find_passes" in cfg:
            self.index.set_find_passes(cfg["find_passes"])
        if "record_tile" in cfg:
            self.index.set_record_tile(cfg["record_tile"])
        if "find_depth" in cfg:
            self.index.set_find_depth(cfg["find_depth"])
        if "grid_percent" in cfg:
            self.index.set_multi_grid_percent(cfg["grid_percent"])
        if "multi_mode" in cfg:
            self.index.set_multi_mode(cfg["multi_mode"])
        if "fanout" in cfg:
            self.index.set_fanout(*cfg["fanout"])

    def apply(self, op, max_key_length):
        """Applies the cached choice for (op, key-length class); False if none."""
        cfg = self.table.get((op, length_class(max_key_length)))
        if cfg is None:
            return False
        self.apply_config(cfg)
        return True

    def best(self, op, max_key_length):
        return self.table.get((op, length_class(max_key_length)))

    # ----------------------------------------------------------------- tuning
    def _pick(self, key, scored):
        self.history[key] = scored
        cfg = min(scored, key=lambda cs: cs[1])[0]
        self.table[key] = cfg
        self.apply_config(cfg)
        return cfg

    def tune_find(self, run_find, max_key_length, warmup=1):
        """run_find() launches one read-only find batch (the same batch every call).
        Every (passes, find tile, depth) candidate is warmed up, then timed `trials` times; the
        minimum counts.  When two passes win, a second stage tries each record tile with
        the winning find tile and depth (the first stage runs with the record tile of the
        first entry of record_tiles)."""
        def measure(cfg):
            self.apply_config(cfg)
            for _ in range(warmup):
                self.timer(run_find)
            return min(self.timer(run_find) for _ in range(self.trials))

        r0 = self.record_tiles[0] if self.record_tiles else 0
        scored = []
        for passes, tile, depth in itertools.product(self.find_passes, self.find_tiles,
                                                     self.find_depths):
            cfg = {"find_passes": passes, "find_tile": tile, "find_depth": depth}
            if passes == 2:
                cfg["record_tile"] = r0
            scored.append((cfg, measure(cfg)))
        best = min(scored, key=lambda cs: cs[1])[0]
        if best["find_passes"] == 2:
            for r in self.record_tiles[1:]:
                cfg = dict(best, record_tile=r)
                scored.append((cfg, measure(cfg)))
        return self._pick(("find", length_class(max_key_length)), scored)

    def tune_online(self, op, run_unit, max_key_length, space=None):
        """run_unit(cfg) runs the next fresh unit of mutating work (a batch, or a group of
        batches) under the knobs already applied and returns its size in operations.
        Each candidate gets `trials` units; the score is milliseconds per operation."""
        if space is None:
            space = [{"tile": t, "grid_percent": g}
                     for t, g in itertools.product(self.tiles, self.grid_percents)]
        scored = []
        for cfg in space:
            self.apply_config(cfg)
            total_ms, total_ops = 0.0, 0
            for _ in range(self.trials):
                box = {}
                total_ms += self.timer(lambda: box.setdefault("n", run_unit(cfg)))
                total_ops += int(box["n"])
            scored.append((dict(cfg), total_ms / max(total_ops, 1)))
        return self._pick((op, length_class(max_key_length)), scored)

    def write_space(self, fanouts=FANOUTS):
        """Candidates for tune_online over tile x grid share x node fanout."""
        return [{"tile": t, "grid_percent": g, "fanout": tuple(f)}
                for t, g, f in itertools.product(self.tiles, self.grid_percents, fanouts)]

    def report(self):
        """Rows (op, length class, config, ms per unit or per op) of every measurement."""
        rows = []
        for (op, lc), scored in sorted(self.history.items()):
            for cfg, ms in scored:
                rows.append((op, lc, cfg, ms))
        return rows
