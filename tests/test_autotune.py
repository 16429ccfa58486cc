"""Host logic of the autotuner: candidate space, scoring, caching per key-length class
and knob application, with a fake index and a fake clock (no GPU)."""
from indexingpu_amd.autotune import Autotuner, length_class


class FakeIndex:
    def __init__(self):
        self.tile_size = 16
        self.depth = 4
        self.grid = 100
        self.ftile = 8
        self.passes = 2
        self.rtile = 0
        self.log = []

    def set_find_tile(self, t):
        self.ftile = t

    def set_find_passes(self, p):
        self.passes = p

    def set_record_tile(self, t):
        self.rtile = t

    def set_find_depth(self, k):
        self.depth = k

    def set_multi_grid_percent(self, p):
        self.grid = p

    def set_fanout(self, b, i):
        self.fanout = (b, i)


class FakeClock:
    """Milliseconds = cost(knobs of the index at the time fn runs)."""

    def __init__(self, index, cost):
        self.index, self.cost = index, cost
        self.calls = 0

    def __call__(self, fn):
        fn()
        self.calls += 1
        i = self.index
        extra = 0.5 if i.passes == 1 else (0.0 if i.rtile == 4 else 0.25)
        return self.cost(i.tile_size, i.depth, i.grid, i.ftile) + extra


def test_length_class():
    assert [length_class(m) for m in (0, 1, 2, 3, 4, 5, 8, 9, 16, 17)] == \
        [1, 1, 2, 4, 4, 8, 8, 16, 16, 32]


def test_tune_find_picks_minimum_and_applies_it():
    idx = FakeIndex()
    cost = lambda t, d, g, f: abs(f - 8) + abs(d - 2) * 3 + 1.0
    clock = FakeClock(idx, cost)
    tu = Autotuner(idx, trials=2, timer=clock)
    runs = []
    cfg = tu.tune_find(lambda: runs.append((idx.ftile, idx.depth)), max_key_length=8)
    assert cfg == {"find_passes": 2, "find_tile": 8, "find_depth": 2, "record_tile": 4}
    assert (idx.passes, idx.ftile, idx.depth, idx.rtile) == (2, 8, 2, 4)
    # 2 pass counts x 5 tiles x 4 depths, then the 3 other record tiles with the best
    # two-pass setting; each 1 warmup + 2 timed runs
    assert len(runs) == 43 * 3 and clock.calls == 43 * 3
    assert tu.best("find", 7) == cfg          # same length class (8)
    assert tu.best("find", 9) is None          # class 16 not tuned
    assert len(tu.report()) == 43


def test_tune_find_skips_record_stage_when_one_pass_wins():
    idx = FakeIndex()
    clock = FakeClock(idx, lambda t, d, g, f: abs(f - 4) + abs(d - 1) + 1.0)
    tu = Autotuner(idx, trials=1, timer=clock, find_passes=(1,))
    cfg = tu.tune_find(lambda: None, max_key_length=4)
    assert cfg == {"find_passes": 1, "find_tile": 4, "find_depth": 1}
    assert len(tu.report()) == 20


def test_tune_online_scores_per_operation_and_never_reuses_units():
    idx = FakeIndex()
    # time per unit grows with unit size; per-op cost best at tile 64 / grid 50
    sizes = iter(range(1000, 100000, 37))
    served = []

    def run_unit(cfg):
        n = next(sizes)
        served.append(n)
        return n

    def cost(t, d, g, f):
        return served[-1] * (1.0 + abs(t - 64) / 64 + abs(g - 50) / 100)

    tu = Autotuner(idx, trials=2, timer=FakeClock(idx, cost))
    cfg = tu.tune_online("multi", run_unit, max_key_length=16)
    assert cfg == {"tile": 64, "grid_percent": 50}
    assert len(served) == len(set(served)) == 9 * 2
    assert (idx.tile_size, idx.grid) == (64, 50)


def test_apply_restores_cached_choice():
    idx = FakeIndex()
    tu = Autotuner(idx, timer=FakeClock(idx, lambda t, d, g, f: f + d))
    tu.tune_find(lambda: None, max_key_length=4)
    assert (idx.passes, idx.ftile, idx.depth) == (2, 1, 1)
    idx.passes, idx.ftile, idx.depth = 1, 16, 8
    assert tu.apply("find", 3)
    assert (idx.passes, idx.ftile, idx.depth) == (2, 1, 1)
    assert not tu.apply("mixed", 3)


def test_tune_online_over_fanout():
    """The write space spans tile x grid share x node fanout; the fanout of the winner is
    applied to the index."""
    idx = FakeIndex()
    idx.fanout = (0, 0)

    def run_unit(cfg):
        assert idx.fanout == cfg["fanout"]
        return 1000

    def cost(t, d, g, f):
        b = idx.fanout[0] or 14
        return 1000 * (1.0 + abs(t - 32) / 64 + abs(g - 100) / 100 + abs(b - 10) / 10)

    tu = Autotuner(idx, trials=1, timer=FakeClock(idx, cost))
    space = tu.write_space()
    assert len(space) == 3 * 3 * 3
    cfg = tu.tune_online("mixed", run_unit, max_key_length=4, space=space)
    assert cfg == {"tile": 32, "grid_percent": 100, "fanout": (10, 12)}
    assert idx.fanout == (10, 12) and idx.tile_size == 32
