"""Host logic of the pair-list autotuner (prism_md.autotune): bucket edges,
bucket lookup, and that the fastest measured candidate wins -- driven through
a stand-in engine whose step time depends on the candidate."""
from __future__ import annotations

import types

import pytest

from prism_md import autotune


def test_log_buckets_cover_the_range():
    b = autotune.log_buckets(500, 20000, 4)
    assert b[0][0] == 500 and b[-1][1] == 20000 and len(b) == 4
    for (lo, hi), (lo2, _) in zip(b, b[1:]):
        assert hi == lo2 and lo < hi
    ratios = [hi / lo for lo, hi in b]
    assert max(ratios) / min(ratios) < 1.01  # log-uniform


def test_choice_for_bucket_and_nearest():
    table = [{"lo": 100, "hi": 1000, "sort_cell_A": 1.5, "resort_every": 4},
             {"lo": 1000, "hi": 10000, "sort_cell_A": 3.0, "resort_every": 16}]
    assert autotune.choice_for(table, 500) == {"sort_cell_A": 1.5, "resort_every": 4}  # a table without units
    assert autotune.choice_for(table, 5000)["sort_cell_A"] == 3.0
    assert autotune.choice_for(table, 50) == {"sort_cell_A": 1.5, "resort_every": 4}
    assert autotune.choice_for(table, 10 ** 6)["resort_every"] == 16


def test_tune_picks_the_fastest_candidate(monkeypatch):
    clock = {"t": 0.0}
    monkeypatch.setattr(autotune.time, "perf_counter", lambda: clock["t"])

    class FakeEnsemble:
        def __init__(self, tops, params):
            self.p = params[0]

        def __enter__(self):
            return self

        def __exit__(self, *a):
            return False

        def step(self, n, graph_steps):
            # cost minimal at (2.0 A, 8, 24 tiles): |cell - 2| + |log2(resort) - 3| + |chunk - 24| / 100
            import math
            chunk = getattr(self.p, "chunk_tiles", 0) or 8  # the engine's default below 100,000 atoms
            cost = 1.0 + abs(self.p.sort_cell_A - 2.0) + abs(math.log2(self.p.resort_every) - 3.0) + \
                abs(chunk - 24) / 100.0
            clock["t"] += n * cost * 1e-6

    fake = types.SimpleNamespace(
        Topology=types.SimpleNamespace(generate=lambda n, seed: ("top", n, seed)),
        default_params=lambda **kw: types.SimpleNamespace(**kw),
        Ensemble=FakeEnsemble)
    table = autotune.tune(fake, [(500, 2000)], steps=100, copies=2)
    assert len(table) == 1 and table[0]["atoms"] == 1000
    assert table[0]["sort_cell_A"] == 2.0 and table[0]["resort_every"] == 8 and table[0]["chunk_tiles"] == 24
    assert len(table[0]["trials"]) == len(autotune.SORT_CELLS) * len(autotune.RESORT_EVERY) + \
        len(autotune.CHUNK_TILES)
    assert table[0]["us_per_step"] == pytest.approx(1.0)
    assert autotune.choice_for(table, 1500) == {"sort_cell_A": 2.0, "resort_every": 8, "chunk_tiles": 24}
