"""Per-size-bucket autotuner of the pair-list layout for batched ensembles.

Three parameters shape the cluster-pair work without touching the physics
(every pair inside the cutoff is evaluated whatever they are):

  sort_cell_A   edge of the Hilbert sort cells that order atoms into 64-atom
                blocks -- finer cells give more compact blocks (fewer, fuller
                tiles) at a longer sort key;
  resort_every  how many pair-list rebuilds reuse the block assignment (only
                refreshing block boxes) before the atoms are sorted again;
  chunk_tiles   row tiles per nonbonded work unit (the shape of a unit: one
                block quarter against chunk_tiles j-blocks) -- longer units
                amortise the quarter's box and force write-back, shorter ones
                balance the persistent grid better.

For each size bucket the tuner builds a small batch of generator systems of
the bucket's representative size and times candidates with the same
captured-graph stepping the ensemble uses, keeping the fastest: first every
(sort_cell_A, resort_every) pair at the engine's default unit length, then
every unit length with the winning pair.
The reference has no counterpart: its production engine runs one system per
engine with a fixed 10 A cell grid (crates/prism-nhs/src/fused_engine.rs:12285-12330).
"""
from __future__ import annotations

import math
import time

SORT_CELLS = (1.5, 2.0, 3.0)
RESORT_EVERY = (4, 8, 16)
CHUNK_TILES = (8, 12, 16, 24, 32)


def representative_size(lo: int, hi: int) -> int:
    return int(round(math.sqrt(lo * hi)))


def _time(pm, tops, seed, steps, graph_steps, **layout):
    params = [pm.default_params(seed=seed + k, stream_id=k, **layout) for k in range(len(tops))]
    with pm.Ensemble(tops, params) as ens:
        ens.step(2 * graph_steps, graph_steps)  # captures the graph
        t0 = time.perf_counter()
        ens.step(steps, graph_steps)
        dt = time.perf_counter() - t0
    return dict(layout, us_per_step=round(dt / steps * 1e6, 2))


def tune(pm, buckets, steps: int = 400, graph_steps: int = 10, copies: int = 4, seed: int = 0x41544E45,
         sort_cells=SORT_CELLS, resort_every=RESORT_EVERY, chunk_tiles=CHUNK_TILES, progress=None):
    """buckets: [(lo, hi)] atom-count ranges.  Returns one entry per bucket:
    {"lo", "hi", "atoms", "sort_cell_A", "resort_every", "chunk_tiles",
    "us_per_step", "trials"}; progress(entry) is called after each bucket."""
    table = []
    for bi, (lo, hi) in enumerate(buckets):
        n = representative_size(lo, hi)
        tops = [pm.Topology.generate(n, seed=seed + 97 * bi + k) for k in range(copies)]
        trials = [_time(pm, tops, seed, steps, graph_steps, sort_cell_A=cell, resort_every=resort)
                  for cell in sort_cells for resort in resort_every]
        best = min(trials, key=lambda t: t["us_per_step"])
        pair = {"sort_cell_A": best["sort_cell_A"], "resort_every": best["resort_every"]}
        units = [_time(pm, tops, seed, steps, graph_steps, chunk_tiles=c, **pair) for c in chunk_tiles]
        trials += units
        best_u = min(units, key=lambda t: t["us_per_step"])
        table.append({"lo": lo, "hi": hi, "atoms": n, **pair, "chunk_tiles": best_u["chunk_tiles"],
                      "us_per_step": best_u["us_per_step"], "trials": trials})
        if progress is not None:
            progress(table[-1])
    return table


def choice_for(table, n_atoms: int) -> dict:
    """The tuned values of the bucket holding n_atoms (nearest bucket outside
    the tuned range)."""
    keys = ("sort_cell_A", "resort_every", "chunk_tiles")
    for e in table:
        if e["lo"] <= n_atoms <= e["hi"]:
            return {k: e[k] for k in keys if k in e}
    e = min(table, key=lambda e: min(abs(n_atoms - e["lo"]), abs(n_atoms - e["hi"])))
    return {k: e[k] for k in keys if k in e}


def log_buckets(lo: int, hi: int, n: int):
    """n buckets with log-uniform edges over [lo, hi]."""
    edges = [lo * (hi / lo) ** (k / n) for k in range(n + 1)]
    return [(int(round(edges[k])), int(round(edges[k + 1]))) for k in range(n)]
