"""Per-size-bucket autotuner of the pair-list layout for batched ensembles.

Two parameters shape the cluster-pair tiles without touching the physics
(every pair inside the cutoff is evaluated whatever they are):

  sort_cell_A   edge of the Hilbert sort cells that order atoms into 64-atom
                blocks -- finer cells give more compact blocks (fewer, fuller
                tiles) at a longer sort key;
  resort_every  how many pair-list rebuilds reuse the block assignment (only
                refreshing block boxes) before the atoms are sorted again.

For each size bucket the tuner builds a small batch of generator systems of
the bucket's representative size, times every candidate pair of values with
the same captured-graph stepping the ensemble uses, and keeps the fastest.
The reference has no counterpart: its production engine runs one system per
engine with a fixed 10 A cell grid (crates/prism-nhs/src/fused_engine.rs:12285-12330).
"""
from __future__ import annotations

import math
import time

SORT_CELLS = (1.5, 2.0, 3.0)
RESORT_EVERY = (4, 8, 16)


def representative_size(lo: int, hi: int) -> int:
    return int(round(math.sqrt(lo * hi)))


def tune(pm, buckets, steps: int = 400, graph_steps: int = 10, copies: int = 4, seed: int = 0x41544E45,
         sort_cells=SORT_CELLS, resort_every=RESORT_EVERY):
    """buckets: [(lo, hi)] atom-count ranges.  Returns one entry per bucket:
    {"lo", "hi", "atoms", "sort_cell_A", "resort_every", "us_per_step", "trials"}."""
    table = []
    for bi, (lo, hi) in enumerate(buckets):
        n = representative_size(lo, hi)
        tops = [pm.Topology.generate(n, seed=seed + 97 * bi + k) for k in range(copies)]
        trials = []
        for cell in sort_cells:
            for resort in resort_every:
                params = [pm.default_params(seed=seed + k, stream_id=k, sort_cell_A=cell, resort_every=resort)
                          for k in range(copies)]
                with pm.Ensemble(tops, params) as ens:
                    ens.step(2 * graph_steps, graph_steps)  # captures the graph
                    t0 = time.perf_counter()
                    ens.step(steps, graph_steps)
                    dt = time.perf_counter() - t0
                trials.append({"sort_cell_A": cell, "resort_every": resort,
                               "us_per_step": round(dt / steps * 1e6, 2)})
        best = min(trials, key=lambda t: t["us_per_step"])
        table.append({"lo": lo, "hi": hi, "atoms": n, "sort_cell_A": best["sort_cell_A"],
                      "resort_every": best["resort_every"], "us_per_step": best["us_per_step"],
                      "trials": trials})
    return table


def choice_for(table, n_atoms: int) -> dict:
    """The tuned values of the bucket holding n_atoms (nearest bucket outside
    the tuned range)."""
    for e in table:
        if e["lo"] <= n_atoms <= e["hi"]:
            return {"sort_cell_A": e["sort_cell_A"], "resort_every": e["resort_every"]}
    e = min(table, key=lambda e: min(abs(n_atoms - e["lo"]), abs(n_atoms - e["hi"])))
    return {"sort_cell_A": e["sort_cell_A"], "resort_every": e["resort_every"]}


def log_buckets(lo: int, hi: int, n: int):
    """n buckets with log-uniform edges over [lo, hi]."""
    edges = [lo * (hi / lo) ** (k / n) for k in range(n + 1)]
    return [(int(round(edges[k])), int(round(edges[k + 1]))) for k in range(n)]
