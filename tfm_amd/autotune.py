"""Autotuner of the force kernel's shape.

The step's pair search has three knobs that change no result beyond FP32
summation order: the cell side, how many cells it reaches on each side
(1: ccp's 3 x 3 x 3 neighbourhood; 2: 5 x 5 x 5 cells of half the side,
i.e. each cell split into 2 x 2 x 2 sub-cells), and how many lanes of a
wavefront share a bead (1; one per z-plane; or 8 over a compacted list of
candidate slots). The best shape depends on density, the het fraction and
the chip, so it is measured: every candidate runs the same state for a few
hundred steps and the fastest whole step wins. Results are cached in a JSON
file keyed by the workload's shape.
"""
import hashlib
import json
import os

import numpy as np

from .ccp import Batch, Simulation


# Batches above this many beads refuse eight lanes per bead over half-cutoff
# cells (reach 2, cells below 1.3 sigma): a larger run of that shape faulted
# the GPU and its cause is not known (the library gives the same limit).
HALF_CELL_BATCH_LIMIT = 524288


def candidates(cutoff, batch_beads=0):
    """Kernel shapes for a longest pair cutoff `cutoff` (sigma); for a batch of
    `batch_beads` beads, without the shapes the library refuses at that size."""
    half = cutoff / 2
    out = [
        {"lanes_per_bead": 3, "cell_reach": 1, "cell_size": cutoff},
        {"lanes_per_bead": 1, "cell_reach": 1, "cell_size": cutoff},
        {"lanes_per_bead": 8, "cell_reach": 1, "cell_size": cutoff},
        {"lanes_per_bead": 5, "cell_reach": 2, "cell_size": half},
        {"lanes_per_bead": 8, "cell_reach": 2, "cell_size": half},
        {"lanes_per_bead": 8, "cell_reach": 2, "cell_size": round(half * 1.1, 4)},
    ]
    if batch_beads > HALF_CELL_BATCH_LIMIT:
        out = [c for c in out if not (c["lanes_per_bead"] == 8 and c["cell_reach"] == 2 and c["cell_size"] < 1.3)]
    return out


def _key(params, n_flags):
    text = json.dumps({k: str(v) for k, v in sorted(params.items())}, sort_keys=True) + f"|{n_flags}"
    return hashlib.sha1(text.encode()).hexdigest()[:16]


def tune(params, positions, flags, cutoff, steps=200, warmup=50, cache=None, log=None):
    """Time every candidate shape on (positions, flags) and return
    (best overrides, [(overrides, ms per step), ...]). `params` must not
    fix the shape keys themselves."""
    key = _key(params, int(np.count_nonzero(flags)))
    if cache and os.path.exists(cache):
        try:
            stored = json.load(open(cache)).get(key)
        except (OSError, ValueError):
            stored = None
        if stored:
            return stored["best"], [(c, t) for c, t in stored["table"]]
    table = []
    for cand in candidates(cutoff):
        sim = Simulation({**params, **cand})
        try:
            sim.flags = flags
            sim.positions = positions
            sim.set_time(0, 0.0, 0)
            sim.run_steps(warmup)
            sim.synchronize()
            ms = float(sim.time_kernels(steps)[4])
        finally:
            sim.close()
        table.append((cand, ms))
        if log:
            log(f"[autotune] {cand} {ms * 1e3:.1f} us/step")
    best = min(table, key=lambda e: e[1])[0]
    if cache:
        data = {}
        if os.path.exists(cache):
            try:
                data = json.load(open(cache))
            except (OSError, ValueError):
                data = {}
        data[key] = {"best": best, "table": table}
        with open(cache, "w") as f:
            json.dump(data, f, indent=1)
    return best, table


def tune_batch(systems, common, cutoff=2.5, steps=100, warmup=20, settle=2000, log=None):
    """Time every candidate shape on a batch of systems (tfm_amd.Batch) from
    the same state: the initial conditions after `settle` steps, closer to
    the long run than the fresh relaxed walks. Returns (best overrides,
    [(overrides, ms per step), ...])."""
    ref = Batch(systems, common)
    try:
        ref.generate_initial_conditions()
        if settle:
            ref.run_steps(settle)
        xyz, fl = ref.state()
    finally:
        ref.close()
    table = []
    for cand in candidates(cutoff, sum(int(s["n_beads"]) for s in systems)):
        b = Batch(systems, {**common, **cand})
        try:
            b.set_state(xyz, fl)
            b.run_steps(warmup)
            ms = b.time_steps(steps)
        finally:
            b.close()
        table.append((cand, ms))
        if log:
            log(f"[autotune] batch {cand} {ms * 1e3:.1f} us/step")
    return min(table, key=lambda e: e[1])[0], table
