"""Ensemble run: independent systems of uneven chain lengths stepped together.

Draws `--systems` systems with one chain each:
- chain length log-uniform in [--min-beads, --max-beads];
- volume fraction U(0.1, 0.35), eps_het U(0.2, 1.0) kT, eps_lamina U(0.2, 1.5) kT;
- het 30 % (geometric blocks of mean 20) and LAD 10 % patterns from the
  system's own seed (`--seed0 + i`).

The kernel shape is picked by the batch autotuner on these systems; then the
initial conditions are generated and the batch runs `--steps` steps of the WCA
+ het LJ set (dt 1e-3) in hipGraph-captured blocks, sampling every system's
128 x 128 contact map (contacts < 1.5 sigma) and 32-bin radial density every
`--interval` steps. One JSON line reports bead-steps/s over the whole batch and
per-system sums of the outputs.

    python -m tfm_amd.ensemble [--systems 256] [--steps 100000] [--out result.json]
"""
import argparse
import json
import sys
import time

import numpy as np

from .autotune import tune_batch
from .ccp import Batch


def draw_systems(n, lo, hi, seed0):
    """The ensemble's systems (dicts for tfm_amd.Batch)."""
    out = []
    for i in range(n):
        rng = np.random.default_rng(seed0 + i)
        out.append({
            "n_beads": int(round(np.exp(rng.uniform(np.log(lo), np.log(hi))))),
            "volume_fraction": float(rng.uniform(0.1, 0.35)),
            "eps_het": float(rng.uniform(0.2, 1.0)),
            "eps_lam": float(rng.uniform(0.2, 1.5)),
            "het_fraction": 0.3,
            "lad_fraction": 0.1,
            "seed": seed0 + i,
        })
    return out


def run(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("--systems", type=int, default=256)
    ap.add_argument("--min-beads", type=int, default=512)
    ap.add_argument("--max-beads", type=int, default=16384)
    ap.add_argument("--steps", type=int, default=100000)
    ap.add_argument("--interval", type=int, default=1000)
    ap.add_argument("--graph-steps", type=int, default=100)
    ap.add_argument("--seed0", type=int, default=4000)
    ap.add_argument("--no-tune", action="store_true",
                    help="use the 8-lane, reach-2, 1.25 sigma shape (refused above 524288 beads)")
    ap.add_argument("--out", default=None, help="also write the JSON line to this file")
    a = ap.parse_args(argv)
    log = lambda m: print(m, file=sys.stderr, flush=True)

    systems = draw_systems(a.systems, a.min_beads, a.max_beads, a.seed0)
    n_total = sum(s["n_beads"] for s in systems)
    common = {"timestep": 0.001, "contact_interval": a.interval, "contact_cutoff": 1.5, "contact_bins": 128,
              "radial_bins": 32, "graph_steps": a.graph_steps}
    log(f"[ensemble] {a.systems} systems, {n_total} beads")
    t0 = time.perf_counter()
    if a.no_tune:
        shape, table = {"lanes_per_bead": 8, "cell_reach": 2, "cell_size": 1.25}, []
    else:
        shape, table = tune_batch(systems, common, log=log)
    t_tune = time.perf_counter() - t0

    b = Batch(systems, {**common, **shape})
    try:
        t0 = time.perf_counter()
        b.generate_initial_conditions()
        b.synchronize()
        t_ic = time.perf_counter() - t0
        b.step = 0
        b.reset_samples()
        t0 = time.perf_counter()
        done = 0
        while done < a.steps:
            chunk = min(10 * a.interval, a.steps - done)
            b.run_steps(chunk)
            b.synchronize()
            done += chunk
            log(f"[ensemble] {done}/{a.steps} steps ({time.perf_counter() - t0:.0f} s)")
        t_run = time.perf_counter() - t0
        maps, n_samples = b.contact_maps()
        rd = b.radial_density()
        xyz, _ = b.state()
    finally:
        b.close()

    lengths = np.array([s["n_beads"] for s in systems])
    out = {
        "workload": f"{a.systems} independent single-chain systems, chain length log-uniform "
                    f"{a.min_beads}-{a.max_beads}, vf U(0.1,0.35), eps_het U(0.2,1.0), eps_lamina U(0.2,1.5), "
                    f"WCA + het LJ, dt 1e-3, contacts < 1.5 sigma into 128x128 per system every {a.interval} steps",
        "systems": a.systems, "beads": int(n_total), "steps": a.steps,
        "bead_steps_per_s": n_total * a.steps / t_run,
        "seconds": {"tune": round(t_tune, 2), "initial_conditions": round(t_ic, 2), "run": round(t_run, 2)},
        "kernel_shape": shape,
        "autotune_us_per_step": [[c, round(t * 1e3, 2)] for c, t in table],
        "chain_length": {"min": int(lengths.min()), "median": float(np.median(lengths)), "max": int(lengths.max())},
        "contact_samples": int(n_samples),
        "finite_state": bool(np.isfinite(xyz).all()),
        "contacts_per_system_mean": float(maps.reshape(a.systems, -1).sum(axis=1, dtype=np.float64).mean()),
        "maps_symmetric": bool(all(np.array_equal(m, m.T) for m in maps)),
        "radial_density_mean": [float(x) for x in rd.mean(axis=0)],
    }
    line = json.dumps(out)
    print(line, flush=True)
    if a.out:
        with open(a.out, "w") as f:
            f.write(line + "\n")
    return out


def main():
    run()


if __name__ == "__main__":
    main()
