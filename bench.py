#!/usr/bin/env python3
"""Throughput of the ccp Brownian-dynamics step on one MI355X.

Workload (the large-nucleus configuration): 8 chains of 16384 beads
(N = 131072) in a sphere at chromatin volume fraction 0.3 (R_n = 44.387 by
ccp's formula), WCA repulsion, heterochromatin LJ attraction eps = 0.6 kT at
2.5 sigma (30 % het in Geometric(20) blocks), 15 % LAD beads clustered at
chain ends with a lamina well, harmonic bonds, bending k_b = 2 kT, dt = 1e-3,
neighbour search over 2.5 sigma cells (the autotuner may split each into
2 x 2 x 2 sub-cells), FP32, contact sampling every 1000 steps into a
1024 x 1024 uint32 map (128 beads per bin), seed 3. One step = one
stochastic-Heun step of all beads (two force evaluations, two cell-list
builds). The force kernel's shape (cell side, reach, lanes per bead) is
picked by tfm_amd.autotune on the workload's initial condition before the
timed run; --no-tune keeps the library default.

Prints one JSON line. `value` is bead-steps/s with the state resident in HBM;
`roofline` is the dominant kernel's algorithmic bytes per launch over its
HIP-event-measured average duration against 8 TB/s; `cpu_baseline` is the
FP64 scalar oracle timed on this host's cores on a bounded sample.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
# vector-ALU issue: 1024 SIMDs x 2.4 GHz / 4 cycles per wave64 instruction
# (no packed FP32), in G wave-instructions/s
VALU_PEAK_GWIPS = 1024 * 2.4 / 4
N_CHAINS, CHAIN_LEN = 8, 16384
N = N_CHAINS * CHAIN_LEN
RCO = float(np.float32(1.154701))
CVF = 0.3
R_N = RCO / 2 + RCO / 2 * (N / CVF) ** (1.0 / 3.0)

PARAMS = {
    "number_of_particles": N,
    "nucleus_radius": f"{R_N:.6f}",
    "number_of_chains": N_CHAINS,
    "force_field": "lj",
    "epsilon_het": 0.6,
    "epsilon_lamina": 1.0,
    "het_fraction": 0.3,
    "het_block_mean": 20,
    "lad_fraction": 0.15,
    "lad_placement": "ends",
    "timestep": 0.001,
    "contact_cutoff": 1.5,
    "contact_bin_size": 128,
    "contact_interval": 1000,
    "seed": 3,
    "graph_steps": 100,
}

# Algorithmic HBM bytes per bead for one launch of each step kernel
# (16 B float4 positions, bucket entries and per-bead terms, 4 B counters;
# neighbour gathers counted once because the sorted copy is read once per
# launch):
#   force stage 1: R sorted pos 16 + R bonded/wall force 16 + R normals 16
#                  + W er 16 + W r+f dt/2+noise 16 + W bucket entry 16
#                  + count atomic 4 = 100
#   force stage 2: R sorted pos 16 + R bonded/wall force 16 + R r+f dt/2+noise 16
#                  + W r 16 + W bucket entry 16 + count atomic 4 = 84
#   order:         R bucket entry 16 + W sorted pos 16 + R bead-ordered pos 16
#                  + W bonded/wall force 16 = 64 per bead, + R count 4
#                  + W count 4 + W start 4 per cell
#   the next step's normals, W 16 per bead once a step: by the second stage's
#   force launch with the type-ordered work list, else by the first order pass
BYTES_FORCE_PER_BEAD = {1: 100, 2: 84}
BYTES_ORDER_PER_BEAD = {1: 64, 2: 64}
BYTES_NORMALS_PER_BEAD = 16
BYTES_ORDER_PER_CELL = 12
# state the timed steps start from: the initial condition advanced by this
# many untimed steps (the relaxed walk's first steps are slower than the
# long run's)
SETTLE_STEPS = 20000
# untimed steps the timed simulation runs on its own kernel shape (continuing
# the settled trajectory) just before the warmup: a long run's step loop
# keeps the GPU busy, and this puts the short timed window in that state
# instead of right after the host-side set-up (measured: a 20-step window
# right after set-up ran 3-4 % slower than the same window after 2000 steps)
RESUME_STEPS = 2000
SETTLE_SHAPE = {"lanes_per_bead": 8, "cell_reach": 2, "cell_size": 1.25}


def sim_shape(params, shape):
    """The kernel-shape keys the timed simulation runs with."""
    out = {k: params[k] for k in ("lanes_per_bead", "cell_reach", "cell_size") if k in params}
    out.update(shape)
    return out


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20000)
    ap.add_argument("--warmup", type=int, default=1000)
    ap.add_argument("--cpu-seconds", type=float, default=20.0, help="budget of the CPU baseline sample (four legs)")
    ap.add_argument("--settle", type=int, default=SETTLE_STEPS,
                    help="untimed steps from the initial condition to the state the run starts from")
    ap.add_argument("--resume", type=int, default=RESUME_STEPS,
                    help="untimed steps of the timed simulation, continuing the settled run, before the warmup")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--traffic", default=os.path.join(ROOT, "profiles", "k_force_typed_lj_n131072_traffic.txt"))
    ap.add_argument("--sq", default=os.path.join(ROOT, "profiles", "k_force_typed_lj_n131072_sq.txt"))
    ap.add_argument("--param", action="append", default=[], metavar="KEY=VALUE",
                    help="override a simulation parameter (e.g. fine_cell_size=1.25)")
    ap.add_argument("--no-tune", action="store_true", help="skip the kernel-shape autotuner")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="run only a plain elementwise kernel of known traffic (256 MiB read + 256 MiB "
                         "written per launch, 10 launches) to scale rocprofv3 FETCH_SIZE/WRITE_SIZE")
    return ap.parse_args()


def copy_calibration():
    """torch.neg over 2^26 float32: 256 MiB read and 256 MiB written per launch."""
    import torch

    a = torch.ones(1 << 26, device="cuda", dtype=torch.float32)
    b = torch.empty_like(a)
    torch.cuda.synchronize()
    for _ in range(10):
        torch.neg(a, out=b)
    torch.cuda.synchronize()
    print(json.dumps({"copy_calibration": {"kernel": "neg (elementwise)", "launches": 10,
                                           "bytes_read_per_launch": 4 << 26, "bytes_written_per_launch": 4 << 26}}),
          flush=True)


def read_traffic(path):
    """HBM bytes per launch of the dominant kernel from a committed rocprofv3
    PMC summary (line 'bytes_per_launch <value>'), or None."""
    try:
        for line in open(path):
            if line.startswith("bytes_per_launch"):
                return float(line.split()[1])
    except OSError:
        pass
    return None


def read_valu(path):
    """Vector-ALU issue of the dominant kernel from a committed rocprofv3 SQ
    summary: (wave-instructions per launch, mean duration of those dispatches
    in us), or None."""
    vals = {}
    try:
        for line in open(path):
            parts = line.split()
            if len(parts) == 2 and parts[0] in ("valu_insts_per_launch", "valu_dispatch_us"):
                vals[parts[0]] = float(parts[1])
    except OSError:
        return None
    if len(vals) != 2:
        return None
    return vals["valu_insts_per_launch"], vals["valu_dispatch_us"]


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown CPU"


def cpu_baseline(sim, budget_s):
    """The scalar oracle on the same state: FP64 and FP32, one thread and all
    of this host's cores (at most 16), each leg a bounded number of steps
    within about a quarter of budget_s. `value` is the fastest leg."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle as po

    info = sim.info
    nproc = len(os.sched_getaffinity(0))
    cores = min(nproc, 16)
    r0 = sim.positions.astype(np.float64)
    osys = po.System(N, float(info.R_n), flags=sim.flags, chain_off=sim.chains.astype(np.int32), ff=po.FF_LJ,
                     csl=2.5, seed=int(info.seed), dt=float(info.dt), bend_all=1, k_b=2.0,
                     eps_het=0.6, eps_lam=1.0, fcap=float(info.force_cap))
    step0 = sim.time[2]
    legs = []
    for prec in (64, 32):
        for threads in (1, cores):
            r = r0.copy()
            r = osys.run(r, step0, 1, prec=prec, threads=threads)  # first step untimed
            t0 = time.perf_counter()
            steps = 0
            while True:
                r = osys.run(r, step0 + 1 + steps, 1, prec=prec, threads=threads)
                steps += 1
                if time.perf_counter() - t0 > budget_s / 4:
                    break
            dt = time.perf_counter() - t0
            legs.append({"prec": f"f{prec}", "threads": threads, "steps": steps, "seconds": round(dt, 3),
                         "value": N * steps / dt})
    best = max(legs, key=lambda e: e["value"])
    return {"value": best["value"], "unit": "bead-steps/s", "cores": best["threads"], "kind": "port",
            "nproc": nproc, "cpu": cpu_model(),
            "sample": f"scalar oracle ({best['prec']}, {best['threads']} threads of {nproc} on {cpu_model()}) on "
                      f"the same N={N} state after the timed run, {best['steps']} steps in {best['seconds']} s; "
                      "four legs (f64/f32 x 1/all threads) in `legs`",
            "legs": legs}


def main():
    args = parse()
    if args.copy_calibration:
        copy_calibration()
        return
    if args.gpus != 1:
        print(f"bench.py measures one GPU only; --gpus {args.gpus} is not supported by this build "
              "(independent replicas shard one per GPU with no exchange)", file=sys.stderr)
        sys.exit(2)
    import tfm_amd

    print(f"[bench] N={N} R_n={R_N:.3f}: initial condition", file=sys.stderr, flush=True)
    params = dict(PARAMS)
    for kv in args.param:
        k, v = kv.split("=", 1)
        params[k] = v
    sim = tfm_amd.Simulation({**params, **SETTLE_SHAPE})
    sim.generate_initial_condition()
    print(f"[bench] settle {args.settle} steps", file=sys.stderr, flush=True)
    sim.run_steps(args.settle)
    sim.synchronize()
    r0, f0, t0_state = sim.positions, sim.flags, sim.time
    sim.close()
    shape, table = {}, []
    if not args.no_tune and not any(k.split("=")[0] in ("cell_size", "cell_reach", "lanes_per_bead")
                                    for k in args.param):
        from tfm_amd import autotune

        shape, table = autotune.tune(params, r0, f0, cutoff=2.5,
                                     log=lambda m: print(m, file=sys.stderr, flush=True))
    sim = tfm_amd.Simulation({**params, **shape})
    sim.flags = f0
    sim.positions = r0
    sim.set_time(*t0_state)
    # every graph block the run replays, captured now
    sim.capture_graphs()
    sim.synchronize()
    print(f"[bench] resume {args.resume} steps", file=sys.stderr, flush=True)
    sim.run_steps(args.resume)
    sim.synchronize()
    sim.contact_reset()
    print(f"[bench] warmup {args.warmup} steps", file=sys.stderr, flush=True)
    sim.run_steps(args.warmup)
    sim.synchronize()

    print(f"[bench] timed {args.steps} steps", file=sys.stderr, flush=True)
    t0 = time.perf_counter()
    sim.run_steps(args.steps)
    sim.synchronize()
    elapsed = time.perf_counter() - t0
    ms_per_step = 1e3 * elapsed / args.steps
    value = N * args.steps / elapsed
    cmap, n_samples = sim.contact_map()

    # per-kernel device time (HIP events on the simulation's stream)
    kt = sim.time_kernels(200)
    stage = 1 if kt[0] >= kt[2] else 2
    force_ms = float(kt[0] if stage == 1 else kt[2])
    eight = int(sim_shape(params, shape).get("lanes_per_bead", 0)) == 8
    reach2 = int(sim_shape(params, shape).get("cell_reach", 1)) == 2
    typed = eight and reach2 and str(params.get("type_order", 1)) != "0"
    kname = "k_force_typed" if typed else "k_force_bal" if eight else "k_force"
    dominant = f"{kname}<LJ,{stage}>"
    normals_force = typed and stage == 2
    normals_order = not typed and stage == 1
    bytes_force = (BYTES_FORCE_PER_BEAD[stage] + (BYTES_NORMALS_PER_BEAD if normals_force else 0)) * N
    achieved = bytes_force / (force_ms * 1e-3) / 1e9
    traffic = read_traffic(args.traffic)
    valu = read_valu(args.sq)
    order_bytes = ((BYTES_ORDER_PER_BEAD[stage] + (BYTES_NORMALS_PER_BEAD if normals_order else 0)) * N
                   + BYTES_ORDER_PER_CELL * sim.info.n_c)
    sim.synchronize()

    cpu = None if args.no_cpu else cpu_baseline(sim, args.cpu_seconds)
    out = {
        "metric": "bead-steps/s (N=131k, 1 GPU)",
        "value": value,
        "unit": "bead-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": "ccp BD step, 8 chains x 16384 beads (N=131072), volume fraction 0.3, WCA + het LJ "
                        "eps 0.6 at 2.5 sigma, LAD 15% at chain ends, dt 1e-3, 2.5 sigma cell neighbourhood, "
                        "contacts every 1000 steps into 1024x1024 uint32",
            "kernel_shape": {"cell_size": float(sim.info.fine_csl), **{k: v for k, v in shape.items()
                                                                        if k != "cell_size"}},
            "autotune_us_per_step": [[c, round(t * 1e3, 2)] for c, t in table],
            "N": N, "chains": N_CHAINS, "R_n": round(R_N, 4), "cells_per_side": sim.info.cps,
            "untimed_steps": {"settle": args.settle, "resume": args.resume, "warmup": args.warmup},
            "contact_samples": n_samples, "contact_total": int(cmap.sum(dtype=np.uint64)),
            "overrides": args.param,
        },
        "roofline": {
            "bound": "hbm", "limiter": "valu issue and load latency (see profiles/k_force_typed_lj_n131072_sq.txt)", "kernel": dominant, "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
            # whole step at 176 algorithmic bytes per bead-step (index-array cell list)
            "step_frac_176B": 176.0 * value / (HBM_PEAK_GBS * 1e9),
            "kernel_ms": {"force_stage1": float(kt[0]), "order_stage1": float(kt[1]),
                          "force_stage2": float(kt[2]), "order_stage2": float(kt[3]), "step": float(kt[4])},
            "bytes_per_launch": {"force": bytes_force, "order": order_bytes},
            # secondary bound: vector-ALU issue of the same kernel (no packed FP32), from the
            # committed SQ counter run (its instruction count over its own dispatch time)
            "valu": None if valu is None else {
                "achieved": valu[0] / (valu[1] * 1e-6) / 1e9, "peak": VALU_PEAK_GWIPS,
                "unit": "G wave-instructions/s", "frac": valu[0] / (valu[1] * 1e-6) / 1e9 / VALU_PEAK_GWIPS,
                "insts_per_launch": valu[0]},
        },
        "cpu_baseline": cpu,
    }
    print(json.dumps(out, default=float), flush=True)
    sim.close()


if __name__ == "__main__":
    main()
