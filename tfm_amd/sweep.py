"""Surrogate-guided parameter sweep over the batch engine.

1. Draws parameter sets: a chain of n beads with a het/LAD pattern from its
   seed, volume fraction U(0.1, 0.35), eps_het U(0.2, 1.0) kT and
   eps_lamina U(0.2, 1.5) kT.
2. Simulates them in batches (tfm_amd.Batch): equilibration, then sampled
   contact maps (n / 16 beads per bin -> bins x bins) and radial densities.
3. Trains the surrogate (tfm_amd.surrogate) on log10 contact probabilities
   and radial densities.
4. Sweeps: scores many candidate parameter sets with the surrogate against
   a target map, re-simulates the best ones with plain BD, and measures the
   surrogate's error on them, mean |dlog10 P| over the pairs i < j.

`python -m tfm_amd.sweep` runs the whole chain on one GPU and prints one JSON
line: parameter sets resolved per hour at the contact-map error bar (only
re-simulated sets whose verified error is within it count), beside plain
one-by-one BD at the same bar and the errors actually achieved.
"""
import argparse
import json
import math
import sys
import time

import numpy as np
import torch

from . import surrogate as sg
from .autotune import tune_batch
from .ccp import Batch, Simulation, sequence_pattern

RCO = float(np.float32(1.154701))
BINS = 128
NEAR = 8  # separation (bins) of the well-sampled band of the maps


def radius(n, vf):
    return RCO / 2 + RCO / 2 * (n / vf) ** (1.0 / 3.0)


def draw(rng, n_sets, seed0):
    return [dict(vf=float(rng.uniform(0.1, 0.35)), eps_het=float(rng.uniform(0.2, 1.0)),
                 eps_lam=float(rng.uniform(0.2, 1.5)), seed=int(seed0 + k)) for k in range(n_sets)]


def single_params(d, n_beads, common):
    return {**common, "number_of_particles": n_beads, "nucleus_radius": f"{radius(n_beads, d['vf']):.6f}",
            "number_of_chains": 1, "force_field": "lj", "epsilon_het": d["eps_het"], "epsilon_lamina": d["eps_lam"],
            "het_fraction": 0.3, "lad_fraction": 0.1, "seed": d["seed"], "number_of_lbs": 0}


def features(draws, n_beads):
    """[5, len(draws) BINS]: per-bin het and LAD fractions and the three scalars (centred)."""
    per = n_beads // BINS
    x = np.zeros((5, len(draws), BINS), np.float32)
    for k, d in enumerate(draws):
        fl = sequence_pattern(single_params(d, n_beads, {}), n_beads)
        x[0, k] = ((fl & 1) > 0).reshape(BINS, per).mean(axis=1)
        x[1, k] = ((fl & 2) > 0).reshape(BINS, per).mean(axis=1)
        x[2, k] = (d["vf"] - 0.225) / 0.072
        x[3, k] = (d["eps_het"] - 0.6) / 0.23
        x[4, k] = (d["eps_lam"] - 0.85) / 0.375
    return x


def log_map(counts, n_samples, n_beads):
    """log10 contact probability per bin pair from summed counts (each pair counted in both
    orders off the diagonal, twice on it) with half a count of pseudo-count."""
    per = n_beads // BINS
    pairs = np.full((BINS, BINS), float(per * per))
    np.fill_diagonal(pairs, float(per * (per - 1)))
    return np.log10((counts.astype(np.float64) + 0.5) / (n_samples * pairs + 1.0)).astype(np.float32)


def simulate(draws, n_beads, warmup, steps, interval, batch_size, shape, log=print, replicas=1, halves=False):
    """Plain BD of every draw, `replicas` independent runs each (same pattern,
    own walk and noise), in batches. Returns (log10 maps of the counts summed
    over replicas [n, BINS, BINS], mean radial densities [n, 32], seconds);
    with halves=True also the maps of the two halves of the replicas (the
    first and second half of them, or of the run when there is one)."""
    common = {"timestep": 0.001, "contact_interval": interval, "contact_cutoff": 1.5, "contact_bins": BINS,
              "radial_bins": 32, "graph_steps": 100, **shape}
    maps, rds, first, second = [], [], [], []
    per_batch = max(1, batch_size // replicas)
    t0 = time.perf_counter()
    for k0 in range(0, len(draws), per_batch):
        part = draws[k0:k0 + per_batch]
        systems = [dict(n_beads=n_beads, volume_fraction=d["vf"], eps_het=d["eps_het"], eps_lam=d["eps_lam"],
                        het_fraction=0.3, lad_fraction=0.1, seed=d["seed"], replica=r)
                   for d in part for r in range(replicas)]
        b = Batch(systems, common)
        b.generate_initial_conditions()
        b.run_steps(warmup)
        b.reset_samples()
        if halves and replicas == 1:
            b.run_steps(steps // 2)
            m1, n1 = b.contact_maps()
            b.run_steps(steps - steps // 2)
            m, ns = b.contact_maps()
            first.extend(log_map(m1[s], n1, n_beads) for s in range(len(part)))
            second.extend(log_map(m[s] - m1[s], ns - n1, n_beads) for s in range(len(part)))
        else:
            b.run_steps(steps)
            m, ns = b.contact_maps()
        rd = b.radial_density().reshape(len(part), replicas, -1).mean(axis=1)
        b.close()
        m = m.reshape(len(part), replicas, BINS, BINS).astype(np.int64)
        maps.extend(log_map(m[s].sum(axis=0), ns * replicas, n_beads) for s in range(len(part)))
        if halves and replicas > 1:
            h = replicas // 2
            first.extend(log_map(m[s, :h].sum(axis=0), ns * h, n_beads) for s in range(len(part)))
            second.extend(log_map(m[s, h:].sum(axis=0), ns * (replicas - h), n_beads) for s in range(len(part)))
        rds.append(rd)
        log(f"[sweep] simulated {k0 + len(part)}/{len(draws)} x {replicas} ({time.perf_counter() - t0:.0f} s)")
    out = (np.stack(maps), np.concatenate(rds), time.perf_counter() - t0)
    return out + (np.stack(first), np.stack(second)) if halves else out


def map_error(a, b, max_sep=None):
    """mean |a - b| over the bin pairs i < j (i < j <= i + max_sep) of each map."""
    iu = np.triu_indices(BINS, 1)
    if max_sep is not None:
        keep = iu[1] - iu[0] <= max_sep
        iu = (iu[0][keep], iu[1][keep])
    return np.abs(a[:, iu[0], iu[1]] - b[:, iu[0], iu[1]]).mean(axis=1)


def ps_error(a, b):
    """mean |dlog10 P(s)| over the separations s = 1 .. BINS - 1 (in bins) of
    each pair of maps, P(s) the contact probability averaged along diagonal s
    (linear average of 10^map, then log10): the map's contact probability by
    genomic separation."""
    def ps(m):
        lin = np.power(10.0, m.astype(np.float64))
        return np.log10(np.stack([np.diagonal(lin, offset=k, axis1=1, axis2=2).mean(axis=1)
                                  for k in range(1, BINS)], axis=1))
    return np.abs(ps(a) - ps(b)).mean(axis=1)


def train(net, X, Y, R, epochs, lr, log=print):
    """Minibatch AdamW with a cosine schedule; X [5, n, BINS], Y [n, BINS, BINS], R [n, 32]."""
    n, B = X.shape[1], net.B
    steps = epochs * (n // B)
    rng = np.random.default_rng(0)
    dev = torch.device("cuda")
    # start from the data's mean map and density: the distance bias takes the
    # mean log10 P of each diagonal, the radial bias the mean density
    L = Y.shape[1]
    dist = np.array([np.mean(np.diagonal(Y, offset=k, axis1=1, axis2=2)) for k in range(L)], np.float32)
    net.p["dist"].copy_(torch.tensor(dist, device=dev))
    net.p["c0"].zero_()
    net.p["br"].copy_(torch.tensor(R.mean(axis=0), dtype=torch.float32, device=dev))
    Xd = torch.tensor(X, dtype=torch.float32, device=dev)
    Yd = torch.tensor(Y, dtype=torch.float32, device=dev)
    Rd = torch.tensor(R, dtype=torch.float32, device=dev)
    t = 0
    for ep in range(epochs):
        perm = torch.tensor(rng.permutation(n), device=dev)
        tot = 0.0
        for k in range(n // B):
            idx = perm[k * B:(k + 1) * B]
            x = Xd[:, idx, :].reshape(5, B * BINS).contiguous()
            y = Yd[idx].contiguous()
            r = Rd[idx].T.contiguous()
            net.zero_grad()
            loss = net.loss_backward(x, y, r, w_pair=1.0, w_rd=1.0)
            cur = lr * 0.5 * (1.0 + math.cos(math.pi * t / max(steps, 1)))
            net.step(lr=cur, wd=1e-4)
            tot += float(loss.item())
            t += 1
        log(f"[sweep] epoch {ep + 1}/{epochs} loss {tot / max(n // B, 1):.4f}")


def predict(net, X):
    n, B = X.shape[1], net.B
    out_m = np.zeros((n, BINS, BINS), np.float32)
    out_r = np.zeros((n, net.rd_bins), np.float32)
    dev = torch.device("cuda")
    for k0 in range(0, n, B):
        k1 = min(n, k0 + B)
        x = np.zeros((5, B, BINS), np.float32)
        x[:, :k1 - k0] = X[:, k0:k1]
        S, rd = net.predict(torch.tensor(x.reshape(5, B * BINS), dtype=torch.float32, device=dev))
        out_m[k0:k1] = S[:k1 - k0].cpu().numpy()
        out_r[k0:k1] = rd[:, :k1 - k0].T.cpu().numpy()
    return out_m, out_r


def one_by_one_rate(draws, n_beads, warmup, steps, interval, shape, log=print):
    """Parameter sets per hour of plain BD run one simulation at a time."""
    common = {"timestep": 0.001, "contact_interval": interval, "contact_cutoff": 1.5, "graph_steps": 100, **shape}
    t0 = time.perf_counter()
    for d in draws:
        sim = Simulation(single_params(d, n_beads, common))
        sim.generate_initial_condition()
        sim.run_steps(warmup)
        sim.contact_reset()
        sim.run_steps(steps)
        sim.contact_map()
        sim.close()
    dt = (time.perf_counter() - t0) / len(draws)
    log(f"[sweep] one-by-one BD: {dt:.2f} s per parameter set")
    return 3600.0 / dt


def main(argv=None):
    print(json.dumps(run(argv)), flush=True)


def run(argv=None):
    """The whole sweep; returns the result record."""
    ap = argparse.ArgumentParser()
    ap.add_argument("--n-beads", type=int, default=2048)
    ap.add_argument("--train", type=int, default=4096)
    ap.add_argument("--warmup", type=int, default=5000)
    ap.add_argument("--steps", type=int, default=20000)
    ap.add_argument("--interval", type=int, default=200)
    ap.add_argument("--batch", type=int, default=256)
    ap.add_argument("--epochs", type=int, default=30)
    ap.add_argument("--candidates", type=int, default=65536)
    ap.add_argument("--top", type=int, default=256)
    ap.add_argument("--single", type=int, default=4, help="one-by-one BD runs timed for the plain rate")
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--replicas", type=int, default=1, help="independent runs averaged per parameter set")
    ap.add_argument("--bar", type=float, default=0.05,
                    help="contact-map error (mean |dlog10 P| over bin pairs i < j) at which a set counts as resolved")
    ap.add_argument("--no-tune", action="store_true",
                    help="skip the batch autotuner (8 lanes per bead, reach 2, 1.25 sigma cells)")
    a = ap.parse_args(argv)
    log = lambda m: print(m, file=sys.stderr, flush=True)
    rng = np.random.default_rng(a.seed)
    t_start = time.perf_counter()

    # data from the batch engine, in the kernel shape the autotuner picks on
    # the first batch of draws
    draws = draw(rng, a.train, 4000)
    shape = {"lanes_per_bead": 8, "cell_reach": 2, "cell_size": 1.25}
    if not a.no_tune:
        first = [dict(n_beads=a.n_beads, volume_fraction=d["vf"], eps_het=d["eps_het"], eps_lam=d["eps_lam"],
                      het_fraction=0.3, lad_fraction=0.1, seed=d["seed"]) for d in draws[:a.batch]]
        common = {"timestep": 0.001, "contact_interval": a.interval, "contact_cutoff": 1.5, "contact_bins": BINS,
                  "radial_bins": 32, "graph_steps": 100}
        shape, _ = tune_batch(first, common, log=log)
    X = features(draws, a.n_beads)
    Y, R, t_sim = simulate(draws, a.n_beads, a.warmup, a.steps, a.interval, a.batch, shape, log, a.replicas)
    n_val = max(a.train // 8, 1)
    net = sg.Surrogate(c_in=5, width=256, n_layers=6, d=64, L=BINS, rd_bins=32, B=64, seed=a.seed)
    t0 = time.perf_counter()
    train(net, X[:, n_val:], Y[n_val:], R[n_val:], a.epochs, a.lr, log)
    t_train = time.perf_counter() - t0
    pm, pr = predict(net, X[:, :n_val])
    mean_map = np.broadcast_to(Y[n_val:].mean(axis=0), Y[:n_val].shape)
    val_err = float(map_error(pm, Y[:n_val]).mean())
    val_near = float(map_error(pm, Y[:n_val], NEAR).mean())
    # the mean map as a no-learning baseline
    base_err = float(map_error(mean_map, Y[:n_val]).mean())
    base_near = float(map_error(mean_map, Y[:n_val], NEAR).mean())
    val_ps = float(ps_error(pm, Y[:n_val]).mean())
    base_ps = float(ps_error(mean_map, Y[:n_val]).mean())
    rd_err = float(np.abs(pr - R[:n_val]).mean())
    rd_base = float(np.abs(R[n_val:].mean(axis=0)[None] - R[:n_val]).mean())
    log(f"[sweep] validation mean |dlog10 P| surrogate {val_err:.4f} (near {val_near:.4f}, P(s) {val_ps:.4f}), "
        f"mean-map baseline {base_err:.4f} (near {base_near:.4f}, P(s) {base_ps:.4f}); radial density "
        f"{rd_err:.4f} vs {rd_base:.4f}")

    # sweep: rank candidates by predicted distance to a target map
    target = Y[0]
    t0 = time.perf_counter()
    cand = draw(rng, a.candidates, 10_000_000)
    Xc = features(cand, a.n_beads)
    t_feat = time.perf_counter() - t0
    t0 = time.perf_counter()
    pc, _ = predict(net, Xc)
    t_score = time.perf_counter() - t0
    dist = map_error(pc, target[None])
    best = np.argsort(dist)[:a.top]
    top = [cand[k] for k in best]
    Yt, _, t_verify, Y1, Y2 = simulate(top, a.n_beads, a.warmup, a.steps, a.interval, a.batch, shape, log,
                                       a.replicas, halves=True)
    top_err = map_error(pc[best], Yt)
    top_near = map_error(pc[best], Yt, NEAR)
    # BD's own sampling noise: the two halves of each verification run
    # differ by about sqrt(2) x sqrt(2) = 2 times the full run's error
    noise_err = float(map_error(Y1, Y2).mean() / 2.0)
    noise_near = float(map_error(Y1, Y2, NEAR).mean() / 2.0)
    top_ps = ps_error(pc[best], Yt)
    noise_ps = ps_error(Y1, Y2) / 2.0
    t_total = time.perf_counter() - t_start

    plain = one_by_one_rate(draw(rng, a.single, 20_000_000), a.n_beads, a.warmup, a.steps, a.interval, shape, log)
    # A set is resolved when its map is known within the bar: for the
    # surrogate, a re-simulated top candidate whose verified error is within
    # it; for plain BD, a run whose own sampling error (half the difference of
    # its two halves) is within it. With one run per set those halves share
    # the chain's conformation, so that error is a lower bound; with replicas
    # the halves are independent runs. Scored candidates with no verification
    # count for nothing.
    resolved = int((top_err <= a.bar).sum())
    guided = resolved / (t_total / 3600.0)
    bd_ok = float((map_error(Y1, Y2) / 2.0 <= a.bar).mean())
    plain_resolved = plain * bd_ok
    # the same count at the bar on P(s), the map's contact probability by separation
    resolved_ps = int((top_ps <= a.bar).sum())
    bd_ok_ps = float((noise_ps <= a.bar).mean())
    out = {
        "metric": "sweep param-sets/hr at fixed contact-map error",
        "value": guided, "unit": "param-sets/hr", "higher_is_better": True,
        "bar": {"contact_map_error": a.bar, "metric": "mean |dlog10 P| over bin pairs i < j"},
        "resolved_verified": resolved, "verified": int(a.top),
        "plain_one_by_one": {"sets_per_hr": plain, "fraction_within_bar": bd_ok, "resolved_per_hr": plain_resolved,
                             "error_estimate": "independent replicas" if a.replicas > 1 else
                             "halves of one run (a lower bound: they share the conformation)"},
        "vs_plain": (guided / plain_resolved) if plain_resolved > 0 else None,
        "scored_per_hr": a.candidates / (t_total / 3600.0),
        "top_errors": [float(e) for e in np.sort(top_err)],
        "contact_map_error": {"surrogate_top": float(top_err.mean()), "surrogate_validation": val_err,
                              "mean_map_baseline": base_err, "bd_sampling_noise": noise_err,
                              "near": {"max_separation_bins": NEAR, "surrogate_top": float(top_near.mean()),
                                       "surrogate_validation": val_near, "mean_map_baseline": base_near,
                                       "bd_sampling_noise": noise_near},
                              "radial_density_mae": {"surrogate": rd_err, "mean_baseline": rd_base},
                              "metric": "mean |dlog10 P| over bin pairs i < j"},
        "contact_probability_error": {
            "metric": "mean |dlog10 P(s)| over separations s = 1..127 bins, P(s) averaged along each diagonal",
            "surrogate_top": float(top_ps.mean()), "surrogate_validation": val_ps, "mean_map_baseline": base_ps,
            "bd_sampling_noise": float(noise_ps.mean()),
            "resolved_verified": resolved_ps, "resolved_per_hr": resolved_ps / (t_total / 3600.0),
            "plain_one_by_one_resolved_per_hr": plain * bd_ok_ps},
        "config": {"n_beads": a.n_beads, "train_sets": a.train, "replicas": a.replicas, "warmup": a.warmup,
                   "steps": a.steps,
                   "sample_interval": a.interval, "candidates": a.candidates, "top": a.top, "epochs": a.epochs,
                   "kernel_shape": shape},
        "seconds": {"simulate_train_sets": t_sim, "train": t_train, "candidate_features": t_feat,
                    "score": t_score, "verify": t_verify, "total": t_total},
    }
    return out


if __name__ == "__main__":
    main()
