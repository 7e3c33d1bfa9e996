"""GPU parity: the HIP path against the CPU oracle on identical inputs."""
import numpy as np
import pytest

import pyoracle as po
import tfm_amd
from gpu_common import make_pair, step
from helpers import lbs_sites, pattern, radius_for, relaxed, walk

pytestmark = pytest.mark.gpu





@pytest.mark.parametrize("N,cvf,seed", [(1, 0.2, 1), (1000, 0.2, 3), (16384, 0.2, 4), (131072, 0.3, 5)])
def test_cell_list_bit_exact(N, cvf, seed):
    step(f"cell list N={N}")
    R_n = radius_for(N, cvf) if N > 1 else 5.0
    r = walk(N, R_n, seed)
    sim, osys, r64 = make_pair(N, R_n, np.zeros(N, np.uint8), r)
    got = sim.cell_list()
    want = po.cell_list(r64.astype(np.float32), np.float32(2.0), sim.info.cps)
    assert want["bad"] == 0
    for k in ("sci", "soi", "beg", "end"):
        assert np.array_equal(got[k], want[k]), k
    sim.close()


def test_noise_matches_oracle():
    step("noise")
    N = 200_000
    sim, _, _ = make_pair(N, radius_for(N, 0.2), np.zeros(N, np.uint8), np.zeros((N, 3)), seed=0xDEADBEEF12345)
    for s in (0, 7, (3 << 32) + 5):
        z, a = sim.noise(s)
        zo, ao = po.noise_batch(0xDEADBEEF12345, s, N)
        assert np.array_equal(a, ao.astype(np.uint32))
        assert np.abs(z - zo).max() < 2e-5
    assert (a > 0).sum() <= 10
    sim.close()


FORCE_CASES = [
    dict(ff=po.FF_CCP, n_l=40),
    dict(ff=po.FF_ICG, n_l=0),
    dict(ff=po.FF_LJ, n_l=30, eps_het=0.6, eps_lam=1.2),
    dict(ff=po.FF_CCP, n_l=20, opening_radius=4.0, bleb_radius=6.0),
    # one lane per bead (the autotuner's other kernel shape)
    dict(ff=po.FF_CCP, n_l=40, lanes_per_bead=1),
    dict(ff=po.FF_LJ, n_l=30, eps_het=0.6, eps_lam=1.2, lanes_per_bead=1),
    # 5 x 5 x 5 cells of half the side, a lane per plane or one lane per bead
    dict(ff=po.FF_CCP, n_l=40, cell_reach=2),
    dict(ff=po.FF_ICG, n_l=0, cell_reach=2),
    dict(ff=po.FF_LJ, n_l=30, eps_het=0.6, eps_lam=1.2, cell_reach=2),
    dict(ff=po.FF_LJ, n_l=30, eps_het=0.6, eps_lam=1.2, cell_reach=2, lanes_per_bead=1),
    # eight lanes per bead over a compacted list of trimmed rows
    dict(ff=po.FF_CCP, n_l=40, lanes_per_bead=8),
    dict(ff=po.FF_ICG, n_l=0, lanes_per_bead=8),
    dict(ff=po.FF_LJ, n_l=30, eps_het=0.6, eps_lam=1.2, lanes_per_bead=8),
    dict(ff=po.FF_CCP, n_l=40, cell_reach=2, lanes_per_bead=8),
    dict(ff=po.FF_LJ, n_l=30, eps_het=0.6, eps_lam=1.2, cell_reach=2, lanes_per_bead=8),
    dict(ff=po.FF_CCP, n_l=20, opening_radius=4.0, bleb_radius=6.0, lanes_per_bead=8),
    # eight lanes for the pairs, bonds / wall / stage in a second one-lane-per-bead pass
    dict(ff=po.FF_LJ, n_l=30, eps_het=0.6, eps_lam=1.2, cell_reach=2, lanes_per_bead=8, integrate_pass=1),
    dict(ff=po.FF_CCP, n_l=40, lanes_per_bead=8, integrate_pass=1),
    # eight lanes at reach 2 without the type-ordered work list (one kernel
    # for every bead, het beads first within each workgroup), and without
    # that reordering either
    dict(ff=po.FF_LJ, n_l=30, eps_het=0.6, eps_lam=1.2, cell_reach=2, lanes_per_bead=8, type_order=0),
    dict(ff=po.FF_LJ, n_l=30, eps_het=0.6, eps_lam=1.2, cell_reach=2, lanes_per_bead=8, type_order=0,
         group_types=0),
    # the type-ordered work list with ccp's force field (1.25 sigma cells keep
    # its short cutoff within one cell)
    dict(ff=po.FF_CCP, n_l=40, cell_reach=2, lanes_per_bead=8, cell_size=1.25),
    dict(ff=po.FF_ICG, n_l=0, cell_reach=2, lanes_per_bead=8, cell_size=1.25),
    # eight lanes per bead, the pair phase of 2, 4 or 8 groups of eight beads
    # per wavefront before the per-bead work
    dict(ff=po.FF_LJ, n_l=30, eps_het=0.6, eps_lam=1.2, cell_reach=2, lanes_per_bead=8, bal_rounds=2),
    dict(ff=po.FF_LJ, n_l=30, eps_het=0.6, eps_lam=1.2, cell_reach=2, lanes_per_bead=8, bal_rounds=4),
    dict(ff=po.FF_LJ, n_l=30, eps_het=0.6, eps_lam=1.2, cell_reach=2, lanes_per_bead=8, bal_rounds=8),
    dict(ff=po.FF_CCP, n_l=40, lanes_per_bead=8, bal_rounds=4),
    dict(ff=po.FF_ICG, n_l=0, lanes_per_bead=8, bal_rounds=8),
    dict(ff=po.FF_CCP, n_l=20, opening_radius=4.0, bleb_radius=6.0, cell_reach=2, lanes_per_bead=8, bal_rounds=4),
    # pair sum split by cutoff over a fine grid and a het grid
    dict(ff=po.FF_CCP, n_l=40, fine_cell_size=1.2),
    dict(ff=po.FF_LJ, n_l=30, eps_het=0.6, eps_lam=1.2, fine_cell_size=1.25),
]


@pytest.mark.parametrize("case", FORCE_CASES)
def test_forces_match_oracle(case):
    case = dict(case)
    step(f"forces {case}")
    N = 2000
    R_n = radius_for(N, 0.3)
    r = relaxed(N, R_n, 7, chain_off=[0, 700, 1400, 2000])
    r *= 1.0 + 0.3 / R_n
    n_l = case.pop("n_l")
    lbs = lbs_sites(R_n, n_l, 3) if n_l else None
    ff = case.pop("ff")
    fl = pattern(N, 2)
    if ff == po.FF_LJ:
        rng = np.random.default_rng(6)
        fl = ((rng.random(N) < 0.3) * po.HET | (rng.random(N) < 0.15) * po.LAD).astype(np.uint8)
    sim_ff = po.FF_CCP if ff == po.FF_ICG else ff
    sim, osys, r64 = make_pair(N, R_n, fl, r, ff=sim_ff, lbs=lbs, chain_off=[0, 700, 1400, 2000], **case)
    osys.p.ff = ff
    got = sim.compute_forces(ff).astype(np.float64)
    want = osys.forces(r64, method=0)
    w32 = osys.forces(r64, method=0, prec=32)
    scale = np.abs(want).max()
    err = np.abs(got - want).max()
    err32 = np.abs(w32 - want).max()
    step(f"  max|f|={scale:.3g} gpu err={err:.3g} oracle fp32 err={err32:.3g}")
    # FP32 bar: no worse than 3x an FP32 evaluation of the same formulas
    assert err <= 3 * err32 + 1e-6 * scale
    sim.close()


@pytest.mark.parametrize("rounds", [1, 4])
def test_forces_partial_workgroup(rounds):
    # N = 1997 leaves the last workgroup of the eight-lane kernel part full
    # (16 beads per workgroup, 64 with four rounds) and its type grouping
    # with fewer het beads than slots
    N = 1997
    step(f"forces, partial last workgroup, N={N}, bal_rounds={rounds}")
    R_n = radius_for(N, 0.3)
    r = relaxed(N, R_n, 11, chain_off=[0, 998, N])
    r *= 1.0 + 0.3 / R_n
    rng = np.random.default_rng(12)
    fl = ((rng.random(N) < 0.3) * po.HET | (rng.random(N) < 0.15) * po.LAD).astype(np.uint8)
    sim, osys, r64 = make_pair(N, R_n, fl, r, ff=po.FF_LJ, lbs=lbs_sites(R_n, 30, 3), chain_off=[0, 998, N],
                               eps_het=0.6, eps_lam=1.2, cell_reach=2, lanes_per_bead=8, bal_rounds=rounds)
    got = sim.compute_forces(po.FF_LJ).astype(np.float64)
    want = osys.forces(r64, method=0)
    w32 = osys.forces(r64, method=0, prec=32)
    err, err32, scale = np.abs(got - want).max(), np.abs(w32 - want).max(), np.abs(want).max()
    step(f"  max|f|={scale:.3g} gpu err={err:.3g} oracle fp32 err={err32:.3g}")
    assert err <= 3 * err32 + 1e-6 * scale
    sim.close()


@pytest.mark.parametrize("ff,dt,tol,reach,lanes,rounds,typed", [
    (po.FF_CCP, 1 / 2048, 1e-3, 1, 3, 1, 1), (po.FF_LJ, 1e-3, 2e-3, 1, 3, 1, 1), (po.FF_LJ, 1e-3, 2e-3, 2, 5, 1, 1),
    (po.FF_CCP, 1 / 2048, 1e-3, 1, 8, 1, 1), (po.FF_LJ, 1e-3, 2e-3, 2, 8, 1, 1), (po.FF_LJ, 1e-3, 2e-3, 2, 8, 1, 0),
    (po.FF_LJ, 1e-3, 2e-3, 2, 8, 4, 1), (po.FF_CCP, 1 / 2048, 1e-3, 1, 8, 8, 1)])
def test_trajectory_1000_steps(ff, dt, tol, reach, lanes, rounds, typed):
    step(f"trajectory ff={ff} reach={reach} lanes={lanes} rounds={rounds} type_order={typed}")
    N = 1024
    R_n = radius_for(N, 0.2)
    r = relaxed(N, R_n, 2)
    if ff == po.FF_LJ:
        rng = np.random.default_rng(4)
        fl = ((rng.random(N) < 0.3) * po.HET | (rng.random(N) < 0.1) * po.LAD).astype(np.uint8)
        lbs = None
    else:
        fl = pattern(N, 1)
        lbs = lbs_sites(R_n, 30, 8)
    sim, osys, r64 = make_pair(N, R_n, fl, r, ff=ff, dt=dt, lbs=lbs, cell_reach=reach, lanes_per_bead=lanes,
                               type_order=typed, **({"bal_rounds": rounds} if rounds > 1 else {}))
    sim.set_time(0, 0.0, 0)
    sim.run_steps(1000)
    sim.synchronize()
    got = sim.positions.astype(np.float64)
    want = osys.run(r64, 0, 1000, prec=64)
    dev = np.linalg.norm(got - want, axis=1)
    moved = np.linalg.norm(want - r64, axis=1)
    step(f"  max dev={dev.max():.3g} p99={np.percentile(dev, 99):.3g} mean displacement={moved.mean():.3g}")
    assert dev.max() < tol
    assert moved.mean() > 50 * tol
    assert sim.time[2] == 1000
    sim.close()


@pytest.mark.parametrize("N,bs,rc,ff", [(1024, 1, 2.0, po.FF_CCP), (16384, 16, 1.5, po.FF_LJ),
                                        (131072, 128, 1.5, po.FF_LJ)])
def test_contacts_bit_exact(N, bs, rc, ff):
    step(f"contacts N={N}")
    R_n = radius_for(N, 0.3)
    r = walk(N, R_n, 11)
    sim, _, r64 = make_pair(N, R_n, np.zeros(N, np.uint8), r, ff=ff)
    assert sim.info.contact_bin_size == bs and sim.info.contact_cutoff == pytest.approx(rc)
    sim.contact_reset()
    sim.contact_sample()
    sim.contact_sample()
    m, ns = sim.contact_map()
    want = po.contacts_grid(r64.astype(np.float32), np.float32(rc), bs, sim.info.contact_bins)
    assert ns == 2
    assert np.array_equal(m, 2 * want)
    assert np.array_equal(m, m.T)
    sim.close()


def test_overlap_count_matches_brute_force():
    step("overlaps")
    N = 3000
    R_n = radius_for(N, 0.3)
    r = walk(N, R_n, 12)
    sim, _, r64 = make_pair(N, R_n, np.zeros(N, np.uint8), r)
    d = np.linalg.norm(r64[:, None, :] - r64[None, :, :], axis=2)
    iu = np.triu_indices(N, k=2)
    want = int((d[iu] < 0.838732).sum())
    assert sim.count_overlaps(0.838732) == want
    sim.close()


def test_step_grid_bit_exact():
    # The step's own cell list (1.25 sigma cells, lookup by multiplying with
    # 1/cell_size, binned in the force kernel's epilogue and ordered by the
    # order pass) against a stable sort by the same rule, on the benchmark's
    # workload: after the initial build and after steps.
    N = 131072
    R_n = radius_for(N, 0.3)
    step(f"step grid N={N}")
    sim = tfm_amd.Simulation({"number_of_particles": N, "nucleus_radius": repr(float(R_n)), "number_of_chains": 8,
                              "force_field": "lj", "epsilon_het": 0.6, "lad_fraction": 0.15, "seed": 3,
                              "timestep": 0.001, "lanes_per_bead": 8, "cell_reach": 2, "cell_size": 1.25})
    sim.generate_initial_condition()
    info = sim.info
    cps, csl = info.fine_cps, np.float32(info.fine_csl)
    icsl = np.float32(1.0) / csl
    for label, n in (("initial build", 0), ("after 7 steps", 7)):
        if n:
            sim.run_steps(n)
        cs, soi = sim.step_grid()
        r = sim.positions
        g = np.floor(r * icsl).astype(np.int64) + cps // 2
        assert g.min() >= 0 and g.max() < cps
        key = (g[:, 2] * cps + g[:, 1]) * cps + g[:, 0]
        order = np.argsort(key, kind="stable")
        start = np.concatenate([[0], np.cumsum(np.bincount(key, minlength=cps ** 3))]).astype(np.uint32)
        step(f"  {label}: {np.count_nonzero(np.diff(start))} occupied cells")
        assert np.array_equal(cs, start), label
        assert np.array_equal(soi, order.astype(np.uint32)), label
    sim.close()


@pytest.mark.parametrize("reach,het", [(2, True), (1, True), (2, False)])
def test_forces_crowded_lists(reach, het):
    # Het beads on a cubic lattice of spacing 1 sigma (volume fraction 0.52):
    # the inner beads' trimmed neighbour rows hold more slots than the
    # eight-lane kernel's 128-entry list, so each of their lanes scans its own
    # rows directly (the crowded path), while the surface beads use the list;
    # at reach 2 (1.25 sigma cells) and reach 1 (2.5 sigma). Other beads
    # (2^(1/6) sigma cutoff, two lanes each over reach 1 of the 1.25 sigma
    # cells, lists of 48) on a lattice of spacing 0.75 sigma overflow theirs.
    step(f"forces, crowded lists, reach {reach}, het {het}")
    k = 12
    N = k ** 3
    spacing, cap, cut = (1.0, 128, 2.5) if het else (0.75, 48, 2.0 ** (1.0 / 6.0))
    rng = np.random.default_rng(5)
    idx = np.arange(N)
    x, y, z = idx % k, (idx // k) % k, idx // (k * k)
    # a snake through the lattice, so consecutive beads are one spacing apart
    x = np.where(y % 2 == 1, k - 1 - x, x)
    y = np.where(z % 2 == 1, k - 1 - y, y)
    r = (np.stack([x, y, z], axis=1).astype(np.float64) - (k - 1) / 2) * spacing
    r += rng.uniform(-0.05, 0.05, (N, 3))
    R_n = 14.0
    fl = np.full(N, po.HET if het else 0, np.uint8)
    sim, osys, r64 = make_pair(N, R_n, fl, r, ff=po.FF_LJ, eps_het=0.6, eps_lam=1.0, cell_reach=reach,
                               lanes_per_bead=8)
    # list length of every bead on the host: the slots of the cells whose
    # nearest point lies within its cutoff (+ 1e-3)
    csl = np.float32(sim.info.fine_csl)
    cps = sim.info.fine_cps
    r32 = r64.astype(np.float32)
    g = np.floor(r32 * (np.float32(1.0) / csl)).astype(np.int64) + cps // 2
    occ = {}
    for c in map(tuple, g):
        occ[c] = occ.get(c, 0) + 1
    lengths = []
    for i in range(0, N, 97):
        tot = 0
        for (cx, cy, cz), m in occ.items():
            gap = 0.0
            for a, c in enumerate((cx, cy, cz)):
                e0 = (c - cps // 2) * csl
                d = max(e0 - r32[i, a], r32[i, a] - (e0 + csl), 0.0)
                gap += d * d
            if gap <= (cut + 1e-3) ** 2:
                tot += m
        lengths.append(tot)
    lengths = np.array(lengths)
    # the lattice's inner beads overflow the list, its surface beads do not
    crowded = float((lengths > cap).mean())
    assert crowded > 0.25 and lengths.min() <= cap, (crowded, lengths.min())
    got = sim.compute_forces(po.FF_LJ).astype(np.float64)
    want = osys.forces(r64, method=0)
    w32 = osys.forces(r64, method=0, prec=32)
    err, err32, scale = np.abs(got - want).max(), np.abs(w32 - want).max(), np.abs(want).max()
    step(f"  lists over {cap} entries: {crowded:.0%} of the sampled beads; max|f|={scale:.3g} gpu err={err:.3g} "
         f"oracle fp32 err={err32:.3g}")
    assert err <= 3 * err32 + 1e-6 * scale
    sim.close()


def test_single_chain_16384_steps():
    # The single-chain workload: one chain of 16384 beads at volume fraction
    # 0.2, WCA + het LJ (eps 0.5) + LAD wall, dt = 1e-3, contacts binned at 16
    # beads per bin. The engine builds the initial condition and the het/LAD
    # pattern from its seed; the GPU and the FP64 oracle then run the same
    # 1000 steps on the same Philox stream from that state, and the final
    # state's contact map and observables are checked against the oracle's.
    import os

    N = 16384
    R_n = radius_for(N, 0.2)
    step(f"single chain N={N}: initial condition")
    params = {"number_of_particles": N, "nucleus_radius": repr(float(R_n)), "number_of_chains": 1,
              "force_field": "lj", "epsilon_het": 0.5, "epsilon_lamina": 1.0, "het_fraction": 0.3,
              "het_block_mean": 20, "lad_fraction": 0.1, "timestep": 0.001, "seed": 21}
    gen = tfm_amd.Simulation(params)
    gen.generate_initial_condition()
    assert gen.count_overlaps(0.838732) == 0
    r0, fl = gen.positions, gen.flags
    gen.close()
    het = (fl & po.HET) != 0
    assert 0.2 < het.mean() < 0.4 and ((fl & po.LAD) != 0).any()
    sim, osys, r64 = make_pair(N, R_n, fl, r0, ff=po.FF_LJ, dt=1e-3, seed=21, eps_het=0.5, eps_lam=1.0)
    assert sim.info.contact_bin_size == 16 and sim.info.contact_bins == 1024
    sim.set_time(0, 0.0, 0)
    sim.run_steps(1000)
    sim.synchronize()
    got = sim.positions.astype(np.float64)
    step(f"single chain N={N}: oracle, 1000 steps")
    threads = max(1, min(16, len(os.sched_getaffinity(0))))
    want = osys.run(r64, 0, 1000, prec=64, threads=threads)
    dev = np.linalg.norm(got - want, axis=1)
    moved = np.linalg.norm(want - r64, axis=1)
    step(f"  max dev={dev.max():.3g} p99={np.percentile(dev, 99):.3g} mean displacement={moved.mean():.3g}")
    # per bead: FP32 against FP64 rounding, amplified by the chaotic dynamics
    # in the few beads of close encounters (measured: max 3.3e-3, p99 1.2e-4,
    # against a mean displacement of 1.0 sigma); a wrong pair set or noise
    # stream moves beads by 0.1 sigma or more
    assert dev.max() < 1e-2
    assert np.percentile(dev, 99) < 5e-4
    assert moved.mean() > 0.5
    assert np.linalg.norm(got, axis=1).max() < sim.info.R_n
    # the final state's contact map (16 beads per bin) bit-exact
    sim.contact_reset()
    sim.contact_sample()
    m, ns = sim.contact_map()
    assert ns == 1 and np.array_equal(m, po.contacts_grid(got.astype(np.float32), np.float32(1.5), 16, 1024))
    # radius of gyration and the 32-bin radial density of the final state
    sc, rcd = sim.observables(32)
    R_nf, R_of = float(np.float32(R_n)), float(sim.info.R_o)
    nod = float(np.float32(np.sqrt(R_nf * R_nf - R_of * R_of)))
    want_sc, want_rcd = po.observables(got, fl, R_nf, nod=nod, n_b=32)
    assert np.allclose(sc, want_sc, rtol=1e-9, atol=1e-9), (sc, want_sc)
    assert np.allclose(rcd, want_rcd, rtol=1e-9, atol=1e-10)
    sim.close()
