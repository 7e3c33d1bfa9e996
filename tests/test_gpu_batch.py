"""GPU tests of the batch engine: every system of a batch evolves bit for bit
as the same system stepped alone (same cells, pair order and noise
counters), its sampled contact map equals the single run's, and its radial
density conserves the bead volume."""
import numpy as np
import pytest

import pyoracle as po
import tfm_amd
from gpu_common import step

pytestmark = pytest.mark.gpu

SYSTEMS = [
    dict(n_beads=600, volume_fraction=0.15, eps_het=0.3, eps_lam=0.5, het_fraction=0.3, lad_fraction=0.1, seed=4001),
    dict(n_beads=1000, volume_fraction=0.3, eps_het=0.9, eps_lam=1.4, het_fraction=0.3, lad_fraction=0.1, seed=4002),
    dict(n_beads=2048, volume_fraction=0.2, eps_het=0.6, eps_lam=1.0, het_fraction=0.3, lad_fraction=0.1, seed=4003),
    dict(n_beads=777, volume_fraction=0.35, eps_het=0.2, eps_lam=0.2, het_fraction=0.3, lad_fraction=0.1, seed=4004),
    dict(n_beads=1500, volume_fraction=0.1, eps_het=1.0, eps_lam=1.5, het_fraction=0.3, lad_fraction=0.1, seed=4005,
         n_chains=2),
]


def single_from(batch, s, xyz, fl):
    sim = tfm_amd.Simulation(batch.system_params(s))
    o0, o1 = batch.offsets[s], batch.offsets[s + 1]
    sim.flags = fl[o0:o1]
    sim.positions = xyz[o0:o1]
    sim.set_time(0, 0.0, 0)
    return sim


@pytest.mark.parametrize("shape", [dict(), dict(lanes_per_bead=8, cell_reach=2, cell_size=1.25)])
def test_batch_systems_match_single_runs(shape):
    step(f"batch of {len(SYSTEMS)} systems, shape {shape}")
    common = dict(timestep=0.001, contact_interval=100, contact_cutoff=1.5, graph_steps=50, **shape)
    b = tfm_amd.Batch(SYSTEMS, common)
    assert b.N == sum(d["n_beads"] for d in SYSTEMS)
    b.generate_initial_conditions()
    xyz0, fl0 = b.state()
    b.run_steps(300)
    b.synchronize()
    xyz, _ = b.state()
    maps, ns = b.contact_maps()
    assert ns == 3 and b.step == 300
    for s in range(len(SYSTEMS)):
        sim = single_from(b, s, xyz0, fl0)
        sim.run_steps(300)
        sim.synchronize()
        o0, o1 = b.offsets[s], b.offsets[s + 1]
        assert np.array_equal(sim.positions, xyz[o0:o1]), s
        m, n1 = sim.contact_map()
        k = m.shape[0]  # the system alone keeps ceil(n / bin size) bins; the batch pads to 128
        assert n1 == 3 and np.array_equal(m, maps[s][:k, :k]), s
        assert maps[s][k:].sum() == 0 and maps[s][:, k:].sum() == 0
        sim.close()
    assert np.isfinite(xyz).all()
    b.close()


def test_batch_initial_conditions_and_radial_density():
    step("batch initial conditions, radial density")
    b = tfm_amd.Batch(SYSTEMS, dict(timestep=0.001, contact_interval=50, graph_steps=25))
    b.generate_initial_conditions()
    xyz, fl = b.state()
    R = []
    for s, d in enumerate(SYSTEMS):
        o0, o1 = b.offsets[s], b.offsets[s + 1]
        # the same initial condition as the system alone
        sim = tfm_amd.Simulation(b.system_params(s))
        sim.generate_initial_condition()
        assert np.array_equal(sim.positions, xyz[o0:o1]) and np.array_equal(sim.flags, fl[o0:o1])
        R.append((float(sim.info.R_n), float(sim.info.R_o)))
        assert np.linalg.norm(xyz[o0:o1].astype(np.float64), axis=1).max() < R[-1][0]
        sim.close()
    # four samples, each checked against the oracle's lens-volume density
    want = np.zeros((len(SYSTEMS), 32))
    for _ in range(4):
        b.run_steps(50)
        xyz, fl = b.state()
        for s in range(len(SYSTEMS)):
            o0, o1 = b.offsets[s], b.offsets[s + 1]
            R_n, R_o = R[s]
            nod = float(np.float32(np.sqrt(R_n * R_n - R_o * R_o)))
            _, rcd = po.observables(xyz[o0:o1].astype(np.float64), fl[o0:o1], R_n, nod=nod, n_b=32)
            want[s] += rcd[2]
    rd = b.radial_density()
    assert rd.shape == (len(SYSTEMS), 32)
    np.testing.assert_allclose(rd, want / 4, rtol=1e-9, atol=1e-12)
    b.close()


def test_batch_rejects_bad_systems():
    with pytest.raises(tfm_amd.CcpError, match="out of range"):
        tfm_amd.Batch([dict(n_beads=100, volume_fraction=0.9, seed=1)])


def test_replicas_share_the_pattern_not_the_walk():
    step("batch replicas")
    base = dict(n_beads=900, volume_fraction=0.2, eps_het=0.5, eps_lam=1.0, seed=77)
    b = tfm_amd.Batch([dict(base, replica=0), dict(base, replica=1), dict(base, replica=2)],
                      dict(timestep=0.001, contact_interval=0))
    b.generate_initial_conditions()
    xyz, fl = b.state()
    o = b.offsets
    assert np.array_equal(fl[o[0]:o[1]], fl[o[1]:o[2]]) and np.array_equal(fl[o[0]:o[1]], fl[o[2]:o[3]])
    assert not np.array_equal(xyz[o[0]:o[1]], xyz[o[1]:o[2]])
    # replica 0 is the plain simulation; the host-only pattern matches its types
    sim = tfm_amd.Simulation(b.system_params(0))
    sim.generate_initial_condition()
    assert np.array_equal(sim.positions, xyz[o[0]:o[1]])
    assert np.array_equal(tfm_amd.ccp.sequence_pattern(b.system_params(0), 900), sim.flags)
    sim.close()
    # replica 1 alone evolves as in the batch (its own noise key)
    b.run_steps(200)
    xyz2, _ = b.state()
    one = tfm_amd.Simulation(b.system_params(1))
    one.flags = fl[o[1]:o[2]]
    one.positions = xyz[o[1]:o[2]]
    one.set_time(0, 0.0, 0)
    one.run_steps(200)
    assert np.array_equal(one.positions, xyz2[o[1]:o[2]])
    one.close()
    b.close()


def test_batch_autotuner_picks_a_measured_shape():
    step("batch autotuner")
    from tfm_amd import autotune

    systems = [dict(n_beads=int(n), volume_fraction=0.25, eps_het=0.6, eps_lam=1.0, seed=900 + k)
               for k, n in enumerate(np.geomspace(512, 4096, 12).astype(int))]
    best, table = autotune.tune_batch(systems, dict(timestep=0.001), steps=50, warmup=10)
    assert len(table) == 6 and all(np.isfinite(t) and t > 0 for _, t in table)
    assert best == min(table, key=lambda e: e[1])[0]


def test_ensemble_runner_small():
    # the ensemble runner end to end on a few uneven systems: finite state,
    # symmetric maps, one sample per interval, densities that integrate
    step("ensemble runner")
    from tfm_amd import ensemble

    out = ensemble.run(["--systems", "6", "--min-beads", "256", "--max-beads", "1024", "--steps", "2000",
                        "--interval", "500", "--no-tune"])
    assert out["systems"] == 6 and 6 * 256 <= out["beads"] <= 6 * 1024
    assert out["finite_state"] and out["maps_symmetric"]
    assert out["contact_samples"] == 4
    assert out["contacts_per_system_mean"] > 0 and out["bead_steps_per_s"] > 0
    assert len(out["radial_density_mean"]) == 32 and min(out["radial_density_mean"]) >= 0


ENSEMBLE_COMMON = dict(timestep=0.001, contact_interval=500, contact_cutoff=1.5, contact_bins=128, radial_bins=32,
                       graph_steps=100)


def test_ensemble_full_size():
    # the ensemble workload at full size: 256 single-chain systems of
    # 512-16384 beads (log-uniform, seeds 4000 + i, 1.1M beads), vf, eps_het
    # and eps_lamina drawn per system, in the shape the batch autotuner picks
    # at this size (3 lanes per bead over 2.5 sigma cells), 2000 graph-replayed
    # steps sampled every 500
    from tfm_amd.ensemble import draw_systems

    systems = draw_systems(256, 512, 16384, 4000)
    step(f"ensemble full size: {len(systems)} systems, {sum(s['n_beads'] for s in systems)} beads")
    b = tfm_amd.Batch(systems, dict(ENSEMBLE_COMMON, lanes_per_bead=3, cell_reach=1, cell_size=2.5))
    b.generate_initial_conditions()
    xyz0, fl0 = b.state()
    lengths = np.array([s["n_beads"] for s in systems])
    check = [int(np.argmin(lengths)), int(np.argmax(lengths)), 17, 200]
    geo = {}
    for s in check:
        sim = tfm_amd.Simulation(b.system_params(s))
        geo[s] = (float(sim.info.R_n), float(sim.info.R_o))
        sim.close()
    want_rd = {s: np.zeros(32) for s in check}
    for k in range(4):
        b.run_steps(500)
        xyz, fl = b.state()
        for s in check:
            o0, o1 = b.offsets[s], b.offsets[s + 1]
            R_n, R_o = geo[s]
            nod = float(np.float32(np.sqrt(R_n * R_n - R_o * R_o)))
            want_rd[s] += po.observables(xyz[o0:o1].astype(np.float64), fl[o0:o1], R_n, nod=nod, n_b=32)[1][2]
        step(f"  {500 * (k + 1)} steps")
    b.synchronize()
    maps, ns = b.contact_maps()
    rd = b.radial_density()
    assert ns == 4 and b.step == 2000
    assert np.isfinite(xyz).all()
    assert maps.shape == (256, 128, 128)
    assert all(np.array_equal(m, m.T) for m in maps)
    assert (maps.reshape(256, -1).sum(axis=1) > 0).all()
    # each system's mean radial density integrates to its beads' volume
    # (less the small part of wall beads' spheres outside the nucleus); the
    # shell volume here comes from the radius recomputed in double, which can
    # differ from the library's FP32 radius in its last place (3e-7 in R^3)
    vb = 4.0 / 3.0 * np.pi * (float(np.float32(1.154701)) / 2) ** 3
    for s in range(256):
        sim_R = float(np.float32(float(f"{0.5773505 + 0.5773505 * (lengths[s] / systems[s]['volume_fraction']) ** (1 / 3):.6f}")))
        shell = 4.0 / 3.0 * np.pi * sim_R ** 3 / 32
        tot = rd[s].sum() * shell / (lengths[s] * vb)
        assert 0.95 < tot <= 1.0 + 1e-6, (s, tot)
    for s in check:
        np.testing.assert_allclose(rd[s], want_rd[s] / 4, rtol=1e-9, atol=1e-12)
    # four systems (the shortest, the longest, two others) alone, bit for bit
    for s in check:
        sim = single_from(b, s, xyz0, fl0)
        sim.run_steps(2000)
        sim.synchronize()
        o0, o1 = b.offsets[s], b.offsets[s + 1]
        assert np.array_equal(sim.positions, xyz[o0:o1]), s
        m, n1 = sim.contact_map()
        k = m.shape[0]
        assert n1 == 4 and np.array_equal(m, maps[s][:k, :k]), s
        sim.close()
    b.close()


@pytest.mark.skip(reason="eight lanes over 1.25 sigma cells on the 1.1M-bead ensemble faulted the GPU once (with "
                         "initial conditions generated on 16 host threads) and the cause is not found; the library "
                         "refuses this shape above 524288 beads until it is")
def test_ensemble_full_size_half_cells():
    from tfm_amd.ensemble import draw_systems

    systems = draw_systems(256, 512, 16384, 4000)
    b = tfm_amd.Batch(systems, dict(ENSEMBLE_COMMON, lanes_per_bead=8, cell_reach=2, cell_size=1.25))
    b.generate_initial_conditions()
    b.run_steps(10000)
    xyz, _ = b.state()
    assert np.isfinite(xyz).all()
    b.close()
