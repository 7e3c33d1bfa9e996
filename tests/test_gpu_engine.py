"""GPU tests of the engine around the step: initial condition, ccp's file
formats, checkpoint resume, the drop-in CLI, error flags and the full-size
(N = 131072) workload's cell list and contacts."""
import os
import struct
import subprocess

import numpy as np
import pytest
from scipy.spatial import cKDTree

import pyoracle as po
import tfm_amd
from gpu_common import step
from helpers import radius_for

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
MIS = 0.838732
RCO = float(np.float32(1.154701))
CHR_LIM = [0, 2284, 4599, 6687, 9338, 12508, 15193, 17477, 19792, 21880, 24531, 27701, 30386]


def min_nonbonded(r, chains=None):
    tree = cKDTree(r)
    pairs = tree.query_pairs(1.2, output_type="ndarray")
    pairs = pairs[np.abs(pairs[:, 0] - pairs[:, 1]) > 1]
    return np.linalg.norm(r[pairs[:, 0]] - r[pairs[:, 1]], axis=1).min() if len(pairs) else np.inf


def laf_sites(R_n, laf):
    return int(laf * 4 / (0.5 / (R_n - RCO)) ** 2)


def test_initial_condition_ccp():
    step("IC ccp N=1024")
    N = 1024
    R_n = radius_for(N, 0.2)
    n_l = laf_sites(R_n, 0.2)
    sim = tfm_amd.Simulation({"number_of_particles": N, "nucleus_radius": f"{R_n:.6f}", "number_of_lbs": n_l,
                              "seed": 11})
    sim.generate_initial_condition()
    r = sim.positions.astype(np.float64)
    assert sim.count_overlaps(MIS) == 0
    assert min_nonbonded(r) >= MIS * (1 - 1e-6)
    assert np.linalg.norm(r, axis=1).max() < R_n
    bl = np.linalg.norm(np.diff(r, axis=0), axis=1)
    assert 0.8 < bl.mean() < 1.2 and bl.max() < 1.8
    lbs = sim.lbs.astype(np.float64)
    assert np.allclose(np.linalg.norm(lbs, axis=1), R_n - RCO, atol=1e-3)
    d = np.linalg.norm(lbs[:, None] - lbs[None], axis=2) + 10 * np.eye(n_l)
    assert d.min() >= 2 * 0.5 - 1e-4
    # ccp's domains: first domain A, exponential lengths (mean 21.1) truncated to [1, 5 mdl]
    pt = sim.types
    assert pt[0] == 0
    edges = np.nonzero(np.diff(pt))[0]
    lens = np.diff(np.concatenate([[-1], edges, [N - 1]]))
    assert lens.max() <= 5 * 21.10139 + 1 and 10 < lens[:-1].mean() < 35
    i_f, t, s = sim.time
    assert (i_f, t, s) == (0, 0.0, 0)
    sim.close()


def test_initial_condition_default_sequence():
    step("IC ccp N=30386")
    seq = os.path.join(HERE, "golden", "sequence.txt")
    R_n = radius_for(30386, 0.2)
    sim = tfm_amd.Simulation({"nucleus_radius": f"{R_n:.6f}", "sequence_file": seq, "seed": 4,
                              "number_of_lbs": laf_sites(R_n, 0.2)})
    assert list(sim.chains) == CHR_LIM
    sim.generate_initial_condition()  # sets the types (set_particle_types, chrsim.cu:634-675)
    expect = np.array([0 if c == "A" else 1 for c in open(seq).read().strip()], np.int32)
    assert np.array_equal(sim.types, expect)
    assert sim.count_overlaps(MIS) == 0
    r = sim.positions.astype(np.float64)
    assert np.linalg.norm(r, axis=1).max() < R_n
    sim.close()


def test_het_lad_generator():
    step("het/LAD generator")
    N, L = 8 * 4096, 4096
    R_n = radius_for(N, 0.3)
    sim = tfm_amd.Simulation({"number_of_particles": N, "nucleus_radius": f"{R_n:.6f}", "number_of_chains": 8,
                              "force_field": "lj", "lad_fraction": 0.15, "het_fraction": 0.3, "seed": 3})
    sim.generate_initial_condition()
    fl = sim.flags
    het, lad = (fl & 1) > 0, (fl & 2) > 0
    assert abs(het.mean() - 0.3) < 0.05
    for c in range(8):
        seg = lad[c * L:(c + 1) * L]
        assert seg.sum() == round(0.15 * L)
        k = np.nonzero(seg)[0]
        assert np.minimum(k, L - 1 - k).mean() < L / 8
    sim.close()


def small_sim(seed=5, **kw):
    N = 1024
    R_n = radius_for(N, 0.2)
    p = {"number_of_particles": N, "nucleus_radius": f"{R_n:.6f}", "number_of_lbs": 30, "seed": seed,
         "frames_per_file": 2, "steps_per_frame": 50}
    p.update(kw)
    return tfm_amd.Simulation(p)


def test_checkpoint_resume_is_bit_exact(tmp_path):
    step("checkpoint resume")
    a = small_sim()
    a.generate_initial_condition()
    a.save_checkpoint(tmp_path / "ic.bin")
    a.run_steps(300)
    a.synchronize()
    whole = a.positions
    b = small_sim(seed=999)  # the trailer restores seed and step
    b.load_checkpoint(tmp_path / "ic.bin")
    b.run_steps(150)
    b.save_checkpoint(tmp_path / "mid.bin")
    c = small_sim(seed=777)
    c.load_checkpoint(tmp_path / "mid.bin")
    c.run_steps(150)
    c.synchronize()
    assert np.array_equal(c.positions, whole)
    assert c.time[2] == 300
    # a plain ccp checkpoint (no trailer) loads: i_f u32, t f32, types, r, lbs
    raw = open(tmp_path / "mid.bin", "rb").read()
    ccp_len = 8 + 4 * 1024 + 12 * 1024 + 12 * 30
    assert raw[ccp_len:ccp_len + 8] == b"CCPMI001"
    (tmp_path / "ccp.bin").write_bytes(raw[:ccp_len])
    d = small_sim(seed=5)
    d.load_checkpoint(tmp_path / "ccp.bin")
    assert np.array_equal(d.positions, b.positions)
    # a handle that already replayed captured graphs under another seed (more
    # than graph_steps = 64 steps) resumes exactly as a fresh one
    e = small_sim(seed=4242)
    e.generate_initial_condition()
    e.run_steps(200)
    e.synchronize()
    e.load_checkpoint(tmp_path / "ic.bin")
    e.run_steps(150)
    e.synchronize()
    f = small_sim(seed=31)
    f.load_checkpoint(tmp_path / "ic.bin")
    f.run_steps(150)
    f.synchronize()
    assert np.array_equal(e.positions, f.positions)
    assert np.array_equal(e.positions, b.positions)
    for s in (a, b, c, d, e, f):
        s.close()


def test_split_runs_and_host_changes_are_bit_exact():
    # run_steps launches no bead-terms pass when the last order pass left the
    # bonded terms and the normals of the current state (the normals drawn in
    # the tail of the second force launch); a host-side change of the
    # positions must discard them. Runs split into pieces, with a force
    # evaluation and synchronizations between them, must match one run, and a
    # handle whose positions were replaced must match a fresh one.
    step("split runs and host changes, LJ set, type-ordered kernel")
    shape = dict(force_field="lj", number_of_lbs=0, lanes_per_bead=8, cell_reach=2, cell_size=1.25)
    a = small_sim(**shape)
    a.generate_initial_condition()
    a.set_time(0, 0.0, 0)
    xyz, fl = a.positions, a.flags
    b = small_sim(**shape)
    b.flags = fl
    b.positions = xyz
    b.set_time(0, 0.0, 0)
    a.run_steps(200)
    a.synchronize()
    b.run_steps(37)
    b.synchronize()
    b.compute_forces()
    b.run_steps(100)
    b.synchronize()
    b.run_steps(63)
    b.synchronize()
    assert np.array_equal(a.positions, b.positions)
    # new positions on a handle that has run, against a fresh handle
    P = a.positions.copy()
    P[::7] += np.float32(0.01)
    T = a.time
    b.positions = P
    b.set_time(*T)
    c = small_sim(**shape)
    c.flags = fl
    c.positions = P
    c.set_time(*T)
    for s in (b, c):
        s.run_steps(100)
        s.synchronize()
    assert np.array_equal(b.positions, c.positions)
    assert not np.array_equal(b.positions, a.positions)
    for s in (a, b, c):
        s.close()


def test_gro_and_trr_formats(tmp_path):
    step("gro/trr formats")
    a = small_sim()
    a.generate_initial_condition()
    a.write_frame_txt(tmp_path / "ic.gro")
    a.write_lbs_txt(tmp_path / "lbs.gro")
    lines = open(tmp_path / "ic.gro").read().split("\n")
    assert lines[0] == "0 0" and lines[1] == " 1024"
    assert lines[2][:10] == "    1A    " or lines[2][:10] == "    1B    "
    assert len(lines[2]) == 44 and lines[1026] == "   0.00000   0.00000   0.00000"
    xyz = np.array([[float(l[20:28]), float(l[28:36]), float(l[36:44])] for l in lines[2:1026]])
    assert np.abs(xyz - a.positions).max() <= 5e-4 + 1e-6
    b = small_sim()
    b.read_frame_txt(tmp_path / "ic.gro")
    b.read_lbs_txt(tmp_path / "lbs.gro")
    assert np.array_equal(b.types, a.types)
    assert np.abs(b.positions - a.positions).max() <= 5e-4 + 1e-6
    assert np.abs(b.lbs - a.lbs).max() <= 5e-4 + 1e-6
    # one trajectory file: frames_per_file frames of steps_per_frame steps
    a.run_simulation(tmp_path / "traj.trr")
    raw = open(tmp_path / "traj.trr", "rb").read()
    fsz = 72 + 12 * 1024
    assert len(raw) == 2 * fsz
    for k in range(2):
        h = struct.unpack("<18I", raw[k * fsz:k * fsz + 72])
        assert h[0] == 1993 and h[1] == 1 and h[10] == 12 * 1024 and h[13] == 1024 and h[14] == k + 1
        t = struct.unpack("<f", struct.pack("<I", h[16]))[0]
        assert t == pytest.approx((k + 1) * 50 / 2048)
    last = np.frombuffer(raw[fsz + 72:], np.float32).reshape(-1, 3)
    assert np.array_equal(last, a.positions)
    assert a.time[:2] == (2, pytest.approx(100 / 2048))
    a.close()
    b.close()


def test_ccp_perform_cli(tmp_path):
    step("ccp-perform CLI")
    exe = os.path.join(ROOT, "tfm_amd", "bin", "ccp-perform")
    R_n = radius_for(1024, 0.2)
    (tmp_path / "adjustable-parameters.dat").write_text(
        f"number_of_particles 01024\nnucleus_radius {R_n:09.6f}\nnumber_of_lbs 00030\n"
        "frames_per_file 3\nsteps_per_frame 20\nseed 8\n")
    cwd = tmp_path / "run"
    cwd.mkdir()
    for args in ([str(tmp_path)], [str(tmp_path), "0"], [str(tmp_path)]):
        out = subprocess.run([exe, *args], cwd=cwd, capture_output=True, text=True, timeout=120)
        assert out.returncode == 0, out.stdout + out.stderr
    names = sorted(os.listdir(tmp_path))
    for f in ("initial-condition-000.gro", "initial-condition-001.gro", "lamina-binding-sites-000.gro",
              "trajectory-000-000.trr", "trajectory-000-001.trr", "trajectory-001-000.trr",
              "checkpoint-000.bin", "checkpoint-001.bin"):
        assert f in names
    assert os.path.getsize(tmp_path / "trajectory-000-001.trr") == 3 * (72 + 12 * 1024)
    assert os.listdir(cwd) == []  # log removed on success
    h = struct.unpack("<18I", open(tmp_path / "trajectory-000-001.trr", "rb").read(72))
    assert h[14] == 4  # frame index continues from the checkpoint


def test_kernel_shape_conflicts_are_rejected():
    # shapes the kernels cannot honour fail at creation with a parameter error
    step("shape conflicts")
    bad = [({"lanes_per_bead": 8, "fine_cell_size": 1.2}, "excludes fine_cell_size"),
           ({"lanes_per_bead": 4}, "lanes_per_bead out of range"),
           ({"cell_reach": 2, "cell_size": 2.0, "fine_cell_size": 1.2}, "needs cell_reach 1"),
           ({"bal_rounds": 3, "lanes_per_bead": 8}, "bal_rounds out of range"),
           ({"bal_rounds": 2, "lanes_per_bead": 8, "integrate_pass": 1}, "excludes integrate_pass")]
    for kw, msg in bad:
        with pytest.raises(tfm_amd.CcpError, match=msg):
            small_sim(**kw)


def test_bead_outside_grid_is_reported():
    step("error flag")
    a = small_sim()
    a.generate_initial_condition()
    r = a.positions
    r[17] = [500.0, 0.0, 0.0]
    a.positions = r
    with pytest.raises(tfm_amd.CcpError, match="bead left the cell grid"):
        a.run_steps(2)
        a.synchronize()
    a.close()


def test_full_size_workload_properties():
    # the benchmark workload (8 x 16384 beads, LJ set): IC, steps, then the
    # cell list and the contact counts of the final state against the oracle
    step("full size N=131072")
    import bench

    sim = tfm_amd.Simulation(dict(bench.PARAMS))
    sim.generate_initial_condition()
    assert sim.count_overlaps(MIS) == 0
    sim.run_steps(1000)
    sim.synchronize()
    m, ns = sim.contact_map()
    assert ns == 1 and np.array_equal(m, m.T) and m.sum() > 0
    r = sim.positions
    assert np.isfinite(r).all() and np.linalg.norm(r, axis=1).max() < sim.info.R_n
    got = sim.cell_list()
    want = po.cell_list(r, np.float32(sim.info.csl), sim.info.cps)
    for k in ("sci", "soi", "beg", "end"):
        assert np.array_equal(got[k], want[k]), k
    sim.contact_reset()
    sim.contact_sample()
    m, _ = sim.contact_map()
    assert np.array_equal(m, po.contacts_grid(r, np.float32(1.5), 128, 1024))
    sim.close()


def test_type_grouping_is_bit_exact():
    # the het-first reordering of each workgroup's beads moves beads between
    # wavefronts but keeps every bead's candidate list and summation order:
    # trajectories of the benchmark workload must agree to the bit (the
    # single kernel for every bead, without the type-ordered work list)
    step("type grouping bit-exact N=131072")
    import bench

    shape = {"lanes_per_bead": 8, "cell_reach": 2, "cell_size": 1.25, "type_order": 0}
    a = tfm_amd.Simulation({**bench.PARAMS, **shape, "group_types": 1})
    a.generate_initial_condition()
    xyz, fl = a.positions, a.flags
    b = tfm_amd.Simulation({**bench.PARAMS, **shape, "group_types": 0})
    b.flags = fl
    b.positions = xyz
    # the type-ordered work list: the het beads keep the same lists and sums,
    # the others run on two lanes over reach 1, another FP32 summation order
    c = tfm_amd.Simulation({**bench.PARAMS, **shape, "type_order": 1})
    c.flags = fl
    c.positions = xyz
    for s in (a, b, c):
        s.set_time(0, 0.0, 0)
        s.run_steps(300)
        s.synchronize()
    assert np.array_equal(a.positions, b.positions)
    # one force evaluation of the same state by both kernels: the het beads'
    # sums are the same bits; the others differ by the FP32 rounding of
    # another order of the same terms (a few tens of terms, each at most
    # max|f|, rounded at 2^-24: a bound of 1e-5 max|f| holds with a margin)
    state = c.positions
    a.positions = state
    fa = a.compute_forces(po.FF_LJ)
    fc = c.compute_forces(po.FF_LJ)
    het = (fl & po.HET) != 0
    assert np.array_equal(fa[het], fc[het])
    scale = float(np.abs(fa).max())
    ferr = float(np.abs(fa[~het].astype(np.float64) - fc[~het]).max())
    step(f"  forces of one state: het beads bit-exact, others max deviation {ferr:.3g} of max|f| {scale:.3g}")
    assert ferr <= 1e-5 * scale
    # the trajectories themselves drift apart as the rounding differences
    # grow in the chaotic dynamics (~1e-3 after 300 steps against steps of
    # ~0.05 sigma); a wrong pair set would move beads by ~0.1 sigma or more
    dev = np.abs(state.astype(np.float64) - b.positions).max()
    step(f"  work list vs one kernel after 300 steps: max deviation {dev:.3g}")
    assert dev < 1e-2
    for s in (a, b, c):
        s.close()


def test_run_frames_callback():
    # frames go to the callback in memory: indices and times advance as in
    # run_simulation, the positions are the frame's, and a true return stops
    step("run_frames")
    a = small_sim()
    a.generate_initial_condition()
    a.set_time(0, 0.0, 0)
    seen = []

    def take(i_f, t, xyz):
        seen.append((i_f, t, xyz))
        return len(seen) == 2

    a.run_frames(5, take)
    assert [s[0] for s in seen] == [1, 2]
    assert abs(seen[1][1] - 2 * 50 * a.info.dt) < 1e-6
    assert np.array_equal(seen[1][2], a.positions)
    assert a.time[2] == 100
    b = small_sim()
    b.generate_initial_condition()
    b.set_time(0, 0.0, 0)
    b.run_steps(50)
    b.synchronize()
    assert np.array_equal(b.positions, seen[0][2])
    a.close()
    b.close()
