"""GPU tests of ccp's analysis path: ccp-perform writes two simulations,
ccp-analyze reads them back, and every number of analysis-XXX.dat,
analysis-fin.dat and contact-map.bin is recomputed from the same frames with
the oracle (observables in FP64, d(s)/P(s) and the contact map in ccp's FP32
arithmetic, statistics by the oracle's restatement of sostat.cu)."""
import os
import subprocess

import numpy as np
import pytest

import pyoracle as po
import sostat
import tfm_amd
from gpu_common import step
from helpers import radius_for

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "tfm_amd", "bin")
N, FPF, SPF = 1024, 6, 40


def blocks(path):
    # ccp's viewer splits the analysis files on blank-line pairs (viewana.py:50-51)
    text = open(path).read()
    return [b for b in text.split("\n\n\n") if b.strip()]


def rows(block):
    return [l.split() for l in block.split("\n")[1:] if l.strip()]


def frames_of(params, d, i_s):
    """Positions (as the library reads them), types, and R_n of every frame of simulation i_s."""
    sim = tfm_amd.Simulation(params)
    sim.read_frame_txt(os.path.join(d, f"initial-condition-{i_s:03d}.gro"))
    out = [sim.positions.copy()]
    flags = sim.flags.copy()
    k = 0
    while os.path.exists(os.path.join(d, f"trajectory-{i_s:03d}-{k:03d}.trr")):
        raw = open(os.path.join(d, f"trajectory-{i_s:03d}-{k:03d}.trr"), "rb").read()
        fsz = 72 + 12 * N
        for f in range(FPF):
            out.append(np.frombuffer(raw[f * fsz + 72:(f + 1) * fsz], np.float32).reshape(N, 3).copy())
        k += 1
    R_n, R_o = float(sim.info.R_n), float(sim.info.R_o)
    sim.close()
    return out, flags, R_n, R_o


def fmt(x):
    return "%12.6f" % x


def test_ccp_analyze_matches_oracle(tmp_path):
    step("ccp-perform x3 + ccp-analyze N=1024")
    d = tmp_path / "sim"
    d.mkdir()
    R_n = radius_for(N, 0.2)
    text = (f"number_of_particles {N}\nnucleus_radius {R_n:.6f}\nnumber_of_lbs 30\n"
            f"frames_per_file {FPF}\nsteps_per_frame {SPF}\nseed 17\n")
    (d / "adjustable-parameters.dat").write_text(text)
    cwd = tmp_path / "run"
    cwd.mkdir()
    for args in ([str(d)], [str(d)], [str(d), "0"]):
        out = subprocess.run([os.path.join(BIN, "ccp-perform"), *args], cwd=cwd, capture_output=True, text=True,
                             timeout=120)
        assert out.returncode == 0, out.stdout + out.stderr
    out = subprocess.run([os.path.join(BIN, "ccp-analyze"), str(d)], cwd=cwd, capture_output=True, text=True,
                         timeout=120)
    assert out.returncode == 0, out.stdout + out.stderr
    assert "added trajectory file 000-001" in out.stdout
    assert os.listdir(cwd) == []

    params = {k: v for k, v in (l.split() for l in text.strip().split("\n"))}
    per_sim = []
    S1_cm = None
    sd_avgs, cp_avgs = [], []
    for i_s in range(2):
        frames, flags, R_nf, R_of = frames_of(params, str(d), i_s)
        assert len(frames) == 1 + FPF * (2 if i_s == 0 else 1)
        nod = float(np.float32(np.sqrt(R_nf * R_nf - R_of * R_of)))
        series = [[] for _ in range(4 + 3 * 64)]
        sd_sum = np.zeros(N // 2)
        cp_sum = np.zeros(N // 2)
        cm_sum = None
        for r in frames:
            sc, rcd = po.observables(r.astype(np.float64), flags, R_nf, nod=nod, n_b=64)
            for k, v in enumerate(list(sc) + list(rcd.ravel())):
                series[k].append(np.float32(v))
            sd, cp = po.sd_cp(r)
            sd_sum += sd.astype(np.float64)
            cp_sum += cp.astype(np.float64)
            cm = po.ccp_contact_map(r).astype(np.float64)
            cm_sum = cm if cm_sum is None else cm_sum + cm
        n_dp = len(frames)
        stats = [sostat.series(np.array(s, np.float32)) for s in series]
        per_sim.append(stats)
        sd_avg, cp_avg, cm_avg = sd_sum / n_dp, cp_sum / n_dp, cm_sum / n_dp
        sd_avgs.append(sd_avg)
        cp_avgs.append(cp_avg)
        S1_cm = cm_avg if S1_cm is None else S1_cm + cm_avg

        b = blocks(d / f"analysis-{i_s:03d}.dat")
        assert len(b) == 1 + 3 + 6 + 6  # scalars, 3 rcd, sd of 6 chromosomes (5 empty), cp of 6
        head = open(d / f"analysis-{i_s:03d}.dat").readline()
        assert head == "         avg   sqrt(var)         sem   f_n_b     i_t ter\n"
        for k, row in enumerate(rows(b[0])):
            a, sv, se, fnb, it, ter = stats[k]
            assert float(row[0]) == pytest.approx(a, abs=2e-6 * max(1, abs(a)))
            assert float(row[2]) == pytest.approx(se, rel=1e-3, abs=2e-6)
            assert (int(row[4]), row[5]) == (it, "yes" if ter else "no")
        for t in range(3):
            rr = rows(b[1 + t])
            assert rr[0] == ["0.000000"] * 4 and len(rr) == 65
            for k in range(64):
                assert float(rr[k + 1][0]) == pytest.approx(R_nf * ((k + 1) / 64) ** (1 / 3), abs=1e-6)
                assert float(rr[k + 1][1]) == pytest.approx(stats[4 + 64 * t + k][0], abs=2e-6)
        # d(s), P(s): ccp's FP32 per-frame values summed in FP64 -> the same text
        sd_rows = rows(b[4])
        assert len(sd_rows) == N // 2 and all(len(rows(x)) == 0 for x in b[5:10])
        assert [r[1] for r in sd_rows] == [fmt(x).strip() for x in sd_avg]
        assert [r[1] for r in rows(b[10])] == [fmt(x).strip() for x in cp_avg]

    # combined statistics
    fin = blocks(d / "analysis-fin.dat")
    for k, row in enumerate(rows(fin[0])):
        a = sostat.weighted([s[k][0] for s in per_sim], [s[k][2] for s in per_sim])
        assert float(row[0]) == pytest.approx(a[0], abs=2e-6 * max(1, abs(a[0])))
    m1 = sd_avgs[0] + sd_avgs[1]
    m2 = sd_avgs[0] * sd_avgs[0] + sd_avgs[1] * sd_avgs[1]
    avg, var = m1 / 2, (m2 / 2 - (m1 / 2) * (m1 / 2)) / 0.5
    got = rows(fin[4])
    assert [r[1] for r in got] == [fmt(x).strip() for x in avg]
    assert [r[2] for r in got] == [fmt(x).strip() for x in np.sqrt(var)]
    cm = np.fromfile(d / "contact-map.bin", np.float64)
    assert cm.shape == (256 * 257 // 2,)
    np.testing.assert_array_equal(cm, S1_cm / 2)


def test_in_situ_analysis_of_a_live_simulation(tmp_path):
    step("in-situ analysis N=1024")
    R_n = radius_for(N, 0.2)
    sim = tfm_amd.Simulation({"number_of_particles": N, "nucleus_radius": f"{R_n:.6f}", "number_of_lbs": 20,
                              "seed": 23})
    sim.generate_initial_condition()
    ana = tfm_amd.Analysis(sim)
    sd_sum = np.zeros(N // 2)
    cm_sum = 0
    for _ in range(5):
        ana.add_frame()
        r = sim.positions
        sd_sum += po.sd_cp(r)[0].astype(np.float64)
        cm_sum = cm_sum + po.ccp_contact_map(r).astype(np.float64)
        sim.run_steps(100)
    ana.end_simulation(tmp_path / "a.dat")
    ana.finish(tmp_path / "fin.dat", tmp_path / "cm.bin")
    b = blocks(tmp_path / "a.dat")
    assert [r[1] for r in rows(b[4])] == [fmt(x).strip() for x in sd_sum / 5]
    np.testing.assert_array_equal(np.fromfile(tmp_path / "cm.bin", np.float64), cm_sum / 5 / 1)
    with pytest.raises(tfm_amd.CcpError, match="no frames added"):
        ana.end_simulation(tmp_path / "b.dat")
    ana.close()
    sim.close()
