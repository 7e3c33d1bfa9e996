"""GPU parity of the analysis kernels (ccp's chrana.cu per-frame work) with
the oracle: frame observables and radial density (FP64 sums, tolerance in
the test), d(s)/P(s) and ccp's packed 4-bead contact map (ccp's FP32
arithmetic, bit-exact), including the two-copy averages at N = 30386."""
import os

import numpy as np
import pytest

import pyoracle as po
import tfm_amd
from gpu_common import step
from helpers import radius_for, walk

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
CHR_LIM = [0, 2284, 4599, 6687, 9338, 12508, 15193, 17477, 19792, 21880, 24531, 27701, 30386]


def sim_with(N, R_n, r, flags, **kw):
    p = {"number_of_particles": N, "nucleus_radius": repr(float(R_n)), "number_of_lbs": 0, "seed": 1}
    p.update(kw)
    sim = tfm_amd.Simulation(p)
    sim.flags = flags
    sim.positions = np.ascontiguousarray(r, np.float32)
    return sim


@pytest.mark.parametrize("N,n_b", [(1024, 64), (16384, 64), (4096, 7)])
def test_observables_match_oracle(N, n_b):
    step(f"observables N={N} n_b={n_b}")
    R_n = radius_for(N, 0.2)
    r = walk(N, R_n, seed=N + n_b)
    flags = (np.random.default_rng(N).random(N) < 0.4).astype(np.uint8)
    sim = sim_with(N, R_n, r, flags)
    r32 = sim.positions.astype(np.float64)
    sc, rcd = sim.observables(n_b)
    R_nf, R_of = float(np.float32(R_n)), float(sim.info.R_o)
    nod = float(np.float32(np.sqrt(R_nf * R_nf - R_of * R_of)))
    want, want_rcd = po.observables(r32, flags, R_nf, nod=nod, n_b=n_b)
    # FP64 sums in another order: relative 1e-9 (rg2 by <r^2> - <r>^2 loses
    # a few digits more to cancellation than the oracle's two-pass form)
    assert np.allclose(sc, want, rtol=1e-9, atol=1e-9), (sc, want)
    assert np.allclose(rcd, want_rcd, rtol=1e-9, atol=1e-10)
    # the radial density integrates to the bead volume
    shell = 4 / 3 * np.pi * float(np.float32(R_n)) ** 3 / n_b
    vb = 4 / 3 * np.pi * (float(np.float32(1.154701)) / 2) ** 3
    assert rcd[2].sum() * shell == pytest.approx(N * vb, rel=1e-9)
    sim.close()


def test_distance_contact_bit_exact():
    step("d(s), P(s) N=2048")
    N = 2048
    R_n = radius_for(N, 0.2)
    sim = sim_with(N, R_n, walk(N, R_n, seed=3, chain_off=[0, 700, N]), np.zeros(N, np.uint8),
                   chain_lengths="700,1348")
    sd, cp = sim.distance_contact(0)
    wsd, wcp = po.sd_cp(sim.positions, 0)
    assert sd.shape == (1024,)
    assert np.array_equal(sd, wsd) and np.array_equal(cp, wcp)
    with pytest.raises(tfm_amd.CcpError, match="no such chromosome"):
        sim.distance_contact(1)
    sim.close()


def test_distance_contact_two_copies():
    step("d(s), P(s) N=30386")
    N = 30386
    R_n = radius_for(N, 0.2)
    sim = tfm_amd.Simulation({"nucleus_radius": repr(float(R_n)), "sequence_file": os.path.join(HERE, "golden",
                              "sequence.txt"), "seed": 4, "number_of_lbs": 0})
    sim.positions = walk(N, R_n, seed=21, chain_off=CHR_LIM).astype(np.float32)
    for c in range(6):
        sd, cp = sim.distance_contact(c)
        wsd, wcp = po.sd_cp(sim.positions, c)
        assert sd.shape == ((CHR_LIM[c + 1] - CHR_LIM[c]) // 2,)
        assert np.array_equal(sd, wsd) and np.array_equal(cp, wcp), c
        assert cp[0] == 1000.0
    sim.close()


def test_frame_contact_map_bit_exact():
    step("ccp frame contact map N=1024")
    N = 1024
    R_n = radius_for(N, 0.3)
    sim = sim_with(N, R_n, walk(N, R_n, seed=8), np.zeros(N, np.uint8))
    r32 = sim.positions
    got = sim.frame_contact_map()
    want = po.ccp_contact_map(r32)
    assert got.shape == want.shape == (256 * 257 // 2,)
    assert np.array_equal(got, want)
    diag = got[np.cumsum(np.arange(1, 257)) - 1]  # a pixel of a bead block with itself
    assert (diag >= 10 * 1000.0 / 16).all()  # 4 self pairs and 6 bonded pairs, both orders
    sim.close()


def test_frame_contact_map_two_copies():
    step("ccp frame contact map N=30386")
    seq = os.path.join(HERE, "golden", "sequence.txt")
    N = 30386
    R_n = radius_for(N, 0.2)
    sim = tfm_amd.Simulation({"nucleus_radius": repr(float(R_n)), "sequence_file": seq, "seed": 4,
                              "number_of_lbs": 0})
    sim.positions = walk(N, R_n, seed=12, chain_off=CHR_LIM).astype(np.float32)
    got = sim.frame_contact_map()
    want = po.ccp_contact_map(sim.positions)
    cms = (N // 2) // 4
    assert got.shape == (cms * (cms + 1) // 2,)
    assert np.array_equal(got, want)
    sim.close()
