"""Pins the oracle's Philox4x32-10 and its normal triples."""
import os

import numpy as np
import pytest
from scipy import special, stats

import pyoracle as po

GOLD = os.path.join(os.path.dirname(__file__), "golden")


def _vectors(name):
    out = []
    for line in open(os.path.join(GOLD, name)):
        if line.startswith("#") or not line.strip():
            continue
        v = [int(x, 16) for x in line.split()]
        out.append((v[:4], v[4:6], v[6:]))
    return out


@pytest.mark.parametrize("name", ["philox_random123_kat.txt", "philox_rocrand.txt"])
def test_philox_known_answers(name):
    vecs = _vectors(name)
    assert len(vecs) >= 3
    for ctr, key, out in vecs:
        assert list(po.philox(ctr, key)) == out


def test_normals_moments_and_independence():
    z, a = po.noise_batch(seed=11, step=7, n=400_000)
    assert np.all(np.abs(z) <= 5.0)
    for k in range(3):
        assert abs(z[:, k].mean()) < 5 * np.sqrt(1 / len(z))
        assert abs(z[:, k].var() - 1.0) < 5 * np.sqrt(2 / len(z))
        assert stats.kstest(z[:20000, k], "norm").pvalue > 1e-4
    c = np.corrcoef(z.T)
    assert np.all(np.abs(c - np.eye(3)) < 5 / np.sqrt(len(z)))


def test_noise_counter_layout():
    # attempt 0 of (seed, bead, step) is philox({bead, step lo, step hi, 0}, {seed lo, seed hi})
    seed, bead, step = 0x0123456789ABCDEF, 77, (5 << 32) + 9
    z, a = po.noise(seed, bead, step)
    assert a == 0
    c = po.philox([bead, step & 0xFFFFFFFF, step >> 32, 0], [seed & 0xFFFFFFFF, seed >> 32])
    u = ((c.astype(np.float64) // 256) + 1) / 2**24
    rad = np.sqrt(-2 * np.log(u[0]))
    assert np.allclose(z, [rad * np.cos(2 * np.pi * u[1]), rad * np.sin(2 * np.pi * u[1]),
                           np.sqrt(-2 * np.log(u[2])) * np.cos(2 * np.pi * u[3])], rtol=0, atol=1e-12)


def test_redraw_rule_and_rate():
    # ccp redraws the whole triple while any |z| > 5 (chrsim.cu:350-356).
    n = 2_000_000
    z, a = po.noise_batch(seed=3, step=1, n=n)
    redrawn = np.nonzero(a > 0)[0]
    # per-triple probability of a redraw: 1 - (1 - erfc(5/sqrt 2))^(3/2 for z0,z1 sharing a radius ...)
    p1 = special.erfc(5 / np.sqrt(2))
    p_triple = 1 - (1 - p1) ** 3  # approximately; z0,z1 share a radius but are uncorrelated
    lam = n * p_triple
    assert stats.poisson.cdf(len(redrawn), lam) > 1e-4 and stats.poisson.sf(len(redrawn), lam) > 1e-4
    # every redrawn bead had |z| > 5 on its first attempt
    for b in redrawn[:20]:
        c = po.philox([int(b), 1, 0, 0], [3, 0])
        u = ((c.astype(np.float64) // 256) + 1) / 2**24
        r01, r23 = np.sqrt(-2 * np.log(u[0])), np.sqrt(-2 * np.log(u[2]))
        zz = [r01 * np.cos(2 * np.pi * u[1]), r01 * np.sin(2 * np.pi * u[1]), r23 * np.cos(2 * np.pi * u[3])]
        assert max(abs(x) for x in zz) > 5
