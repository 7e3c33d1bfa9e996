"""Pins the oracle's stochastic Heun step with closed forms."""
import numpy as np
from scipy import integrate

import pyoracle as po


def test_free_diffusion_variance():
    # no forces (single-bead chains, e_p = 0, far wall): each component
    # performs a random walk with per-step variance sd^2 = 2 kT dt.
    N, steps = 4000, 64
    sysm = po.System(N, 80.0, chain_off=np.arange(N + 1), e_p=0.0, seed=5)
    r0 = np.random.default_rng(3).uniform(-20, 20, size=(N, 3))
    r = sysm.run(r0, 0, steps)
    var = (r - r0).var(axis=0)
    expect = steps * sysm.p.sd**2
    assert np.allclose(sysm.p.sd**2, 2 * po.K_B * 298.0 / 2048, rtol=1e-6)
    assert np.all(np.abs(var / expect - 1) < 5 * np.sqrt(2 / N))


def test_harmonic_dimer_bond_distribution():
    # p(d) ~ d^2 exp(-k_e (d - 1)^2 / (2 kT)) for the stationary bond length
    N = 4000
    sysm = po.System(N, 80.0, chain_off=np.arange(0, N + 1, 2), e_p=0.0, seed=7,
                     flags=np.zeros(N, np.uint8))
    rng = np.random.default_rng(1)
    r = rng.uniform(-30, 30, size=(N, 3))
    r[1::2] = r[0::2] + [1.0, 0, 0]
    r = sysm.run(r, 0, 400)
    samples = []
    for k in range(10):
        r = sysm.run(r, 400 + 40 * k, 40)
        samples.append(np.linalg.norm(r[1::2] - r[0::2], axis=1))
    d = np.concatenate(samples)
    kT = po.K_B * 298.0
    w = lambda x: x * x * np.exp(-128.0 * (x - 1) ** 2 / (2 * kT))
    Z = integrate.quad(w, 0, 3)[0]
    m1 = integrate.quad(lambda x: x * w(x), 0, 3)[0] / Z
    m2 = integrate.quad(lambda x: x * x * w(x), 0, 3)[0] / Z
    var = m2 - m1 * m1
    assert abs(d.mean() - m1) < 5 * np.sqrt(var / (N // 2))
    assert abs(d.var() / var - 1) < 0.05


def test_fp32_and_fp64_agree_short():
    from helpers import relaxed, radius_for, pattern
    N = 256
    R_n = radius_for(N, 0.2)
    r = relaxed(N, R_n, 2)
    sysm = po.System(N, R_n, flags=pattern(N, 1), seed=9)
    a = sysm.run(r, 0, 50, prec=64)
    b = sysm.run(r, 0, 50, prec=32)
    assert np.abs(a - b).max() < 1e-3
