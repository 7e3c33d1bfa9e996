"""Pins the oracle's forces: finite differences of the closed-form
potentials (Bibliography/forces.md), cutoff zeros, and cell scan vs brute
force."""
import numpy as np
import pytest

import pyoracle as po
from helpers import pattern, relaxed, radius_for


def fd_forces(sysm, r, h=1e-6):
    f = np.zeros_like(r)
    for i in range(r.shape[0]):
        for a in range(3):
            rp, rm = r.copy(), r.copy()
            rp[i, a] += h
            rm[i, a] -= h
            f[i, a] = -(sysm.energy(rp).sum() - sysm.energy(rm).sum()) / (2 * h)
    return f


def small_system(ff, seed, N=48, cvf=0.3, flags=None, n_l=0, **kw):
    R_n = radius_for(N, cvf)
    r = relaxed(N, R_n, seed)
    # push a few beads close to the wall so wall terms are active
    rng = np.random.default_rng(seed)
    lbs = None
    if n_l:
        d = rng.normal(size=(n_l, 3))
        lbs = (R_n - po.CCP_CONST["rco"]) * d / np.linalg.norm(d, axis=1, keepdims=True)
        # put bead k right next to site k
        for k in range(n_l):
            r[4 * k + 2] = lbs[k] * (1 - 0.3 / np.linalg.norm(lbs[k]))
    r *= 1.0 + 0.25 / R_n  # compress outward: some beads within rco of the wall
    sysm = po.System(N, R_n, flags=flags, lbs=lbs, ff=ff, **kw)
    return sysm, r


@pytest.mark.parametrize("case", [
    dict(ff=po.FF_CCP, flags="A", n_l=4),
    dict(ff=po.FF_CCP, flags="mixed", n_l=4, k_b=0.0),
    dict(ff=po.FF_ICG, flags="A"),
    dict(ff=po.FF_LJ, flags="hetlad", bend_all=1, n_l=3, eps_het=0.6, eps_lam=1.2),
    dict(ff=po.FF_LJ, flags="hetlad", bend_all=1, bond_type=1, fene_r0=2.0),
])
def test_forces_are_minus_gradient(case):
    case = dict(case)
    N = 48
    kind = case.pop("flags")
    if kind == "A":
        fl = np.full(N, po.HET | po.LAD, np.uint8)
    elif kind == "mixed":
        fl = pattern(N, 5, mean=6)
    else:
        rng = np.random.default_rng(9)
        fl = (rng.random(N) < 0.5).astype(np.uint8) * po.HET | (rng.random(N) < 0.4).astype(np.uint8) * po.LAD
    sysm, r = small_system(seed=21, flags=fl, **case)
    f = sysm.forces(r, method=1)
    g = fd_forces(sysm, r)
    scale = np.abs(g).max()
    assert scale > 1.0
    assert np.abs(f - g).max() < 2e-5 * scale + 1e-6
    # every term family is exercised
    e = sysm.energy(r)
    assert e[0] != 0 and e[2] != 0 and e[4] != 0
    if sysm.p.n_l and case["ff"] != po.FF_ICG:
        assert e[3] < 0


def test_ccp_bending_quirk_is_not_a_gradient():
    # ccp applies all three angle terms of bead i only when i is type A
    # (chrsim.cu:65-72): for mixed chains the force differs from -grad V.
    N = 48
    fl = pattern(N, 5, mean=4)
    sysm, r = small_system(po.FF_CCP, 21, flags=fl)
    f = sysm.forces(r, method=1)
    g = fd_forces(sysm, r)
    assert np.abs(f - g).max() > 1e-3


@pytest.mark.parametrize("chains", [[0, 48], [0, 20, 21, 48]])
def test_ccp_bending_on_mixed_chains_is_the_masked_gradient(chains):
    # ccp's bending term on bead i (chrsim.cu:65-72) holds the three angle
    # terms of -grad V_all, V_all = k_b sum_j (1 - cos theta_j) over every
    # interior angle of the chains, and is applied only where i is type A:
    # f(k_b) - f(k_b = 0) = 1[i is A] (-grad V_all) (the other force terms
    # cancel in the difference). A one-bead chain and chain limits included.
    N = 48
    fl = pattern(N, 5, mean=4)
    assert 0 < (fl & po.HET).astype(bool).sum() < N
    sysm, r = small_system(po.FF_CCP, 21, flags=fl, chain_off=chains)
    kb = sysm.p.k_b
    assert kb > 0
    flat, _ = small_system(po.FF_CCP, 21, flags=fl, chain_off=chains, k_b=0.0)
    fb = sysm.forces(r, method=1) - flat.forces(r, method=1)
    g = np.zeros_like(r)
    h = 1e-6
    for i in range(N):
        for a in range(3):
            rp, rm = r.copy(), r.copy()
            rp[i, a] += h
            rm[i, a] -= h
            g[i, a] = -(sysm.energy(rp)[1] - sysm.energy(rm)[1]) / (2 * h)
    mask = ((fl & po.HET) != 0)[:, None]
    scale = np.abs(g).max()
    assert scale > 0.1
    assert np.abs(fb - mask * g).max() < 2e-5 * scale + 1e-6
    # the B beads' part of -grad V_all is what ccp leaves out, and it is not small
    assert np.abs((~mask) * g).max() > 0.05 * scale


@pytest.mark.parametrize("ff", [po.FF_CCP, po.FF_ICG, po.FF_LJ])
def test_cell_scan_matches_brute_force(ff):
    N = 1500
    R_n = radius_for(N, 0.2)
    r = relaxed(N, R_n, 4)
    fl = pattern(N, 8)
    rng = np.random.default_rng(2)
    d = rng.normal(size=(60, 3))
    lbs = (R_n - po.CCP_CONST["rco"]) * d / np.linalg.norm(d, axis=1, keepdims=True)
    sysm = po.System(N, R_n, flags=fl, lbs=lbs, ff=ff, bend_all=int(ff == po.FF_LJ),
                     chain_off=[0, 500, 1100, 1500])
    fc = sysm.forces(r, method=0)
    fb = sysm.forces(r, method=1)
    assert np.abs(fc - fb).max() < 1e-9 * max(1.0, np.abs(fb).max())


def test_wang_frenkel_zeros_and_alpha():
    # WF with rc = 2: alpha = 1, force zero at its minimum 2/sqrt(3) (= rco)
    # and at rc = 2; both cutoffs of ccp sit on force zeros.
    for d in (2 / np.sqrt(3), 2.0):
        d2 = d * d
        assert abs((18 * d2 * d2 - 96 * d2 + 96) / d2**4) < 1e-12
    rc, mu, nu = 2.0, 1, 1
    alpha = 2 * nu * rc**(2 * mu) * ((1 + 2 * nu) / (2 * nu * (rc**(2 * mu) - 1)))**(2 * nu + 1)
    assert alpha == pytest.approx(1.0)


def test_two_bead_pair_forces_exact():
    # a 4-bead system where only beads 0 and 3 interact: check the closed form
    for ff, hh, d in ((po.FF_CCP, True, 1.6), (po.FF_CCP, False, 1.05), (po.FF_ICG, True, 1.0),
                      (po.FF_LJ, True, 1.8), (po.FF_LJ, False, 1.0)):
        fl = np.array([po.HET if hh else 0, 0, 0, po.HET if hh else 0], np.uint8)
        sysm = po.System(4, 30.0, chain_off=[0, 1, 2, 3, 4], flags=fl, ff=ff, eps_het=0.6)
        r = np.array([[0, 0, 0], [10, 0, 0], [-10, 0, 0], [d, 0, 0]], float)
        f = sysm.forces(r, method=1)
        d2 = d * d
        if ff == po.FF_CCP:
            expect = 0.5 * (18 * d2 * d2 - 96 * d2 + 96) / d2**4 * (-d)
        elif ff == po.FF_ICG:
            expect = 128 * (3 * po.CCP_CONST["rco"] - 3 * d) * (-d)
        else:
            s6 = (1 / d2)**3
            eps = 1.0 if d2 < 2**(1 / 3) else 0.6
            expect = 24 * eps * s6 * (2 * s6 - 1) / d2 * (-d)
        assert f[0, 0] == pytest.approx(expect, rel=1e-12)
        assert f[3, 0] == pytest.approx(-expect, rel=1e-12)
