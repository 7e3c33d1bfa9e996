"""Pins the oracle's observables (ccp's chrana.cu:13-45, 372-480) with
closed forms."""
import numpy as np
import pytest

import pyoracle as po

A = float(np.float32(1.154701)) / 2


def test_straight_chain_distance_and_contacts():
    N = 40
    r = np.zeros((N, 3))
    r[:, 0] = np.arange(N) * 1.0
    sd, cp = po.sd_cp(r)
    assert sd.shape == (20,) and sd.dtype == np.float32
    assert np.array_equal(sd, np.arange(1, 21, dtype=np.float32))
    assert cp[0] == 1000.0 and np.all(cp[1:] == 0.0)  # d = 2 is not a contact
    with pytest.raises(ValueError):
        po.sd_cp(r, 1)  # one chromosome unless N = 30386


LIM = [0, 2284, 4599, 6687, 9338, 12508, 15193, 17477, 19792, 21880, 24531, 27701, 30386]


@pytest.mark.parametrize("c", [0, 2, 5])
def test_two_copy_average_matches_brute_force(c):
    # N = 30386: chromosome c and its homologue c + 6 are averaged; s runs to
    # half the chromosome length
    rng = np.random.default_rng(c)
    v = rng.normal(size=(30386, 3))
    r = np.cumsum(v / np.linalg.norm(v, axis=1, keepdims=True), axis=0).astype(np.float32)
    sd, cp = po.sd_cp(r, c)
    L = LIM[c + 1] - LIM[c]
    assert sd.shape == (L // 2,)
    r64 = r.astype(np.float64)
    for s in (1, 2, 7, 100, L // 2):
        d = np.concatenate([np.linalg.norm(r64[b + s:b + L] - r64[b:b + L - s], axis=1)
                            for b in (LIM[c], LIM[c + 6])])
        assert sd[s - 1] == pytest.approx(d.mean(), rel=2e-5)
        assert abs(cp[s - 1] - 1000 * (d < 2).mean()) <= 1000 / len(d) + 1e-3


def test_shell_and_radial_bond_order():
    # beads on a sphere of radius 5 centred at 0: dcm ~ 0, rg2 = 25
    rng = np.random.default_rng(1)
    v = rng.normal(size=(2000, 3))
    v /= np.linalg.norm(v, axis=1, keepdims=True)
    r = 5 * np.concatenate([v, -v])
    out, _ = po.observables(r, np.zeros(len(r), np.uint8), 20.0)
    assert out[0] < 1e-9 and out[1] == pytest.approx(25.0)
    # a radial line of beads: every bond parallel to r_i -> P2 = 1; ncf counts z < nod
    r = np.zeros((10, 3))
    r[:, 2] = 1.0 + np.arange(10)
    out, _ = po.observables(r, np.zeros(10, np.uint8), 20.0, nod=4.5)
    assert out[2] == pytest.approx(1.0)
    assert out[3] == pytest.approx(0.4)
    # tangential bonds on a circle of many beads -> P2 = -1/2
    th = np.linspace(0, 2 * np.pi, 2001)[:-1]
    r = np.stack([10 * np.cos(th), 10 * np.sin(th), np.zeros_like(th)], 1)
    out, _ = po.observables(r, np.zeros(len(r), np.uint8), 20.0)
    assert out[2] == pytest.approx(-0.5, abs=1e-4)


def test_radial_density_volume_conservation_and_bins():
    R_n, n_b = 12.0, 32
    rng = np.random.default_rng(2)
    d = rng.uniform(0, R_n - A, 3000) ** 1.0
    v = rng.normal(size=(3000, 3))
    r = v / np.linalg.norm(v, axis=1, keepdims=True) * d[:, None]
    fl = (rng.random(3000) < 0.4).astype(np.uint8)
    _, rcd = po.observables(r, fl, R_n, n_b=n_b)
    shell = 4 / 3 * np.pi * R_n**3 / n_b
    vb = 4 / 3 * np.pi * A**3
    assert rcd[2].sum() * shell == pytest.approx(len(r) * vb, rel=1e-9)
    assert rcd[0].sum() * shell == pytest.approx(fl.sum() * vb, rel=1e-9)
    assert np.allclose(rcd[0] + rcd[1], rcd[2])
    # a bead well inside bin k contributes only to bin k (wide bins)
    n_b = 4
    shell = 4 / 3 * np.pi * R_n**3 / n_b
    rb = R_n * ((np.arange(n_b) + 1) / n_b) ** (1 / 3)
    k = 1
    x = 0.5 * (rb[k - 1] + rb[k])
    assert rb[k] - rb[k - 1] > 2 * A
    _, rcd = po.observables(np.array([[x, 0, 0]]), np.zeros(1, np.uint8), R_n, n_b=n_b)
    want = np.zeros(n_b)
    want[k] = vb / shell
    assert np.allclose(rcd[2], want)
