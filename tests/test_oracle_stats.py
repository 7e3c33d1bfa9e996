"""Pins the oracle's restatement of ccp's statistics (sostat.cu) with closed
forms, then checks the library's host statistics against it bit for bit."""
import numpy as np
import pytest

import sostat
import tfm_amd


def ar1(n, phi, seed):
    rng = np.random.default_rng(seed)
    e = rng.normal(size=n)
    x = np.empty(n)
    x[0] = e[0] / np.sqrt(1 - phi * phi)
    for i in range(1, n):
        x[i] = phi * x[i - 1] + e[i]
    return x.astype(np.float32)


def test_white_noise_error_of_mean():
    for seed in range(3):
        v = np.random.default_rng(seed).normal(3.0, 2.0, size=4096).astype(np.float32)
        avg, var, sem, f_n_b = sostat.blocking(v)
        assert avg == pytest.approx(v.astype(np.float64).mean(), rel=1e-12)
        assert var == pytest.approx(4.0, rel=0.1)
        assert sem == pytest.approx(2.0 / 64, rel=0.25)
        assert f_n_b <= 4096


@pytest.mark.parametrize("phi", [0.8, 0.95])
def test_ar1_blocking_error_matches_closed_form(phi):
    # Var(mean) -> sigma_x^2 (1 + phi) / ((1 - phi) n), sigma_x^2 = 1 / (1 - phi^2)
    n = 1 << 16
    v = ar1(n, phi, seed=int(phi * 100))
    _, var, sem, f_n_b = sostat.blocking(v)
    true = np.sqrt((1 + phi) / (1 - phi) / (1 - phi * phi) / n)
    assert var == pytest.approx(1 / (1 - phi * phi), rel=0.15)
    assert sem == pytest.approx(true, rel=0.3)
    assert sem > 2.5 * np.sqrt(var / n)  # correlation raises the error above the naive one
    assert f_n_b < n // 4


def test_thermalisation_cut():
    # a transient over the first quarter: the n/4 cut has the smallest marginal error
    rng = np.random.default_rng(5)
    n = 4096
    v = rng.normal(size=n).astype(np.float32)
    v[: n // 4] += np.linspace(20, 5, n // 4).astype(np.float32)
    avg, var, sem, f_n_b, i_t, thm = sostat.series(v)
    assert i_t == n // 4 and thm
    assert abs(avg) < 5 * sem
    # a transient over the whole first half: the cut stays at n/2, which
    # counts as not thermalised
    w = rng.normal(size=n).astype(np.float32)
    w[: n // 2] += 50
    assert sostat.series(w)[4:] == (n // 2, False)


def test_constant_series_and_weighted_mean():
    assert sostat.independent(np.full(100, 0.5, np.float32)) == (0.5, 0.0, 0.0)
    assert sostat.weighted([1.0, 3.0], [1.0, 1.0]) == pytest.approx((2.0, 2.0, 1.0))
    avg, var, sem = sostat.weighted([1.0, 3.0], [1.0, 2.0])
    assert avg == pytest.approx((1 + 0.75) / 1.25)
    # a zero error gives an infinite weight, which counts as 1
    assert sostat.weighted([1.0, 3.0], [0.0, 1.0])[0] == pytest.approx(2.0)


@pytest.mark.parametrize("n", [1, 2, 3, 4, 5, 7, 64, 100, 1000, 4099])
def test_library_series_stats_bit_exact(n):
    rng = np.random.default_rng(n)
    v = (ar1(n, 0.7, n) * 3 + 10).astype(np.float32) if n > 1 else np.array([4.25], np.float32)
    got = tfm_amd.series_stats(v)
    want = sostat.series(v)
    np.testing.assert_array_equal(got[:3], np.array(want[:3], np.float64))
    assert (got[3], got[4], bool(got[5])) == (want[3], want[4], want[5])


def test_library_weighted_stats_bit_exact():
    rng = np.random.default_rng(9)
    avg, sem = rng.normal(size=17), rng.random(17) + 0.1
    sem[3] = 0.0
    got = tfm_amd.weighted_stats(avg, sem)
    np.testing.assert_array_equal(got, np.array(sostat.weighted(avg, sem)))
