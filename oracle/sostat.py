"""Oracle (test infrastructure only): ccp's statistics of per-frame
observables (Program/src/sostat.cu) restated in numpy, with ccp's FP32/FP64
rounding points. Only tests import this module.

Pinned in tests/test_oracle_stats.py by closed forms: the blocking error of
white noise and of an AR(1) series against sigma^2 (1 + phi) / ((1 - phi) n),
the thermalisation cut of a series with a transient, and the inverse-variance
weighted mean of results with known errors.
"""
import numpy as np


def _seq_sum(x):
    # sequential FP64 sum (ccp's loop order); np.sum would pair-sum
    return float(np.cumsum(np.asarray(x, np.float64))[-1]) if len(x) else 0.0


def independent(v):
    """calc_stats(vector<float>, idstat&), sostat.cu:72-93."""
    v = np.asarray(v, np.float32)
    n = v.size
    with np.errstate(divide="ignore", invalid="ignore"):
        m1 = np.float64(_seq_sum(v)) / n
        m2 = np.float64(_seq_sum((v * v).astype(np.float32))) / n  # FP32 squares
        var = (m2 - m1 * m1) / (1.0 - 1.0 / np.float64(n))
        return float(m1), float(var), float(np.sqrt(var / n))


def blocking(v):
    """calc_stats(vector<float>, cdstat&), sostat.cu:96-138: (avg, var, sem, f_n_b)."""
    av = np.asarray(v, np.float32).copy()
    avg, var, sem = independent(av)
    f_n_b = av.size
    n_e = av.size
    ivm = var / n_e
    ulivm = ivm * (1.0 + np.sqrt(2.0 / (n_e - 1.0))) if n_e > 1 else np.inf
    while n_e >= 4:
        nb = n_e // 2
        av = ((av[0:2 * nb:2] + av[1:2 * nb:2]).astype(np.float32) * np.float32(0.5)).astype(np.float32)
        n_e = nb
        _, v2, _ = independent(av)
        ivm = v2 / n_e
        sem = np.sqrt(ivm)
        f_n_b = n_e
        if ivm > ulivm:
            ulivm = ivm * (1.0 + np.sqrt(2.0 / (n_e - 1.0)))
        else:
            break
    return avg, var, sem, f_n_b


def series(v):
    """calc_stats(vector<float>, tdstat&), sostat.cu:141-183:
    (avg, var, sem, f_n_b, i_t, thermalised)."""
    v = np.asarray(v, np.float32)
    n = v.size
    best, i_t = np.inf, n // 2
    d = 2
    while d < 128:
        cut = n // d
        m = n - cut
        _, var, _ = independent(v[cut:])
        # m * m in 32-bit unsigned arithmetic, as ccp
        mse = var * float((m - 1) & 0xFFFFFFFF) / float((m * m) & 0xFFFFFFFF)
        if mse < best:
            best, i_t = mse, cut
        d *= 2
    avg, var, sem, f_n_b = blocking(v[i_t:])
    return avg, var, sem, f_n_b, i_t, i_t != n // 2


def weighted(avg, sem):
    """calc_stats(vector<tdstat>, idstat&), sostat.cu:186-211: (avg, var, sem)."""
    m1 = m2 = w1 = w2 = 0.0
    for a, s in zip(avg, sem):
        with np.errstate(divide="ignore", invalid="ignore", over="ignore"):
            w = float(np.float64(1.0) / (np.float64(s) * np.float64(s)))
        if not np.isfinite(w):
            w = 1.0
        m1 += w * a
        m2 += w * a * a
        w1 += w
        w2 += w * w
    m1 /= w1
    m2 /= w1
    var = (m2 - m1 * m1) / (1.0 - w2 / (w1 * w1))
    return m1, var, np.sqrt(var * w2 / (w1 * w1))
