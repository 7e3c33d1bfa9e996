"""Pins the numpy oracle of the contact-map surrogate: its backward pass
against central finite differences in float64, and its bfloat16 rounding
against the format's definition."""
import numpy as np
import pytest

import surrogate_oracle as so


def small(seed=0):
    B, L, c_in, width, n_layers, d, rdb = 2, 12, 5, 8, 3, 4, 6
    rng = np.random.default_rng(seed)
    p = so.init_params(c_in, width, n_layers, d, L, rdb, seed=seed)
    for k in p:
        p[k] = p[k] + 0.1 * rng.normal(size=p[k].shape)  # non-zero biases and head offsets
    x = rng.normal(size=(c_in, B * L))
    T = rng.normal(size=(B, L, L))
    T = 0.5 * (T + T.transpose(0, 2, 1))
    T_rd = rng.normal(size=(rdb, B))
    return p, x, T, T_rd, B, L, n_layers


def test_backward_matches_finite_differences():
    p, x, T, T_rd, B, L, n = small()
    loss, g, _, _ = so.loss_and_grads(p, x, T, T_rd, B, L, n, w_pair=1.0, w_rd=0.7)
    rng = np.random.default_rng(1)
    h = 1e-6
    for name in p:
        flat = p[name].ravel()
        for idx in rng.choice(flat.size, size=min(6, flat.size), replace=False):
            keep = flat[idx]
            flat[idx] = keep + h
            lp = so.loss_and_grads(p, x, T, T_rd, B, L, n, w_pair=1.0, w_rd=0.7)[0]
            flat[idx] = keep - h
            lm = so.loss_and_grads(p, x, T, T_rd, B, L, n, w_pair=1.0, w_rd=0.7)[0]
            flat[idx] = keep
            fd = (lp - lm) / (2 * h)
            assert g[name].ravel()[idx] == pytest.approx(fd, rel=1e-5, abs=1e-8), (name, idx)


def test_symmetry_and_distance_bias():
    p, x, T, T_rd, B, L, n = small(2)
    S, rd, _ = so.forward(p, x, B, L, n)
    assert np.allclose(S, S.transpose(0, 2, 1))
    # a constant shift of every pair through c0 and of each diagonal through dist
    p["c0"] = p["c0"] + 1.5
    p["dist"][3] += 2.0
    S2, _, _ = so.forward(p, x, B, L, n)
    D = S2 - S
    i, j = np.indices((L, L))
    assert np.allclose(D, 1.5 + 2.0 * (np.abs(i - j) == 3))


def test_bf16_rounding():
    vals = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 2 ** -7, 1.0 + 3 * 2 ** -8, -2.5, 1e-40, 3.0e38], np.float32)
    got = so.bf16(vals)
    # ties go to even: 1 + 2^-8 -> 1, 1 + 3 2^-8 -> 1 + 2^-6
    assert got[0] == 1.0 and got[1] == 1.0 and got[2] == 1.0 + 2 ** -7 and got[3] == 1.0 + 2 ** -6
    assert got[4] == -2.5
    r = np.random.default_rng(0).normal(size=1000).astype(np.float32)
    assert np.all(np.abs(so.bf16(r) - r) <= np.abs(r) * 2 ** -8)
