"""GPU parity of the surrogate's kernels: the strided batched BF16 GEMM on the
matrix cores against float64 products of the same bfloat16 values, and the
network's forward pass and every gradient against the numpy oracle with the
same bfloat16 rounding points; then a short training run."""
import numpy as np
import pytest
import torch

import surrogate_oracle as so
from gpu_common import step
from tfm_amd import surrogate as sg

pytestmark = pytest.mark.gpu


def to_bf16_dev(a):
    """float array -> (device int16 words of its bfloat16 rounding, float64 values)."""
    r = so.bf16(a)
    words = (r.astype(np.float32).view(np.uint32) >> 16).astype(np.uint16).view(np.int16)
    return torch.tensor(words, device="cuda"), r


@pytest.mark.parametrize("M,N,K,batch,layout", [(64, 64, 16, 1, "nn"), (70, 33, 19, 1, "nn"), (128, 96, 200, 1, "nt"),
                                                (45, 130, 77, 1, "tn"), (32, 40, 24, 5, "nn"), (17, 17, 9, 3, "tt")])
def test_gemm_matches_float64(M, N, K, batch, layout):
    step(f"gemm {M}x{N}x{K} batch {batch} {layout}")
    rng = np.random.default_rng(M * N + K)
    a = rng.normal(size=(batch, M, K)) if layout[0] == "n" else rng.normal(size=(batch, K, M))
    b = rng.normal(size=(batch, K, N)) if layout[1] == "n" else rng.normal(size=(batch, N, K))
    A, av = to_bf16_dev(a)
    B, bv = to_bf16_dev(b)
    Am = av if layout[0] == "n" else av.transpose(0, 2, 1)
    Bm = bv if layout[1] == "n" else bv.transpose(0, 2, 1)
    want = 0.5 * np.einsum("bmk,bkn->bmn", Am, Bm) + 2.0
    sa = (M * K, K, 1) if layout[0] == "n" else (M * K, 1, M)
    sb = (K * N, N, 1) if layout[1] == "n" else (K * N, 1, K)
    C = torch.full((batch, M, N), 1.0, dtype=torch.float32, device="cuda")
    Cb = torch.zeros((batch, M, N), dtype=torch.int16, device="cuda")
    sg.gemm(A, B, M, N, K, sa, sb, (M * N, N, 1), C=C, Cb=Cb, batch=batch, alpha=0.5, beta=2.0)
    torch.cuda.synchronize()
    got = C.cpu().numpy().astype(np.float64)
    scale = 0.5 * np.einsum("bmk,bkn->bmn", np.abs(Am), np.abs(Bm)) + 2.0
    assert np.all(np.abs(got - want) <= 1e-6 * scale + 1e-6)
    gotb = (Cb.cpu().numpy().view(np.uint16).astype(np.uint32) << 16).view(np.float32)
    assert np.array_equal(gotb, so.bf16(got).astype(np.float32))


def cfg():
    return dict(c_in=5, width=32, n_layers=3, d=8, L=16, rd_bins=8, B=3)


def problem(seed=0):
    c = cfg()
    rng = np.random.default_rng(seed)
    p = so.init_params(c["c_in"], c["width"], c["n_layers"], c["d"], c["L"], c["rd_bins"], seed=seed)
    for k in p:
        p[k] = (p[k] + 0.05 * rng.normal(size=p[k].shape)).astype(np.float32).astype(np.float64)
    x = rng.normal(size=(c["c_in"], c["B"] * c["L"])).astype(np.float32).astype(np.float64)
    T = rng.normal(size=(c["B"], c["L"], c["L"]))
    T = (0.5 * (T + T.transpose(0, 2, 1))).astype(np.float32).astype(np.float64)
    T_rd = rng.normal(size=(c["rd_bins"], c["B"])).astype(np.float32).astype(np.float64)
    return c, p, x, T, T_rd


def rel(a, b):
    return np.linalg.norm(np.ravel(a) - np.ravel(b)) / max(np.linalg.norm(np.ravel(b)), 1e-30)


def test_forward_and_gradients_match_oracle():
    step("surrogate forward/backward vs oracle")
    c, p, x, T, T_rd = problem()
    net = sg.Surrogate(params=p, **c)
    dev = lambda a: torch.tensor(np.asarray(a, np.float32), device="cuda")
    net.zero_grad()
    loss = net.loss_backward(dev(x), dev(T), dev(T_rd), w_pair=1.0, w_rd=0.5)
    torch.cuda.synchronize()
    want_loss, g, S, rd = so.loss_and_grads(p, x, T, T_rd, c["B"], c["L"], c["n_layers"], 1.0, 0.5,
                                            round_bf16=True)
    assert rel(net.S.cpu().numpy(), S) < 1e-3
    assert rel(net.rd.cpu().numpy(), rd) < 1e-3
    assert float(loss.item()) == pytest.approx(want_loss, rel=1e-3)
    for k in p:
        err = rel(net.g[k].cpu().numpy(), g[k])
        step(f"  grad {k}: rel err {err:.2e}")
        assert err < 2e-2, k


def test_training_reduces_loss():
    step("surrogate training")
    c, p, x, T, T_rd = problem(3)
    net = sg.Surrogate(params=p, **c)
    dev = lambda a: torch.tensor(np.asarray(a, np.float32), device="cuda")
    X, TT, TR = dev(x), dev(T), dev(T_rd)
    losses = []
    for it in range(300):
        net.zero_grad()
        losses.append(float(net.loss_backward(X, TT, TR).item()))
        net.step(lr=3e-3)
    assert np.isfinite(losses).all()
    assert losses[-1] < 0.2 * losses[0]
