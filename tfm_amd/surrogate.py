"""Contact-map surrogate: a 1-D convolutional encoder over a chain's binned
het/LAD pattern and the run's (volume fraction, eps_het, eps_lamina), with a
pairwise outer-product head for the log10 contact-probability map and a
pooled head for the radial density.

Every forward and backward step runs on the library's kernels
(include/ccpmi_nn.h): BF16 GEMMs on the matrix cores with FP32
accumulation, FP32 element-wise work, AdamW on FP32 master weights with BF16
copies. torch only allocates the device buffers and provides the stream.

Network (activations x[c, b L + l]):
  h1 = gelu(W1 * x + b1), h_k = h_{k-1} + gelu(W_k * h_{k-1} + b_k)  (conv k=3)
  U = Wu h, V = Wv h; raw[b](i, j) = U_b(:, i) . V_b(:, j) / sqrt(d)
  S = (raw + raw^T) / 2 + dist[|i - j|] + c0          (log10 P map)
  rd = Wr mean_l h + br                                (radial density)
"""
import ctypes
import math

import numpy as np
import torch

from .ccp import lib

NN_EXPORTS = (
    "ccp_nn_gemm", "ccp_nn_im2col", "ccp_nn_col2im", "ccp_nn_bias_gelu", "ccp_nn_gelu_back", "ccp_nn_pair_map",
    "ccp_nn_pair_loss", "ccp_nn_mean_pool", "ccp_nn_pool_back", "ccp_nn_rd_loss", "ccp_nn_cast_bf16", "ccp_nn_adamw",
)

_bound = False


def _nn():
    global _bound
    L = lib()
    if not _bound:
        vp, f, i, lg = ctypes.c_void_p, ctypes.c_float, ctypes.c_int, ctypes.c_long
        L.ccp_nn_gemm.argtypes = [vp, vp, vp, vp, i, i, i, i] + [lg] * 9 + [f, f, vp]
        L.ccp_nn_im2col.argtypes = [vp, vp, i, i, i, vp]
        L.ccp_nn_col2im.argtypes = [vp, vp, i, i, i, f, vp]
        L.ccp_nn_bias_gelu.argtypes = [vp, vp, vp, vp, vp, i, lg, vp]
        L.ccp_nn_gelu_back.argtypes = [vp, vp, vp, vp, vp, i, lg, vp]
        L.ccp_nn_pair_map.argtypes = [vp, vp, vp, vp, i, i, vp]
        L.ccp_nn_pair_loss.argtypes = [vp, vp, vp, vp, vp, vp, vp, vp, i, i, f, vp]
        L.ccp_nn_mean_pool.argtypes = [vp, vp, vp, i, i, i, vp]
        L.ccp_nn_pool_back.argtypes = [vp, vp, i, i, i, vp]
        L.ccp_nn_rd_loss.argtypes = [vp, vp, vp, vp, vp, vp, vp, i, i, f, vp]
        L.ccp_nn_cast_bf16.argtypes = [vp, vp, lg, vp]
        L.ccp_nn_adamw.argtypes = [vp, vp, vp, vp, vp, lg, f, f, f, f, f, i, vp]
        _bound = True
    return L


def _p(t):
    return ctypes.c_void_p(t.data_ptr()) if t is not None else None


def _stream():
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


def _ok(rc, what):
    if rc != 0:
        raise RuntimeError(f"{what} failed ({rc})")


def gemm(A, B, M, N, K, sa, sb, sc, C=None, Cb=None, batch=1, alpha=1.0, beta=0.0):
    """Strided batched BF16 GEMM; sa = (batch, m, k) strides of A, sb = (batch,
    k, n) of B, sc = (batch, m, n) of C/Cb."""
    _ok(_nn().ccp_nn_gemm(_p(A), _p(B), _p(C), _p(Cb), M, N, K, batch, *sa, *sb, *sc, alpha, beta, _stream()),
        "gemm")


def bf16_like(t):
    return torch.empty(t.shape, dtype=torch.int16, device=t.device)


class Surrogate:
    """Weights (FP32 masters, BF16 copies) and buffers for batches of B
    sequences of L bins with c_in input channels."""

    def __init__(self, c_in=5, width=256, n_layers=6, d=64, L=128, rd_bins=32, B=64, seed=0, params=None):
        self.c_in, self.width, self.n_layers, self.d, self.L, self.rd_bins, self.B = \
            c_in, width, n_layers, d, L, rd_bins, B
        dev = torch.device("cuda")
        if params is None:
            import_seed = np.random.default_rng(seed)
            params = {"W1": import_seed.normal(0, 1 / math.sqrt(3 * c_in), (width, 3 * c_in)), "b1": np.zeros(width)}
            for k in range(2, n_layers + 1):
                params[f"W{k}"] = import_seed.normal(0, 0.5 / math.sqrt(3 * width), (width, 3 * width))
                params[f"b{k}"] = np.zeros(width)
            params["Wu"] = import_seed.normal(0, 1 / math.sqrt(width), (d, width))
            params["Wv"] = import_seed.normal(0, 1 / math.sqrt(width), (d, width))
            params["Wr"] = import_seed.normal(0, 1 / math.sqrt(width), (rd_bins, width))
            params["br"] = np.zeros(rd_bins)
            params["dist"] = np.zeros(L)
            params["c0"] = np.zeros(1)
        self.names = list(params.keys())
        self.p = {k: torch.tensor(np.asarray(v, np.float32), device=dev).contiguous() for k, v in params.items()}
        self.pb = {k: bf16_like(v) for k, v in self.p.items() if k.startswith("W")}
        self.g = {k: torch.zeros_like(v) for k, v in self.p.items()}
        self.m = {k: torch.zeros_like(v) for k, v in self.p.items()}
        self.v = {k: torch.zeros_like(v) for k, v in self.p.items()}
        self.t = 0
        self._cast_weights()
        self._alloc(B)

    def _alloc(self, B):
        self.B = B
        dev = torch.device("cuda")
        P, W, L = B * self.L, self.width, self.L
        f = lambda *s: torch.zeros(s, dtype=torch.float32, device=dev)
        h = lambda *s: torch.zeros(s, dtype=torch.int16, device=dev)
        self.xb = h(self.c_in, P)
        self.cols = {1: h(3 * self.c_in, P)}
        for k in range(2, self.n_layers + 1):
            self.cols[k] = h(3 * W, P)
        self.pre = {k: f(W, P) for k in range(1, self.n_layers + 1)}
        self.hs = {k: f(W, P) for k in range(1, self.n_layers + 1)}
        self.hb = {k: h(W, P) for k in range(1, self.n_layers + 1)}
        self.Ub, self.Vb = h(self.d, P), h(self.d, P)
        self.raw = f(B, L, L)
        self.S = f(B, L, L)
        self.pooled, self.pooledb = f(W, B), h(W, B)
        self.rd = f(self.rd_bins, B)
        # backward
        self.gS, self.graw, self.grawb = f(B, L, L), f(B, L, L), h(B, L, L)
        self.dU, self.dUb, self.dV, self.dVb = f(self.d, P), h(self.d, P), f(self.d, P), h(self.d, P)
        self.dh = f(W, P)
        self.gp, self.gpb = f(W, P), h(W, P)
        self.dcols = f(3 * W, P)
        self.grd, self.grdb = f(self.rd_bins, B), h(self.rd_bins, B)
        self.dpooled = f(W, B)
        self.loss = f(1)

    def _cast_weights(self):
        L = _nn()
        for k, t in self.pb.items():
            _ok(L.ccp_nn_cast_bf16(_p(self.p[k]), _p(t), t.numel(), _stream()), "cast")

    # ------------------------------------------------------------ forward
    @staticmethod
    def _check_input(t, shape, name):
        if t.dtype != torch.float32 or not t.is_cuda or not t.is_contiguous() or tuple(t.shape) != tuple(shape):
            raise ValueError(f"{name} must be a contiguous float32 cuda tensor of shape {tuple(shape)}")

    def forward(self, x):
        """x: [c_in, B L] float32 (cuda). Returns (S [B, L, L], rd [rd_bins, B]) on the device."""
        self._check_input(x, (self.c_in, self.B * self.L), "x")
        N = _nn()
        B, L, W, d = self.B, self.L, self.width, self.d
        P = B * L
        s = _stream()
        _ok(N.ccp_nn_cast_bf16(_p(x), _p(self.xb), x.numel(), s), "cast")
        inp_b, inp_f, c_in = self.xb, None, self.c_in
        for k in range(1, self.n_layers + 1):
            _ok(N.ccp_nn_im2col(_p(inp_b), _p(self.cols[k]), c_in, B, L, s), "im2col")
            K = 3 * c_in
            gemm(self.pb[f"W{k}"], self.cols[k], W, P, K, (0, K, 1), (0, P, 1), (0, P, 1), C=self.pre[k])
            _ok(N.ccp_nn_bias_gelu(_p(self.pre[k]), _p(self.p[f"b{k}"]), _p(inp_f), _p(self.hs[k]),
                                   _p(self.hb[k]), W, P, s), "bias_gelu")
            inp_b, inp_f, c_in = self.hb[k], self.hs[k], W
        hb, hf = inp_b, inp_f
        gemm(self.pb["Wu"], hb, d, P, W, (0, W, 1), (0, P, 1), (0, P, 1), Cb=self.Ub)
        gemm(self.pb["Wv"], hb, d, P, W, (0, W, 1), (0, P, 1), (0, P, 1), Cb=self.Vb)
        # raw[b](i, j) = sum_c U[c, b L + i] V[c, b L + j] / sqrt(d)
        gemm(self.Ub, self.Vb, L, L, d, (L, 1, P), (L, P, 1), (L * L, L, 1), C=self.raw, batch=B,
             alpha=1.0 / math.sqrt(d))
        _ok(N.ccp_nn_pair_map(_p(self.raw), _p(self.p["dist"]), _p(self.p["c0"]), _p(self.S), L, B, s), "pair_map")
        _ok(N.ccp_nn_mean_pool(_p(hf), _p(self.pooled), _p(self.pooledb), W, B, L, s), "pool")
        gemm(self.pb["Wr"], self.pooledb, self.rd_bins, B, W, (0, W, 1), (0, B, 1), (0, B, 1), C=self.rd)
        return self.S, self.rd

    def predict(self, x):
        S, rd = self.forward(x)
        rd = rd + self.p["br"][:, None]
        return S, rd

    # ------------------------------------------------------------ backward
    def zero_grad(self):
        for t in self.g.values():
            t.zero_()

    def loss_backward(self, x, T, T_rd, w_pair=1.0, w_rd=1.0):
        """Forward, loss and gradients into self.g; returns the loss (device scalar)."""
        N = _nn()
        B, L, W, d = self.B, self.L, self.width, self.d
        P = B * L
        s = _stream()
        self._check_input(T, (B, L, L), "T")
        self._check_input(T_rd, (self.rd_bins, B), "T_rd")
        self.forward(x)
        self.loss.zero_()
        _ok(N.ccp_nn_pair_loss(_p(self.S), _p(T), _p(self.gS), _p(self.graw), _p(self.grawb), _p(self.g["dist"]),
                               _p(self.g["c0"]), _p(self.loss), L, B, w_pair, s), "pair_loss")
        _ok(N.ccp_nn_rd_loss(_p(self.rd), _p(self.p["br"]), _p(T_rd), _p(self.grd), _p(self.grdb), _p(self.g["br"]),
                             _p(self.loss), self.rd_bins, B, w_rd, s), "rd_loss")
        a = 1.0 / math.sqrt(d)
        # dU_b = a V_b graw_b^T, dV_b = a U_b graw_b
        gemm(self.Vb, self.grawb, d, L, L, (L, P, 1), (L * L, 1, L), (L, P, 1), C=self.dU, Cb=self.dUb, batch=B,
             alpha=a)
        gemm(self.Ub, self.grawb, d, L, L, (L, P, 1), (L * L, L, 1), (L, P, 1), C=self.dV, Cb=self.dVb, batch=B,
             alpha=a)
        hb, hf = self.hb[self.n_layers], self.hs[self.n_layers]
        gemm(self.dUb, hb, d, W, P, (0, P, 1), (0, 1, P), (0, W, 1), C=self.g["Wu"], beta=1.0)
        gemm(self.dVb, hb, d, W, P, (0, P, 1), (0, 1, P), (0, W, 1), C=self.g["Wv"], beta=1.0)
        gemm(self.pb["Wu"], self.dUb, W, P, d, (0, 1, W), (0, P, 1), (0, P, 1), C=self.dh)
        gemm(self.pb["Wv"], self.dVb, W, P, d, (0, 1, W), (0, P, 1), (0, P, 1), C=self.dh, beta=1.0)
        # radial head
        gemm(self.grdb, self.pooledb, self.rd_bins, W, B, (0, B, 1), (0, 1, B), (0, W, 1), C=self.g["Wr"], beta=1.0)
        gemm(self.pb["Wr"], self.grdb, W, B, self.rd_bins, (0, 1, W), (0, B, 1), (0, B, 1), C=self.dpooled)
        _ok(N.ccp_nn_pool_back(_p(self.dpooled), _p(self.dh), W, B, L, s), "pool_back")
        for k in range(self.n_layers, 0, -1):
            _ok(N.ccp_nn_gelu_back(_p(self.pre[k]), _p(self.dh), _p(self.gp), _p(self.gpb), _p(self.g[f"b{k}"]), W,
                                   P, s), "gelu_back")
            c_in = self.c_in if k == 1 else W
            K = 3 * c_in
            gemm(self.gpb, self.cols[k], W, K, P, (0, P, 1), (0, 1, P), (0, K, 1), C=self.g[f"W{k}"], beta=1.0)
            if k > 1:
                gemm(self.pb[f"W{k}"], self.gpb, K, P, W, (0, 1, K), (0, P, 1), (0, P, 1), C=self.dcols)
                _ok(N.ccp_nn_col2im(_p(self.dcols), _p(self.dh), W, B, L, 1.0, s), "col2im")
        return self.loss

    def step(self, lr=1e-3, b1=0.9, b2=0.999, eps=1e-8, wd=1e-4):
        N = _nn()
        self.t += 1
        s = _stream()
        for k in self.names:
            decay = wd if k.startswith("W") else 0.0
            _ok(N.ccp_nn_adamw(_p(self.p[k]), _p(self.g[k]), _p(self.m[k]), _p(self.v[k]), _p(self.pb.get(k)),
                               self.p[k].numel(), lr, b1, b2, eps, decay, self.t, s), "adamw")

    def state_dict(self):
        return {k: v.detach().cpu().numpy() for k, v in self.p.items()}
