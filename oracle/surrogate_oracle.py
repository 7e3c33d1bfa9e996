"""Oracle (test infrastructure only): numpy restatement of the contact-map
surrogate that tfm_amd/surrogate.py runs on the GPU, forward and backward.

Network (activations channel-major, x[c, b L + l]):
  h1 = gelu(W1 * x + b1)                      conv kernel 3, padding 1
  h_k = h_{k-1} + gelu(W_k * h_{k-1} + b_k)   k = 2..n_layers
  U = Wu h, V = Wv h                          1x1 projections to d channels
  raw[b](i, j) = sum_c U[c, b L + i] V[c, b L + j] / sqrt(d)
  S[b](i, j) = (raw(i, j) + raw(j, i)) / 2 + dist[|i - j|] + c0
  rd[:, b] = Wr mean_l h[:, b L + l] + br
  loss = w_pair mean_{i <= j} (S - T)^2 + w_rd mean (rd - T_rd)^2

With bf16=True every GEMM operand is rounded to bfloat16 (round to nearest
even) where the GPU path rounds it; the sums stay in float64. The backward
pass is pinned by central finite differences in tests/test_oracle_surrogate.py.
"""
import numpy as np


def bf16(a):
    """Round float values to bfloat16 (nearest even), returned as float64."""
    a32 = np.ascontiguousarray(a, np.float32)
    u = a32.view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    out = (r & 0xFFFFFFFF).astype(np.uint32).view(np.float32).astype(np.float64)
    nan = np.isnan(a32)
    out[nan] = np.nan
    return out


def _id(a):
    return np.asarray(a, np.float64)


K_GELU = np.sqrt(2.0 / np.pi)


def gelu(x):
    return 0.5 * x * (1.0 + np.tanh(K_GELU * (x + 0.044715 * x ** 3)))


def gelu_grad(x):
    u = K_GELU * (x + 0.044715 * x ** 3)
    t = np.tanh(u)
    return 0.5 * (1.0 + t) + 0.5 * x * (1.0 - t * t) * K_GELU * (1.0 + 3 * 0.044715 * x * x)


def im2col(x, B, L):
    """x [C, B L] -> cols [3 C, B L] with cols[3 c + t, b L + l] = x[c, b L + l + t - 1]."""
    C = x.shape[0]
    xs = x.reshape(C, B, L)
    cols = np.zeros((C, 3, B, L))
    cols[:, 0, :, 1:] = xs[:, :, :-1]
    cols[:, 1] = xs
    cols[:, 2, :, :-1] = xs[:, :, 1:]
    return cols.reshape(3 * C, B * L)


def col2im(d, B, L):
    C3 = d.shape[0]
    ds = d.reshape(C3 // 3, 3, B, L)
    out = ds[:, 1].copy()
    out[:, :, :-1] += ds[:, 0, :, 1:]
    out[:, :, 1:] += ds[:, 2, :, :-1]
    return out.reshape(C3 // 3, B * L)


def init_params(c_in, width, n_layers, d, L, rd_bins, seed=0):
    rng = np.random.default_rng(seed)
    p = {"W1": rng.normal(0, 1 / np.sqrt(3 * c_in), (width, 3 * c_in)), "b1": np.zeros(width)}
    for k in range(2, n_layers + 1):
        p[f"W{k}"] = rng.normal(0, 0.5 / np.sqrt(3 * width), (width, 3 * width))
        p[f"b{k}"] = np.zeros(width)
    p["Wu"] = rng.normal(0, 1 / np.sqrt(width), (d, width))
    p["Wv"] = rng.normal(0, 1 / np.sqrt(width), (d, width))
    p["Wr"] = rng.normal(0, 1 / np.sqrt(width), (rd_bins, width))
    p["br"] = np.zeros(rd_bins)
    p["dist"] = np.zeros(L)
    p["c0"] = np.zeros(1)
    return p


def forward(p, x, B, L, n_layers, round_bf16=False):
    """Returns (S [B, L, L], rd [rd_bins, B], cache)."""
    r = bf16 if round_bf16 else _id
    cache = {"x": x}
    h = None
    inp = x
    for k in range(1, n_layers + 1):
        cols = im2col(r(inp), B, L)
        pre = r(p[f"W{k}"]) @ cols + p[f"b{k}"][:, None]
        g = gelu(pre)
        h = g if k == 1 else inp + g
        cache[f"cols{k}"], cache[f"pre{k}"] = cols, pre
        inp = h
    cache["h"] = h
    hb = r(h)
    d = p["Wu"].shape[0]
    U = r(r(p["Wu"]) @ hb)
    V = r(r(p["Wv"]) @ hb)
    Ub = U.reshape(d, B, L).transpose(1, 2, 0)  # [B, L, d]
    Vb = V.reshape(d, B, L).transpose(1, 2, 0)
    raw = np.einsum("bic,bjc->bij", Ub, Vb) / np.sqrt(d)
    idx = np.abs(np.arange(L)[:, None] - np.arange(L)[None, :])
    S = 0.5 * (raw + raw.transpose(0, 2, 1)) + p["dist"][idx][None] + p["c0"][0]
    pooled = h.reshape(h.shape[0], B, L).mean(axis=2)
    rd = r(p["Wr"]) @ r(pooled) + p["br"][:, None]
    cache.update(U=U, V=V, hb=hb, pooled=pooled)
    return S, rd, cache


def loss_and_grads(p, x, T, T_rd, B, L, n_layers, w_pair=1.0, w_rd=1.0, round_bf16=False):
    r = bf16 if round_bf16 else _id
    S, rd, c = forward(p, x, B, L, n_layers, round_bf16)
    iu = np.triu_indices(L)
    count = B * len(iu[0])
    e = S - T
    mask = np.zeros((L, L), bool)
    mask[iu] = True
    loss = w_pair * np.sum((e * mask) ** 2) / count
    e_rd = rd - T_rd
    loss += w_rd * np.sum(e_rd ** 2) / e_rd.size
    g = {}
    gS = 2 * w_pair * e * mask / count  # dL/dS on i <= j
    idx = np.abs(np.arange(L)[:, None] - np.arange(L)[None, :])
    g["dist"] = np.array([gS[:, np.arange(L - dd), np.arange(dd, L)].sum() for dd in range(L)])
    g["c0"] = np.array([gS.sum()])
    # S(lo, hi) = (raw(lo, hi) + raw(hi, lo)) / 2: graw(i, j) = gS(lo, hi) / 2, diagonal gS
    gsym = gS + gS.transpose(0, 2, 1)
    graw = 0.5 * gsym
    di = np.arange(L)
    graw[:, di, di] = gS[:, di, di]
    d = p["Wu"].shape[0]
    grawr = r(graw)
    Ub = c["U"].reshape(d, B, L).transpose(1, 2, 0)
    Vb = c["V"].reshape(d, B, L).transpose(1, 2, 0)
    dUb = np.einsum("bij,bjc->bic", grawr, Vb) / np.sqrt(d)  # [B, L, d]
    dVb = np.einsum("bij,bic->bjc", grawr, Ub) / np.sqrt(d)
    dU = dUb.transpose(2, 0, 1).reshape(d, B * L)
    dV = dVb.transpose(2, 0, 1).reshape(d, B * L)
    g["Wu"] = r(dU) @ c["hb"].T
    g["Wv"] = r(dV) @ c["hb"].T
    dh = r(p["Wu"]).T @ r(dU) + r(p["Wv"]).T @ r(dV)
    g_rd = 2 * w_rd * e_rd / e_rd.size
    g["br"] = g_rd.sum(axis=1)
    g["Wr"] = r(g_rd) @ r(c["pooled"]).T
    dpooled = r(p["Wr"]).T @ r(g_rd)
    dh += np.repeat(dpooled, L, axis=1) / L
    for k in range(n_layers, 0, -1):
        gp = dh * gelu_grad(c[f"pre{k}"])
        g[f"b{k}"] = gp.sum(axis=1)
        g[f"W{k}"] = r(gp) @ c[f"cols{k}"].T
        if k > 1:
            dcols = r(p[f"W{k}"]).T @ r(gp)
            dh = dh + col2im(dcols, B, L)
    return loss, g, S, rd
