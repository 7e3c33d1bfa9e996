// gfx950 kernels of the contact-map surrogate: a strided, batched BF16 GEMM
// on the matrix cores (v_mfma_f32_32x32x16 bf16, FP32 accumulation) and the
// element-wise pieces of a 1-D convolutional encoder with a pairwise
// outer-product head (im2col / col2im, bias + GELU and its derivative, the
// symmetric pair map, pooling, losses, AdamW).
//
// Matrices are addressed by element strides, so one GEMM covers every
// transpose the forward and backward passes need:
//   C[b](m, n) = alpha * sum_k A[b](m, k) B[b](k, n) + beta * C[b](m, n)
// with A[b](m, k) = A[b * sab + m * sam + k * sak] and likewise for B and C.

#include "ccpmi_nn.h"

#include <hip/hip_bf16.h>
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));

__device__ __forceinline__ uint16_t f2bf(float f) {
  // round to nearest even; NaN stays NaN
  uint32_t u = __float_as_uint(f);
  if ((u & 0x7fffffffu) > 0x7f800000u) return (uint16_t)((u >> 16) | 0x40u);
  u += 0x7fffu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}
__device__ __forceinline__ float bf2f(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }

// Block tile 64 x 64 (four wavefronts of 32 x 32), K tile 32: two
// v_mfma_f32_32x32x16_bf16 per wavefront per K tile.
constexpr int TM = 64, TN = 64, TK = 32, GEMM_THREADS = 256;
// LDS rows: 32 K values plus 8 of padding (80 bytes), so the 16-byte reads
// of the 32 rows one MFMA operand takes spread over the banks
constexpr int TKP = TK + 8;

// Eight elements along K of one row (A) or column (B) of the tile, for the
// staging thread: one 16-byte load when they are contiguous and aligned in
// memory, else eight loads (zero outside the matrix).
__device__ __forceinline__ uint4 load8(const uint16_t *__restrict__ P, long base, long sk, int row_ok, int k,
                                       int K) {
  if (row_ok && sk == 1 && k + 8 <= K && ((uintptr_t)(P + base + k) & 15u) == 0)
    return *(const uint4 *)(P + base + k);
  uint16_t v[8];
#pragma unroll
  for (int j = 0; j < 8; ++j) v[j] = (row_ok && k + j < K) ? P[base + (long)(k + j) * sk] : (uint16_t)0;
  uint4 r;
  r.x = v[0] | (uint32_t)v[1] << 16;
  r.y = v[2] | (uint32_t)v[3] << 16;
  r.z = v[4] | (uint32_t)v[5] << 16;
  r.w = v[6] | (uint32_t)v[7] << 16;
  return r;
}

__global__ __launch_bounds__(GEMM_THREADS) void k_gemm(const uint16_t *__restrict__ A, const uint16_t *__restrict__ B,
                                                       float *C, uint16_t *Cb, int M, int N, int K, long sab,
                                                       long sam, long sak, long sbb, long sbk, long sbn, long scb,
                                                       long scm, long scn, float alpha, float beta) {
  // A tile by row (m, k) and B tile by column (n, k): a lane's eight K values
  // of either operand are one 16-byte LDS read
  __shared__ __align__(16) uint16_t As[TM][TKP];
  __shared__ __align__(16) uint16_t Bs[TN][TKP];
  const int b = blockIdx.z;
  const int m0 = blockIdx.y * TM, n0 = blockIdx.x * TN;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wm = (wave >> 1) * 32, wn = (wave & 1) * 32;
  A += (long)b * sab;
  B += (long)b * sbb;
  // staging: each thread one row (column) and one group of eight K values.
  // When the matrix is contiguous along M (N), consecutive threads take
  // consecutive rows, so each of the eight element loads is coalesced;
  // otherwise four threads cover one row's 32 K values.
  const bool a_rows = sam == 1 && sak != 1;
  const int am = a_rows ? tid & 63 : tid >> 2, ak = a_rows ? (tid >> 6) * 8 : (tid & 3) * 8;
  const bool b_cols = sbn == 1 && sbk != 1;
  const int bn = b_cols ? tid & 63 : tid >> 2, bk = b_cols ? (tid >> 6) * 8 : (tid & 3) * 8;
  const int gm = m0 + am, hn = n0 + bn;
  const long abase = (long)gm * sam, bbase = (long)hn * sbn;
  f32x16 acc;
#pragma unroll
  for (int v = 0; v < 16; ++v) acc[v] = 0.0f;
  for (int k0 = 0; k0 < K; k0 += TK) {
    const uint4 av = load8(A, abase, sak, gm < M, k0 + ak, K);
    const uint4 bv = load8(B, bbase, sbk, hn < N, k0 + bk, K);
    *(uint4 *)&As[am][ak] = av;
    *(uint4 *)&Bs[bn][bk] = bv;
    __syncthreads();
    // lane l holds A(row l % 32, k 8 (l / 32) + j) and B(k 8 (l / 32) + j, col l % 32), j = 0..7
    const int r = lane & 31, kq = 8 * (lane >> 5);
#pragma unroll
    for (int ks = 0; ks < TK; ks += 16) {
      const bf16x8 af = *(const bf16x8 *)&As[wm + r][ks + kq];
      const bf16x8 bfr = *(const bf16x8 *)&Bs[wn + r][ks + kq];
      acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(af, bfr, acc, 0, 0, 0);
    }
    __syncthreads();
  }
  // D element v of lane l: row 8 (v / 4) + 4 (l / 32) + v % 4, column l % 32
  const int col = n0 + wn + (lane & 31);
  if (col >= N) return;
#pragma unroll
  for (int v = 0; v < 16; ++v) {
    const int row = m0 + wm + 8 * (v >> 2) + 4 * (lane >> 5) + (v & 3);
    if (row >= M) continue;
    const long o = (long)b * scb + (long)row * scm + (long)col * scn;
    float out = alpha * acc[v];
    if (beta != 0.0f) out += beta * C[o];
    if (C) C[o] = out;
    if (Cb) Cb[o] = f2bf(out);
  }
}

constexpr int EW = 256;
inline unsigned grid_for(long n) { return (unsigned)((n + EW - 1) / EW); }

// cols[(c 3 + t), b L + l] = x[c, b L + l + t - 1] (zero outside the sequence)
__global__ void k_im2col(const uint16_t *x, uint16_t *cols, int C, int Bn, int L) {
  const long P = (long)Bn * L, n = (long)C * 3 * P;
  const long i = (long)blockIdx.x * EW + threadIdx.x;
  if (i >= n) return;
  const long p = i % P;
  const int row = (int)(i / P), c = row / 3, t = row % 3;
  const int l = (int)(p % L), src = l + t - 1;
  cols[i] = (src >= 0 && src < L) ? x[(long)c * P + p + (t - 1)] : (uint16_t)0;
}

// dx[c, b L + l] = sum_t dcols[(c 3 + t), b L + l - t + 1]
__global__ void k_col2im(const float *dcols, float *dx, int C, int Bn, int L, float beta) {
  const long P = (long)Bn * L, n = (long)C * P;
  const long i = (long)blockIdx.x * EW + threadIdx.x;
  if (i >= n) return;
  const int c = (int)(i / P);
  const long p = i % P;
  const int l = (int)(p % L);
  float s = 0.0f;
#pragma unroll
  for (int t = 0; t < 3; ++t) {
    const int dst = l - t + 1;
    if (dst >= 0 && dst < L) s += dcols[(long)(c * 3 + t) * P + p - t + 1];
  }
  dx[i] = s + (beta != 0.0f ? beta * dx[i] : 0.0f);
}

__device__ __forceinline__ float gelu(float x) {
  const float k = 0.7978845608f; // sqrt(2 / pi)
  return 0.5f * x * (1.0f + tanhf(k * (x + 0.044715f * x * x * x)));
}
__device__ __forceinline__ float gelu_grad(float x) {
  const float k = 0.7978845608f;
  const float u = k * (x + 0.044715f * x * x * x), t = tanhf(u);
  return 0.5f * (1.0f + t) + 0.5f * x * (1.0f - t * t) * k * (1.0f + 3.0f * 0.044715f * x * x);
}

// pre[c, p] += bias[c]; y = gelu(pre) (+ res when given); y in FP32 and BF16
__global__ void k_bias_gelu(float *pre, const float *bias, const float *res, float *y, uint16_t *yb, int C, long P) {
  const long i = (long)blockIdx.x * EW + threadIdx.x;
  if (i >= (long)C * P) return;
  const float v = pre[i] + bias[i / P];
  pre[i] = v;
  float o = gelu(v);
  if (res) o += res[i];
  if (y) y[i] = o;
  yb[i] = f2bf(o);
}

// g = dy * gelu'(pre), in FP32 and BF16; dbias[c] accumulates over p
__global__ void k_gelu_back(const float *pre, const float *dy, float *g, uint16_t *gb, float *dbias, int C, long P) {
  const int c = blockIdx.y;
  float part = 0.0f;
  for (long p = (long)blockIdx.x * EW + threadIdx.x; p < P; p += (long)gridDim.x * EW) {
    const long i = (long)c * P + p;
    const float v = dy[i] * gelu_grad(pre[i]);
    g[i] = v;
    gb[i] = f2bf(v);
    part += v;
  }
  __shared__ float red[EW];
  red[threadIdx.x] = part;
  __syncthreads();
  for (int o = EW / 2; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0) atomicAdd(dbias + c, red[0]);
}

// Pair map of sample b: S = (M + M^T) / 2 + dist[|i - j|] + c0, with
// M = raw[b] (L x L); pred written in FP32.
__global__ void k_pair_map(const float *raw, const float *dist, const float *c0, float *pred, int L, int Bn) {
  const long i = (long)blockIdx.x * EW + threadIdx.x;
  if (i >= (long)Bn * L * L) return;
  const int b = (int)(i / ((long)L * L)), r = (int)((i / L) % L), c = (int)(i % L);
  const float *M = raw + (long)b * L * L;
  const int d = r > c ? r - c : c - r;
  pred[i] = 0.5f * (M[(long)r * L + c] + M[(long)c * L + r]) + dist[d] + c0[0];
}

// Loss of the pair maps over the pairs r <= c (mean squared error against
// target, weight w) and its gradient. gS(r, c) = dLoss/dS for r <= c; the
// raw map M enters S(lo, hi) as (M(lo, hi) + M(hi, lo)) / 2, so
// graw(r, c) = gS(lo, hi) / 2 off the diagonal and gS(r, r) on it. One block
// per row (b, r); the block's squared errors are summed into loss[0].
__global__ void k_pair_loss(const float *pred, const float *target, float *gS, float *graw, uint16_t *grawb,
                            float *loss, int L, int Bn, float w) {
  const long row = blockIdx.x; // b L + r
  const int r = (int)(row % L);
  const long base = (row - r) * L; // start of sample b
  const float count = (float)Bn * L * (L + 1) / 2;
  float sq = 0.0f;
  for (int c = threadIdx.x; c < L; c += blockDim.x) {
    const int lo = r < c ? r : c, hi = r < c ? c : r;
    const long u = base + (long)lo * L + hi;
    const float e = pred[u] - target[u];
    const float g = 2.0f * w * e / count;
    const float gr = (r == c) ? g : 0.5f * g;
    const long i = row * L + c;
    graw[i] = gr;
    grawb[i] = f2bf(gr);
    if (r <= c) {
      gS[i] = g;
      sq += w * e * e / count;
    } else {
      gS[i] = 0.0f;
    }
  }
  __shared__ float red[EW];
  red[threadIdx.x] = sq;
  __syncthreads();
  for (int o = EW / 2; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0) atomicAdd(loss, red[0]);
}

// gdist[d] = sum over samples and rows of gS(r, r + d); gc0 = sum of all.
__global__ void k_pair_reduce(const float *gS, float *gdist, float *gc0, int L, int Bn) {
  const int d = blockIdx.x;
  float s = 0.0f;
  const long n = (long)Bn * (L - d);
  for (long k = threadIdx.x; k < n; k += blockDim.x) {
    const long b = k / (L - d);
    const int r = (int)(k % (L - d));
    s += gS[(b * L + r) * L + r + d];
  }
  __shared__ float red[EW];
  red[threadIdx.x] = s;
  __syncthreads();
  for (int o = EW / 2; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    gdist[d] += red[0];
    atomicAdd(gc0, red[0]);
  }
}

// pooled[c, b] = mean over l of h[c, b L + l], FP32 and BF16
__global__ void k_mean_pool(const float *h, float *pooled, uint16_t *pooledb, int C, int Bn, int L) {
  const long i = (long)blockIdx.x * EW + threadIdx.x;
  if (i >= (long)C * Bn) return;
  const int c = (int)(i / Bn), b = (int)(i % Bn);
  const float *row = h + (long)c * Bn * L + (long)b * L;
  float s = 0.0f;
  for (int l = 0; l < L; ++l) s += row[l];
  s /= (float)L;
  pooled[i] = s;
  pooledb[i] = f2bf(s);
}

// dh[c, b L + l] += dpooled[c, b] / L
__global__ void k_pool_back(const float *dpooled, float *dh, int C, int Bn, int L) {
  const long i = (long)blockIdx.x * EW + threadIdx.x;
  if (i >= (long)C * Bn * L) return;
  const long P = (long)Bn * L;
  const int c = (int)(i / P), b = (int)((i % P) / L);
  dh[i] += dpooled[(long)c * Bn + b] / (float)L;
}

// Radial head: pred[o, b] += bias[o]; loss of w * mean squared error
// against target; g = dLoss/dpred (FP32, BF16); dbias accumulates. One block.
__global__ void k_rd_loss(float *pred, const float *bias, const float *target, float *g, uint16_t *gb, float *dbias,
                          float *loss, int O, int Bn, float w) {
  float sq = 0.0f;
  const float count = (float)O * Bn;
  for (int i = threadIdx.x; i < O * Bn; i += blockDim.x) {
    const int o = i / Bn;
    const float v = pred[i] + bias[o];
    pred[i] = v;
    const float e = v - target[i];
    const float gi = 2.0f * w * e / count;
    g[i] = gi;
    gb[i] = f2bf(gi);
    sq += w * e * e / count;
    atomicAdd(dbias + o, gi);
  }
  __shared__ float red[EW];
  red[threadIdx.x] = sq;
  __syncthreads();
  for (int o = EW / 2; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0) atomicAdd(loss, red[0]);
}

__global__ void k_cast(const float *x, uint16_t *y, long n) {
  const long i = (long)blockIdx.x * EW + threadIdx.x;
  if (i < n) y[i] = f2bf(x[i]);
}

// AdamW (Loshchilov & Hutter) with bias correction; writes the BF16 copy.
__global__ void k_adamw(float *p, const float *g, float *m, float *v, uint16_t *pb, long n, float lr, float b1,
                        float b2, float eps, float wd, float c1, float c2) {
  const long i = (long)blockIdx.x * EW + threadIdx.x;
  if (i >= n) return;
  const float gi = g[i];
  const float mi = b1 * m[i] + (1.0f - b1) * gi;
  const float vi = b2 * v[i] + (1.0f - b2) * gi * gi;
  m[i] = mi;
  v[i] = vi;
  float pi = p[i] * (1.0f - lr * wd);
  pi -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
  p[i] = pi;
  if (pb) pb[i] = f2bf(pi);
}

} // namespace

extern "C" {

int ccp_nn_gemm(const void *A, const void *B, float *C, void *Cb, int M, int N, int K, int batch, long sab, long sam,
                long sak, long sbb, long sbk, long sbn, long scb, long scm, long scn, float alpha, float beta,
                void *stream) {
  if (M <= 0 || N <= 0 || K <= 0 || batch <= 0 || (!C && !Cb) || (beta != 0.0f && !C)) return 1;
  const dim3 grid((unsigned)((N + TN - 1) / TN), (unsigned)((M + TM - 1) / TM), (unsigned)batch);
  k_gemm<<<grid, GEMM_THREADS, 0, (hipStream_t)stream>>>((const uint16_t *)A, (const uint16_t *)B, C,
                                                         (uint16_t *)Cb, M, N, K, sab, sam, sak, sbb, sbk, sbn, scb,
                                                         scm, scn, alpha, beta);
  return hipGetLastError() == hipSuccess ? 0 : 2;
}

int ccp_nn_im2col(const void *x, void *cols, int C, int Bn, int L, void *stream) {
  const long n = (long)C * 3 * Bn * L;
  k_im2col<<<grid_for(n), EW, 0, (hipStream_t)stream>>>((const uint16_t *)x, (uint16_t *)cols, C, Bn, L);
  return hipGetLastError() == hipSuccess ? 0 : 2;
}

int ccp_nn_col2im(const float *dcols, float *dx, int C, int Bn, int L, float beta, void *stream) {
  const long n = (long)C * Bn * L;
  k_col2im<<<grid_for(n), EW, 0, (hipStream_t)stream>>>(dcols, dx, C, Bn, L, beta);
  return hipGetLastError() == hipSuccess ? 0 : 2;
}

int ccp_nn_bias_gelu(float *pre, const float *bias, const float *res, float *y, void *yb, int C, long P,
                     void *stream) {
  k_bias_gelu<<<grid_for((long)C * P), EW, 0, (hipStream_t)stream>>>(pre, bias, res, y, (uint16_t *)yb, C, P);
  return hipGetLastError() == hipSuccess ? 0 : 2;
}

int ccp_nn_gelu_back(const float *pre, const float *dy, float *g, void *gb, float *dbias, int C, long P,
                     void *stream) {
  const unsigned bx = (unsigned)((P + EW - 1) / EW < 64 ? (P + EW - 1) / EW : 64);
  k_gelu_back<<<dim3(bx, (unsigned)C), EW, 0, (hipStream_t)stream>>>(pre, dy, g, (uint16_t *)gb, dbias, C, P);
  return hipGetLastError() == hipSuccess ? 0 : 2;
}

int ccp_nn_pair_map(const float *raw, const float *dist, const float *c0, float *pred, int L, int Bn, void *stream) {
  k_pair_map<<<grid_for((long)Bn * L * L), EW, 0, (hipStream_t)stream>>>(raw, dist, c0, pred, L, Bn);
  return hipGetLastError() == hipSuccess ? 0 : 2;
}

int ccp_nn_pair_loss(const float *pred, const float *target, float *gS, float *graw, void *grawb, float *gdist,
                     float *gc0, float *loss, int L, int Bn, float w, void *stream) {
  k_pair_loss<<<(unsigned)(Bn * L), EW, 0, (hipStream_t)stream>>>(pred, target, gS, graw, (uint16_t *)grawb, loss, L,
                                                                  Bn, w);
  k_pair_reduce<<<(unsigned)L, EW, 0, (hipStream_t)stream>>>(gS, gdist, gc0, L, Bn);
  return hipGetLastError() == hipSuccess ? 0 : 2;
}

int ccp_nn_mean_pool(const float *h, float *pooled, void *pooledb, int C, int Bn, int L, void *stream) {
  k_mean_pool<<<grid_for((long)C * Bn), EW, 0, (hipStream_t)stream>>>(h, pooled, (uint16_t *)pooledb, C, Bn, L);
  return hipGetLastError() == hipSuccess ? 0 : 2;
}

int ccp_nn_pool_back(const float *dpooled, float *dh, int C, int Bn, int L, void *stream) {
  k_pool_back<<<grid_for((long)C * Bn * L), EW, 0, (hipStream_t)stream>>>(dpooled, dh, C, Bn, L);
  return hipGetLastError() == hipSuccess ? 0 : 2;
}

int ccp_nn_rd_loss(float *pred, const float *bias, const float *target, float *g, void *gb, float *dbias, float *loss,
                   int O, int Bn, float w, void *stream) {
  k_rd_loss<<<1, EW, 0, (hipStream_t)stream>>>(pred, bias, target, g, (uint16_t *)gb, dbias, loss, O, Bn, w);
  return hipGetLastError() == hipSuccess ? 0 : 2;
}

int ccp_nn_cast_bf16(const float *x, void *y, long n, void *stream) {
  k_cast<<<grid_for(n), EW, 0, (hipStream_t)stream>>>(x, (uint16_t *)y, n);
  return hipGetLastError() == hipSuccess ? 0 : 2;
}

int ccp_nn_adamw(float *p, const float *g, float *m, float *v, void *pb, long n, float lr, float b1, float b2,
                 float eps, float wd, int t, void *stream) {
  const float c1 = 1.0f - powf(b1, (float)t), c2 = 1.0f - powf(b2, (float)t);
  k_adamw<<<grid_for(n), EW, 0, (hipStream_t)stream>>>(p, g, m, v, (uint16_t *)pb, n, lr, b1, b2, eps, wd, c1, c2);
  return hipGetLastError() == hipSuccess ? 0 : 2;
}

} // extern "C"
