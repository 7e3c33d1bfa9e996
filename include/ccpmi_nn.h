/*
 * ccpmi_nn — kernels of the contact-map surrogate (no counterpart in ccp).
 *
 * A 1-D convolutional encoder over the binned het/LAD sequence with a
 * pairwise outer-product head, trained and evaluated with these kernels:
 * a strided batched BF16 GEMM on the matrix cores with FP32 accumulation,
 * and the element-wise steps around it. Activations are stored
 * channel-major: x[c, b L + l] for channel c, sample b, bin l. BF16 values
 * are passed as raw 16-bit words (void pointers). Every call is
 * asynchronous on `stream` (a hipStream_t) and returns 0, 1 for bad sizes,
 * or 2 for a launch failure.
 */
#ifndef CCPMI_NN_H
#define CCPMI_NN_H

#ifdef __cplusplus
extern "C" {
#endif

/* C[b](m, n) = alpha sum_k A[b](m, k) B[b](k, n) + beta C[b](m, n), with
 * X[b](i, j) at X + b sxb + i sxi + j sxj; A, B BF16. The result is written
 * to C (FP32, may be NULL when beta = 0) and/or Cb (BF16). */
int ccp_nn_gemm(const void *A, const void *B, float *C, void *Cb, int M, int N, int K, int batch, long sab, long sam,
                long sak, long sbb, long sbk, long sbn, long scb, long scm, long scn, float alpha, float beta,
                void *stream);
/* kernel-3, padding-1 convolution as a GEMM: cols[(c 3 + t), p] = x[c, p + t - 1] within each sample */
int ccp_nn_im2col(const void *x, void *cols, int C, int Bn, int L, void *stream);
/* its adjoint: dx = col2im(dcols) + beta dx */
int ccp_nn_col2im(const float *dcols, float *dx, int C, int Bn, int L, float beta, void *stream);
/* pre += bias (per channel); y = gelu(pre) + res (res may be NULL); y FP32 (may be NULL) and yb BF16 */
int ccp_nn_bias_gelu(float *pre, const float *bias, const float *res, float *y, void *yb, int C, long P,
                     void *stream);
/* g = dy gelu'(pre) (FP32 and BF16); dbias += sum over positions */
int ccp_nn_gelu_back(const float *pre, const float *dy, float *g, void *gb, float *dbias, int C, long P,
                     void *stream);
/* pred[b](i, j) = (raw[b](i, j) + raw[b](j, i)) / 2 + dist[|i - j|] + c0[0] */
int ccp_nn_pair_map(const float *raw, const float *dist, const float *c0, float *pred, int L, int Bn, void *stream);
/* w * mean squared error over i <= j; gradients of raw (FP32, BF16), dist and c0 (accumulated); loss += */
int ccp_nn_pair_loss(const float *pred, const float *target, float *gS, float *graw, void *grawb, float *gdist,
                     float *gc0, float *loss, int L, int Bn, float w, void *stream);
int ccp_nn_mean_pool(const float *h, float *pooled, void *pooledb, int C, int Bn, int L, void *stream);
int ccp_nn_pool_back(const float *dpooled, float *dh, int C, int Bn, int L, void *stream);
/* pred += bias; w * mean squared error; g = dLoss/dpred (FP32, BF16); dbias += ; loss += */
int ccp_nn_rd_loss(float *pred, const float *bias, const float *target, float *g, void *gb, float *dbias, float *loss,
                   int O, int Bn, float w, void *stream);
int ccp_nn_cast_bf16(const float *x, void *y, long n, void *stream);
/* AdamW step t (1-based) on FP32 parameters; pb receives the BF16 copy (may be NULL) */
int ccp_nn_adamw(float *p, const float *g, float *m, float *v, void *pb, long n, float lr, float b1, float b2,
                 float eps, float wd, int t, void *stream);

#ifdef __cplusplus
}
#endif

#endif
