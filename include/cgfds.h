/* cgfds: kernels of the learned surrogate of surface seismograms, part of
 * libcgfdh.so.
 *
 * No counterpart in the reference: the surrogate maps source parameters
 * (x, y, depth, strike, dip, rake, f0) to 3-component surface seismograms for
 * one fixed grid and medium, trained on runs of the batched forward solver
 * (cgfdh_batch_*), which stays the fallback and the measure of its misfit.
 *
 * Storage conventions: bf16 is the upper 16 bits of an IEEE binary32 value
 * (conversion rounds to nearest, ties to even); matrices are row-major with
 * leading dimensions in elements; every pointer is device memory; `stream`
 * is a hipStream_t (NULL = the default stream).  Calls only enqueue work.
 * Return codes and cgfdh_last_error() follow include/cgfdh.h.
 */
#ifndef CGFDS_H
#define CGFDS_H

#ifdef __cplusplus
extern "C" {
#endif

/* Epilogues of cgfds_gemm_bf16 */
#define CGFDS_EPI_NONE 0 /* v = acc (+ bias[n]) */
#define CGFDS_EPI_RELU 1 /* v = max(acc + bias[n], 0) */
#define CGFDS_EPI_MASK 2 /* v = acc if bf16 mask[m][n] > 0 else 0 (ReLU backward) */

/* C[M][N] = sum_k opA[m][k] opB[k][n] on BF16 matrix cores, FP32
 * accumulation.  opA = A stored [M][lda] or, transA, A^T of A stored
 * [K][lda]; opB = B stored [K][ldb] or, transB, B^T of B stored [N][ldb].
 * The epilogue's v goes to C32 (FP32, optional) and/or C16 (bf16, optional).
 * M, N, K, lda, ldb must be multiples of 8 and A, B 16-byte aligned.
 * splitk > 1 splits the K range over that many workgroup layers whose FP32
 * partials (work: splitk*M*N floats) are summed in a fixed order into C32;
 * it requires EPI_NONE, no bias and no C16. */
int cgfds_gemm_bf16(int M, int N, int K, const void *A, long lda, int transA, const void *B,
                    long ldb, int transB, const float *bias, int epilogue, const void *mask,
                    long ldm, float *C32, long ldc32, void *C16, long ldc16, int splitk,
                    float *work, void *stream);

/* out[n] = sum over m of G[m][n] (bf16 G, FP32 sums), rows cut into `splits`
 * ranges whose partials (work: splits*N floats) are summed in a fixed order. */
int cgfds_colsum_bf16(int M, int N, const void *G, long ldg, float *out, int splits, float *work,
                      void *stream);

/* Per-sample relative squared misfit of nsamp samples of n values each:
 * loss[b] = |Y_b - T_b|^2 / (tnorm2[b] + eps), and (G16 optional) its
 * gradient with respect to Y of the mean over samples, stored as bf16. */
int cgfds_loss_grad(int nsamp, long n, const float *Y, const float *T, const float *tnorm2,
                    float eps, void *G16, float *loss, void *stream);

/* One Adam step on FP32 master values w with moments m, v and gradient g:
 * m = b1 m + c1 g; v = b2 v + c2 g^2; w -= lr (m/bc1) / (sqrt(v/bc2) + eps),
 * with c1 = 1-b1, c2 = 1-b2, bc1 = 1-b1^t, bc2 = 1-b2^t given by the caller;
 * w16 (optional) receives the bf16 copy of w. */
int cgfds_adam(long n, float *w, float *m, float *v, const float *g, void *w16, float lr,
               float b1, float b2, float c1, float c2, float bc1, float bc2, float eps,
               void *stream);

/* Moment-tensor contraction of basis waveforms (the seismograms are linear
 * in the moment tensor): y[b][t][c] = sum_k m[b][k] g[b][t][k][c] for
 * nsamp samples, nt times, nc channels, nk <= 8 basis terms. */
int cgfds_mt_contract(int nsamp, int nt, int nc, int nk, const float *m, const float *g,
                      float *y, void *stream);

/* The contraction fused with cgfds_loss_grad's per-sample misfit against T
 * [nsamp][nt][nc]; dg16 (optional) receives the bf16 gradient with respect
 * to g: m[b][k] times the misfit gradient with respect to y. */
int cgfds_mt_loss_grad(int nsamp, int nt, int nc, int nk, const float *m, const float *g,
                       const float *T, const float *tnorm2, float eps, void *dg16, float *loss,
                       void *stream);

/* dst[i] = bf16(src[i]) */
int cgfds_f32_to_bf16(long n, const float *src, void *dst, void *stream);

#ifdef __cplusplus
}
#endif

#endif
