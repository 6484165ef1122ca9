/* grand_mlp.h -- BF16 MFMA kernels of the learned search order's MLP (gfx950).
 *
 * No counterpart in the reference: the learned search order is an addition.
 * Pointers are device pointers; `stream` is a hipStream_t (NULL = the default
 * stream).  Each call returns 0 (hipSuccess) or a HIP error code.
 */
#ifndef GRAND_MLP_H
#define GRAND_MLP_H

#ifdef __cplusplus
extern "C" {
#endif

/* C[M][N] = op(A)[M][K] * op(B)[K][N] on bf16 operands with fp32 accumulation.
 * a_layout: 0 = A[m*lda+k], 1 = A[k*lda+m]; b_layout: 0 = B[n*ldb+k],
 * 1 = B[k*ldb+n].  epilogue: 0 = store (+bias[N], relu, fp32 C and/or bf16 Cb),
 * 1 = bf16 Cb = C where the bf16 mask[m*ldm+n] > 0 else 0, 2 = atomic add into
 * fp32 C with K split into `splits` ranges. */
int grand_mlp_gemm(int a_layout, int b_layout, int epilogue, const void *A, int lda, const void *B, int ldb, int M,
                   int N, int K, float *C, int ldc, void *Cb, int ldcb, const float *bias, int relu, const void *mask,
                   int ldm, int splits, void *stream);

/* Masked binary cross-entropy with logits z (fp32 [B][ldz]) against 0/1 targets
 * y and mask (uint8 [B][ldy]; mask NULL = every element counts): adds the masked loss sum to *loss and writes
 * the bf16 gradient scale * (sigmoid(z) - y) * mask to dz [B][lddz]. */
int grand_mlp_bce(const float *z, int ldz, const unsigned char *y, const unsigned char *mask, int ldy, int B, int N,
                  float scale, void *dz, int lddz, float *loss, void *stream);

/* out[n] += sum over r < R of the bf16 X[r*ldx+n] (bias gradients). */
int grand_mlp_colsum(const void *X, int ldx, int R, int N, float *out, void *stream);

/* Adam step `step` (1-based) on fp32 weights w with moments m, v and gradient g;
 * also writes the bf16 copy of the weights to wb when wb is not NULL. */
int grand_mlp_adam(float *w, float *m, float *v, const float *g, int n, float lr, float b1, float b2, float eps,
                   int step, void *wb, void *stream);

/* int32 0/1 rows [rows][cols] -> bf16 rows [rows][ld_out] (padding columns 0). */
int grand_mlp_rows_to_bf16(const int *in, int cols, int rows, void *out, int ld_out, void *stream);

#ifdef __cplusplus
}
#endif

#endif /* GRAND_MLP_H */
