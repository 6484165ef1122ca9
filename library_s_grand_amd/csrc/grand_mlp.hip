// grand_mlp.hip -- BF16 MFMA kernels of the learned search order (gfx950).
//
// A small MLP maps a shot's detector bits to scores of the active positions;
// the search then tries positions in score order (see the learned pass in
// grand_kernels.hip).  This file holds its training and inference kernels:
//
//  * mlp_gemm_kernel   C = op(A) * op(B) on v_mfma_f32_16x16x32_bf16, bf16
//                      operands staged through LDS, fp32 accumulation.  One
//                      256-thread workgroup computes a 64 x 64 tile, each wave
//                      a 32 x 32 quarter (2 x 2 MFMA tiles), 32 K per step.
//                      Operand layouts: A as [M][K] (K contiguous) or [K][M];
//                      B as [N][K] (a weight matrix used transposed) or [K][N].
//                      Epilogues: bias + optional ReLU to fp32 and/or bf16;
//                      multiply by a ReLU mask (backward through ReLU); or
//                      fp32 atomic add (split-K weight gradients).
//  * bce_grad_kernel   masked binary cross-entropy on logits: loss sum and
//                      dL/dz (bf16) in one pass.
//  * colsum_kernel     bias gradients (column sums of a bf16 matrix).
//  * adam_kernel       Adam on fp32 master weights, refreshing the bf16 copy.

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace grand_mlp {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ float bf2f(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }
// round to nearest even (inputs here are finite)
__device__ __forceinline__ uint16_t f2bf(float f) {
    uint32_t u = __float_as_uint(f);
    u += 0x7FFFu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}

enum { A_MK = 0, A_KM = 1 };          // A element (m, k) at A[m*lda + k] or A[k*lda + m]
enum { B_NK = 0, B_KN = 1 };          // B element (k, n) at B[n*ldb + k] or B[k*ldb + n]
enum { EPI_STORE = 0, EPI_MASK = 1, EPI_ATOMIC = 2 };

constexpr int BM = 64, BN = 64, BK = 32;
constexpr int LDK = BK + 8;           // padded LDS row (bf16 elements): 80 bytes, 16-byte aligned

struct GemmArgs {
    const uint16_t *A; int lda;
    const uint16_t *B; int ldb;
    int M, N, K;
    int k_per_split;                  // K range of one blockIdx.z
    float *C; int ldc;                // fp32 output (STORE: optional; ATOMIC: accumulated)
    uint16_t *Cb; int ldcb;           // bf16 output (STORE / MASK)
    const float *bias;                // [N] (STORE, optional)
    const uint16_t *mask; int ldm;    // bf16 [M][N]: output kept where mask > 0 (MASK)
    int relu;                         // STORE: apply ReLU
};

// Loads the 64 x 32 tile of op(A) (rows m0.., k0..) into As[m][k].
template <int LA>
__device__ __forceinline__ void load_a(const GemmArgs &g, int m0, int k0, int kend, uint16_t *As) {
    const int t = threadIdx.x;
    if (LA == A_MK) {
        const int m = t >> 2, kc = (t & 3) * 8;           // 64 rows x 4 chunks of 8
        const int gm = m0 + m, gk = k0 + kc;
        uint16_t v[8];
        if (gm < g.M && gk + 8 <= kend && (g.lda & 7) == 0) {
            const uint4 q = *reinterpret_cast<const uint4 *>(g.A + (size_t)gm * g.lda + gk);
            *reinterpret_cast<uint4 *>(v) = q;
        } else {
#pragma unroll
            for (int i = 0; i < 8; ++i) v[i] = (gm < g.M && gk + i < kend) ? g.A[(size_t)gm * g.lda + gk + i] : 0;
        }
        *reinterpret_cast<uint4 *>(As + m * LDK + kc) = *reinterpret_cast<uint4 *>(v);
    } else {
        const int k = t >> 3, mc = (t & 7) * 8;           // 32 k-rows x 8 chunks of 8 m
        const int gk = k0 + k, gm = m0 + mc;
        uint16_t v[8];
        if (gk < kend && gm + 8 <= g.M && (g.lda & 7) == 0) {
            const uint4 q = *reinterpret_cast<const uint4 *>(g.A + (size_t)gk * g.lda + gm);
            *reinterpret_cast<uint4 *>(v) = q;
        } else {
#pragma unroll
            for (int i = 0; i < 8; ++i) v[i] = (gk < kend && gm + i < g.M) ? g.A[(size_t)gk * g.lda + gm + i] : 0;
        }
#pragma unroll
        for (int i = 0; i < 8; ++i) As[(mc + i) * LDK + k] = v[i];
    }
}

// Loads the 32 x 64 tile of op(B) (k0.., cols n0..) into Bs[n][k].
template <int LB>
__device__ __forceinline__ void load_b(const GemmArgs &g, int n0, int k0, int kend, uint16_t *Bs) {
    const int t = threadIdx.x;
    if (LB == B_NK) {
        const int n = t >> 2, kc = (t & 3) * 8;
        const int gn = n0 + n, gk = k0 + kc;
        uint16_t v[8];
        if (gn < g.N && gk + 8 <= kend && (g.ldb & 7) == 0) {
            const uint4 q = *reinterpret_cast<const uint4 *>(g.B + (size_t)gn * g.ldb + gk);
            *reinterpret_cast<uint4 *>(v) = q;
        } else {
#pragma unroll
            for (int i = 0; i < 8; ++i) v[i] = (gn < g.N && gk + i < kend) ? g.B[(size_t)gn * g.ldb + gk + i] : 0;
        }
        *reinterpret_cast<uint4 *>(Bs + n * LDK + kc) = *reinterpret_cast<uint4 *>(v);
    } else {
        const int k = t >> 3, nc = (t & 7) * 8;
        const int gk = k0 + k, gn = n0 + nc;
        uint16_t v[8];
        if (gk < kend && gn + 8 <= g.N && (g.ldb & 7) == 0) {
            const uint4 q = *reinterpret_cast<const uint4 *>(g.B + (size_t)gk * g.ldb + gn);
            *reinterpret_cast<uint4 *>(v) = q;
        } else {
#pragma unroll
            for (int i = 0; i < 8; ++i) v[i] = (gk < kend && gn + i < g.N) ? g.B[(size_t)gk * g.ldb + gn + i] : 0;
        }
#pragma unroll
        for (int i = 0; i < 8; ++i) Bs[(nc + i) * LDK + k] = v[i];
    }
}

// MFMA operand of 16 rows (row = lane % 16) x 8 consecutive k starting at
// (lane / 16) * 8, from a [row][k] LDS tile.
__device__ __forceinline__ bf16x8 frag(const uint16_t *S, int row0) {
    const int l = threadIdx.x & 63;
    const uint4 q = *reinterpret_cast<const uint4 *>(S + (row0 + (l & 15)) * LDK + (l >> 4) * 8);
    return __builtin_bit_cast(bf16x8, q);
}

template <int LA, int LB, int EPI>
__global__ __launch_bounds__(256) void mlp_gemm_kernel(const GemmArgs g) {
    __shared__ __attribute__((aligned(16))) uint16_t As[BM * LDK];
    __shared__ __attribute__((aligned(16))) uint16_t Bs[BN * LDK];
    const int m0 = blockIdx.y * BM, n0 = blockIdx.x * BN;
    const int kbeg = blockIdx.z * g.k_per_split;
    const int kend = (kbeg + g.k_per_split < g.K) ? kbeg + g.k_per_split : g.K;
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int wm = (wave >> 1) * 32, wn = (wave & 1) * 32;
    f32x4 acc[2][2];
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j) acc[i][j] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
    for (int k0 = kbeg; k0 < kend; k0 += BK) {
        load_a<LA>(g, m0, k0, kend, As);
        load_b<LB>(g, n0, k0, kend, Bs);
        __syncthreads();
        const bf16x8 a0 = frag(As, wm), a1 = frag(As, wm + 16);
        const bf16x8 b0 = frag(Bs, wn), b1 = frag(Bs, wn + 16);
        acc[0][0] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a0, b0, acc[0][0], 0, 0, 0);
        acc[0][1] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a0, b1, acc[0][1], 0, 0, 0);
        acc[1][0] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a1, b0, acc[1][0], 0, 0, 0);
        acc[1][1] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a1, b1, acc[1][1], 0, 0, 0);
        __syncthreads();
    }
    // D layout: column = lane % 16, rows (lane / 16) * 4 + r
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int row = m0 + wm + i * 16 + (lane >> 4) * 4 + r;
                const int col = n0 + wn + j * 16 + (lane & 15);
                if (row >= g.M || col >= g.N) continue;
                float v = acc[i][j][r];
                if (EPI == EPI_ATOMIC) {
                    atomicAdd(&g.C[(size_t)row * g.ldc + col], v);
                } else if (EPI == EPI_MASK) {
                    const bool on = bf2f(g.mask[(size_t)row * g.ldm + col]) > 0.0f;
                    g.Cb[(size_t)row * g.ldcb + col] = f2bf(on ? v : 0.0f);
                } else {
                    if (g.bias) v += g.bias[col];
                    if (g.relu) v = v > 0.0f ? v : 0.0f;
                    if (g.C) g.C[(size_t)row * g.ldc + col] = v;
                    if (g.Cb) g.Cb[(size_t)row * g.ldcb + col] = f2bf(v);
                }
            }
}

// Masked BCE on logits z [B][N] (fp32) with 0/1 labels y and 0/1 mask:
// dz = mask (sigmoid(z) - y) * scale  (bf16), loss += mask (softplus(z) - y z).
__global__ __launch_bounds__(256) void bce_grad_kernel(const float *z, int ldz, const uint8_t *y, const uint8_t *mask,
                                                       int ldy, int B, int N, float scale, uint16_t *dz, int lddz,
                                                       float *loss) {
    __shared__ float red[256];
    float acc = 0.0f;
    const size_t total = (size_t)B * N;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < total; i += (size_t)gridDim.x * blockDim.x) {
        const int b = (int)(i / N), c = (int)(i - (size_t)b * N);
        const float v = z[(size_t)b * ldz + c];
        const float on = mask ? (float)mask[(size_t)b * ldy + c] : 1.0f;
        const float t = (float)y[(size_t)b * ldy + c];
        const float sp = (v > 0.0f ? v : 0.0f) + log1pf(expf(-fabsf(v)));   // softplus(v)
        acc += on * (sp - t * v);
        const float sg = 1.0f / (1.0f + expf(-v));
        dz[(size_t)b * lddz + c] = f2bf(on * (sg - t) * scale);
    }
    red[threadIdx.x] = acc;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) atomicAdd(loss, red[0]);
}

// out[c] += sum over rows of X[r][c] (bf16 [R][N]); block = 64 columns x 4 row lanes.
__global__ __launch_bounds__(256) void colsum_kernel(const uint16_t *X, int ldx, int R, int N, int rows_per_block,
                                                     float *out) {
    __shared__ float red[4][64];
    const int c = blockIdx.x * 64 + (threadIdx.x & 63);
    const int ty = threadIdx.x >> 6;
    const int r0 = blockIdx.y * rows_per_block;
    const int r1 = (r0 + rows_per_block < R) ? r0 + rows_per_block : R;
    float s = 0.0f;
    if (c < N)
        for (int r = r0 + ty; r < r1; r += 4) s += bf2f(X[(size_t)r * ldx + c]);
    red[ty][threadIdx.x & 63] = s;
    __syncthreads();
    if (ty == 0 && c < N) atomicAdd(&out[c], red[0][threadIdx.x] + red[1][threadIdx.x] + red[2][threadIdx.x] +
                                                 red[3][threadIdx.x]);
}

__global__ __launch_bounds__(256) void adam_kernel(float *w, float *m, float *v, const float *g, int n, float lr,
                                                   float b1, float b2, float eps, float bc1, float bc2,
                                                   uint16_t *wb) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float gi = g[i];
    const float mi = b1 * m[i] + (1.0f - b1) * gi;
    const float vi = b2 * v[i] + (1.0f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    const float wi = w[i] - lr * (mi / bc1) / (sqrtf(vi / bc2) + eps);
    w[i] = wi;
    if (wb) wb[i] = f2bf(wi);
}

// int32 0/1 rows -> bf16 rows (ld_out >= cols; padding columns set to 0).
__global__ __launch_bounds__(256) void rows_to_bf16_kernel(const int *in, int cols, int rows, uint16_t *out,
                                                           int ld_out) {
    const size_t total = (size_t)rows * ld_out;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < total; i += (size_t)gridDim.x * blockDim.x) {
        const int r = (int)(i / ld_out), c = (int)(i - (size_t)r * ld_out);
        out[i] = (c < cols && in[(size_t)r * cols + c] != 0) ? (uint16_t)0x3F80u : (uint16_t)0u;
    }
}

}  // namespace grand_mlp

// ------------------------------------------------------------------ C ABI
extern "C" {
using namespace grand_mlp;

// C[M][N] = op(A)[M][K] * op(B)[K][N].  a_layout: 0 = A[m*lda+k], 1 = A[k*lda+m];
// b_layout: 0 = B[n*ldb+k], 1 = B[k*ldb+n].  epilogue: 0 = store (bias, relu,
// fp32 C and/or bf16 Cb), 1 = bf16 Cb = C where mask > 0 else 0, 2 = atomic
// add into fp32 C with K split into `splits` ranges.
hipError_t grand_mlp_gemm(int a_layout, int b_layout, int epilogue, const void *A, int lda, const void *B, int ldb,
                          int M, int N, int K, float *C, int ldc, void *Cb, int ldcb, const float *bias, int relu,
                          const void *mask, int ldm, int splits, hipStream_t st) {
    if (M <= 0 || N <= 0) return hipSuccess;
    if (splits < 1 || epilogue != EPI_ATOMIC) splits = 1;
    GemmArgs g;
    g.A = (const uint16_t *)A; g.lda = lda; g.B = (const uint16_t *)B; g.ldb = ldb;
    g.M = M; g.N = N; g.K = K;
    int kps = (K + splits - 1) / splits;
    kps = (kps + BK - 1) / BK * BK;
    g.k_per_split = kps > 0 ? kps : BK;
    splits = (K + g.k_per_split - 1) / g.k_per_split;
    if (splits < 1) splits = 1;
    g.C = C; g.ldc = ldc; g.Cb = (uint16_t *)Cb; g.ldcb = ldcb; g.bias = bias; g.mask = (const uint16_t *)mask;
    g.ldm = ldm; g.relu = relu;
    dim3 grid((N + BN - 1) / BN, (M + BM - 1) / BM, splits);
#define GRAND_MLP_LAUNCH(LA, LB, EP) hipLaunchKernelGGL((mlp_gemm_kernel<LA, LB, EP>), grid, dim3(256), 0, st, g)
    if (epilogue == EPI_STORE) {
        if (a_layout == A_MK && b_layout == B_NK) GRAND_MLP_LAUNCH(A_MK, B_NK, EPI_STORE);
        else return hipErrorInvalidValue;
    } else if (epilogue == EPI_MASK) {
        if (a_layout == A_MK && b_layout == B_KN) GRAND_MLP_LAUNCH(A_MK, B_KN, EPI_MASK);
        else if (a_layout == A_MK && b_layout == B_NK) GRAND_MLP_LAUNCH(A_MK, B_NK, EPI_MASK);
        else return hipErrorInvalidValue;
    } else {
        if (a_layout == A_KM && b_layout == B_KN) GRAND_MLP_LAUNCH(A_KM, B_KN, EPI_ATOMIC);
        else if (a_layout == A_MK && b_layout == B_NK) GRAND_MLP_LAUNCH(A_MK, B_NK, EPI_ATOMIC);
        else return hipErrorInvalidValue;
    }
#undef GRAND_MLP_LAUNCH
    return hipGetLastError();
}

hipError_t grand_mlp_bce(const float *z, int ldz, const uint8_t *y, const uint8_t *mask, int ldy, int B, int N,
                         float scale, void *dz, int lddz, float *loss, hipStream_t st) {
    const size_t total = (size_t)B * N;
    int grid = (int)((total + 255) / 256);
    if (grid > 4096) grid = 4096;
    if (grid < 1) grid = 1;
    hipLaunchKernelGGL(bce_grad_kernel, dim3(grid), dim3(256), 0, st, z, ldz, y, mask, ldy, B, N, scale,
                       (uint16_t *)dz, lddz, loss);
    return hipGetLastError();
}

hipError_t grand_mlp_colsum(const void *X, int ldx, int R, int N, float *out, hipStream_t st) {
    const int rpb = 256;
    dim3 grid((N + 63) / 64, (R + rpb - 1) / rpb);
    hipLaunchKernelGGL(colsum_kernel, grid, dim3(256), 0, st, (const uint16_t *)X, ldx, R, N, rpb, out);
    return hipGetLastError();
}

hipError_t grand_mlp_adam(float *w, float *m, float *v, const float *g, int n, float lr, float b1, float b2,
                          float eps, int step, void *wb, hipStream_t st) {
    float bc1 = 1.0f, bc2 = 1.0f;
    for (int i = 0; i < step; ++i) { bc1 *= b1; bc2 *= b2; }
    bc1 = 1.0f - bc1;
    bc2 = 1.0f - bc2;
    hipLaunchKernelGGL(adam_kernel, dim3((n + 255) / 256), dim3(256), 0, st, w, m, v, g, n, lr, b1, b2, eps, bc1, bc2,
                       (uint16_t *)wb);
    return hipGetLastError();
}

hipError_t grand_mlp_rows_to_bf16(const int *in, int cols, int rows, void *out, int ld_out, hipStream_t st) {
    const size_t total = (size_t)rows * ld_out;
    int grid = (int)((total + 255) / 256);
    if (grid > 8192) grid = 8192;
    if (grid < 1) grid = 1;
    hipLaunchKernelGGL(rows_to_bf16_kernel, dim3(grid), dim3(256), 0, st, in, cols, rows, (uint16_t *)out, ld_out);
    return hipGetLastError();
}
}
