// General strided, batched GEMM on MFMA (fp64 16x16x4, fp32 16x16x4) for the dense
// contractions of the path: filtering V*RF^T (replaces cublasDgemm at vtem/vrf.py:291),
// spatial projection RF*B, temporal projection, synthesis and the structured
// Kronecker mat-vec used by refinement and the RNN.
//
//   C[b][m, n] = alpha * sum_k A[b](m, k) B[b](k, n) + beta * C[b][m, n]
//   A(m,k) = A[m*sam + k*sak], B(k,n) = B[k*sbk + n*sbn], C[m*ldc + n].
//
// Tile 64x64x16 per 256-thread workgroup; 4 waves in a 2x2 arrangement, each wave a
// 32x32 sub-tile = 2x2 MFMA 16x16 tiles.  Operands staged through LDS k-major so each
// lane's MFMA operand is one LDS read.
#include "vt_common.hpp"

template <typename T> struct MfmaOp;
template <> struct MfmaOp<double> {
    typedef v4d acc_t;
    static __device__ __forceinline__ v4d op(double a, double b, v4d c) {
        return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
    }
};
template <> struct MfmaOp<float> {
    typedef v4f acc_t;
    static __device__ __forceinline__ v4f op(float a, float b, v4f c) {
        return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
    }
};

#define GBM 64
#define GBN 64
#define GBK 16

template <typename T>
__global__ void __launch_bounds__(256)
gemm_kernel(int M, int N, int K, T alpha, const T *__restrict__ A, long long sam, long long sak,
            long long bsa, const T *__restrict__ B, long long sbk, long long sbn, long long bsb,
            T beta, T *__restrict__ C, long long ldc, long long bsc)
{
    typedef typename MfmaOp<T>::acc_t acc_t;
    __shared__ T As[GBK][GBM + 4];
    __shared__ T Bs[GBK][GBN + 4];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int m0 = blockIdx.y * GBM, n0 = blockIdx.x * GBN;
    A += (long long)blockIdx.z * bsa;
    B += (long long)blockIdx.z * bsb;
    C += (long long)blockIdx.z * bsc;
    const int wm = (wave >> 1) * 32, wn = (wave & 1) * 32;
    acc_t acc[2][2];
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j) acc[i][j] = (acc_t){0, 0, 0, 0};

    const bool a_kfast = (sak == 1);   // row-major A: consecutive k contiguous
    const bool b_nfast = (sbn == 1);   // row-major B: consecutive n contiguous
    for (int k0 = 0; k0 < K; k0 += GBK) {
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            int e = tid + i * 256;
            int m, k;
            if (a_kfast) { k = e & 15; m = e >> 4; } else { m = e & 63; k = e >> 6; }
            int gm = m0 + m, gk = k0 + k;
            As[k][m] = (gm < M && gk < K) ? A[(long long)gm * sam + (long long)gk * sak] : (T)0;
            int n, kb;
            if (b_nfast) { n = e & 63; kb = e >> 6; } else { kb = e & 15; n = e >> 4; }
            int gn = n0 + n, gkb = k0 + kb;
            Bs[kb][n] = (gn < N && gkb < K) ? B[(long long)gkb * sbk + (long long)gn * sbn] : (T)0;
        }
        __syncthreads();
#pragma unroll
        for (int kk = 0; kk < GBK; kk += 4) {
            const int kr = kk + (lane >> 4);
            T a0 = As[kr][wm + (lane & 15)];
            T a1 = As[kr][wm + 16 + (lane & 15)];
            T b0 = Bs[kr][wn + (lane & 15)];
            T b1 = Bs[kr][wn + 16 + (lane & 15)];
            acc[0][0] = MfmaOp<T>::op(a0, b0, acc[0][0]);
            acc[0][1] = MfmaOp<T>::op(a0, b1, acc[0][1]);
            acc[1][0] = MfmaOp<T>::op(a1, b0, acc[1][0]);
            acc[1][1] = MfmaOp<T>::op(a1, b1, acc[1][1]);
        }
        __syncthreads();
    }
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                int gm = m0 + wm + i * 16 + acc_row<T>(lane, r);
                int gn = n0 + wn + j * 16 + mfma_col(lane);
                if (gm < M && gn < N) {
                    T v = alpha * acc[i][j][r];
                    T *p = C + (long long)gm * ldc + gn;
                    *p = (beta == (T)0) ? v : v + beta * *p;
                }
            }
}

template <typename T>
static int gemm_launch(int M, int N, int K, T alpha, const T *A, long long sam, long long sak,
                       long long bsa, const T *B, long long sbk, long long sbn, long long bsb,
                       T beta, T *C, long long ldc, long long bsc, int batch, hipStream_t s)
{
    VT_CHECK_ARG(M >= 0 && N >= 0 && K >= 0 && batch >= 1 && C && ldc >= N);
    VT_CHECK_ARG((A && B) || K == 0);
    if (M == 0 || N == 0) return VT_OK;
    dim3 grid(vt_blocks(N, GBN), vt_blocks(M, GBM), batch);
    VT_CHECK_ARG(grid.y <= 65535 && grid.z <= 65535);
    gemm_kernel<T><<<grid, 256, 0, s>>>(M, N, K, alpha, A, sam, sak, bsa, B, sbk, sbn, bsb, beta, C,
                                        ldc, bsc);
    VT_LAUNCH_CHECK();
    return VT_OK;
}

VT_EXPORT int vt_gemm_f64(int M, int N, int K, double alpha, const double *A, long long sam,
                          long long sak, long long bsa, const double *B, long long sbk,
                          long long sbn, long long bsb, double beta, double *C, long long ldc,
                          long long bsc, int batch, void *stream)
{
    return gemm_launch<double>(M, N, K, alpha, A, sam, sak, bsa, B, sbk, sbn, bsb, beta, C, ldc,
                               bsc, batch, (hipStream_t)stream);
}

VT_EXPORT int vt_gemm_f32(int M, int N, int K, float alpha, const float *A, long long sam,
                          long long sak, long long bsa, const float *B, long long sbk,
                          long long sbn, long long bsb, float beta, float *C, long long ldc,
                          long long bsc, int batch, void *stream)
{
    return gemm_launch<float>(M, N, K, alpha, A, sam, sak, bsa, B, sbk, sbn, bsb, beta, C, ldc,
                              bsc, batch, (hipStream_t)stream);
}

// ------------------------------------------------------------------------------------
// Lane-layout probe: one MFMA of each kind on caller-provided inputs; writes the 16x16
// result as (row, col) from the accumulator mapping that the kernels above assume, so a
// test compares it with A @ B on the host.
//   kind 0: fp64 16x16x4   A (16x4)  B (4x16)   out fp64 16x16
//   kind 1: fp32 16x16x4   A (16x4)  B (4x16)   out fp32 16x16
//   kind 2: bf16 16x16x32  A (16x32) B (32x16)  out fp32 16x16  (inputs given as fp32)
//   kind 3: f16  16x16x32  A (16x32) B (32x16)  out fp32 16x16  (inputs given as fp32)
// ------------------------------------------------------------------------------------
__global__ void mfma_probe_kernel(int kind, const void *A, const void *B, void *out)
{
    const int lane = threadIdx.x;
    if (kind == 0) {
        const double *a = (const double *)A, *b = (const double *)B;
        v4d c = {0, 0, 0, 0};
        c = __builtin_amdgcn_mfma_f64_16x16x4f64(a[(lane & 15) * 4 + (lane >> 4)],
                                                 b[(lane >> 4) * 16 + (lane & 15)], c, 0, 0, 0);
        for (int r = 0; r < 4; ++r)
            ((double *)out)[mfma_row_f64(lane, r) * 16 + mfma_col(lane)] = c[r];
    } else if (kind == 1) {
        const float *a = (const float *)A, *b = (const float *)B;
        v4f c = {0, 0, 0, 0};
        c = __builtin_amdgcn_mfma_f32_16x16x4f32(a[(lane & 15) * 4 + (lane >> 4)],
                                                 b[(lane >> 4) * 16 + (lane & 15)], c, 0, 0, 0);
        for (int r = 0; r < 4; ++r)
            ((float *)out)[mfma_row(lane, r) * 16 + mfma_col(lane)] = c[r];
    } else {
        const float *a = (const float *)A, *b = (const float *)B;
        v4f c = {0, 0, 0, 0};
        const int row = lane & 15, kb = (lane >> 4) * 8;
        if (kind == 2) {
            v8bf av, bv;
            for (int i = 0; i < 8; ++i) {
                av[i] = (__bf16)a[row * 32 + kb + i];
                bv[i] = (__bf16)b[(kb + i) * 16 + row];
            }
            c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(av, bv, c, 0, 0, 0);
        } else {
            v8h av, bv;
            for (int i = 0; i < 8; ++i) {
                av[i] = (_Float16)a[row * 32 + kb + i];
                bv[i] = (_Float16)b[(kb + i) * 16 + row];
            }
            c = __builtin_amdgcn_mfma_f32_16x16x32_f16(av, bv, c, 0, 0, 0);
        }
        for (int r = 0; r < 4; ++r)
            ((float *)out)[mfma_row(lane, r) * 16 + mfma_col(lane)] = c[r];
    }
}

VT_EXPORT int vt_mfma_probe(int kind, const void *A, const void *B, void *out, void *stream)
{
    VT_CHECK_ARG(kind >= 0 && kind <= 3 && A && B && out);
    mfma_probe_kernel<<<1, 64, 0, (hipStream_t)stream>>>(kind, A, B, out);
    VT_LAUNCH_CHECK();
    return VT_OK;
}

// ------------------------------------------------------------------------------------
// Skinny transposed product for K-long reductions with a narrow output:
//   C[m, n] = alpha * sum_k A[k*lda + m] * B[k*ldb + n] + beta * C[m, n],  n < Nc <= 32.
// Used for r = S^T t and S^T Z (K = neurons, output Ds x Dt).  Pass 1: grid
// (ceil(M/16), TN_KS); a workgroup owns 16 rows m and one K chunk, thread t reads
// A[k, m0 + t%16] (128-byte rows) for its k-group and keeps Nc partial sums; the 16
// k-groups are summed in LDS and written to workspace[ks][m][n].  Pass 2 sums the TN_KS
// chunks in a fixed order, so results are deterministic (no atomics).
// workspace: TN_KS * M * Nc doubles.
// ------------------------------------------------------------------------------------
#define TN_KS 16

__global__ void __launch_bounds__(256)
gemm_tn_small_partial(int M, int Nc, int K, const double *__restrict__ A, long long lda,
                      const double *__restrict__ B, long long ldb, double *__restrict__ ws)
{
    __shared__ double red[16][16][33];
    const int tid = threadIdx.x, ms = tid & 15, kg = tid >> 4;
    const int m = blockIdx.x * 16 + ms;
    const int chunk = (K + TN_KS - 1) / TN_KS;
    const int kbeg = blockIdx.y * chunk, kend = min(K, kbeg + chunk);
    double acc[32];
#pragma unroll
    for (int n = 0; n < 32; ++n) acc[n] = 0.0;
#pragma unroll 4
    for (int k = kbeg + kg; k < kend; k += 16) {
        const double a = (m < M) ? A[(long long)k * lda + m] : 0.0;
        const double *b = B + (long long)k * ldb;
#pragma unroll
        for (int n = 0; n < 32; ++n)
            if (n < Nc) acc[n] += a * b[n];
    }
#pragma unroll
    for (int n = 0; n < 32; ++n)
        if (n < Nc) red[kg][ms][n] = acc[n];
    __syncthreads();
    for (int e = tid; e < 16 * Nc; e += 256) {
        const int mm = e / Nc, n = e % Nc;
        const int gm = blockIdx.x * 16 + mm;
        if (gm < M) {
            double s = 0.0;
            for (int g = 0; g < 16; ++g) s += red[g][mm][n];
            ws[((long long)blockIdx.y * M + gm) * Nc + n] = s;
        }
    }
}

__global__ void gemm_tn_small_reduce(int M, int Nc, double alpha, const double *__restrict__ ws,
                                     double beta, double *__restrict__ C, long long ldc)
{
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= M * Nc) return;
    const int m = e / Nc, n = e % Nc;
    double s = 0.0;
    for (int k = 0; k < TN_KS; ++k) s += ws[((long long)k * M + m) * Nc + n];
    double *p = C + (long long)m * ldc + n;
    *p = (beta == 0.0) ? alpha * s : alpha * s + beta * *p;
}

VT_EXPORT int vt_gemm_tn_small(int M, int Nc, int K, double alpha, const double *A, long long lda,
                               const double *B, long long ldb, double beta, double *C,
                               long long ldc, double *workspace, void *stream)
{
    VT_CHECK_ARG(M >= 0 && Nc >= 1 && Nc <= 32 && K >= 0 && A && B && C && ldc >= Nc && workspace);
    if (M == 0) return VT_OK;
    hipStream_t s = (hipStream_t)stream;
    gemm_tn_small_partial<<<dim3(vt_blocks(M, 16), TN_KS), 256, 0, s>>>(M, Nc, K, A, lda, B, ldb,
                                                                      workspace);
    gemm_tn_small_reduce<<<vt_blocks((long long)M * Nc, 256), 256, 0, s>>>(M, Nc, alpha, workspace,
                                                                         beta, C, ldc);
    VT_LAUNCH_CHECK();
    return VT_OK;
}
