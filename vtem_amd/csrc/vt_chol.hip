// Regularised solve (A + lam I) X = r  by blocked Cholesky on the lower triangle.
//
// Replaces the reference's truncated-SVD solve of the spike-space G (CULA gesvd,
// vtem/utils/linalg.py:591-631) by a Tikhonov-regularised Cholesky of the D-space Gram;
// both are the minimum-norm solution family of the same system (push-through identity).
//
// Right-looking, block size 64.  Per 64-column step kb:
//   diag   : one workgroup per matrix factors the 64x64 diagonal block in LDS and forms
//            its inverse (kept in a workspace for the panel and the triangular solves);
//   panel  : X = A[kb+64:, kb:kb+64] * Linv^T on MFMA (one workgroup per 64 rows);
//   update : A[I, J] -= X_I X_J^T for the lower 64x64 tiles of the trailing matrix (MFMA).
// Matrices are described by {pointer, lda, D} descriptors; a launch serves a whole batch
// (matrices whose D is already exhausted skip the step), so a batch of B matrices costs
// the same number of launches as one.
#include "vt_common.hpp"

#define NB 64

template <typename T> struct VtMat {
    T *A;
    long long lda;
    int D;
    int pad;
};

template <typename T> struct Mfma16;
template <> struct Mfma16<double> {
    typedef v4d acc_t;
    static __device__ __forceinline__ v4d op(double a, double b, v4d c) {
        return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
    }
};
template <> struct Mfma16<float> {
    typedef v4f acc_t;
    static __device__ __forceinline__ v4f op(float a, float b, v4f c) {
        return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
    }
};

// ------------------------------------------------------------------------------------
// diag: factor + invert the 64x64 diagonal block (in LDS, 256 threads)
// ------------------------------------------------------------------------------------
template <typename T>
__global__ void __launch_bounds__(256)
potrf_diag_kernel(const VtMat<T> *__restrict__ mats, int kb, T *__restrict__ inv_ws,
                  long long inv_stride, int *__restrict__ info)
{
    __shared__ T L[NB][NB + 1];
    __shared__ T X[NB][NB + 1];
    const VtMat<T> m = mats[blockIdx.x];
    if (kb >= m.D) return;
    const int nb = min(NB, m.D - kb);
    const int tid = threadIdx.x;
    T *Ad = m.A + (long long)kb * m.lda + kb;
    for (int e = tid; e < NB * NB; e += 256) {
        int r = e / NB, c = e % NB;
        L[r][c] = (r < nb && c < nb && c <= r) ? Ad[(long long)r * m.lda + c] : (T)0;
        X[r][c] = (r == c) ? (T)1 : (T)0;
    }
    __syncthreads();
    // right-looking unblocked factorisation
    for (int c = 0; c < nb; ++c) {
        if (tid == 0) {
            T d = L[c][c];
            if (!(d > (T)0)) {
                if (info && info[blockIdx.x] == 0) info[blockIdx.x] = kb + c + 1;
                d = (T)1e-300 > (T)0 ? (T)1e-30 : (T)1e-30;
            }
            L[c][c] = sqrt(d);
        }
        __syncthreads();
        const T dc = L[c][c];
        for (int r = c + 1 + tid; r < nb; r += 256) L[r][c] /= dc;
        __syncthreads();
        const int w = nb - c - 1;   // trailing size
        for (int e = tid; e < w * w; e += 256) {
            int r = c + 1 + e / w, rp = c + 1 + e % w;
            if (rp <= r) L[r][rp] -= L[r][c] * L[rp][c];
        }
        __syncthreads();
    }
    // inverse: X = L^{-1} (forward substitution on the identity, columns in parallel)
    for (int r = 0; r < nb; ++r) {
        const T d = L[r][r];
        for (int c = tid; c <= r; c += 256) X[r][c] /= d;
        __syncthreads();
        const int rows = nb - r - 1;
        for (int e = tid; e < rows * (r + 1); e += 256) {
            int rr = r + 1 + e / (r + 1), c = e % (r + 1);
            X[rr][c] -= L[rr][r] * X[r][c];
        }
        __syncthreads();
    }
    // write back L (lower) and Linv
    T *W = inv_ws + (long long)blockIdx.x * inv_stride + (long long)(kb / NB) * NB * NB;
    for (int e = tid; e < NB * NB; e += 256) {
        int r = e / NB, c = e % NB;
        if (r < nb && c <= r) Ad[(long long)r * m.lda + c] = L[r][c];
        W[e] = (r < nb && c < nb) ? X[r][c] : (T)0;
    }
}

// ------------------------------------------------------------------------------------
// panel: rows below the diagonal block, X = A_chunk * Linv^T (MFMA 64x64x64)
// ------------------------------------------------------------------------------------
template <typename T>
__global__ void __launch_bounds__(256)
potrf_panel_kernel(const VtMat<T> *__restrict__ mats, int kb, const T *__restrict__ inv_ws,
                   long long inv_stride)
{
    typedef typename Mfma16<T>::acc_t acc_t;
    __shared__ T As[NB][NB + 1];   // [row][k]
    __shared__ T Bs[NB][NB + 1];   // Linv [c][k]
    const VtMat<T> m = mats[blockIdx.y];
    const int r0 = kb + NB + blockIdx.x * NB;
    if (kb >= m.D || r0 >= m.D) return;
    const int nb = min(NB, m.D - kb);
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const T *W = inv_ws + (long long)blockIdx.y * inv_stride + (long long)(kb / NB) * NB * NB;
    T *Ap = m.A + (long long)r0 * m.lda + kb;
    for (int e = tid; e < NB * NB; e += 256) {
        int r = e / NB, k = e % NB;
        As[r][k] = (r0 + r < m.D && k < nb) ? Ap[(long long)r * m.lda + k] : (T)0;
        Bs[r][k] = W[e];
    }
    __syncthreads();
    const int wm = (wave >> 1) * 32, wn = (wave & 1) * 32;
    acc_t acc[2][2];
    for (int i = 0; i < 2; ++i)
        for (int j = 0; j < 2; ++j) acc[i][j] = (acc_t){0, 0, 0, 0};
    for (int k0 = 0; k0 < NB; k0 += 4) {
        const int k = k0 + (lane >> 4);
        T a0 = As[wm + (lane & 15)][k], a1 = As[wm + 16 + (lane & 15)][k];
        T b0 = Bs[wn + (lane & 15)][k], b1 = Bs[wn + 16 + (lane & 15)][k];
        acc[0][0] = Mfma16<T>::op(a0, b0, acc[0][0]);
        acc[0][1] = Mfma16<T>::op(a0, b1, acc[0][1]);
        acc[1][0] = Mfma16<T>::op(a1, b0, acc[1][0]);
        acc[1][1] = Mfma16<T>::op(a1, b1, acc[1][1]);
    }
    for (int i = 0; i < 2; ++i)
        for (int j = 0; j < 2; ++j)
            for (int q = 0; q < 4; ++q) {
                int r = wm + i * 16 + mfma_row(lane, q), c = wn + j * 16 + mfma_col(lane);
                if (r0 + r < m.D && c < nb) Ap[(long long)r * m.lda + c] = acc[i][j][q];
            }
}

// ------------------------------------------------------------------------------------
// update: A[I,J] -= X_I X_J^T for lower tiles I >= J of the trailing matrix
// ------------------------------------------------------------------------------------
template <typename T>
__global__ void __launch_bounds__(256)
potrf_update_kernel(const VtMat<T> *__restrict__ mats, int kb)
{
    typedef typename Mfma16<T>::acc_t acc_t;
    __shared__ T Xi[NB][NB + 1];
    __shared__ T Xj[NB][NB + 1];
    const VtMat<T> m = mats[blockIdx.y];
    const int t0 = kb + NB;
    if (t0 >= m.D) return;
    const int nt = (m.D - t0 + NB - 1) / NB;
    const int pid = blockIdx.x;
    if (pid >= nt * (nt + 1) / 2) return;
    int bi = (int)((sqrt(8.0 * pid + 1.0) - 1.0) * 0.5);
    while ((bi + 1) * (bi + 2) / 2 <= pid) ++bi;
    while (bi * (bi + 1) / 2 > pid) --bi;
    const int bj = pid - bi * (bi + 1) / 2;
    const int ri = t0 + bi * NB, rj = t0 + bj * NB;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    for (int e = tid; e < NB * NB; e += 256) {
        int r = e / NB, k = e % NB;
        Xi[r][k] = (ri + r < m.D) ? m.A[(long long)(ri + r) * m.lda + kb + k] : (T)0;
        Xj[r][k] = (rj + r < m.D) ? m.A[(long long)(rj + r) * m.lda + kb + k] : (T)0;
    }
    __syncthreads();
    const int wm = (wave >> 1) * 32, wn = (wave & 1) * 32;
    acc_t acc[2][2];
    for (int i = 0; i < 2; ++i)
        for (int j = 0; j < 2; ++j) acc[i][j] = (acc_t){0, 0, 0, 0};
    for (int k0 = 0; k0 < NB; k0 += 4) {
        const int k = k0 + (lane >> 4);
        T a0 = Xi[wm + (lane & 15)][k], a1 = Xi[wm + 16 + (lane & 15)][k];
        T b0 = Xj[wn + (lane & 15)][k], b1 = Xj[wn + 16 + (lane & 15)][k];
        acc[0][0] = Mfma16<T>::op(a0, b0, acc[0][0]);
        acc[0][1] = Mfma16<T>::op(a0, b1, acc[0][1]);
        acc[1][0] = Mfma16<T>::op(a1, b0, acc[1][0]);
        acc[1][1] = Mfma16<T>::op(a1, b1, acc[1][1]);
    }
    for (int i = 0; i < 2; ++i)
        for (int j = 0; j < 2; ++j)
            for (int q = 0; q < 4; ++q) {
                int r = ri + wm + i * 16 + mfma_row(lane, q);
                int c = rj + wn + j * 16 + mfma_col(lane);
                if (r < m.D && c < m.D && c <= r) m.A[(long long)r * m.lda + c] -= acc[i][j][q];
            }
}

template <typename T>
static int potrf_run(const void *mats_dev, int batch, int maxD, T *inv_ws, long long inv_stride,
                     int *info, hipStream_t s)
{
    VT_CHECK_ARG(mats_dev && batch >= 1 && maxD >= 1 && inv_ws && inv_stride >= 0);
    const VtMat<T> *mats = (const VtMat<T> *)mats_dev;
    for (int kb = 0; kb < maxD; kb += NB) {
        potrf_diag_kernel<T><<<batch, 256, 0, s>>>(mats, kb, inv_ws, inv_stride, info);
        const int rows = maxD - kb - NB;
        if (rows > 0) {
            const int nch = (rows + NB - 1) / NB;
            potrf_panel_kernel<T><<<dim3(nch, batch), 256, 0, s>>>(mats, kb, inv_ws, inv_stride);
            potrf_update_kernel<T><<<dim3(nch * (nch + 1) / 2, batch), 256, 0, s>>>(mats, kb);
        }
    }
    VT_LAUNCH_CHECK();
    return VT_OK;
}

// mats_dev: device array of {A, lda, D} (16 bytes each).  inv_ws: per matrix
// ceil(maxD/64)*64*64 elements (inv_stride apart) receiving the diagonal-block inverses.
// info (device, batch ints, zeroed by the caller): first non-positive pivot + 1.
VT_EXPORT int vt_potrf_f64(const void *mats_dev, int batch, int maxD, double *inv_ws,
                           long long inv_stride, int *info, void *stream)
{
    return potrf_run<double>(mats_dev, batch, maxD, inv_ws, inv_stride, info, (hipStream_t)stream);
}

VT_EXPORT int vt_potrf_f32(const void *mats_dev, int batch, int maxD, float *inv_ws,
                           long long inv_stride, int *info, void *stream)
{
    return potrf_run<float>(mats_dev, batch, maxD, inv_ws, inv_stride, info, (hipStream_t)stream);
}

// ------------------------------------------------------------------------------------
// Triangular solves with the factor: x = L^-T L^-1 r (in place on x), one workgroup per
// matrix (batch via the descriptor array).  The factor may be fp32 (preconditioner of the
// mixed-precision path) while the vector is always fp64.
// ------------------------------------------------------------------------------------
template <typename T>
__global__ void __launch_bounds__(1024)
potrs_kernel(const VtMat<T> *__restrict__ mats, const T *__restrict__ inv_ws, long long inv_stride,
             double *__restrict__ x, long long x_stride)
{
    __shared__ double yk[NB];
    __shared__ double part[1024];
    const VtMat<T> m = mats[blockIdx.x];
    const int D = m.D;
    const int tid = threadIdx.x;
    double *xv = x + (long long)blockIdx.x * x_stride;
    const T *Wb = inv_ws + (long long)blockIdx.x * inv_stride;
    const int nblk = (D + NB - 1) / NB;
    // forward: L y = r
    for (int b = 0; b < nblk; ++b) {
        const int kb = b * NB, nb = min(NB, D - kb);
        const T *W = Wb + (long long)b * NB * NB;
        // y_b = Linv_b * r_b : 16 threads per row
        {
            const int r = tid >> 4, l = tid & 15;
            double s = 0.0;
            if (r < nb)
                for (int c = l; c <= r; c += 16) s += (double)W[r * NB + c] * xv[kb + c];
            part[tid] = s;
        }
        __syncthreads();
        if (tid < nb) {
            double s = 0.0;
            for (int l = 0; l < 16; ++l) s += part[tid * 16 + l];
            yk[tid] = s;
        }
        __syncthreads();
        if (tid < nb) xv[kb + tid] = yk[tid];
        // r_i -= L[i, kb:kb+nb] . y_b  for i >= kb + nb   (16 lanes per row)
        for (int i0 = kb + nb + (tid >> 4); i0 < D; i0 += 64) {
            const int l = tid & 15;
            const T *Lr = m.A + (long long)i0 * m.lda + kb;
            double s = 0.0;
            for (int c = l; c < nb; c += 16) s += (double)Lr[c] * yk[c];
            for (int o = 8; o > 0; o >>= 1) s += __shfl_xor(s, o, 16);
            if (l == 0) xv[i0] -= s;
        }
        __syncthreads();
    }
    // backward: L^T x = y
    for (int b = nblk - 1; b >= 0; --b) {
        const int kb = b * NB, nb = min(NB, D - kb);
        const T *W = Wb + (long long)b * NB * NB;
        // x_b = Linv_b^T * y_b
        {
            const int r = tid >> 4, l = tid & 15;
            double s = 0.0;
            if (r < nb)
                for (int c = r + l; c < nb; c += 16) s += (double)W[c * NB + r] * xv[kb + c];
            part[tid] = s;
        }
        __syncthreads();
        if (tid < nb) {
            double s = 0.0;
            for (int l = 0; l < 16; ++l) s += part[tid * 16 + l];
            yk[tid] = s;
        }
        __syncthreads();
        if (tid < nb) xv[kb + tid] = yk[tid];
        // y_i -= sum_c L[kb+c, i] x_b[c]  for i < kb  (coalesced over i)
        for (int i = tid; i < kb; i += 1024) {
            double s = 0.0;
            for (int c = 0; c < nb; ++c) s += (double)m.A[(long long)(kb + c) * m.lda + i] * yk[c];
            xv[i] -= s;
        }
        __syncthreads();
    }
}

VT_EXPORT int vt_potrs_f64(const void *mats_dev, int batch, const double *inv_ws,
                           long long inv_stride, double *x, long long x_stride, void *stream)
{
    VT_CHECK_ARG(mats_dev && batch >= 1 && inv_ws && x);
    potrs_kernel<double><<<batch, 1024, 0, (hipStream_t)stream>>>(
        (const VtMat<double> *)mats_dev, inv_ws, inv_stride, x, x_stride);
    VT_LAUNCH_CHECK();
    return VT_OK;
}

VT_EXPORT int vt_potrs_f32(const void *mats_dev, int batch, const float *inv_ws,
                           long long inv_stride, double *x, long long x_stride, void *stream)
{
    VT_CHECK_ARG(mats_dev && batch >= 1 && inv_ws && x);
    potrs_kernel<float><<<batch, 1024, 0, (hipStream_t)stream>>>(
        (const VtMat<float> *)mats_dev, inv_ws, inv_stride, x, x_stride);
    VT_LAUNCH_CHECK();
    return VT_OK;
}

// fp64 -> fp32 copy of the lower triangle (with lam added to the diagonal), for the
// mixed-precision factorisation of an fp64 Gram.
__global__ void lower_to_f32_kernel(float *__restrict__ dst, long long ldd, const double *__restrict__ src,
                                    long long lds, int D, const double *__restrict__ lam)
{
    const int i = blockIdx.y;
    for (int j = blockIdx.x * blockDim.x + threadIdx.x; j <= i && j < D; j += gridDim.x * blockDim.x) {
        double v = src[(long long)i * lds + j];
        if (i == j && lam) v += lam[0];
        dst[(long long)i * ldd + j] = (float)v;
    }
}

VT_EXPORT int vt_lower_to_f32(float *dst, long long ldd, const double *src, long long lds, int D,
                              const double *lam_dev, void *stream)
{
    VT_CHECK_ARG(dst && src && D >= 1 && ldd >= D && lds >= D && D <= 65535);
    lower_to_f32_kernel<<<dim3(vt_blocks(D, 256) < 64 ? vt_blocks(D, 256) : 64, D), 256, 0,
                          (hipStream_t)stream>>>(dst, ldd, src, lds, D, lam_dev);
    VT_LAUNCH_CHECK();
    return VT_OK;
}
