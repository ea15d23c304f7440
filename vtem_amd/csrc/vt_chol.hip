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

// 1 / sqrt(d) for d > 0: the hardware estimate and Newton steps r <- r (3 - d r^2) / 2
// (two for FP64, one for FP32), each error squared: full precision in fewer dependent
// operations than a correctly rounded square root and a division (the diagonal factor's
// column steps are a chain of these).
__device__ __forceinline__ double chol_rsqrt(double d)
{
    double r = __builtin_amdgcn_rsq(d);
    const double h = 0.5 * d;
#pragma unroll
    for (int it = 0; it < 2; ++it) r = __builtin_fma(r, __builtin_fma(-h * r, r, 0.5), r);
    return r;
}
__device__ __forceinline__ float chol_rsqrt(float d)
{
    float r = __builtin_amdgcn_rsqf(d);
    const float h = 0.5f * d;
    return __builtin_fmaf(r, __builtin_fmaf(-h * r, r, 0.5f), r);
}

// ------------------------------------------------------------------------------------
// diag: factor + invert the 64x64 diagonal block, one 256-thread workgroup per matrix.
// Thread (i, q) = (tid / 4, tid % 4) keeps elements [16q, 16q + 16) of row i in registers.
// Column step c: the owner of (i, c) publishes it to LDS (double-buffered, one barrier
// per step); every thread reads the pivot and the column back and applies
//   A[i][j] -= L[i][c] L[j][c]   for c < j <= i,   L[i][c] = A[i][c] / sqrt(A[c][c])
// to its own elements (1/sqrt by chol_rsqrt, L[c][c] = A[c][c] / sqrt(A[c][c])).  The inverse X = L^-1 solves L X = I the same way, right-looking:
// row c's owners publish X[c][:] = B[c][:] / L[c][c], then every row below subtracts
// L[i][c] X[c][:], so each X[k][j] is (delta_kj - sum_{j'<k} L[k][j'] X[j'][j]) / L[k][k]
// summed in j' order.  The column loop is unrolled 16 at a time (static register
// indices), the 64 steps are not: a short instruction stream.
// ------------------------------------------------------------------------------------
template <typename T>
__global__ void __launch_bounds__(256)
potrf_diag_kernel(const VtMat<T> *__restrict__ mats, int kb, T *__restrict__ inv_ws,
                  long long inv_stride, int *__restrict__ info)
{
    __shared__ T col[2][NB];
    __shared__ T Ls[NB][NB + 1];
    __shared__ T rdiag[NB];
    const VtMat<T> m = mats[blockIdx.x];
    if (kb >= m.D) return;
    const int nb = min(NB, m.D - kb);
    const int tid = threadIdx.x, i = tid >> 2, q = tid & 3, j0 = 16 * q;
    T *Ad = m.A + (long long)kb * m.lda + kb;
    T a[16];
#pragma unroll
    for (int t = 0; t < 16; ++t) {
        const int j = j0 + t;
        T v = (T)0;
        if (i < nb && j <= i) v = Ad[(long long)i * m.lda + j];
        if (i >= nb && j == i) v = (T)1;      // identity rows pad a partial last block
        a[t] = v;
    }
    int bad = 0;
    for (int c0 = 0; c0 < NB; c0 += 16) {
        const bool own = q == (c0 >> 4);     // this thread holds column c of its row
#pragma unroll
        for (int cc = 0; cc < 16; ++cc) {
            const int c = c0 + cc, bf = cc & 1;
            if (own) col[bf][i] = a[cc];      // A[i][c] after the previous updates
            __syncthreads();
            T d = col[bf][c];
            if (!(d > (T)0)) {
                if (c < nb && !bad) bad = kb + c + 1;
                d = (T)1e-30;    // keep going without NaNs; info reports the failure
            }
            const T inv = chol_rsqrt(d);
            const T sd = d * inv;
            const T lic = (i > c) ? col[bf][i] * inv : (T)0;   // L[i][c]
            if (own && i >= c) a[cc] = (i == c) ? sd : lic;
#pragma unroll
            for (int t = 0; t < 16; ++t) {
                const int j = j0 + t;
                if (j > c && j <= i) a[t] -= lic * (col[bf][j] * inv);
            }
        }
    }
    if (tid == 0 && bad && info && info[blockIdx.x] == 0) info[blockIdx.x] = bad;
#pragma unroll
    for (int t = 0; t < 16; ++t) {
        const int j = j0 + t;
        Ls[i][j] = (j <= i) ? a[t] : (T)0;
        if (i < nb && j <= i) Ad[(long long)i * m.lda + j] = a[t];
        if (j == i) rdiag[i] = (T)1 / a[t];   // the inverse's pivots, off its dependent chain
    }
    __syncthreads();
    // inverse: b = row i of B (the identity, then eliminated), X rows published through col
    T b[16];
#pragma unroll
    for (int t = 0; t < 16; ++t) b[t] = (j0 + t == i) ? (T)1 : (T)0;
    for (int c0 = 0; c0 < NB; c0 += 16) {
#pragma unroll
        for (int cc = 0; cc < 16; ++cc) {
            const int c = c0 + cc, bf = cc & 1;
            if (i == c) {
                const T rinv = rdiag[c];
#pragma unroll
                for (int t = 0; t < 16; ++t) {
                    b[t] = b[t] * rinv;       // X[c][j]
                    col[bf][j0 + t] = b[t];
                }
            }
            __syncthreads();
            if (i > c) {
                const T lic = Ls[i][c];
#pragma unroll
                for (int t = 0; t < 16; ++t) b[t] -= lic * col[bf][j0 + t];
            }
        }
    }
    T *W = inv_ws + (long long)blockIdx.x * inv_stride + (long long)(kb / NB) * NB * NB;
#pragma unroll
    for (int t = 0; t < 16; ++t) {
        const int j = j0 + t;
        W[i * NB + j] = (i < nb && j < nb) ? b[t] : (T)0;
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
                int r = wm + i * 16 + acc_row<T>(lane, q), c = wn + j * 16 + mfma_col(lane);
                if (r0 + r < m.D && c < nb) Ap[(long long)r * m.lda + c] = acc[i][j][q];
            }
}

// ------------------------------------------------------------------------------------
// panel update: A[I,J] -= X_I X_J^T for the lower 64x64 tiles I >= J whose columns lie in
// [kb + 64, jend) -- inside the current 256-column outer panel.  grid (ntg * njg, batch):
// tile pid <-> (bi, bj) = (pid % ntg, pid / ntg); tiles outside a matrix exit.
// ------------------------------------------------------------------------------------
template <typename T>
__global__ void __launch_bounds__(256)
potrf_update_kernel(const VtMat<T> *__restrict__ mats, int kb, int jend, int ntg)
{
    typedef typename Mfma16<T>::acc_t acc_t;
    __shared__ T Xi[NB][NB + 1];
    __shared__ T Xj[NB][NB + 1];
    const VtMat<T> m = mats[blockIdx.y];
    const int t0 = kb + NB;
    if (t0 >= m.D) return;
    const int nt = (m.D - t0 + NB - 1) / NB;
    const int nj = (min(jend, m.D) - t0 + NB - 1) / NB;
    const int bi = blockIdx.x % ntg, bj = blockIdx.x / ntg;
    if (bi >= nt || bj >= nj || bi < bj) return;
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
                int r = ri + wm + i * 16 + acc_row<T>(lane, q);
                int c = rj + wn + j * 16 + mfma_col(lane);
                if (r < m.D && c < m.D && c <= r) m.A[(long long)r * m.lda + c] -= acc[i][j][q];
            }
}

// ------------------------------------------------------------------------------------
// trailing update of an outer panel: A[I,J] -= sum_{k in [K, K+kw)} L[I,k] L[J,k] for the
// lower 128x128 tiles I >= J of A[K+kw:, K+kw:].  Four waves, each a 64x64 quadrant (4x4
// MFMA tiles); the panel's rows stream through LDS 16 columns at a time.  grid
// (nt (nt+1) / 2, batch) for the largest matrix; tiles outside a matrix exit.
// ------------------------------------------------------------------------------------
#define TB 128
#define KC 16

template <typename T>
__global__ void __launch_bounds__(256)
potrf_trail_kernel(const VtMat<T> *__restrict__ mats, int K, int kw)
{
    typedef typename Mfma16<T>::acc_t acc_t;
    __shared__ T Xi[TB][KC + 1];
    __shared__ T Xj[TB][KC + 1];
    const VtMat<T> m = mats[blockIdx.y];
    const int t0 = K + kw;
    if (t0 >= m.D) return;
    const int nt = (m.D - t0 + TB - 1) / TB;
    const int pid = blockIdx.x;
    if (pid >= nt * (nt + 1) / 2) return;
    int bi = (int)((sqrt(8.0 * pid + 1.0) - 1.0) * 0.5);
    while ((bi + 1) * (bi + 2) / 2 <= pid) ++bi;
    while (bi * (bi + 1) / 2 > pid) --bi;
    const int bj = pid - bi * (bi + 1) / 2;
    const int ri = t0 + bi * TB, rj = t0 + bj * TB;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wm = (wave >> 1) * 64, wn = (wave & 1) * 64;
    acc_t acc[4][4];
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) acc[i][j] = (acc_t){0, 0, 0, 0};
    // the panel's rows: lane pair (2 fr, 2 fr + 1) loads the two 8-column halves of row fr
    // of both tiles (KC = 16 columns); the next chunk is loaded into registers while this
    // one is used
    static_assert(TB * KC == 256 * 8, "one row half per thread");
    typedef T t2 __attribute__((ext_vector_type(2), aligned(sizeof(T))));
    typedef const t2 __attribute__((address_space(1))) *gt2;
    const int fr = tid >> 1, fh = (tid & 1) * 8;
    // rows past the matrix read its last row and are zeroed: no branch around the loads
    const bool in_i = ri + fr < m.D, in_j = rj + fr < m.D;
    const T *pa = m.A + (long long)min(ri + fr, m.D - 1) * m.lda + K + fh;
    const T *pb = m.A + (long long)min(rj + fr, m.D - 1) * m.lda + K + fh;
    const t2 zero2 = {(T)0, (T)0};
    t2 pi[4], pj[4];
    auto fetch = [&](int k0) {
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const t2 x = *(gt2)(pa + k0 + 2 * i), y = *(gt2)(pb + k0 + 2 * i);
            pi[i] = in_i ? x : zero2;
            pj[i] = in_j ? y : zero2;
        }
    };
    auto stash = [&]() {
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            Xi[fr][fh + 2 * i] = pi[i][0];
            Xi[fr][fh + 2 * i + 1] = pi[i][1];
            Xj[fr][fh + 2 * i] = pj[i][0];
            Xj[fr][fh + 2 * i + 1] = pj[i][1];
        }
    };
    fetch(0);
    stash();
    __syncthreads();
    for (int k0 = 0; k0 < kw; k0 += KC) {
        // the last chunk fetches itself again: the loop body has no branch
        fetch(k0 + KC < kw ? k0 + KC : k0);
#pragma unroll
        for (int ks = 0; ks < KC; ks += 4) {
            const int k = ks + (lane >> 4);
            T a[4], b[4];
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                a[i] = Xi[wm + 16 * i + (lane & 15)][k];
                b[i] = Xj[wn + 16 * i + (lane & 15)][k];
            }
#pragma unroll
            for (int i = 0; i < 4; ++i)
#pragma unroll
                for (int j = 0; j < 4; ++j) acc[i][j] = Mfma16<T>::op(a[i], b[j], acc[i][j]);
        }
        __syncthreads();
        stash();
        __syncthreads();
    }
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j)
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const int r = ri + wm + 16 * i + acc_row<T>(lane, q);
                const int c = rj + wn + 16 * j + mfma_col(lane);
                if (r < m.D && c <= r) m.A[(long long)r * m.lda + c] -= acc[i][j][q];
            }
}

// Two-level blocking: 256-column outer panels, each factored in 64-column steps (diag,
// panel solve of every row below, update of the tiles inside the outer panel); the rest
// of the matrix receives the outer panel's 256 columns in one trailing update, so the
// trailing matrix is read and written once per 256 columns instead of once per 64.
#define KW 256

template <typename T>
static int potrf_run(const void *mats_dev, int batch, int maxD, T *inv_ws, long long inv_stride,
                     int *info, hipStream_t s)
{
    VT_CHECK_ARG(mats_dev && batch >= 1 && batch <= 65535 && maxD >= 1 && inv_ws && inv_stride >= 0);
    const VtMat<T> *mats = (const VtMat<T> *)mats_dev;
    for (int K = 0; K < maxD; K += KW) {
        const int kend = min(K + KW, maxD);
        for (int kb = K; kb < kend; kb += NB) {
            potrf_diag_kernel<T><<<batch, 256, 0, s>>>(mats, kb, inv_ws, inv_stride, info);
            const int rows = maxD - kb - NB;
            if (rows > 0) {
                const int nch = (rows + NB - 1) / NB;
                potrf_panel_kernel<T><<<dim3(nch, batch), 256, 0, s>>>(mats, kb, inv_ws, inv_stride);
                const int njg = (kend - kb - NB + NB - 1) / NB;     // tile columns left in the panel
                if (njg > 0)
                    potrf_update_kernel<T><<<dim3(nch * njg, batch), 256, 0, s>>>(mats, kb, K + KW, nch);
            }
        }
        if (K + KW < maxD) {
            const int nt = (maxD - K - KW + TB - 1) / TB;
            potrf_trail_kernel<T><<<dim3(nt * (nt + 1) / 2, batch), 256, 0, s>>>(mats, K, KW);
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
// Triangular solves with the factor: x <- L^-T L^-1 x (x fp64; the factor may be fp32,
// the preconditioner of the mixed-precision path).
//
// One workgroup per 64-row block and matrix, one launch per direction.  Workgroup b of
// the forward sweep folds in L[b, j] y_j for j < b as soon as block j publishes y_j, then
// applies the stored diagonal-block inverse and publishes y_b; the backward sweep mirrors
// it with the bottom block first.  The XCDs' L2 caches are not coherent with each other,
// and a fence or acquire at agent scope writes back or invalidates a whole L2 (every
// spinning poll would): so the solution blocks are stored and read with relaxed
// agent-scope atomics (written through to, and read from, the device coherence point),
// and a block's flag (= epoch, relaxed, agent scope) is stored after the workgroup's
// stores have completed and polled with relaxed loads; L is only read.  A workgroup only ever waits on lower workgroup IDs, which the
// in-order dispatcher has already started, so the chain always drains; each wait is
// additionally bounded (VT_SPIN_LIMIT polls) and sets the error word instead of hanging.
//
// sync workspace (ints, zeroed once by the caller): [0] epoch, [1] error, then
// 2*batch*nblk flags.  The epoch is read on device and bumped by a one-thread kernel at
// the end of each solve, so the sequence can be captured in a graph and replayed.
// ------------------------------------------------------------------------------------
#define VT_SPIN_LIMIT 20000000

__device__ __forceinline__ bool wait_flag(const int *flag, int want, int *err)
{
    long it = 0;
    while (__hip_atomic_load(flag, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != want) {
        if (++it > VT_SPIN_LIMIT) {
            atomicExch(err, 1);
            return false;
        }
        __builtin_amdgcn_s_sleep(1);
    }
    return true;
}

template <typename T>
__global__ void __launch_bounds__(256)
trsv_fwd_kernel(const VtMat<T> *__restrict__ mats, const T *__restrict__ inv_ws,
                long long inv_stride, double *__restrict__ x, long long x_stride, int *sync,
                int nblk_max)
{
    __shared__ double yj[NB];
    __shared__ double part[4][NB];
    __shared__ double acc[NB];
    __shared__ int ok;
    const VtMat<T> m = mats[blockIdx.y];
    const int b = blockIdx.x;
    const int kb = b * NB;
    if (kb >= m.D) return;
    const int nb = min(NB, m.D - kb);
    const int tid = threadIdx.x, r = tid & 63, q = tid >> 6;
    const int epoch = sync[0] + 1;
    int *err = sync + 1;
    int *flags = sync + 2 + (long long)blockIdx.y * nblk_max;
    double *xv = x + (long long)blockIdx.y * x_stride;
    if (tid < NB) acc[tid] = tid < nb ? xv[kb + tid] : 0.0;
    if (tid == 0) ok = 1;
    __syncthreads();
    for (int j = 0; j < b; ++j) {
        if (tid == 0 && !wait_flag(flags + j, epoch, err)) ok = 0;
        __syncthreads();
        if (!ok) return;
        if (tid < NB)
            yj[tid] = __hip_atomic_load(xv + j * NB + tid, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        __syncthreads();
        // part[q][r] = sum_{c in quarter q} L[kb+r, j*64+c] y_j[c]
        double s = 0.0;
        if (r < nb) {
            const T *Lr = m.A + (long long)(kb + r) * m.lda + j * NB + q * 16;
#pragma unroll
            for (int c = 0; c < 16; ++c) s += (double)Lr[c] * yj[q * 16 + c];
        }
        part[q][r] = s;
        __syncthreads();
        if (tid < NB) acc[tid] -= part[0][tid] + part[1][tid] + part[2][tid] + part[3][tid];
        __syncthreads();
    }
    // y_b = Linv_b acc
    const T *W = inv_ws + (long long)blockIdx.y * inv_stride + (long long)b * NB * NB;
    double s = 0.0;
    if (r < nb)
        for (int c = q * 16; c < q * 16 + 16; ++c) s += (double)W[r * NB + c] * acc[c];
    part[q][r] = s;
    __syncthreads();
    if (tid < nb) {
        double y = part[0][tid] + part[1][tid] + part[2][tid] + part[3][tid];
        __hip_atomic_store(xv + kb + tid, y, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    __builtin_amdgcn_s_waitcnt(0);                // this wave's block stores have completed
    __syncthreads();
    if (tid == 0) __hip_atomic_store(flags + b, epoch, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

template <typename T>
__global__ void __launch_bounds__(256)
trsv_bwd_kernel(const VtMat<T> *__restrict__ mats, const T *__restrict__ inv_ws,
                long long inv_stride, double *__restrict__ x, long long x_stride, int *sync,
                int nblk_max, int batch)
{
    __shared__ double xj[NB];
    __shared__ double part[4][NB];
    __shared__ double acc[NB];
    __shared__ int ok;
    const VtMat<T> m = mats[blockIdx.y];
    const int nblk = (m.D + NB - 1) / NB;
    if ((int)blockIdx.x >= nblk) return;
    const int b = nblk - 1 - blockIdx.x;      // bottom block has the lowest ID
    const int kb = b * NB;
    const int nb = min(NB, m.D - kb);
    const int tid = threadIdx.x, r = tid & 63, q = tid >> 6;
    const int epoch = sync[0] + 1;
    int *err = sync + 1;
    int *flags = sync + 2 + (long long)(batch + blockIdx.y) * nblk_max;
    double *xv = x + (long long)blockIdx.y * x_stride;
    if (tid < NB) acc[tid] = tid < nb ? xv[kb + tid] : 0.0;
    if (tid == 0) ok = 1;
    __syncthreads();
    for (int j = nblk - 1; j > b; --j) {
        const int nbj = min(NB, m.D - j * NB);
        if (tid == 0 && !wait_flag(flags + j, epoch, err)) ok = 0;
        __syncthreads();
        if (!ok) return;
        if (tid < NB)
            xj[tid] = tid < nbj ? __hip_atomic_load(xv + j * NB + tid, __ATOMIC_RELAXED,
                                                    __HIP_MEMORY_SCOPE_AGENT) : 0.0;
        __syncthreads();
        // part[q][r] = sum_{c in quarter q} L[j*64+c, kb+r] x_j[c]   (coalesced over r)
        double s = 0.0;
        if (r < nb)
            for (int c = q * 16; c < q * 16 + 16 && c < nbj; ++c)
                s += (double)m.A[(long long)(j * NB + c) * m.lda + kb + r] * xj[c];
        part[q][r] = s;
        __syncthreads();
        if (tid < NB) acc[tid] -= part[0][tid] + part[1][tid] + part[2][tid] + part[3][tid];
        __syncthreads();
    }
    // x_b = Linv_b^T acc
    const T *W = inv_ws + (long long)blockIdx.y * inv_stride + (long long)b * NB * NB;
    double s = 0.0;
    if (r < nb)
        for (int c = q * 16; c < q * 16 + 16; ++c) s += (double)W[c * NB + r] * acc[c];
    part[q][r] = s;
    __syncthreads();
    if (tid < nb) {
        double v = part[0][tid] + part[1][tid] + part[2][tid] + part[3][tid];
        __hip_atomic_store(xv + kb + tid, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    __builtin_amdgcn_s_waitcnt(0);                // this wave's block stores have completed
    __syncthreads();
    if (tid == 0) __hip_atomic_store(flags + b, epoch, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__global__ void epoch_bump_kernel(int *sync) { sync[0] += 1; }

template <typename T>
static int potrs_run(const void *mats_dev, int batch, int maxD, const T *inv_ws,
                     long long inv_stride, double *x, long long x_stride, int *sync, hipStream_t s)
{
    VT_CHECK_ARG(mats_dev && batch >= 1 && maxD >= 1 && inv_ws && x && sync);
    VT_CHECK_ARG(batch <= 65535);
    const int nblk = (maxD + NB - 1) / NB;
    const VtMat<T> *mats = (const VtMat<T> *)mats_dev;
    trsv_fwd_kernel<T><<<dim3(nblk, batch), 256, 0, s>>>(mats, inv_ws, inv_stride, x, x_stride,
                                                        sync, nblk);
    trsv_bwd_kernel<T><<<dim3(nblk, batch), 256, 0, s>>>(mats, inv_ws, inv_stride, x, x_stride,
                                                        sync, nblk, batch);
    epoch_bump_kernel<<<1, 1, 0, s>>>(sync);
    VT_LAUNCH_CHECK();
    return VT_OK;
}

// x <- (L L^T)^-1 x for each matrix.  sync: 2 + 2*batch*ceil(maxD/64) ints, zeroed once
// by the caller and reused across calls; sync[1] != 0 after a solve means a bounded wait
// expired (results invalid).
VT_EXPORT int vt_potrs_f64(const void *mats_dev, int batch, int maxD, const double *inv_ws,
                           long long inv_stride, double *x, long long x_stride, int *sync,
                           void *stream)
{
    return potrs_run<double>(mats_dev, batch, maxD, inv_ws, inv_stride, x, x_stride, sync,
                             (hipStream_t)stream);
}

VT_EXPORT int vt_potrs_f32(const void *mats_dev, int batch, int maxD, const float *inv_ws,
                           long long inv_stride, double *x, long long x_stride, int *sync,
                           void *stream)
{
    return potrs_run<float>(mats_dev, batch, maxD, inv_ws, inv_stride, x, x_stride, sync,
                            (hipStream_t)stream);
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
