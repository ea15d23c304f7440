// Symmetric pseudoinverse and eigendecomposition by one-sided (Hestenes) Jacobi, FP64.
//
// The reference solves G c = q by the SVD of G (CULA gesvd, vtem/utils/linalg.py:366-523)
// and forms c = U diag(1/s_i, s_i > rcond s_0) U^H q (:591-631, sq_Kernel :790-831), and
// offers the symmetric eigendecomposition eig_sym (:634-718).  Here both come from one
// iteration: the rows of A start as the rows of the symmetric G (= its columns), and each
// rotation of a row pair (p, q) makes a_p and a_q orthogonal.  At convergence A = (G V)^T
// with orthogonal columns a_j = s_j u_j: s_j = |a_j| are the singular values and u_j the
// left singular vectors, so
//     c = sum_{s_j > rcond s_max} u_j (u_j . q) / s_j = sum a_j (a_j . q) / s_j^3.
// For the eigendecomposition the iteration runs on G + shift I, positive definite for
// shift > |G|, with V accumulated (rows of V start as the identity's): then a_j = s_j v_j,
// v_j is an eigenvector and w_j = a_j . v_j - shift its eigenvalue.
//
// A sweep visits every pair once: n' - 1 rounds (n' = n rounded up to even) of a
// round-robin tournament, each round n'/2 disjoint pairs, one workgroup per pair.  A pair
// is rotated when |a_p . a_q| > tol |a_p| |a_q|; the rotation count tells the caller when a
// sweep changed nothing.  Every reduction runs in a fixed order, so results are
// reproducible.
#include "vt_common.hpp"

#define SV_THREADS 256

__device__ __forceinline__ void tournament_pair(int m, int r, int i, int &p, int &q)
{
    if (i == 0) {
        p = r;
        q = m - 1;
    } else {
        p = (r + i) % (m - 1);
        q = (r - i + (m - 1)) % (m - 1);
    }
    if (p > q) {
        const int t = p;
        p = q;
        q = t;
    }
}

// three sums over the workgroup, tree order fixed
__device__ __forceinline__ void sum3(double &a, double &b, double &c, double (*red)[SV_THREADS])
{
    const int t = threadIdx.x;
    red[0][t] = a;
    red[1][t] = b;
    red[2][t] = c;
    __syncthreads();
    for (int o = SV_THREADS / 2; o > 0; o >>= 1) {
        if (t < o) {
            red[0][t] += red[0][t + o];
            red[1][t] += red[1][t + o];
            red[2][t] += red[2][t + o];
        }
        __syncthreads();
    }
    a = red[0][0];
    b = red[1][0];
    c = red[2][0];
}

__global__ void __launch_bounds__(SV_THREADS)
jacobi_round_kernel(double *__restrict__ A, long long lda, double *__restrict__ V, long long ldv, int n,
                    int m, int r, double tol, double small2, int *__restrict__ rotations)
{
    __shared__ double red[3][SV_THREADS];
    int p, q;
    tournament_pair(m, r, blockIdx.x, p, q);
    if (q >= n) return;                      // the padding player of an odd n
    double *ap = A + (long long)p * lda, *aq = A + (long long)q * lda;
    double spp = 0.0, sqq = 0.0, spq = 0.0;
    for (int k = threadIdx.x; k < n; k += SV_THREADS) {
        const double x = ap[k], y = aq[k];
        spp += x * x;
        sqq += y * y;
        spq += x * y;
    }
    sum3(spp, sqq, spq, red);
    // a numerically zero row (|a|^2 <= small2) is left alone: its direction is rounding
    // noise, and rotating it against others would never settle
    if (!(spp > small2 && sqq > small2)) return;
    if (!(fabs(spq) > tol * sqrt(spp) * sqrt(sqq))) return;
    // t = tan of the angle that zeroes the inner product (the smaller root of
    // t^2 + 2 zeta t - 1 = 0), c = cos, s = sin
    const double zeta = (sqq - spp) / (2.0 * spq);
    const double t = (zeta >= 0.0 ? 1.0 : -1.0) / (fabs(zeta) + sqrt(1.0 + zeta * zeta));
    const double c = 1.0 / sqrt(1.0 + t * t), s = c * t;
    for (int k = threadIdx.x; k < n; k += SV_THREADS) {
        const double x = ap[k], y = aq[k];
        ap[k] = c * x - s * y;
        aq[k] = s * x + c * y;
    }
    if (V) {
        double *vp = V + (long long)p * ldv, *vq = V + (long long)q * ldv;
        for (int k = threadIdx.x; k < n; k += SV_THREADS) {
            const double x = vp[k], y = vq[k];
            vp[k] = c * x - s * y;
            vq[k] = s * x + c * y;
        }
    }
    if (threadIdx.x == 0) atomicAdd(rotations, 1);
}

// One sweep (every row pair once).  V (optional, may be null) accumulates the rotations.
// Rows with |a|^2 <= small2 count as zero.  rotations (device int) is incremented by the
// number of rotations applied; the caller zeroes it and stops when a sweep leaves it at 0.
VT_EXPORT int vt_svd_jacobi_sweep(double *A, long long lda, double *V, long long ldv, int n, double tol,
                                  double small2, int *rotations, void *stream)
{
    VT_CHECK_ARG(A && rotations && n >= 1 && lda >= n && (!V || ldv >= n) && tol >= 0.0 && small2 >= 0.0);
    if (n == 1) return VT_OK;
    const int m = n + (n & 1);
    hipStream_t s = (hipStream_t)stream;
    for (int r = 0; r < m - 1; ++r)
        jacobi_round_kernel<<<m / 2, SV_THREADS, 0, s>>>(A, lda, V, ldv, n, m, r, tol, small2, rotations);
    VT_LAUNCH_CHECK();
    return VT_OK;
}

__global__ void __launch_bounds__(SV_THREADS) zero_int_kernel(int *x) { *x = 0; }

VT_EXPORT int vt_zero_int(int *x, void *stream)
{
    VT_CHECK_ARG(x);
    zero_int_kernel<<<1, 1, 0, (hipStream_t)stream>>>(x);
    VT_LAUNCH_CHECK();
    return VT_OK;
}

// row j: ws[j] = |a_j|, ws[n + j] = a_j . q
__global__ void __launch_bounds__(SV_THREADS)
row_norm_dot_kernel(const double *__restrict__ A, long long lda, int n, const double *__restrict__ q,
                    double *__restrict__ ws)
{
    __shared__ double red[3][SV_THREADS];
    const double *a = A + (long long)blockIdx.x * lda;
    double ss = 0.0, sq = 0.0, z = 0.0;
    for (int k = threadIdx.x; k < n; k += SV_THREADS) {
        const double x = a[k];
        ss += x * x;
        sq += x * q[k];
    }
    sum3(ss, sq, z, red);
    if (threadIdx.x == 0) {
        ws[blockIdx.x] = sqrt(ss);
        ws[n + blockIdx.x] = sq;
    }
}

// y_j = (a_j . q) / s_j^3 for s_j > rcond max s, else 0 (into ws[n + j]); one workgroup
__global__ void __launch_bounds__(SV_THREADS) pinv_weights_kernel(double *__restrict__ ws, int n, double rcond)
{
    __shared__ double red[SV_THREADS];
    double mx = 0.0;
    for (int j = threadIdx.x; j < n; j += SV_THREADS) mx = fmax(mx, ws[j]);
    red[threadIdx.x] = mx;
    __syncthreads();
    for (int o = SV_THREADS / 2; o > 0; o >>= 1) {
        if (threadIdx.x < o) red[threadIdx.x] = fmax(red[threadIdx.x], red[threadIdx.x + o]);
        __syncthreads();
    }
    const double cut = red[0] * rcond;
    for (int j = threadIdx.x; j < n; j += SV_THREADS) {
        const double s = ws[j];
        ws[n + j] = s > cut ? ws[n + j] / (s * s * s) : 0.0;
    }
}

// c[k] = sum_j y_j A[j][k]  (j in order)
__global__ void __launch_bounds__(SV_THREADS)
combine_rows_kernel(const double *__restrict__ A, long long lda, int n, const double *__restrict__ y,
                    double *__restrict__ c)
{
    const int k = blockIdx.x * SV_THREADS + threadIdx.x;
    if (k >= n) return;
    double acc = 0.0;
    for (int j = 0; j < n; ++j) acc += y[j] * A[(long long)j * lda + k];
    c[k] = acc;
}

// After the sweeps: c = sum_{s_j > rcond s_max} a_j (a_j . q) / s_j^3.  ws: 2n doubles.
VT_EXPORT int vt_svd_pinv_apply(const double *A, long long lda, int n, const double *q, double rcond,
                                double *c, double *ws, void *stream)
{
    VT_CHECK_ARG(A && q && c && ws && n >= 1 && lda >= n && rcond >= 0.0);
    hipStream_t s = (hipStream_t)stream;
    row_norm_dot_kernel<<<n, SV_THREADS, 0, s>>>(A, lda, n, q, ws);
    pinv_weights_kernel<<<1, SV_THREADS, 0, s>>>(ws, n, rcond);
    combine_rows_kernel<<<vt_blocks(n, SV_THREADS), SV_THREADS, 0, s>>>(A, lda, n, ws + n, c);
    VT_LAUNCH_CHECK();
    return VT_OK;
}

// w_j = a_j . v_j - shift (the eigenvalue of v_j: (G + shift I) v_j = a_j)
__global__ void __launch_bounds__(SV_THREADS)
rayleigh_kernel(const double *__restrict__ A, long long lda, const double *__restrict__ V, long long ldv,
                int n, double shift, double *__restrict__ w)
{
    __shared__ double red[3][SV_THREADS];
    const double *a = A + (long long)blockIdx.x * lda, *v = V + (long long)blockIdx.x * ldv;
    double d = 0.0, z1 = 0.0, z2 = 0.0;
    for (int k = threadIdx.x; k < n; k += SV_THREADS) d += a[k] * v[k];
    sum3(d, z1, z2, red);
    if (threadIdx.x == 0) w[blockIdx.x] = d - shift;
}

// ascending order: rank of w_j = #{i : w_i < w_j or (w_i == w_j and i < j)}; the sorted
// eigenvalues and the eigenvector rows placed at their ranks
__global__ void __launch_bounds__(SV_THREADS)
rank_place_kernel(const double *__restrict__ wj, const double *__restrict__ V, long long ldv, int n,
                  double *__restrict__ w_out, double *__restrict__ E, long long lde)
{
    __shared__ int rank_s;
    const int j = blockIdx.x;
    const double x = wj[j];
    if (threadIdx.x == 0) rank_s = 0;
    __syncthreads();
    int cnt = 0;
    for (int i = threadIdx.x; i < n; i += SV_THREADS) {
        const double y = wj[i];
        cnt += (y < x || (y == x && i < j)) ? 1 : 0;
    }
    atomicAdd(&rank_s, cnt);
    __syncthreads();
    const int rk = rank_s;
    if (threadIdx.x == 0) w_out[rk] = x;
    for (int k = threadIdx.x; k < n; k += SV_THREADS) E[(long long)rk * lde + k] = V[(long long)j * ldv + k];
}

// After the sweeps with V accumulated on A = G + shift I (positive definite, so its
// singular vectors are eigenvectors and no two opposite eigenvalues share a singular
// value): eigenvalues of G ascending in w (n), eigenvector k in row k of E (n x n, row
// stride lde).  ws: n doubles.
VT_EXPORT int vt_svd_eig_finish(const double *A, long long lda, const double *V, long long ldv, int n,
                                double shift, double *w, double *E, long long lde, double *ws,
                                void *stream)
{
    VT_CHECK_ARG(A && V && w && E && ws && n >= 1 && lda >= n && ldv >= n && lde >= n);
    hipStream_t s = (hipStream_t)stream;
    rayleigh_kernel<<<n, SV_THREADS, 0, s>>>(A, lda, V, ldv, n, shift, ws);
    rank_place_kernel<<<n, SV_THREADS, 0, s>>>(ws, V, ldv, n, w, E, lde);
    VT_LAUNCH_CHECK();
    return VT_OK;
}
