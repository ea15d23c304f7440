// Kronecker-preconditioned conjugate gradients for the coefficient-space Tikhonov system
//
//     (A + lam I) X = r,     A = sum_j S_j S_j^T (x) T_j        (D = Ds * Dt unknowns)
//
// solved for a batch of independent blocks without forming A.  The operator is applied
// through its structure (Y = S P, Z_j = T_j Y_j, Q = S^T Z: 4 N D flops, S read twice);
// the preconditioner is the nearest separable operator
//
//     P = (S^T C S + lam I) (x) Tbar,   c_j = tr(T_j)/Dt,   Tbar = sum_j T_j / sum_j c_j,
//
// whose inverse is Mi (x) Ti with two small explicit inverses (Ds x Ds, Dt x Dt).  The
// spatial part is inverted exactly, so the spatial ill-conditioning of A (modes weakly
// covered by the receptive fields) does not slow the iteration; what remains is the
// spread of each neuron's temporal moments about c_j Tbar, set by how much the interval
// lengths vary within a neuron.
//
// Every reduction runs inside one workgroup per block in a fixed order, so a block's
// iterates do not depend on the other blocks of the batch.
//
// The reference solves the spike-space system with a truncated SVD
// (vtem/utils/linalg.py:591-631) and has no iterative Tikhonov solver.
#include "vt_common.hpp"
#include "vt_skinny.hpp"
#include "vt_graph.hpp"

#define KP_CHUNKS 16
#define KP_MAX_DT 33

// ------------------------------------------------------------------------------------
// Tbar, Ti and c_j.  Pass 1: grid (KP_CHUNKS, batch); each workgroup sums T_j over its
// share of the block's neurons into ws[(b*KP_CHUNKS + g)*Dt^2 + e] and writes c_j.
// ------------------------------------------------------------------------------------
__global__ void __launch_bounds__(256)
kpcg_tbar_partial(const double *__restrict__ Tm, const int *__restrict__ nbase, int Dt,
                  double *__restrict__ ws, double *__restrict__ c)
{
    const int b = blockIdx.y, g = blockIdx.x;
    const int n0 = nbase[b], N = nbase[b + 1] - n0;
    const int E = Dt * Dt;
    const int chunk = (N + KP_CHUNKS - 1) / KP_CHUNKS;
    const int j0 = g * chunk, j1 = min(N, j0 + chunk);
    for (int e = threadIdx.x; e < E; e += blockDim.x) {
        double s = 0.0;
        for (int j = j0; j < j1; ++j) s += Tm[(long long)(n0 + j) * E + e];
        ws[((long long)b * KP_CHUNKS + g) * E + e] = s;
    }
    for (int j = j0 + threadIdx.x; j < j1; j += blockDim.x) {
        const double *T = Tm + (long long)(n0 + j) * E;
        double t = 0.0;
        for (int a = 0; a < Dt; ++a) t += T[a * Dt + a];
        c[n0 + j] = t / Dt;
    }
}

// Pass 2: one workgroup per block sums the partials (fixed order), scales to trace Dt and
// inverts Tbar by Gauss-Jordan elimination in LDS (Tbar is SPD and well conditioned: it is
// the interval-averaged temporal moment of the whole window).
__global__ void __launch_bounds__(256)
kpcg_tbar_final(const double *__restrict__ ws, int Dt, double *__restrict__ Ti)
{
    __shared__ double G[KP_MAX_DT][2 * KP_MAX_DT + 1];
    __shared__ double piv[KP_MAX_DT];
    __shared__ double tr;
    const int b = blockIdx.x, E = Dt * Dt, W = 2 * Dt;
    for (int e = threadIdx.x; e < E; e += blockDim.x) {
        double s = 0.0;
        for (int g = 0; g < KP_CHUNKS; ++g) s += ws[((long long)b * KP_CHUNKS + g) * E + e];
        G[e / Dt][e % Dt] = s;
        G[e / Dt][Dt + e % Dt] = (e / Dt == e % Dt) ? 1.0 : 0.0;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        double t = 0.0;
        for (int a = 0; a < Dt; ++a) t += G[a][a];
        tr = t;
    }
    __syncthreads();
    // Tbar = Tsum * Dt / tr(Tsum)   (sum_j c_j = tr(Tsum)/Dt)
    const double scale = Dt / tr;
    for (int e = threadIdx.x; e < E; e += blockDim.x) G[e / Dt][e % Dt] *= scale;
    __syncthreads();
    for (int k = 0; k < Dt; ++k) {
        const double inv = 1.0 / G[k][k];
        __syncthreads();
        for (int col = threadIdx.x; col < W; col += blockDim.x) G[k][col] *= inv;
        __syncthreads();
        if (threadIdx.x < Dt) piv[threadIdx.x] = G[threadIdx.x][k];
        __syncthreads();
        for (int e = threadIdx.x; e < Dt * W; e += blockDim.x) {
            const int i = e / W, col = e % W;
            if (i != k) G[i][col] -= piv[i] * G[k][col];
        }
        __syncthreads();
    }
    // transposed, so that a row vector y gives (y Ti)[c] = sum_k TiT[c][k] y[k]
    for (int e = threadIdx.x; e < E; e += blockDim.x)
        Ti[(long long)b * E + e] = G[e % Dt][Dt + e / Dt];
}

// ------------------------------------------------------------------------------------
// Kronecker-sum preconditioner P = (S^T C S) (x) Tbar + lam I.  For blocks decoded in a
// period longer than their window, Tbar is badly conditioned (the basis functions that are
// small inside the window have tiny moments), and lam, not (S^T C S) (x) Tbar, dominates A
// there; the Kronecker-product preconditioner above then misjudges those directions by up
// to 1/theta_min.  With Tbar = U diag(theta) U^T, P = (I (x) U) blockdiag_k(theta_k K +
// lam I) (I (x) U)^T, so P^-1 R = sum_k (theta_k K + lam I)^-1 (R u_k) u_k^T for the
// residual R (Ds x Dt).  The eigendecomposition of Tbar (Dt <= 33) is a two-sided cyclic
// Jacobi in LDS, one workgroup per block, rounds of disjoint pairs, fixed order.
// ------------------------------------------------------------------------------------
#define KP_EIG_SWEEPS 30

__global__ void __launch_bounds__(256)
kpcg_tbar_eig_kernel(const double *__restrict__ ws, int Dt, double *__restrict__ theta,
                     double *__restrict__ U, double *__restrict__ UT)
{
    __shared__ double G[KP_MAX_DT + 1][KP_MAX_DT + 1];
    __shared__ double V[KP_MAX_DT + 1][KP_MAX_DT + 1];
    __shared__ double cs[KP_MAX_DT + 1][2];
    __shared__ int rotated;
    const int b = blockIdx.x, E = Dt * Dt, tid = threadIdx.x;
    for (int e = tid; e < E; e += blockDim.x) {
        double t = 0.0;
        for (int g = 0; g < KP_CHUNKS; ++g) t += ws[((long long)b * KP_CHUNKS + g) * E + e];
        G[e / Dt][e % Dt] = t;
        V[e / Dt][e % Dt] = (e / Dt == e % Dt) ? 1.0 : 0.0;
    }
    __syncthreads();
    __shared__ double tr;
    if (tid == 0) {
        double t = 0.0;
        for (int a = 0; a < Dt; ++a) t += G[a][a];
        tr = t;
    }
    __syncthreads();
    const double scale = Dt / tr;             // trace Dt, as the Kronecker-product form
    for (int e = tid; e < E; e += blockDim.x) G[e / Dt][e % Dt] *= scale;
    __syncthreads();
    const int m = Dt + (Dt & 1), half = m / 2;
    const double tol = 4.0 * Dt * 2.220446049250313e-16;
    for (int sweep = 0; sweep < KP_EIG_SWEEPS; ++sweep) {
        if (tid == 0) rotated = 0;
        __syncthreads();
        for (int r = 0; r < m - 1; ++r) {
            // angles of the round's disjoint pairs (c = 1, s = 0: not rotated)
            if (tid < half) {
                int p, q;
                if (tid == 0) {
                    p = r;
                    q = m - 1;
                } else {
                    p = (r + tid) % (m - 1);
                    q = (r - tid + (m - 1)) % (m - 1);
                }
                if (p > q) { const int t = p; p = q; q = t; }
                double c = 1.0, sn = 0.0;
                if (q < Dt) {
                    const double app = G[p][p], aqq = G[q][q], apq = G[p][q];
                    if (fabs(apq) > tol * sqrt(fabs(app) * fabs(aqq)) && apq != 0.0) {
                        const double zeta = (aqq - app) / (2.0 * apq);
                        const double t = (zeta >= 0.0 ? 1.0 : -1.0) / (fabs(zeta) + sqrt(1.0 + zeta * zeta));
                        c = 1.0 / sqrt(1.0 + t * t);
                        sn = c * t;
                        rotated = 1;
                    }
                }
                cs[tid][0] = c;
                cs[tid][1] = sn;
            }
            __syncthreads();
            // rows p, q of G <- J^T G ; then columns (G J) and V J
            for (int e = tid; e < half * Dt; e += blockDim.x) {
                const int i = e / Dt, k = e % Dt;
                int p, q;
                if (i == 0) { p = r; q = m - 1; } else { p = (r + i) % (m - 1); q = (r - i + (m - 1)) % (m - 1); }
                if (p > q) { const int t = p; p = q; q = t; }
                if (q >= Dt) continue;
                const double c = cs[i][0], sn = cs[i][1], x = G[p][k], y = G[q][k];
                G[p][k] = c * x - sn * y;
                G[q][k] = sn * x + c * y;
            }
            __syncthreads();
            for (int e = tid; e < half * Dt; e += blockDim.x) {
                const int i = e / Dt, k = e % Dt;
                int p, q;
                if (i == 0) { p = r; q = m - 1; } else { p = (r + i) % (m - 1); q = (r - i + (m - 1)) % (m - 1); }
                if (p > q) { const int t = p; p = q; q = t; }
                if (q >= Dt) continue;
                const double c = cs[i][0], sn = cs[i][1];
                const double x = G[k][p], y = G[k][q];
                G[k][p] = c * x - sn * y;
                G[k][q] = sn * x + c * y;
                const double vx = V[k][p], vy = V[k][q];
                V[k][p] = c * vx - sn * vy;
                V[k][q] = sn * vx + c * vy;
            }
            __syncthreads();
        }
        const int done = !rotated;
        __syncthreads();
        if (done) break;
    }
    // theta_k = G[k][k]; U[i][k] = V[i][k] (eigenvector k in column k), UT = U^T
    for (int e = tid; e < E; e += blockDim.x) {
        const int i = e / Dt, k = e % Dt;
        U[(long long)b * E + e] = V[i][k];
        UT[(long long)b * E + e] = V[k][i];
    }
    for (int k = tid; k < Dt; k += blockDim.x) theta[(long long)b * Dt + k] = G[k][k];
}

// c_j, and per block Tbar's eigenvalues theta (batch x Dt) and eigenvectors U (batch x Dt x
// Dt, eigenvector k in column k; UT = U^T), Tbar scaled to trace Dt.
VT_EXPORT int vt_kpcg_tbar_eig(const double *Tm, const int *nbase, int Dt, int batch, double *ws,
                               double *c, double *theta, double *U, double *UT, void *stream)
{
    VT_CHECK_ARG(Tm && nbase && ws && c && theta && U && UT && batch >= 1 && batch <= 65535 && Dt >= 1);
    if (Dt > KP_MAX_DT) return VT_ERANGE;
    hipStream_t s = (hipStream_t)stream;
    kpcg_tbar_partial<<<dim3(KP_CHUNKS, batch), 256, 0, s>>>(Tm, nbase, Dt, ws, c);
    kpcg_tbar_eig_kernel<<<batch, 256, 0, s>>>(ws, Dt, theta, U, UT);
    VT_LAUNCH_CHECK();
    return VT_OK;
}

// out_{b,k} = theta[b*Dt + k] K_b + lam_b I (Ds_b x Ds_b each) for k < Dt: K_b at
// kpool + kbase[b], out_{b,k} at opool + obase[b] + k Ds_b^2.  Grid (chunks, batch*Dt).
__global__ void __launch_bounds__(256)
kpcg_shift_pool_kernel(const double *__restrict__ kpool, const long long *__restrict__ kbase,
                       const long long *__restrict__ obase, const int *__restrict__ ds,
                       const double *__restrict__ theta, const double *__restrict__ lam, int Dt,
                       double *__restrict__ opool)
{
    const int b = blockIdx.y / Dt, k = blockIdx.y % Dt;
    const int D = ds[b];
    const long long n = (long long)D * D;
    const double th = theta[(long long)b * Dt + k], l = lam[2 * b];
    const double *K = kpool + kbase[b];
    double *O = opool + obase[b] + (long long)k * n;
    for (long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x; e < n;
         e += (long long)gridDim.x * blockDim.x) {
        const long long i = e / D, j = e % D;
        O[e] = th * K[e] + (i == j ? l : 0.0);
    }
}

VT_EXPORT int vt_kpcg_shift_pool(const double *kpool, const long long *kbase, const long long *obase,
                                 const int *ds, const double *theta, const double *lam, int Dt,
                                 int batch, int max_Ds, double *opool, void *stream)
{
    VT_CHECK_ARG(kpool && kbase && obase && ds && theta && lam && opool && Dt >= 1 && batch >= 1 &&
                 max_Ds >= 1);
    VT_CHECK_ARG((long long)batch * Dt <= 65535);
    long long g = ((long long)max_Ds * max_Ds + 255) / 256;
    if (g > 64) g = 64;
    kpcg_shift_pool_kernel<<<dim3((unsigned)g, batch * Dt), 256, 0, (hipStream_t)stream>>>(
        kpool, kbase, obase, ds, theta, lam, Dt, opool);
    VT_LAUNCH_CHECK();
    return VT_OK;
}

VT_EXPORT int vt_kpcg_tbar(const double *Tm, const int *nbase, int Dt, int batch, double *ws,
                           double *c, double *Ti, void *stream)
{
    VT_CHECK_ARG(Tm && nbase && ws && c && Ti && batch >= 1 && batch <= 65535 && Dt >= 1);
    if (Dt > KP_MAX_DT) return VT_ERANGE;
    hipStream_t s = (hipStream_t)stream;
    kpcg_tbar_partial<<<dim3(KP_CHUNKS, batch), 256, 0, s>>>(Tm, nbase, Dt, ws, c);
    kpcg_tbar_final<<<batch, 256, 0, s>>>(ws, Dt, Ti);
    VT_LAUNCH_CHECK();
    return VT_OK;
}

// ------------------------------------------------------------------------------------
// Spatial preconditioner Gram  M_b = S_b^T diag(c_b) S_b  (Ds x Ds, K = N_b), FP64 MFMA
// 16x16x4 on 64x64 tiles; grid (Ds tiles, Ds tiles, batch).
// ------------------------------------------------------------------------------------
struct KpMDesc {
    const double *S;
    const double *c;
    double *M;
    long long lds, ldm;
    int N, Ds;
};

__global__ void __launch_bounds__(256) kpcg_mgram_kernel(const KpMDesc *__restrict__ descs)
{
    __shared__ double As[16][68];
    __shared__ double Bs[16][68];
    const KpMDesc d = descs[blockIdx.z];
    const int m0 = blockIdx.y * 64, n0 = blockIdx.x * 64;
    if (m0 >= d.Ds || n0 >= d.Ds) return;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wm = (wave >> 1) * 32, wn = (wave & 1) * 32;
    v4d acc[2][2];
    for (int i = 0; i < 2; ++i)
        for (int j = 0; j < 2; ++j) acc[i][j] = (v4d){0, 0, 0, 0};
    for (int k0 = 0; k0 < d.N; k0 += 16) {
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const int e = tid + i * 256;
            const int col = e & 63, k = e >> 6;
            const int gk = k0 + k;
            const bool kin = gk < d.N;
            const double *row = d.S + (long long)gk * d.lds;
            As[k][col] = (kin && m0 + col < d.Ds) ? row[m0 + col] : 0.0;
            Bs[k][col] = (kin && n0 + col < d.Ds) ? d.c[gk] * row[n0 + col] : 0.0;
        }
        __syncthreads();
#pragma unroll
        for (int kk = 0; kk < 16; kk += 4) {
            const int kr = kk + (lane >> 4);
            double a0 = As[kr][wm + (lane & 15)], a1 = As[kr][wm + 16 + (lane & 15)];
            double b0 = Bs[kr][wn + (lane & 15)], b1 = Bs[kr][wn + 16 + (lane & 15)];
            acc[0][0] = __builtin_amdgcn_mfma_f64_16x16x4f64(a0, b0, acc[0][0], 0, 0, 0);
            acc[0][1] = __builtin_amdgcn_mfma_f64_16x16x4f64(a0, b1, acc[0][1], 0, 0, 0);
            acc[1][0] = __builtin_amdgcn_mfma_f64_16x16x4f64(a1, b0, acc[1][0], 0, 0, 0);
            acc[1][1] = __builtin_amdgcn_mfma_f64_16x16x4f64(a1, b1, acc[1][1], 0, 0, 0);
        }
        __syncthreads();
    }
    for (int i = 0; i < 2; ++i)
        for (int j = 0; j < 2; ++j)
            for (int r = 0; r < 4; ++r) {
                const int gm = m0 + wm + i * 16 + mfma_row_f64(lane, r);
                const int gn = n0 + wn + j * 16 + mfma_col(lane);
                if (gm < d.Ds && gn < d.Ds) d.M[(long long)gm * d.ldm + gn] = acc[i][j][r];
            }
}

VT_EXPORT int vt_kpcg_mgram(const void *descs, int batch, int max_Ds, void *stream)
{
    VT_CHECK_ARG(descs && batch >= 1 && batch <= 65535 && max_Ds >= 1);
    const unsigned t = vt_blocks(max_Ds, 64);
    kpcg_mgram_kernel<<<dim3(t, t, batch), 256, 0, (hipStream_t)stream>>>((const KpMDesc *)descs);
    VT_LAUNCH_CHECK();
    return VT_OK;
}

// Diagonal blocks of L^-1: X_b[64i.., 64i..] = Linv_ii (the 64x64 block inverses the
// Cholesky left at inv_ws + b*inv_stride + i*4096).  grid (blocks of the largest
// factor, batch).  X_b starts at X + xoff[b] with leading dimension dim[b].
__global__ void __launch_bounds__(256)
kpcg_place_diag_kernel(double *__restrict__ X, const long long *__restrict__ xoff,
                       const int *__restrict__ dim, const double *__restrict__ inv_ws,
                       long long inv_stride)
{
    const int b = blockIdx.y, i = blockIdx.x, D = dim[b];
    if (i * 64 >= D) return;
    const int nb = min(64, D - i * 64);
    const double *W = inv_ws + (long long)b * inv_stride + (long long)i * 4096;
    double *Xb = X + xoff[b] + (long long)i * 64 * D + i * 64;
    for (int e = threadIdx.x; e < 4096; e += 256) {
        const int rr = e >> 6, cc = e & 63;
        if (rr < nb && cc < nb) Xb[(long long)rr * D + cc] = W[e];
    }
}

VT_EXPORT int vt_kpcg_place_diag(double *X, const long long *xoff, const int *dim, int batch,
                                 int max_dim, const double *inv_ws, long long inv_stride,
                                 void *stream)
{
    VT_CHECK_ARG(X && xoff && dim && inv_ws && batch >= 1 && batch <= 65535 && max_dim >= 1);
    kpcg_place_diag_kernel<<<dim3(vt_blocks(max_dim, 64), batch), 256, 0, (hipStream_t)stream>>>(
        X, xoff, dim, inv_ws, inv_stride);
    VT_LAUNCH_CHECK();
    return VT_OK;
}

// ------------------------------------------------------------------------------------
// CG recurrences over concatenated per-block vectors (block b owns [seg[b], seg[b+1])).
// State per block (st + 4b): rz, rr0, active (0/1), iterations done.  hist[b*hs + k]
// holds |r_k|^2 of the recursively updated residual.
// ------------------------------------------------------------------------------------
__device__ __forceinline__ double block_sum(double v, double *red)
{
    red[threadIdx.x] = v;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if (threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
        __syncthreads();
    }
    const double s = red[0];
    __syncthreads();
    return s;
}

// p's copy in the forward product's tiled B layout (SK_B_TILED: the tiles of block b's
// P^T, Dt x Ds), optional: ptab[2b] = block b's offset in pt, ptab[2b+1] = its Dt; loc is
// the element's place in the block's n = Ds Dt coefficients (row-major Ds x Dt)
__device__ __forceinline__ void kp_store_tiled(double *pt, const long long *ptab, int b, long long loc,
                                               long long n, double v)
{
    const long long Dt = ptab[2 * b + 1];
    pt[ptab[2 * b] + sk_tile_index(loc % Dt, loc / Dt, (n / Dt + 15) / 16)] = v;
}

// x = 0, p = z, rz = <r,z>, rr0 = <r,r>
__global__ void __launch_bounds__(256)
kpcg_init_kernel(const double *__restrict__ r, const double *__restrict__ z, double *__restrict__ p,
                 double *__restrict__ x, const long long *__restrict__ seg, double *__restrict__ st,
                 double *__restrict__ hist, int hs, double *__restrict__ pt, const long long *__restrict__ ptab)
{
    __shared__ double red[256];
    const long long a = seg[blockIdx.x], e = seg[blockIdx.x + 1];
    double rz = 0.0, rr = 0.0;
    for (long long i = a + threadIdx.x; i < e; i += 256) {
        rz += r[i] * z[i];
        rr += r[i] * r[i];
        p[i] = z[i];
        if (pt) kp_store_tiled(pt, ptab, blockIdx.x, i - a, e - a, z[i]);
        x[i] = 0.0;
    }
    rz = block_sum(rz, red);
    rr = block_sum(rr, red);
    if (threadIdx.x == 0) {
        double *s = st + 4 * blockIdx.x;
        s[0] = rz;
        s[1] = rr;
        s[2] = (rr > 0.0 && rz > 0.0) ? 1.0 : 0.0;
        s[3] = 0.0;
        hist[(long long)blockIdx.x * hs] = rr;
    }
}

VT_EXPORT int vt_kpcg_init(const double *r, const double *z, double *p, double *x,
                           const long long *seg, int batch, double *st, double *hist, int hist_stride,
                           double *pt, const long long *ptab, void *stream)
{
    VT_CHECK_ARG(r && z && p && x && seg && st && hist && batch >= 1 && hist_stride >= 1 && (!pt || ptab));
    kpcg_init_kernel<<<batch, 256, 0, (hipStream_t)stream>>>(r, z, p, x, seg, st, hist, hist_stride, pt, ptab);
    VT_LAUNCH_CHECK();
    return VT_OK;
}

// q <- q + lam p; alpha = rz / <p, q>; x += alpha p; r -= alpha q; record |r|^2 and stop
// the block once |r|^2 <= stop_rel2 * rr0.  prune > 0: from iteration `prune` on, every
// KP_PRUNE_SPAN iterations, a block whose residual decay over the last KP_PRUNE_SPAN
// iterations, continued geometrically, would not reach the stop within `cap` iterations
// is stopped as well (it then fails the caller's residual guard and is solved another
// way); the test reads only the block's own history, so it does not depend on the batch.
#define KP_PRUNE_SPAN 10

__global__ void zero_words_kernel(unsigned *__restrict__ w, long long n)
{
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (long long)gridDim.x * blockDim.x)
        w[i] = 0u;
}

// fixed-order sum over the NTH threads of a workgroup (red: NTH doubles)
template <int NTH>
__device__ __forceinline__ double block_sum_n(double v, double *red)
{
    red[threadIdx.x] = v;
    __syncthreads();
    for (int o = NTH / 2; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
        __syncthreads();
    }
    const double s = red[0];
    __syncthreads();
    return s;
}

// step_a of block b by a workgroup of NTH threads; COHQ: q was written inside the same
// launch by other workgroups (read with agent-scope loads)
template <int NTH, bool COHQ>
__device__ void kpcg_step_a_block(int b, double *__restrict__ x, double *__restrict__ r,
                                  const double *__restrict__ p, double *__restrict__ q, double l, long long a,
                                  long long e, double *__restrict__ s, double *__restrict__ hist, int hs, int it,
                                  double stop_rel2, int prune, int cap, double *red)
{
    double pq = 0.0;
    for (long long i = a + threadIdx.x; i < e; i += NTH) {
        const double q0 = COHQ ? __hip_atomic_load(q + i, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) : q[i];
        const double qi = q0 + l * p[i];
        q[i] = qi;
        pq += p[i] * qi;
    }
    pq = block_sum_n<NTH>(pq, red);
    const double rz = s[0];
    if (!(pq > 0.0)) {
        if (threadIdx.x == 0) s[2] = 0.0;
        return;
    }
    const double alpha = rz / pq;
    double rr = 0.0;
    for (long long i = a + threadIdx.x; i < e; i += NTH) {
        x[i] += alpha * p[i];
        const double ri = r[i] - alpha * q[i];
        r[i] = ri;
        rr += ri * ri;
    }
    rr = block_sum_n<NTH>(rr, red);
    if (threadIdx.x == 0) {
        // it < 0: the block's own count of completed iterations (s[3], which equals the
        // loop index while the block is active), so every iteration launches the same
        // kernel arguments and a loop of them can be replayed as a graph
        const int itx = it >= 0 ? it : (int)s[3];
        if (itx + 1 < hs) hist[(long long)b * hs + itx + 1] = rr;
        s[3] = itx + 1;
        if (rr <= stop_rel2 * s[1]) {
            s[2] = 0.0;
        } else if (prune > 0 && itx + 1 >= prune && (itx + 1 - prune) % KP_PRUNE_SPAN == 0 &&
                   itx + 1 < hs) {
            const double old = hist[(long long)b * hs + itx + 1 - KP_PRUNE_SPAN];
            const double rate = old > 0.0 ? rr / old : 1.0;     // squared-norm ratio per span
            bool hopeless = !(rate < 1.0);
            if (!hopeless) {
                const double spans = log(stop_rel2 * s[1] / rr) / log(rate);
                hopeless = itx + 1 + KP_PRUNE_SPAN * spans > (double)cap;
            }
            if (hopeless) s[2] = 0.0;
        }
    }
}

__global__ void __launch_bounds__(256)
kpcg_step_a_kernel(double *__restrict__ x, double *__restrict__ r, const double *__restrict__ p,
                   double *__restrict__ q, const double *__restrict__ lam,
                   const long long *__restrict__ seg, double *__restrict__ st,
                   double *__restrict__ hist, int hs, int it, double stop_rel2, int prune, int cap)
{
    __shared__ double red[256];
    double *s = st + 4 * blockIdx.x;
    if (s[2] == 0.0) return;
    kpcg_step_a_block<256, false>(blockIdx.x, x, r, p, q, lam[2 * blockIdx.x], seg[blockIdx.x],
                                  seg[blockIdx.x + 1], s, hist, hs, it, stop_rel2, prune, cap, red);
}

// The CG operator's backward product with step_a fused: problem b (block order, skipped
// while the block is inactive) is S_b^T Z_b, this workgroup its row tile and K split.
// The last of a row tile's splits to finish (a counter per row tile, taken after the
// workgroup's stores have completed) adds the partials in split order into q; the last
// row tile of the block to finish (a counter per block) runs step_a for the whole block,
// reading q with agent-scope loads.  The backward's second pass and step_a are gone as
// launches; their sums are those of a 512-thread step_a.
template <int NT, bool X16>
__device__ __forceinline__ void kpcg_bwd_tile(const SkDesc &d, int m0, int z, int b, unsigned *tcnt,
                                              unsigned *bcnt, double *__restrict__ x, double *__restrict__ r,
                                              const double *__restrict__ p, double *__restrict__ q, double l,
                                              long long a, long long e, double *__restrict__ s,
                                              double *__restrict__ hist, int hs, double stop_rel2, int prune,
                                              int cap, double (&red)[SK_WAVES / 2][16][33], double *red1)
{
    __shared__ int last;
    const int ks = skinny_splits(d.K), Nc = d.Nc;
    skinny_tile<NT, 1, X16, true>(d, m0, z, red);
    __builtin_amdgcn_s_waitcnt(0);                // this wave's partial or q stores have completed
    __syncthreads();
    if (ks > 1) {
        if (threadIdx.x == 0) {
            const unsigned o = __hip_atomic_fetch_add(tcnt, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            last = o == (unsigned)ks - 1;
            if (last) __hip_atomic_store(tcnt, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
        __syncthreads();
        if (!last) return;
        const int rows = min(16, d.M - m0);
        const long long mn = (long long)d.M * Nc;
        for (int t = threadIdx.x; t < rows * Nc; t += 64 * SK_WAVES) {
            const long long loc = (long long)(m0 + t / Nc) * Nc + t % Nc;
            double qi = 0.0;
            for (int zz = 0; zz < ks; ++zz)
                qi += __hip_atomic_load(d.ws + zz * mn + loc, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            __hip_atomic_store(d.C + loc, qi, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
        __builtin_amdgcn_s_waitcnt(0);            // the tile's q stores have completed
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        const unsigned o = __hip_atomic_fetch_add(bcnt, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        last = o == (unsigned)((d.M + 15) / 16) - 1;
        if (last) __hip_atomic_store(bcnt, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    __syncthreads();
    if (!last) return;
    kpcg_step_a_block<64 * SK_WAVES, true>(b, x, r, p, q, l, a, e, s, hist, hs, -1, stop_rel2, prune, cap, red1);
}

// grid (row tiles, batch, max splits); cnt: row-tile counters (batch * row tiles), then a
// counter per block, zero before the first launch and left zero by every launch
template <int NT, bool X16>
__global__ void __launch_bounds__(64 * SK_WAVES)
kpcg_bwd_kernel(const SkDesc *__restrict__ descs, double *__restrict__ x, double *__restrict__ r,
                const double *__restrict__ p, double *__restrict__ q, const double *__restrict__ lam,
                const long long *__restrict__ seg, double *__restrict__ st, double *__restrict__ hist, int hs,
                double stop_rel2, int prune, int cap, unsigned *__restrict__ cnt)
{
    __shared__ double red[SK_WAVES / 2][16][33];
    __shared__ double red1[64 * SK_WAVES];
    const int b = blockIdx.y;
    const SkDesc d = descs[b];
    const int m0 = blockIdx.x * 16;
    if (m0 >= d.M || (int)blockIdx.z >= skinny_splits(d.K)) return;
    if (d.active && *d.active == 0.0) return;
    double *sb = st + 4 * b;
    if (sb[2] == 0.0) return;                    // the block has stopped (as step_a would)
    const int ntile = gridDim.x;
    unsigned *tc = cnt + (long long)b * ntile + blockIdx.x;
    unsigned *bc = cnt + (long long)gridDim.y * ntile + b;
    const double l = lam[2 * b];
    const long long a = seg[b], e = seg[b + 1];
    const int z = blockIdx.z;
    if constexpr (NT == 1) {
        if (X16 && d.Nc == 17)
            kpcg_bwd_tile<1, true>(d, m0, z, b, tc, bc, x, r, p, q, l, a, e, sb, hist, hs, stop_rel2, prune, cap, red, red1);
        else
            kpcg_bwd_tile<1, false>(d, m0, z, b, tc, bc, x, r, p, q, l, a, e, sb, hist, hs, stop_rel2, prune, cap, red, red1);
    } else {
        if (d.Nc == 17)
            kpcg_bwd_tile<1, true>(d, m0, z, b, tc, bc, x, r, p, q, l, a, e, sb, hist, hs, stop_rel2, prune, cap, red, red1);
        else if (d.Nc <= 16)
            kpcg_bwd_tile<1, false>(d, m0, z, b, tc, bc, x, r, p, q, l, a, e, sb, hist, hs, stop_rel2, prune, cap, red, red1);
        else
            kpcg_bwd_tile<2, false>(d, m0, z, b, tc, bc, x, r, p, q, l, a, e, sb, hist, hs, stop_rel2, prune, cap, red, red1);
    }
}

VT_EXPORT int vt_kpcg_step_a(double *x, double *r, const double *p, double *q, const double *lam,
                             const long long *seg, int batch, double *st, double *hist,
                             int hist_stride, int it, double stop_rel2, int prune, int cap,
                             void *stream)
{
    VT_CHECK_ARG(x && r && p && q && lam && seg && st && hist && batch >= 1 && it >= -1 &&
                 it + 1 < hist_stride && (prune == 0 || prune >= KP_PRUNE_SPAN));
    kpcg_step_a_kernel<<<batch, 256, 0, (hipStream_t)stream>>>(x, r, p, q, lam, seg, st, hist,
                                                               hist_stride, it, stop_rel2, prune, cap);
    VT_LAUNCH_CHECK();
    return VT_OK;
}

// rz' = <r, z>; beta = rz'/rz; p = z + beta p
__global__ void __launch_bounds__(256)
kpcg_step_b_kernel(const double *__restrict__ r, const double *__restrict__ z, double *__restrict__ p,
                   const long long *__restrict__ seg, double *__restrict__ st, double *__restrict__ pt,
                   const long long *__restrict__ ptab)
{
    __shared__ double red[256];
    double *s = st + 4 * blockIdx.x;
    if (s[2] == 0.0) return;
    const long long a = seg[blockIdx.x], e = seg[blockIdx.x + 1];
    double rz = 0.0;
    for (long long i = a + threadIdx.x; i < e; i += 256) rz += r[i] * z[i];
    rz = block_sum(rz, red);
    const double beta = rz / s[0];
    for (long long i = a + threadIdx.x; i < e; i += 256) {
        const double pi = z[i] + beta * p[i];
        p[i] = pi;
        if (pt) kp_store_tiled(pt, ptab, blockIdx.x, i - a, e - a, pi);
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        s[0] = rz;
        if (!(rz > 0.0)) s[2] = 0.0;
    }
}

VT_EXPORT int vt_kpcg_step_b(const double *r, const double *z, double *p, const long long *seg,
                             int batch, double *st, double *pt, const long long *ptab, void *stream)
{
    VT_CHECK_ARG(r && z && p && seg && st && batch >= 1 && (!pt || ptab));
    kpcg_step_b_kernel<<<batch, 256, 0, (hipStream_t)stream>>>(r, z, p, seg, st, pt, ptab);
    VT_LAUNCH_CHECK();
    return VT_OK;
}

// ------------------------------------------------------------------------------------
// Whole iteration loops on the host side of the library (one C call enqueues every
// launch of the solve, with no per-launch round trip through the caller).
// The operator and preconditioner are skinny grouped products (vt_gemm_skinny).
// ------------------------------------------------------------------------------------
extern "C" int vt_gemm_skinny(const void *descs, int batch, int max_M, int max_K, int max_Nc,
                              void *stream);

struct VtSkinnyLaunch {
    const void *descs;
    int max_M, max_K;
};

static int skinny_run(const VtSkinnyLaunch &l, int batch, int max_Nc, void *stream)
{
    return vt_gemm_skinny(l.descs, batch, l.max_M, l.max_K, max_Nc, stream);
}

// `iters` CG iterations (after vt_kpcg_init): launches = {fwd, bwd} as {descs, max_M, max_K}
// triples for the operator halves (batch problems each, max_Nc columns), and the
// preconditioner as up to three skinny launches run in order, each {descs, count, max_M,
// max_K, max_Nc} (a null descs skips it).
VT_EXPORT int vt_kpcg_iterate(const void *fwd, int fwd_M, int fwd_K, const void *bwd, int bwd_M,
                              int bwd_K, const void *pre, int pre_n, int pre_M, int pre_K, int pre_Nc,
                              const void *pre2, int pre2_n, int pre2_M, int pre2_K, int pre2_Nc,
                              const void *pre3, int pre3_n, int pre3_M, int pre3_K, int pre3_Nc,
                              int batch, int max_Nc, double *x, double *r, double *p, double *q,
                              double *z, const double *lam, const long long *seg, double *st,
                              double *hist, int hist_stride, int iters, double stop_rel2,
                              int prune, int graph_chunk, unsigned *cnt, double *pt, const long long *ptab,
                              void *stream)
{
    VT_CHECK_ARG(fwd && bwd && pre && cnt && iters >= 0 && iters < hist_stride && graph_chunk >= 0 &&
                 batch >= 1 && batch <= 65535 && bwd_M >= 1 && max_Nc >= 1 && max_Nc <= 32 &&
                 (prune == 0 || prune >= KP_PRUNE_SPAN));
    const VtSkinnyLaunch F{fwd, fwd_M, fwd_K}, Bk{bwd, bwd_M, bwd_K};
    struct PreLaunch { const void *d; int n, M, K, Nc; };
    const PreLaunch P[3] = {{pre, pre_n, pre_M, pre_K, pre_Nc}, {pre2, pre2_n, pre2_M, pre2_K, pre2_Nc},
                            {pre3, pre3_n, pre3_M, pre3_K, pre3_Nc}};
    hipStream_t s0 = (hipStream_t)stream;
    const int ntile = (int)vt_blocks(bwd_M, 16);
    const long long words = (long long)batch * (ntile + 1);   // row-tile counters, then one per block
    const unsigned gw = vt_blocks(words, 256);
    zero_words_kernel<<<gw < 1024 ? gw : 1024, 256, 0, s0>>>(cnt, words);   // a kernel, not a memset node
    VT_LAUNCH_CHECK();
    const dim3 bgrid(ntile, batch, skinny_splits(bwd_K));
    const SkDesc *bd = (const SkDesc *)bwd;
    // the iteration index comes from each block's count: identical iterations
    auto step = [&](hipStream_t sq) -> int {
        VT_TRY(skinny_run(F, batch, max_Nc, sq));
        if (max_Nc <= 16)
            kpcg_bwd_kernel<1, false><<<bgrid, 64 * SK_WAVES, 0, sq>>>(bd, x, r, p, q, lam, seg, st, hist, hist_stride,
                                                                   stop_rel2, prune, iters, cnt);
        else if (max_Nc == 17)
            kpcg_bwd_kernel<1, true><<<bgrid, 64 * SK_WAVES, 0, sq>>>(bd, x, r, p, q, lam, seg, st, hist, hist_stride,
                                                                  stop_rel2, prune, iters, cnt);
        else
            kpcg_bwd_kernel<2, false><<<bgrid, 64 * SK_WAVES, 0, sq>>>(bd, x, r, p, q, lam, seg, st, hist, hist_stride,
                                                                   stop_rel2, prune, iters, cnt);
        VT_LAUNCH_CHECK();
        for (int i = 0; i < 3; ++i)
            if (P[i].d) VT_TRY(vt_gemm_skinny(P[i].d, P[i].n, P[i].M, P[i].K, P[i].Nc, sq));
        VT_TRY(vt_kpcg_step_b(r, z, p, seg, batch, st, pt, ptab, sq));
        return VT_OK;
    };
    return vt_graph_loop(s0, iters, graph_chunk, step);
}

// ------------------------------------------------------------------------------------
// Gradient-flow (RNN) decoder in coefficient space: X <- X + eta_b (r - (A + lam I) X),
// the image under u = Phi^T c of the reference's c <- c + eta (q - G c).  eta_b = 1/theta_b
// with theta_b the Lanczos estimate of the largest eigenvalue of A + lam I.
// ------------------------------------------------------------------------------------
__global__ void zero_kernel(double *__restrict__ x, long long n)
{
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (long long)gridDim.x * blockDim.x)
        x[i] = 0.0;
}

// Backward product of the RNN step with the step itself fused: problem b (block order) is
// S_b^T Z_b, this workgroup its row tile m0 and K split z.  An unsplit problem's workgroup
// updates its rows at once, x += eta_b (r - q - lam_b x).  Split problems store their
// partials agent-coherent (the L2 caches of the XCDs are separate); the last of a row
// tile's ks splits to finish (a counter per row tile, taken after the workgroup's stores
// have completed) reads the partials back with agent-scope loads, adds them in split
// order, as the product's own second pass does, updates the rows and resets the counter
// for the next iteration.  One launch per iteration fewer than a separate update, same
// values: the counters only pick which workgroup does the update.
// xt (optional): the block's iterate also stored as the forward product's tiled B (the
// tiles of x^T, SK_B_TILED, ceil(M / 16) tiles per row), each element where the update
// writes it.
template <int NT, bool X16>
__device__ __forceinline__ void rnn_bwd_tile(const SkDesc &d, int m0, int z, double *__restrict__ x,
                                             const double *__restrict__ r, double lam, double eta,
                                             long long s0, unsigned *cnt, double *__restrict__ xt,
                                             double (&red)[SK_WAVES / 2][16][33])
{
    __shared__ int last;
    const int ks = skinny_splits(d.K), Nc = d.Nc;
    skinny_tile<NT, 1, X16, true>(d, m0, z, red);
    __builtin_amdgcn_s_waitcnt(0);                // this wave's partial or q stores have completed
    __syncthreads();
    if (ks > 1) {
        if (threadIdx.x == 0) {
            const unsigned o = __hip_atomic_fetch_add(cnt, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            last = o == (unsigned)ks - 1;
            if (last) __hip_atomic_store(cnt, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
        __syncthreads();
        if (!last) return;
    }
    const int rows = min(16, d.M - m0);
    const long long mn = (long long)d.M * Nc;
    for (int e = threadIdx.x; e < rows * Nc; e += 64 * SK_WAVES) {
        const long long loc = (long long)(m0 + e / Nc) * Nc + e % Nc;
        double qi;
        if (ks > 1) {
            qi = 0.0;
            for (int zz = 0; zz < ks; ++zz)
                qi += __hip_atomic_load(d.ws + zz * mn + loc, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        } else {
            qi = __hip_atomic_load(d.C + loc, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
        const long long i = s0 + loc;
        const double xn = rnn_step(x[i], r[i], qi, lam, eta);
        x[i] = xn;
        if (xt) xt[sk_tile_index(e % Nc, m0 + e / Nc, (d.M + 15) / 16)] = xn;
    }
}

// grid (row tiles, batch, max splits); the route by the problem's own Nc, as skinny_kernel
template <int NT, bool X16>
__global__ void __launch_bounds__(64 * SK_WAVES)
rnn_bwd_update_kernel(const SkDesc *__restrict__ descs, double *__restrict__ x, const double *__restrict__ r,
                      const double *__restrict__ lam, const double *__restrict__ eta,
                      const long long *__restrict__ seg, unsigned *__restrict__ cnt,
                      double *__restrict__ xt, const long long *__restrict__ xtoff)
{
    __shared__ double red[SK_WAVES / 2][16][33];
    const int b = blockIdx.y;
    const SkDesc d = descs[b];
    const int m0 = blockIdx.x * 16;
    if (m0 >= d.M || (int)blockIdx.z >= skinny_splits(d.K)) return;
    unsigned *c = cnt + (long long)b * gridDim.x + blockIdx.x;
    const double l = lam[2 * b], et = eta[b];
    const long long s0 = seg[b];
    double *xb = xt ? xt + xtoff[b] : nullptr;
    if constexpr (NT == 1) {
        if (X16 && d.Nc == 17)
            rnn_bwd_tile<1, true>(d, m0, blockIdx.z, x, r, l, et, s0, c, xb, red);
        else
            rnn_bwd_tile<1, false>(d, m0, blockIdx.z, x, r, l, et, s0, c, xb, red);
    } else {
        if (d.Nc == 17)
            rnn_bwd_tile<1, true>(d, m0, blockIdx.z, x, r, l, et, s0, c, xb, red);
        else if (d.Nc <= 16)
            rnn_bwd_tile<1, false>(d, m0, blockIdx.z, x, r, l, et, s0, c, xb, red);
        else
            rnn_bwd_tile<2, false>(d, m0, blockIdx.z, x, r, l, et, s0, c, xb, red);
    }
}

static int rnn_bwd_update(const VtSkinnyLaunch &Bk, int batch, int max_Nc, double *x, const double *r,
                          const double *lam, const double *eta, const long long *seg, unsigned *cnt,
                          double *xt, const long long *xtoff, hipStream_t s)
{
    if (Bk.max_M <= 0) return VT_OK;
    const dim3 grid(vt_blocks(Bk.max_M, 16), batch, skinny_splits(Bk.max_K));
    const SkDesc *d = (const SkDesc *)Bk.descs;
    if (max_Nc <= 16)
        rnn_bwd_update_kernel<1, false><<<grid, 64 * SK_WAVES, 0, s>>>(d, x, r, lam, eta, seg, cnt, xt, xtoff);
    else if (max_Nc == 17)
        rnn_bwd_update_kernel<1, true><<<grid, 64 * SK_WAVES, 0, s>>>(d, x, r, lam, eta, seg, cnt, xt, xtoff);
    else
        rnn_bwd_update_kernel<2, false><<<grid, 64 * SK_WAVES, 0, s>>>(d, x, r, lam, eta, seg, cnt, xt, xtoff);
    VT_LAUNCH_CHECK();
    return VT_OK;
}




// fixed-order sum over the 64 * SK_WAVES threads of a workgroup (red: as many doubles)
__device__ __forceinline__ double wg_sum(double v, double *red)
{
    red[threadIdx.x] = v;
    __syncthreads();
    for (int o = 32 * SK_WAVES; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
        __syncthreads();
    }
    const double r = red[0];
    __syncthreads();
    return r;
}

// Lanczos estimate of lambda_max(A_b + lam_b I), the RNN step eta_b = 1/lambda_max.  Per
// step j, v_j of unit norm (v_0 = ones / sqrt(D_b)):
//   u = (A + lam I) v_j,  alpha_j = <u, v_j>,  w = u - alpha_j v_j - beta_{j-1} v_{j-1},
//   beta_j = |w|,  v_{j+1} = w / beta_j,
// and eta_b = 1/theta with theta the largest eigenvalue of the tridiagonal matrix of the
// alphas and betas.  Fused into the backward launch as the RNN step is: the last split of
// a row tile forms u for its rows (in q's place) and the tile's sum <u, v>; the last row
// tile of the block adds the tiles' sums in tile order (alpha), then makes two passes over
// the block: |w|^2 (a fixed-order reduction), then v_{j-1} <- v_j, v_j <- w / beta_j.  A
// zero or non-finite beta (an invariant subspace) stops the block's recurrence (v = 0).
// The step index comes from each block's own counter, so the steps are identical launches
// (graph replay).
template <int NT, bool X16>
__device__ __forceinline__ void lanczos_bwd_tile(const SkDesc &d, int m0, int z, int b, int tile, int ntile,
                                                 double *__restrict__ v, double *__restrict__ vp, double l,
                                                 long long s0, long long s1, unsigned *tcnt, unsigned *bcnt,
                                                 unsigned *jstep, double *part, double *__restrict__ tri,
                                                 int tri_ld, double *__restrict__ vt, const long long *__restrict__ vtab,
                                                 double (&red)[SK_WAVES / 2][16][33], double *red1)
{
    __shared__ int last;
    __shared__ double ab[3];
    const int ks = skinny_splits(d.K), Nc = d.Nc;
    skinny_tile<NT, 1, X16, true>(d, m0, z, red);
    __builtin_amdgcn_s_waitcnt(0);                // this wave's partial or q stores have completed
    __syncthreads();
    if (ks > 1) {
        if (threadIdx.x == 0) {
            const unsigned o = __hip_atomic_fetch_add(tcnt, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            last = o == (unsigned)ks - 1;
            if (last) __hip_atomic_store(tcnt, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
        __syncthreads();
        if (!last) return;
    }
    const int rows = min(16, d.M - m0);
    const long long mn = (long long)d.M * Nc;
    double uv = 0.0;
    for (int e = threadIdx.x; e < rows * Nc; e += 64 * SK_WAVES) {
        const long long loc = (long long)(m0 + e / Nc) * Nc + e % Nc;
        double qi;
        if (ks > 1) {
            qi = 0.0;
            for (int zz = 0; zz < ks; ++zz)
                qi += __hip_atomic_load(d.ws + zz * mn + loc, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        } else {
            qi = __hip_atomic_load(d.C + loc, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
        const double vi = v[s0 + loc];
        const double ui = qi + l * vi;
        __hip_atomic_store(d.C + loc, ui, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        uv += ui * vi;
    }
    uv = wg_sum(uv, red1);
    if (threadIdx.x == 0)
        __hip_atomic_store(part + (long long)b * ntile + tile, uv, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __builtin_amdgcn_s_waitcnt(0);                // the u and partial-sum stores have completed
    __syncthreads();
    const int nt = (d.M + 15) / 16;
    if (threadIdx.x == 0) {
        const unsigned o = __hip_atomic_fetch_add(bcnt, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        last = o == (unsigned)nt - 1;
        if (last) {
            __hip_atomic_store(bcnt, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            double a = 0.0;
            for (int t = 0; t < nt; ++t)
                a += __hip_atomic_load(part + (long long)b * ntile + t, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            const unsigned j = *jstep;
            ab[0] = a;
            ab[1] = j > 0 ? tri[(long long)b * tri_ld + 2 * (j - 1) + 1] : 0.0;
        }
    }
    __syncthreads();
    if (!last) return;
    const double alpha = ab[0], bprev = ab[1];
    const double *u = d.C;                        // the block's u, at q + s0
    double ww = 0.0;
    for (long long i = s0 + threadIdx.x; i < s1; i += 64 * SK_WAVES) {
        const double w = __hip_atomic_load(u + (i - s0), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) -
                         alpha * v[i] - bprev * vp[i];
        ww += w * w;
    }
    ww = wg_sum(ww, red1);
    double beta = sqrt(ww);
    if (!(beta > 0.0) || !(beta < INFINITY)) beta = 0.0;
    const double inv = beta > 0.0 ? 1.0 / beta : 0.0;
    for (long long i = s0 + threadIdx.x; i < s1; i += 64 * SK_WAVES) {
        const double vi = v[i];
        const double w = __hip_atomic_load(u + (i - s0), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) -
                         alpha * vi - bprev * vp[i];
        vp[i] = vi;
        v[i] = w * inv;
        if (vt) kp_store_tiled(vt, vtab, b, i - s0, s1 - s0, w * inv);
    }
    if (threadIdx.x == 0) {
        const unsigned j = *jstep;
        tri[(long long)b * tri_ld + 2 * j] = alpha;
        tri[(long long)b * tri_ld + 2 * j + 1] = beta;
        *jstep = j + 1;
    }
}

// grid (row tiles, batch, max splits); cnt: row-tile counters (batch * row tiles), a
// counter per block, then a step index per block; part: a double per block and row tile
template <int NT, bool X16>
__global__ void __launch_bounds__(64 * SK_WAVES)
lanczos_bwd_kernel(const SkDesc *__restrict__ descs, double *__restrict__ v, double *__restrict__ vp,
                   const double *__restrict__ lam, const long long *__restrict__ seg, unsigned *__restrict__ cnt,
                   double *__restrict__ part, double *__restrict__ tri, int tri_ld, double *__restrict__ vt,
                   const long long *__restrict__ vtab)
{
    __shared__ double red[SK_WAVES / 2][16][33];
    __shared__ double red1[64 * SK_WAVES];
    const int b = blockIdx.y;
    const SkDesc d = descs[b];
    const int m0 = blockIdx.x * 16;
    if (m0 >= d.M || (int)blockIdx.z >= skinny_splits(d.K)) return;
    const int ntile = gridDim.x;
    const long long B = gridDim.y;
    unsigned *tc = cnt + (long long)b * ntile + blockIdx.x;
    unsigned *bc = cnt + B * ntile + b;
    unsigned *js = cnt + B * ntile + B + b;
    const double l = lam[2 * b];
    const long long s0 = seg[b], s1 = seg[b + 1];
    const int z = blockIdx.z, t = blockIdx.x;
    if constexpr (NT == 1) {
        if (X16 && d.Nc == 17)
            lanczos_bwd_tile<1, true>(d, m0, z, b, t, ntile, v, vp, l, s0, s1, tc, bc, js, part, tri, tri_ld, vt, vtab, red, red1);
        else
            lanczos_bwd_tile<1, false>(d, m0, z, b, t, ntile, v, vp, l, s0, s1, tc, bc, js, part, tri, tri_ld, vt, vtab, red, red1);
    } else {
        if (d.Nc == 17)
            lanczos_bwd_tile<1, true>(d, m0, z, b, t, ntile, v, vp, l, s0, s1, tc, bc, js, part, tri, tri_ld, vt, vtab, red, red1);
        else if (d.Nc <= 16)
            lanczos_bwd_tile<1, false>(d, m0, z, b, t, ntile, v, vp, l, s0, s1, tc, bc, js, part, tri, tri_ld, vt, vtab, red, red1);
        else
            lanczos_bwd_tile<2, false>(d, m0, z, b, t, ntile, v, vp, l, s0, s1, tc, bc, js, part, tri, tri_ld, vt, vtab, red, red1);
    }
}

// v_0 = ones / sqrt(D_b), v_{-1} = 0 per block
__global__ void __launch_bounds__(256)
lanczos_init_kernel(double *__restrict__ v, double *__restrict__ vp, const long long *__restrict__ seg,
                    double *__restrict__ vt, const long long *__restrict__ vtab)
{
    const long long a = seg[blockIdx.x], e = seg[blockIdx.x + 1];
    const double x = e > a ? 1.0 / sqrt((double)(e - a)) : 0.0;
    for (long long i = a + threadIdx.x; i < e; i += 256) {
        v[i] = x;
        vp[i] = 0.0;
        if (vt) kp_store_tiled(vt, vtab, blockIdx.x, i - a, e - a, x);
    }
}

// Sturm count: eigenvalues of the k x k symmetric tridiagonal (alpha_i = t[2i], off-diagonal
// beta_i = t[2i+1] between i and i+1) below x
__device__ int tri_count_below(const double *t, int k, double x)
{
    int c = 0;
    double dd = 1.0;
    for (int i = 0; i < k; ++i) {
        const double b2 = i > 0 ? t[2 * (i - 1) + 1] * t[2 * (i - 1) + 1] : 0.0;
        dd = (t[2 * i] - x) - (i > 0 ? b2 / dd : 0.0);
        if (dd == 0.0) dd = -1e-300;
        if (dd < 0.0) ++c;
    }
    return c;
}

// eta_b = 1 / (largest eigenvalue of block b's tridiagonal), by bisection on the Sturm count
__global__ void lanczos_eta_kernel(const double *__restrict__ tri, int tri_ld, int k, int batch,
                                   double *__restrict__ eta)
{
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= batch) return;
    const double *t = tri + (long long)b * tri_ld;
    double lo = INFINITY, hi = -INFINITY;
    for (int i = 0; i < k; ++i) {           // Gershgorin interval
        const double r = (i > 0 ? fabs(t[2 * (i - 1) + 1]) : 0.0) + (i + 1 < k ? fabs(t[2 * i + 1]) : 0.0);
        lo = fmin(lo, t[2 * i] - r);
        hi = fmax(hi, t[2 * i] + r);
    }
    for (int it = 0; it < 200 && hi - lo > 1e-15 * fmax(fabs(lo), fabs(hi)); ++it) {
        const double mid = 0.5 * (lo + hi);
        if (tri_count_below(t, k, mid) < k) lo = mid;    // the largest eigenvalue is >= mid
        else hi = mid;
    }
    const double th = 0.5 * (lo + hi);
    eta[b] = th > 0.0 ? 1.0 / th : 0.0;
}

static int lanczos_bwd(const VtSkinnyLaunch &Bk, int batch, int max_Nc, double *v, double *vp,
                       const double *lam, const long long *seg, unsigned *cnt, double *part, double *tri,
                       int tri_ld, double *vt, const long long *vtab, hipStream_t s)
{
    const dim3 grid(vt_blocks(Bk.max_M, 16), batch, skinny_splits(Bk.max_K));
    const SkDesc *d = (const SkDesc *)Bk.descs;
    if (max_Nc <= 16)
        lanczos_bwd_kernel<1, false><<<grid, 64 * SK_WAVES, 0, s>>>(d, v, vp, lam, seg, cnt, part, tri, tri_ld, vt, vtab);
    else if (max_Nc == 17)
        lanczos_bwd_kernel<1, true><<<grid, 64 * SK_WAVES, 0, s>>>(d, v, vp, lam, seg, cnt, part, tri, tri_ld, vt, vtab);
    else
        lanczos_bwd_kernel<2, false><<<grid, 64 * SK_WAVES, 0, s>>>(d, v, vp, lam, seg, cnt, part, tri, tri_ld, vt, vtab);
    VT_LAUNCH_CHECK();
    return VT_OK;
}

// eta (batch doubles) = 1/theta_b from `steps` Lanczos steps per block (v, vp: the joint
// vector's two work copies; q: the backward's output, holds u and w); per step the forward
// operator launch and the backward launch with the step fused.  bwd's descriptors in block
// order; cnt: batch * ceil(bwd_M / 16) + 2 batch words (zeroed here); part: batch * ceil(bwd_M / 16)
// doubles; tri: batch * 2 * steps doubles (alpha_j, beta_j of every block).
VT_EXPORT int vt_kron_lanczos(const void *fwd, int fwd_M, int fwd_K, const void *bwd, int bwd_M, int bwd_K,
                              int batch, int max_Nc, double *v, double *vp, double *q, const double *lam,
                              const long long *seg, int steps, double *eta, int graph_chunk, unsigned *cnt,
                              double *part, double *tri, double *vt, const long long *vtab, void *stream)
{
    VT_CHECK_ARG(fwd && bwd && v && vp && q && lam && seg && eta && cnt && part && tri && batch >= 1 &&
                 batch <= 65535 && steps >= 1 && graph_chunk >= 0 && bwd_M >= 1 && max_Nc >= 1 &&
                 max_Nc <= 32 && (!vt || vtab));
    const VtSkinnyLaunch F{fwd, fwd_M, fwd_K}, Bk{bwd, bwd_M, bwd_K};
    hipStream_t s = (hipStream_t)stream;
    const long long ntile = vt_blocks(bwd_M, 16);
    const long long words = (long long)batch * ntile + 2LL * batch;
    const unsigned gw = vt_blocks(words, 256);
    zero_words_kernel<<<gw < 1024 ? gw : 1024, 256, 0, s>>>(cnt, words);
    lanczos_init_kernel<<<batch, 256, 0, s>>>(v, vp, seg, vt, vtab);
    VT_LAUNCH_CHECK();
    const int tri_ld = 2 * steps;
    auto step = [&](hipStream_t st) -> int {
        VT_TRY(skinny_run(F, batch, max_Nc, st));
        VT_TRY(lanczos_bwd(Bk, batch, max_Nc, v, vp, lam, seg, cnt, part, tri, tri_ld, vt, vtab, st));
        return VT_OK;
    };
    VT_TRY(vt_graph_loop(s, steps, graph_chunk, step));
    lanczos_eta_kernel<<<vt_blocks(batch, 64), 64, 0, s>>>(tri, tri_ld, steps, batch, eta);
    VT_LAUNCH_CHECK();
    return VT_OK;
}

// steps+1 gradient-flow iterations from x = 0: per iteration the forward operator launch
// (reads x) and the backward launch with the update fused (writes x); graph_chunk >= 2
// replays them as HIP graphs of that many iterations (vt_graph_loop), 0 launches each.  cnt: batch * ceil(bwd_M / 16) words of row-tile counters (zeroed here).
// xt (optional, n_xt doubles, zeroed here): the iterate's tiled copy that forward
// descriptors with SK_B_TILED read, block b's at xt + xtoff[b]; null: x row-major.
VT_EXPORT int vt_kron_rnn(const void *fwd, int fwd_M, int fwd_K, const void *bwd, int bwd_M, int bwd_K,
                          int batch, int max_Nc, double *x, const double *r, double *q,
                          const double *lam, const double *eta, const long long *seg, long long n,
                          int steps, int graph_chunk, unsigned *cnt, double *xt, const long long *xtoff,
                          long long n_xt, void *stream)
{
    VT_CHECK_ARG(fwd && bwd && x && r && q && lam && eta && seg && cnt && n >= 0 && steps >= 0 &&
                 graph_chunk >= 0 && batch >= 1 && batch <= 65535 && max_Nc >= 1 && max_Nc <= 32 &&
                 (!xt || (xtoff && n_xt > 0)));
    const VtSkinnyLaunch F{fwd, fwd_M, fwd_K}, Bk{bwd, bwd_M, bwd_K};
    hipStream_t s = (hipStream_t)stream;
    unsigned g = vt_blocks(n, 256);
    if (g > 4096) g = 4096;
    zero_kernel<<<g, 256, 0, s>>>(x, n);   // kernels, not memset nodes, if captured
    if (xt) {
        unsigned gx = vt_blocks(n_xt, 256);
        zero_kernel<<<gx > 4096 ? 4096 : gx, 256, 0, s>>>(xt, n_xt);
    }
    const long long words = (long long)batch * vt_blocks(bwd_M > 0 ? bwd_M : 1, 16);
    const unsigned gw = vt_blocks(words, 256);
    zero_words_kernel<<<gw < 1024 ? gw : 1024, 256, 0, s>>>(cnt, words);
    VT_LAUNCH_CHECK();
    auto step = [&](hipStream_t st) -> int {
        VT_TRY(skinny_run(F, batch, max_Nc, st));
        VT_TRY(rnn_bwd_update(Bk, batch, max_Nc, x, r, lam, eta, seg, cnt, xt, xtoff, st));
        return VT_OK;
    };
    return vt_graph_loop(s, steps + 1, graph_chunk, step);
}
