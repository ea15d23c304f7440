// Decode side, element-wise and segmented-reduction kernels.
//
// Coefficient space (D-space) factors, index conventions:
//   temporal basis  Theta_0 = 1, Theta_{2m-1} = sqrt2 cos(m W t), Theta_{2m} = sqrt2 sin(m W t)
//   alpha[k, pt]    = integral of Theta_pt over [tk1_k, tk2_k]
//   T_j[pt, pt']    = sum_{k in j} norm_j^2 alpha[k,pt] alpha[k,pt']
//   t_j[pt]         = sum_{k in j} norm_j alpha[k,pt] q_k
// Spike-space kernels restate vtem/dirichlet.py (compute_q :204-222, G :272-330,
// diag_add :334-350, reconstruct :354-420) and the RNN update of
// vtem/neural_network.py:102-128.
#include "vt_common.hpp"
#include "vt_skinny.hpp"

#define SQRT2 1.4142135623730950488

// q_k = (kappa_j delta_j - b_j (tk2 - tk1)) norm_j   (dirichlet.py:216)
__global__ void q_kernel(double *__restrict__ q, const double *__restrict__ tk1,
                         const double *__restrict__ tk2, const int *__restrict__ nid,
                         const double *__restrict__ kappa, const double *__restrict__ delta,
                         const double *__restrict__ bias, const double *__restrict__ norm,
                         const int *__restrict__ n_dev, int n_host)
{
    const int n = n_dev ? *n_dev : n_host;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        int j = nid[i];
        q[i] = (kappa[j] * delta[j] - bias[j] * (tk2[i] - tk1[i])) * norm[j];
    }
}

VT_EXPORT int vt_q(double *q, const double *tk1, const double *tk2, const int *neuron_ind,
                   const double *kappa, const double *delta, const double *bias,
                   const double *norm, const int *n_dev, int n, void *stream)
{
    VT_CHECK_ARG(q && tk1 && tk2 && neuron_ind && kappa && delta && bias && norm && n >= 0);
    if (n == 0) return VT_OK;
    q_kernel<<<vt_blocks(n, 256) < 2048 ? vt_blocks(n, 256) : 2048, 256, 0, (hipStream_t)stream>>>(
        q, tk1, tk2, neuron_ind, kappa, delta, bias, norm, n_dev, n);
    VT_LAUNCH_CHECK();
    return VT_OK;
}

// alpha for one interval into a[0..Dt)
__device__ __forceinline__ void temporal_factor(double t1, double t2, int Lt, double W, double w,
                                                double *a)
{
    a[0] = (t2 - t1) * w;
    const double mid = 0.5 * (t1 + t2), half = 0.5 * (t2 - t1);
    for (int m = 1; m <= Lt; ++m) {
        const double om = m * W;
        double s = 2.0 * SQRT2 * sin(om * half) / om;
        double sm, cm;
        sincos(om * mid, &sm, &cm);
        a[2 * m - 1] = s * cm * w;
        a[2 * m] = s * sm * w;
    }
}

__global__ void temporal_factor_kernel(double *__restrict__ alpha, const double *__restrict__ tk1,
                                       const double *__restrict__ tk2, int Lt, double W,
                                       const int *__restrict__ n_dev, int n_host)
{
    const int n = n_dev ? *n_dev : n_host;
    const int Dt = 2 * Lt + 1;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
        temporal_factor(tk1[i], tk2[i], Lt, W, 1.0, alpha + (long long)i * Dt);
}

VT_EXPORT int vt_temporal_factors(double *alpha, const double *tk1, const double *tk2, int Lt,
                                  double W, const int *n_dev, int n, void *stream)
{
    VT_CHECK_ARG(alpha && tk1 && tk2 && Lt >= 0 && Lt <= 32 && n >= 0 && W > 0);
    if (n == 0) return VT_OK;
    temporal_factor_kernel<<<vt_blocks(n, 256) < 2048 ? vt_blocks(n, 256) : 2048, 256, 0,
                             (hipStream_t)stream>>>(alpha, tk1, tk2, Lt, W, n_dev, n);
    VT_LAUNCH_CHECK();
    return VT_OK;
}

// One wave per neuron, TJ_NPW neurons per workgroup: intervals in chunks of 64, alpha
// staged in LDS, each lane owns a strided set of (pt, pt') entries of T_j.  All waves of a
// workgroup run the largest chunk count among its neurons (a wave past its own intervals
// adds nothing), so every barrier is reached by the whole workgroup.  Batched over
// independent problems through per-problem descriptors (offsets into shared arrays).
struct TjDesc {
    int neuron_base;      // first neuron of this problem in the global neuron numbering
    int num_neurons;
    long long t_base;     // offset of this problem's intervals (tk1/tk2/q arrays)
    int Lt;
    double W;
    int T_base;           // offset (in neurons) into T/t outputs
};

#define TJ_NPW 4          // neurons (waves) per workgroup

__global__ void __launch_bounds__(64 * TJ_NPW)
neuron_T_kernel(double *__restrict__ Tm, double *__restrict__ tv, const double *__restrict__ tk1,
                const double *__restrict__ tk2, const double *__restrict__ q,
                const double *__restrict__ norm, const int *__restrict__ offsets,
                const TjDesc *__restrict__ descs, TjDesc d0)
{
    __shared__ double al[TJ_NPW][64][18];    // Dt <= 17
    __shared__ double qs[TJ_NPW][64];
    __shared__ int trips[TJ_NPW];
    // locate problem: blockIdx.y = problem, blockIdx.x = group of TJ_NPW neurons
    const TjDesc d = descs ? descs[blockIdx.y] : d0;
    if (blockIdx.x * TJ_NPW >= d.num_neurons) return;        // uniform over the workgroup
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int jl = blockIdx.x * TJ_NPW + wave;
    const bool valid = jl < d.num_neurons;
    const int j = d.neuron_base + jl;
    const int Dt = 2 * d.Lt + 1;
    const int k0 = valid ? offsets[j] : 0, k1 = valid ? offsets[j + 1] : 0;
    const double w = (norm && valid) ? norm[j] : 1.0;
    if (lane == 0) trips[wave] = (k1 - k0 + 63) / 64;
    // T_j is symmetric: each lane sums the upper-triangle entries u = lane, lane + 64, ...
    // (p <= pp; at most 3 for Dt <= 17) and writes both (p, pp) and (pp, p).  The product
    // al[r][p] * al[r][pp] and the order over r are those of the mirrored entry.
    const int DTU = Dt * (Dt + 1) / 2;
    int tp[3], tq[3];
#pragma unroll
    for (int e = 0; e < 3; ++e) {
        int u = lane + 64 * e, p = 0;
        while (p < Dt && u >= Dt - p) { u -= Dt - p; ++p; }
        tp[e] = p;
        tq[e] = p + u;
    }
    double accT[3] = {0, 0, 0};
    double acct = 0.0;
    __syncthreads();
    int nmax = 0;
#pragma unroll
    for (int v = 0; v < TJ_NPW; ++v) nmax = max(nmax, trips[v]);
    double (*a)[18] = al[wave];
    for (int it = 0; it < nmax; ++it) {
        const int base = k0 + 64 * it;
        const int k = base + lane;
        if (k < k1) {
            const long long kk = d.t_base + k;
            temporal_factor(tk1[kk], tk2[kk], d.Lt, d.W, w, &a[lane][0]);
            qs[wave][lane] = q[kk];
        } else {
            for (int p = 0; p < Dt; ++p) a[lane][p] = 0.0;
            qs[wave][lane] = 0.0;
        }
        __syncthreads();
        const int cnt = min(64, k1 - base);
        if (cnt > 0) {
#pragma unroll
            for (int e = 0; e < 3; ++e) {
                if (lane + 64 * e < DTU) {
                    const int p = tp[e], pp = tq[e];
                    double s = 0.0;
                    for (int r = 0; r < cnt; ++r) s += a[r][p] * a[r][pp];
                    accT[e] += s;
                }
            }
            if (lane < Dt) {
                double s = 0.0;
                for (int r = 0; r < cnt; ++r) s += a[r][lane] * qs[wave][r];
                acct += s;
            }
        }
        __syncthreads();
    }
    if (!valid) return;
    const long long ob = (long long)(d.T_base + jl);
#pragma unroll
    for (int e = 0; e < 3; ++e) {
        if (lane + 64 * e < DTU) {
            const int p = tp[e], pp = tq[e];
            Tm[ob * Dt * Dt + p * Dt + pp] = accT[e];
            if (pp != p) Tm[ob * Dt * Dt + pp * Dt + p] = accT[e];
        }
    }
    if (lane < Dt) tv[ob * Dt + lane] = acct;
}

// Single problem: N neurons with CSR offsets (N+1), intervals tk1/tk2/q.
VT_EXPORT int vt_neuron_T(double *Tm, double *tv, const double *tk1, const double *tk2,
                          const double *q, const double *norm, const int *offsets, int N,
                          int Lt, double W, void *stream)
{
    VT_CHECK_ARG(Tm && tv && tk1 && tk2 && q && offsets && N >= 1 && Lt >= 0 && Lt <= 8 && W > 0);
    TjDesc d{0, N, 0, Lt, W, 0};
    hipStream_t s = (hipStream_t)stream;
    neuron_T_kernel<<<dim3((N + TJ_NPW - 1) / TJ_NPW, 1), 64 * TJ_NPW, 0, s>>>(Tm, tv, tk1, tk2, q, norm,
                                                                            offsets, nullptr, d);
    VT_LAUNCH_CHECK();
    return VT_OK;
}

// Batched: descs (device) has n_desc problems; neurons numbered globally, offsets global
// (offsets[j] indexes the shared interval arrays directly, so t_base = 0 for all).
VT_EXPORT int vt_neuron_T_batched(double *Tm, double *tv, const double *tk1, const double *tk2,
                                  const double *q, const double *norm, const int *offsets,
                                  const void *descs_dev, int n_desc, int max_neurons,
                                  void *stream)
{
    VT_CHECK_ARG(Tm && tv && tk1 && tk2 && q && offsets && descs_dev && n_desc >= 1 &&
                 max_neurons >= 1);
    TjDesc d0{0, 0, 0, 0, 1.0, 0};
    neuron_T_kernel<<<dim3((max_neurons + TJ_NPW - 1) / TJ_NPW, n_desc), 64 * TJ_NPW, 0,
                      (hipStream_t)stream>>>(Tm, tv, tk1, tk2, q, norm, offsets,
                                             (const TjDesc *)descs_dev, d0);
    VT_LAUNCH_CHECK();
    return VT_OK;
}

// ------------------------------------------------------------------------------------
// Structured operator helpers: Z[j, :] = T_j Y[j, :]  (T_j symmetric Dt x Dt)
// ------------------------------------------------------------------------------------
__global__ void tj_apply_kernel(double *__restrict__ Z, const double *__restrict__ Tm,
                                const double *__restrict__ Y, int N, int Dt)
{
    const long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= (long long)N * Dt) return;
    const long long j = idx / Dt;
    const int p = (int)(idx % Dt);
    const double *T = Tm + j * Dt * Dt + (long long)p * Dt;
    const double *y = Y + j * Dt;
    double s = 0.0;
    for (int pp = 0; pp < Dt; ++pp) s += T[pp] * y[pp];
    Z[idx] = s;
}

VT_EXPORT int vt_tj_apply(double *Z, const double *Tm, const double *Y, int N, int Dt,
                          void *stream)
{
    VT_CHECK_ARG(Z && Tm && Y && N >= 0 && Dt >= 1);
    if (N == 0) return VT_OK;
    tj_apply_kernel<<<vt_blocks((long long)N * Dt, 256), 256, 0, (hipStream_t)stream>>>(Z, Tm, Y,
                                                                                      N, Dt);
    VT_LAUNCH_CHECK();
    return VT_OK;
}

// out = a*x + b*y + c*z (element-wise, fp64), used for residuals and RNN updates.
__global__ void axpbypcz_kernel(double *__restrict__ out, double a, const double *__restrict__ x,
                                double b, const double *__restrict__ y, double c,
                                const double *__restrict__ z, long long n)
{
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (long long)gridDim.x * blockDim.x) {
        double v = a * x[i];
        if (y) v += b * y[i];
        if (z) v += c * z[i];
        out[i] = v;
    }
}

VT_EXPORT int vt_axpbypcz(double *out, double a, const double *x, double b, const double *y,
                          double c, const double *z, long long n, void *stream)
{
    VT_CHECK_ARG(out && x && n >= 0);
    if (n == 0) return VT_OK;
    unsigned g = vt_blocks(n, 256);
    axpbypcz_kernel<<<g < 4096 ? g : 4096, 256, 0, (hipStream_t)stream>>>(out, a, x, b, y, c, z, n);
    VT_LAUNCH_CHECK();
    return VT_OK;
}

// ------------------------------------------------------------------------------------
// Tikhonov lambda on device: lam = lam_rel * trace(A) / n, n read from offsets[N].
// Batched by descriptor arrays of (A pointer, lda, D, offsets pointer index).
// ------------------------------------------------------------------------------------
template <typename T>
__global__ void __launch_bounds__(256)
trace_lambda_kernel(const T *A, long long lda, int D, const int *n_ptr, double lam_rel,
                    double *lam_out)
{
    __shared__ double red[256];
    double s = 0.0;
    for (int i = threadIdx.x; i < D; i += 256) s += (double)A[(long long)i * lda + i];
    red[threadIdx.x] = s;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if (threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        int n = *n_ptr;
        lam_out[0] = n > 0 ? lam_rel * red[0] / (double)n : 0.0;
        lam_out[1] = red[0];
    }
}

VT_EXPORT int vt_trace_lambda(const void *A, int dtype, long long lda, int D, const int *n_dev,
                              double lam_rel, double *lam_out, void *stream)
{
    VT_CHECK_ARG(A && n_dev && lam_out && D >= 1 && lda >= D);
    hipStream_t s = (hipStream_t)stream;
    if (dtype == 0)
        trace_lambda_kernel<double><<<1, 256, 0, s>>>((const double *)A, lda, D, n_dev, lam_rel, lam_out);
    else
        trace_lambda_kernel<float><<<1, 256, 0, s>>>((const float *)A, lda, D, n_dev, lam_rel, lam_out);
    VT_LAUNCH_CHECK();
    return VT_OK;
}

// ------------------------------------------------------------------------------------
// Spike-space reference kernels
// ------------------------------------------------------------------------------------

// Spike-space Gram G_kl = Gt(k,l) Dsw[j(k), j(l)], the matrix the reference builds at
// vtem/dirichlet.py:62-85 (its kernels :274-285, :316-323).  Gt(k,l) is the double integral
// of the temporal reproducing kernel 1 + 2 sum_{m<=M} cos(m W (t - s)) over the two
// intervals.  That kernel is sum_p theta_p(t) theta_p(s) for the real orthonormal basis
// theta_0 = 1, theta_{2m-1} = sqrt2 cos(m W t), theta_{2m} = sqrt2 sin(m W t), so
// Gt(k,l) = alpha_k . alpha_l with alpha_k[p] the integral of theta_p over interval k:
//   alpha[0] = length,  alpha[2m-1] + i alpha[2m] = 2 sqrt2 sin(m W h) e^{i m W c} / (m W)
// (h the half-length, c the midpoint).  One workgroup per row: the row's alpha in LDS,
// each thread forms its column's alpha term by term.  W = Wt / Mt.
#define VT_GSPIKE_MAX_M 1024

__global__ void __launch_bounds__(256)
G_spike_kernel(double *__restrict__ G, long long ldg, const double *__restrict__ tk1,
               const double *__restrict__ tk2, double Wt, int Mt, const double *__restrict__ Dsw,
               long long ldw, const int *__restrict__ nid, int n)
{
    __shared__ double ac[VT_GSPIKE_MAX_M], as[VT_GSPIKE_MAX_M];
    const int row = blockIdx.y;
    const double W = Wt / Mt;
    const double hk = 0.5 * (tk2[row] - tk1[row]), ck = 0.5 * (tk1[row] + tk2[row]);
    for (int m = threadIdx.x + 1; m <= Mt; m += blockDim.x) {
        const double w = m * W;
        const double amp = 2.0 * sin(w * hk) / w;     // the sqrt2 of both factors: a 2 below
        ac[m - 1] = amp * cos(w * ck);
        as[m - 1] = amp * sin(w * ck);
    }
    __syncthreads();
    const int col = blockIdx.x * blockDim.x + threadIdx.x;
    if (row >= n || col >= n) return;
    const double hl = 0.5 * (tk2[col] - tk1[col]), cl = 0.5 * (tk1[col] + tk2[col]);
    double dot = 0.0;
    for (int m = 1; m <= Mt; ++m) {
        const double w = m * W;
        const double amp = 2.0 * sin(w * hl) / w;
        dot += amp * (ac[m - 1] * cos(w * cl) + as[m - 1] * sin(w * cl));
    }
    const double gt = 4.0 * hk * hl + 2.0 * dot;
    G[(long long)row * ldg + col] = gt * Dsw[(long long)nid[row] * ldw + nid[col]];
}

VT_EXPORT int vt_G_spike(double *G, long long ldg, const double *tk1, const double *tk2, double Wt,
                         int Mt, const double *Dsw, long long ldw, const int *neuron_ind, int n,
                         void *stream)
{
    VT_CHECK_ARG(G && tk1 && tk2 && Dsw && neuron_ind && n >= 1 && ldg >= n && Mt >= 1);
    if (Mt > VT_GSPIKE_MAX_M) return VT_ERANGE;
    VT_CHECK_ARG(n <= 65535);
    G_spike_kernel<<<dim3(vt_blocks(n, 256), n), 256, 0, (hipStream_t)stream>>>(G, ldg, tk1, tk2, Wt,
                                                                              Mt, Dsw, ldw,
                                                                              neuron_ind, n);
    VT_LAUNCH_CHECK();
    return VT_OK;
}

template <typename T>
__global__ void diag_add_kernel(T *A, long long lda, int D, double lam, const double *lam_dev)
{
    const double l = lam_dev ? lam_dev[0] : lam;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < D; i += gridDim.x * blockDim.x)
        A[(long long)i * lda + i] += (T)l;
}

VT_EXPORT int vt_diag_add(void *A, int dtype, long long lda, int D, double lam,
                          const double *lam_dev, void *stream)
{
    VT_CHECK_ARG(A && D >= 0 && lda >= D);
    if (D == 0) return VT_OK;
    hipStream_t s = (hipStream_t)stream;
    if (dtype == 0)
        diag_add_kernel<double><<<vt_blocks(D, 256), 256, 0, s>>>((double *)A, lda, D, lam, lam_dev);
    else
        diag_add_kernel<float><<<vt_blocks(D, 256), 256, 0, s>>>((float *)A, lda, D, lam, lam_dev);
    VT_LAUNCH_CHECK();
    return VT_OK;
}

// u(t,p) = sum_k c_k norm_j psi_j(p) [2/WM sum_m (sin mWM(t-tk1) - sin mWM(t-tk2))/m + tk2-tk1]
// (dirichlet.py:356-416).  Intervals staged 256 at a time in LDS; the temporal factor of
// interval k at time t is computed once per block and shared by the block's pixels.
__global__ void __launch_bounds__(256)
reconstruct_spike_kernel(double *__restrict__ u, long long ldu, const double *__restrict__ psi,
                         long long ldpsi, const double *__restrict__ tk1,
                         const double *__restrict__ tk2, const double *__restrict__ c,
                         const double *__restrict__ times, const int *__restrict__ nid,
                         const double *__restrict__ norm, int M, double WM, int n, int P)
{
    __shared__ double wk[256];
    __shared__ int jk[256];
    const int pix = blockIdx.x * blockDim.x + threadIdx.x;
    const double t = times[blockIdx.y];
    double acc = 0.0;
    for (int base = 0; base < n; base += 256) {
        const int k = base + threadIdx.x;
        if (k < n) {
            double a1 = tk1[k], a2 = tk2[k];
            double phi = 0;
            for (int m = 1; m <= M; ++m)
                phi += (sin(m * WM * (t - a1)) - sin(m * WM * (t - a2))) / m;
            int j = nid[k];
            wk[threadIdx.x] = c[k] * (phi * 2 / WM + a2 - a1) * norm[j];
            jk[threadIdx.x] = j;
        }
        __syncthreads();
        const int cnt = min(256, n - base);
        if (pix < P)
            for (int r = 0; r < cnt; ++r) acc += wk[r] * psi[(long long)jk[r] * ldpsi + pix];
        __syncthreads();
    }
    if (pix < P) u[(long long)blockIdx.y * ldu + pix] = acc;
}

VT_EXPORT int vt_reconstruct_spike(double *u, long long ldu, const double *psi, long long ldpsi,
                                   const double *tk1, const double *tk2, const double *c,
                                   const double *times, int T, const int *neuron_ind,
                                   const double *norm, int Mt, double WM, int n, int P,
                                   void *stream)
{
    VT_CHECK_ARG(u && psi && tk1 && tk2 && c && times && neuron_ind && norm && T >= 1 && n >= 0 &&
                 P >= 1 && ldu >= P && ldpsi >= P && T <= 65535);
    reconstruct_spike_kernel<<<dim3(vt_blocks(P, 256), T), 256, 0, (hipStream_t)stream>>>(
        u, ldu, psi, ldpsi, tk1, tk2, c, times, neuron_ind, norm, Mt, WM, n, P);
    VT_LAUNCH_CHECK();
    return VT_OK;
}

// Symmetric mat-vec y = G c for the RNN (one wave per row, fp64).
__global__ void __launch_bounds__(256)
gemv_kernel(double *__restrict__ y, const double *__restrict__ G, long long ldg,
            const double *__restrict__ x, int n)
{
    const int row = blockIdx.x * 4 + (threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    if (row >= n) return;
    const double *g = G + (long long)row * ldg;
    double s = 0.0;
    for (int i = lane; i < n; i += 64) s += g[i] * x[i];
    for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o);
    if (lane == 0) y[row] = s;
}

// c += eta (q - Gc); y = q - Gc  (neural_network.py:111-121, groupsize 1)
__global__ void rnn_update_kernel(double *__restrict__ c, double eta, const double *__restrict__ q,
                                  const double *__restrict__ Gc, double *__restrict__ y, int n)
{
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        double dy = q[i] - Gc[i];
        c[i] += eta * dy;
        y[i] = dy;
    }
}

VT_EXPORT int vt_gemv(double *y, const double *G, long long ldg, const double *x, int n, void *stream)
{
    VT_CHECK_ARG(y && G && x && n >= 0 && ldg >= n);
    if (n == 0) return VT_OK;
    gemv_kernel<<<vt_blocks(n, 4), 256, 0, (hipStream_t)stream>>>(y, G, ldg, x, n);
    VT_LAUNCH_CHECK();
    return VT_OK;
}

// steps+1 iterations of {Gc = G c; c += eta (q - Gc)} on the stream (rnn3's loop).
VT_EXPORT int vt_rnn_spike(double *c, const double *G, long long ldg, const double *q, int n,
                           double eta, int steps, double *work_Gc, double *work_y, void *stream)
{
    VT_CHECK_ARG(c && G && q && work_Gc && work_y && n >= 1 && steps >= 0 && ldg >= n);
    hipStream_t s = (hipStream_t)stream;
    for (int i = 0; i <= steps; ++i) {
        gemv_kernel<<<vt_blocks(n, 4), 256, 0, s>>>(work_Gc, G, ldg, c, n);
        unsigned g = vt_blocks(n, 256);
        rnn_update_kernel<<<g < 1024 ? g : 1024, 256, 0, s>>>(c, eta, q, work_Gc, work_y, n);
    }
    VT_LAUNCH_CHECK();
    return VT_OK;
}

// res = r - AX - lam[0] * X  (lam on device, no host round trip)
__global__ void residual_kernel(double *__restrict__ res, const double *__restrict__ r,
                                const double *__restrict__ AX, const double *__restrict__ X,
                                const double *__restrict__ lam, long long n)
{
    const double l = lam ? lam[0] : 0.0;
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (long long)gridDim.x * blockDim.x)
        res[i] = r[i] - AX[i] - l * X[i];
}

VT_EXPORT int vt_residual(double *res, const double *r, const double *AX, const double *X,
                          const double *lam_dev, long long n, void *stream)
{
    VT_CHECK_ARG(res && r && AX && X && n >= 0);
    if (n == 0) return VT_OK;
    unsigned g = vt_blocks(n, 256);
    residual_kernel<<<g < 1024 ? g : 1024, 256, 0, (hipStream_t)stream>>>(res, r, AX, X, lam_dev, n);
    VT_LAUNCH_CHECK();
    return VT_OK;
}

// out[slot] = sum x^2 over n doubles (one workgroup; small vectors)
__global__ void __launch_bounds__(1024) sumsq_kernel(const double *__restrict__ x, long long n,
                                                    double *__restrict__ out)
{
    __shared__ double red[1024];
    double s = 0.0;
    for (long long i = threadIdx.x; i < n; i += 1024) s += x[i] * x[i];
    red[threadIdx.x] = s;
    __syncthreads();
    for (int o = 512; o > 0; o >>= 1) {
        if (threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[0] = red[0];
}

VT_EXPORT int vt_sumsq(const double *x, long long n, double *out, void *stream)
{
    VT_CHECK_ARG(x && out && n >= 0);
    sumsq_kernel<<<1, 1024, 0, (hipStream_t)stream>>>(x, n, out);
    VT_LAUNCH_CHECK();
    return VT_OK;
}

// Row norms s2[j] = |S_j|^2 (computed once per neuron set).
__global__ void row_sumsq_kernel(const double *__restrict__ S, long long lds, int N, int Ds,
                                 double *__restrict__ s2)
{
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= N) return;
    double s = 0.0;
    for (int p = 0; p < Ds; ++p) s += S[(long long)j * lds + p] * S[(long long)j * lds + p];
    s2[j] = s;
}

VT_EXPORT int vt_row_sumsq(const double *S, long long lds, int N, int Ds, double *s2, void *stream)
{
    VT_CHECK_ARG(S && s2 && N >= 1 && Ds >= 1 && lds >= Ds);
    row_sumsq_kernel<<<vt_blocks(N, 256), 256, 0, (hipStream_t)stream>>>(S, lds, N, Ds, s2);
    VT_LAUNCH_CHECK();
    return VT_OK;
}

// dst (cols x rows, row stride ldd) = src^T (src rows x cols, row stride lds): 32 x 32
// tiles through LDS (padded rows), coalesced reads and writes.
__global__ void __launch_bounds__(256)
transpose_kernel(const double *__restrict__ src, long long lds, int rows, int cols,
                 double *__restrict__ dst, long long ldd)
{
    __shared__ double tile[32][33];
    const int c0 = blockIdx.x * 32, r0 = blockIdx.y * 32;
    const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;   // 32 x 8 threads
    for (int k = ty; k < 32; k += 8) {
        const int r = r0 + k, c = c0 + tx;
        if (r < rows && c < cols) tile[k][tx] = src[(long long)r * lds + c];
    }
    __syncthreads();
    for (int k = ty; k < 32; k += 8) {
        const int c = c0 + k, r = r0 + tx;
        if (r < rows && c < cols) dst[(long long)c * ldd + r] = tile[tx][k];
    }
}

VT_EXPORT int vt_transpose_f64(const double *src, long long lds, int rows, int cols, double *dst,
                               long long ldd, void *stream)
{
    VT_CHECK_ARG(src && dst && rows >= 0 && cols >= 0 && lds >= cols && ldd >= rows);
    if (rows == 0 || cols == 0) return VT_OK;
    VT_CHECK_ARG(vt_blocks(rows, 32) <= 65535);
    dim3 g(vt_blocks(cols, 32), vt_blocks(rows, 32));
    transpose_kernel<<<g, 256, 0, (hipStream_t)stream>>>(src, lds, rows, cols, dst, ldd);
    VT_LAUNCH_CHECK();
    return VT_OK;
}

// dst = A in the skinny product's MFMA tile layout (sk_tile_index), A = src (rows x cols,
// row stride lds) or its transpose (trans): ceil(M / 16) * ceil(K / 16) tiles of 256 doubles
// for A of M x K, entries beyond M or K zero.  One thread per destination element.
__global__ void __launch_bounds__(256)
tile_kernel(const double *__restrict__ src, long long lds, int M, int K, int trans,
            double *__restrict__ dst, long long total)
{
    const long long kc = (K + 15) / 16;
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < total;
         i += (long long)gridDim.x * blockDim.x) {
        const long long tile = i >> 8;
        const int w = (int)(i & 255), half = w >> 7, lane = (w & 127) >> 1, s = 2 * half + (w & 1);
        const long long m = 16 * (tile / kc) + (lane & 15), k = 16 * (tile % kc) + 4 * (lane >> 4) + s;
        double v = 0.0;
        if (m < M && k < K) v = trans ? src[k * lds + m] : src[m * lds + k];
        dst[i] = v;
    }
}

VT_EXPORT int vt_tile_f64(const double *src, long long lds, int rows, int cols, int trans, double *dst,
                          void *stream)
{
    VT_CHECK_ARG(src && dst && rows >= 1 && cols >= 1 && lds >= cols && (trans == 0 || trans == 1));
    const int M = trans ? cols : rows, K = trans ? rows : cols;
    const long long total = (long long)vt_blocks(M, 16) * vt_blocks(K, 16) * 256;
    const long long g = vt_blocks(total, 256);
    tile_kernel<<<g > 8192 ? 8192 : g, 256, 0, (hipStream_t)stream>>>(src, lds, M, K, trans, dst, total);
    VT_LAUNCH_CHECK();
    return VT_OK;
}

// Exact trace of the Kronecker Gram without forming it: tr(A) = sum_j s2_j tr(T_j) with
// s2_j = |S_j|^2.  lam_out[0] = lam_rel * tr(A) / n (n from n_dev), lam_out[1] = tr(A).
// Batched: problem b uses s2 + s2_off[b], Tm + T_off[b] (in neurons), N[b], n_dev[b].
__global__ void __launch_bounds__(256)
kron_trace_kernel(const double *__restrict__ s2, const double *__restrict__ Tm, const int *Nb,
                  const long long *s2_off, const long long *T_off, int N0, int Dt,
                  const int *__restrict__ n_dev, int n_stride, double lam_rel,
                  double *__restrict__ lam_out)
{
    __shared__ double red[256];
    const int b = blockIdx.x;
    const int N = Nb ? Nb[b] : N0;
    const double *s = s2 + (s2_off ? s2_off[b] : 0);
    const double *T = Tm + (T_off ? T_off[b] : 0) * (long long)Dt * Dt;
    double acc = 0.0;
    for (int j = threadIdx.x; j < N; j += 256) {
        double tt = 0.0;
        for (int p = 0; p < Dt; ++p) tt += T[((long long)j * Dt + p) * Dt + p];
        acc += s[j] * tt;
    }
    red[threadIdx.x] = acc;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if (threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        int n = n_dev[(long long)b * n_stride];
        lam_out[2 * b] = n > 0 ? lam_rel * red[0] / (double)n : 0.0;
        lam_out[2 * b + 1] = red[0];
    }
}

VT_EXPORT int vt_kron_trace(const double *s2, const double *Tm, int N, int Dt, const int *n_dev,
                            double lam_rel, double *lam_out, void *stream)
{
    VT_CHECK_ARG(s2 && Tm && n_dev && lam_out && N >= 1 && Dt >= 1);
    kron_trace_kernel<<<1, 256, 0, (hipStream_t)stream>>>(s2, Tm, nullptr, nullptr, nullptr, N, Dt,
                                                         n_dev, 0, lam_rel, lam_out);
    VT_LAUNCH_CHECK();
    return VT_OK;
}

// Batched: per problem b: N[b] neurons at s2 + s2_off[b], T at Tm + T_off[b]*Dt*Dt,
// n at n_dev[b * n_stride]; lam_out[2b..2b+1].  All arrays on device.
VT_EXPORT int vt_kron_trace_batched(const double *s2, const double *Tm, const int *Nb,
                                    const long long *s2_off, const long long *T_off, int Dt,
                                    const int *n_dev, int n_stride, double lam_rel,
                                    double *lam_out, int batch, void *stream)
{
    VT_CHECK_ARG(s2 && Tm && Nb && s2_off && T_off && n_dev && lam_out && batch >= 1 && Dt >= 1);
    kron_trace_kernel<<<batch, 256, 0, (hipStream_t)stream>>>(s2, Tm, Nb, s2_off, T_off, 0, Dt, n_dev,
                                                             n_stride, lam_rel, lam_out);
    VT_LAUNCH_CHECK();
    return VT_OK;
}
