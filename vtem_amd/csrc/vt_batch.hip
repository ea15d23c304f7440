// Batched forms of the decode stages: many independent blocks (windows / tiles / served
// requests) per launch, each described by a device-side descriptor.  Blocks keep their
// own sizes (neurons, spatial order, spike counts); the kernels skip work beyond a
// block's own extent, so one launch sequence serves a ragged batch.
//
// Neuron numbering is global across the batch (block b owns neurons
// [neuron_base[b], neuron_base[b] + N_b)), so interval arrays, q, T_j and t_j of all
// blocks live in shared arrays and the single-problem kernels (vt_q, vt_neuron_T_batched,
// vt_gram_split_T, vt_tj_apply) run over the whole batch unchanged.
#include "vt_common.hpp"
#include "vt_skinny.hpp"

// ------------------------------------------------------------------------------------
// Intervals over a batch.  job_of[j] maps a global neuron to its block; descriptor per
// block: spike buffer, leading dimension, rows, counts, neuron base, window [t0, t0+Tw].
// ------------------------------------------------------------------------------------
struct SpikeDesc {
    const double *spikes;
    const int *counts;
    long long ld;
    int max_rows, neuron_base;
    double t0, Tw;
};

__global__ void batch_interval_count(const SpikeDesc *__restrict__ sd, const int *__restrict__ job_of,
                                     int Ntot, int *__restrict__ lo_out, int *__restrict__ cnt_out)
{
    const int g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= Ntot) return;
    const SpikeDesc d = sd[job_of[g]];
    const int j = g - d.neuron_base;
    const int c = min(d.counts[j], d.max_rows);
    int first = -1, last = -1;
    for (int i = 0; i < c; ++i) {
        double r = d.spikes[(long long)i * d.ld + j] - d.t0;
        if (r > 0 && r < d.Tw) {
            if (first < 0) first = i;
            last = i;
        }
    }
    int lo = 0, cnt = 0;
    if (first >= 0) {
        lo = first > 0 ? first - 1 : 0;
        int hi = (last + 2 < c) ? last + 2 : c;
        cnt = max(0, hi - lo - 1);
    }
    lo_out[g] = lo;
    cnt_out[g] = cnt;
}

__global__ void batch_interval_write(const SpikeDesc *__restrict__ sd, const int *__restrict__ job_of,
                                     const int *__restrict__ lo, const int *__restrict__ offsets,
                                     double *__restrict__ tk1, double *__restrict__ tk2,
                                     int *__restrict__ nid, int cap)
{
    const int g = blockIdx.x;
    const SpikeDesc d = sd[job_of[g]];
    const int j = g - d.neuron_base;
    const int base = offsets[g], cnt = offsets[g + 1] - base, l = lo[g];
    for (int k = threadIdx.x; k < cnt; k += blockDim.x) {
        const int o = base + k;
        if (o >= cap) break;
        tk1[o] = d.spikes[(long long)(l + k) * d.ld + j] - d.t0;
        tk2[o] = d.spikes[(long long)(l + k + 1) * d.ld + j] - d.t0;
        nid[o] = g;
    }
}

// Exclusive scan of n counts in one workgroup: thread t owns a contiguous chunk (a multiple
// of 4 elements, so its reads are 16-byte vectors), the chunk sums are scanned in LDS, then
// each thread writes its chunk's offsets.
__global__ void __launch_bounds__(1024) batch_scan_kernel(const int *__restrict__ in, int n,
                                                         int *__restrict__ offsets)
{
    __shared__ long long part[1024];
    const int t = threadIdx.x;
    const int chunk = ((n + 1023) / 1024 + 3) & ~3;
    const int b = min(n, t * chunk), e = min(n, b + chunk);
    typedef int i4a4 __attribute__((ext_vector_type(4), aligned(4)));   // 4-byte aligned
    long long s = 0;
    int i = b;
    for (; i + 4 <= e; i += 4) {
        const i4a4 v = *(const i4a4 *)(in + i);
        s += (long long)v[0] + v[1] + v[2] + v[3];
    }
    for (; i < e; ++i) s += in[i];
    part[t] = s;
    __syncthreads();
    for (int off = 1; off < 1024; off <<= 1) {
        long long v = t >= off ? part[t - off] : 0;
        __syncthreads();
        part[t] += v;
        __syncthreads();
    }
    long long run = t > 0 ? part[t - 1] : 0;
    i = b;
    for (; i + 4 <= e; i += 4) {
        const i4a4 v = *(const i4a4 *)(in + i);
        i4a4 o;
        o[0] = (int)run;
        run += v[0];
        o[1] = (int)run;
        run += v[1];
        o[2] = (int)run;
        run += v[2];
        o[3] = (int)run;
        run += v[3];
        *(i4a4 *)(offsets + i) = o;
    }
    for (; i < e; ++i) {
        offsets[i] = (int)run;
        run += in[i];
    }
    if (t == 1023) offsets[n] = (int)part[1023];
}

// offsets (Ntot+1) global CSR; workspace 2*Ntot ints; nid holds GLOBAL neuron indices.
VT_EXPORT int vt_intervals_batched(const void *spike_descs, const int *job_of, int Ntot,
                                   int *workspace, int *offsets, double *tk1, double *tk2,
                                   int *neuron_ind, int capacity, void *stream)
{
    VT_CHECK_ARG(spike_descs && job_of && Ntot >= 1 && workspace && offsets && tk1 && tk2 &&
                 neuron_ind && capacity >= 0);
    hipStream_t s = (hipStream_t)stream;
    const SpikeDesc *sd = (const SpikeDesc *)spike_descs;
    int *lo = workspace, *cnt = workspace + Ntot;
    batch_interval_count<<<vt_blocks(Ntot, 256), 256, 0, s>>>(sd, job_of, Ntot, lo, cnt);
    batch_scan_kernel<<<1, 1024, 0, s>>>(cnt, Ntot, offsets);
    batch_interval_write<<<Ntot, 64, 0, s>>>(sd, job_of, lo, offsets, tk1, tk2, neuron_ind,
                                             capacity);
    VT_LAUNCH_CHECK();
    return VT_OK;
}

// ------------------------------------------------------------------------------------
// Grouped GEMM (fp64, MFMA 16x16x4): problem b computes
//   C_b = alpha_b A_b(m,k) B_b(k,n) + beta_b C_b  with its own M, N, K and strides.
// grid (max N tiles, max M tiles, batch); tiles outside a problem exit at once.
// ------------------------------------------------------------------------------------
struct GemmDesc {
    const double *A, *B;
    double *C;
    long long sam, sak, sbk, sbn, ldc;
    int M, N, K, pad;
    double alpha, beta;
};

__global__ void __launch_bounds__(256) gemm_grouped_kernel(const GemmDesc *__restrict__ descs)
{
    __shared__ double As[16][68];
    __shared__ double Bs[16][68];
    const GemmDesc d = descs[blockIdx.z];
    const int m0 = blockIdx.y * 64, n0 = blockIdx.x * 64;
    if (m0 >= d.M || n0 >= d.N) return;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wm = (wave >> 1) * 32, wn = (wave & 1) * 32;
    v4d acc[2][2];
    for (int i = 0; i < 2; ++i)
        for (int j = 0; j < 2; ++j) acc[i][j] = (v4d){0, 0, 0, 0};
    const bool a_kfast = (d.sak == 1), b_nfast = (d.sbn == 1);
    for (int k0 = 0; k0 < d.K; k0 += 16) {
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            int e = tid + i * 256;
            int m, k;
            if (a_kfast) { k = e & 15; m = e >> 4; } else { m = e & 63; k = e >> 6; }
            int gm = m0 + m, gk = k0 + k;
            As[k][m] = (gm < d.M && gk < d.K) ? d.A[(long long)gm * d.sam + (long long)gk * d.sak] : 0.0;
            int n, kb;
            if (b_nfast) { n = e & 63; kb = e >> 6; } else { kb = e & 15; n = e >> 4; }
            int gn = n0 + n, gkb = k0 + kb;
            Bs[kb][n] = (gn < d.N && gkb < d.K) ? d.B[(long long)gkb * d.sbk + (long long)gn * d.sbn] : 0.0;
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
                int gm = m0 + wm + i * 16 + mfma_row_f64(lane, r);
                int gn = n0 + wn + j * 16 + mfma_col(lane);
                if (gm < d.M && gn < d.N) {
                    double v = d.alpha * acc[i][j][r];
                    double *p = d.C + (long long)gm * d.ldc + gn;
                    *p = (d.beta == 0.0) ? v : v + d.beta * *p;
                }
            }
}

VT_EXPORT int vt_gemm_grouped(const void *descs, int batch, int max_M, int max_N, void *stream)
{
    VT_CHECK_ARG(descs && batch >= 1 && batch <= 65535 && max_M >= 0 && max_N >= 0);
    if (max_M == 0 || max_N == 0) return VT_OK;
    dim3 grid(vt_blocks(max_N, 64), vt_blocks(max_M, 64), batch);
    VT_CHECK_ARG(grid.y <= 65535);
    gemm_grouped_kernel<<<grid, 256, 0, (hipStream_t)stream>>>((const GemmDesc *)descs);
    VT_LAUNCH_CHECK();
    return VT_OK;
}

// ------------------------------------------------------------------------------------
// Batched skinny A^T B (K = neurons): problem b: C_b[m, n] = sum_k A_b[k*lda_b + m] B_b[k*ldb + n]
// with M_b = Ds_b, Nc common (= Dt), K_b = N_b.  Two passes through ws (16 partials per
// problem at ws + ws_off[b]); deterministic.
// ------------------------------------------------------------------------------------
struct TnDesc {
    const double *A, *B;
    double *C, *ws;
    long long lda, ldb, ldc;
    int M, K;
};

__global__ void __launch_bounds__(256)
tn_batched_partial(const TnDesc *__restrict__ descs, int Nc)
{
    __shared__ double red[16][16][33];
    const TnDesc d = descs[blockIdx.z];
    if ((int)blockIdx.x * 16 >= d.M) return;
    const int tid = threadIdx.x, ms = tid & 15, kg = tid >> 4;
    const int m = blockIdx.x * 16 + ms;
    const int chunk = (d.K + 15) / 16;
    const int kbeg = blockIdx.y * chunk, kend = min(d.K, kbeg + chunk);
    double acc[32];
#pragma unroll
    for (int n = 0; n < 32; ++n) acc[n] = 0.0;
#pragma unroll 4
    for (int k = kbeg + kg; k < kend; k += 16) {
        const double a = (m < d.M) ? d.A[(long long)k * d.lda + m] : 0.0;
        const double *b = d.B + (long long)k * d.ldb;
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
        if (gm < d.M) {
            double s = 0.0;
            for (int g = 0; g < 16; ++g) s += red[g][mm][n];
            d.ws[((long long)blockIdx.y * d.M + gm) * Nc + n] = s;
        }
    }
}

__global__ void tn_batched_reduce(const TnDesc *__restrict__ descs, int Nc)
{
    const TnDesc d = descs[blockIdx.y];
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= d.M * Nc) return;
    const int m = e / Nc, n = e % Nc;
    double s = 0.0;
    for (int k = 0; k < 16; ++k) s += d.ws[((long long)k * d.M + m) * Nc + n];
    d.C[(long long)m * d.ldc + n] = s;
}

VT_EXPORT int vt_gemm_tn_batched(const void *descs, int batch, int max_M, int Nc, void *stream)
{
    VT_CHECK_ARG(descs && batch >= 1 && batch <= 65535 && max_M >= 1 && Nc >= 1 && Nc <= 32);
    hipStream_t s = (hipStream_t)stream;
    const TnDesc *d = (const TnDesc *)descs;
    tn_batched_partial<<<dim3(vt_blocks(max_M, 16), 16, batch), 256, 0, s>>>(d, Nc);
    tn_batched_reduce<<<dim3(vt_blocks((long long)max_M * Nc, 256), batch), 256, 0, s>>>(d, Nc);
    VT_LAUNCH_CHECK();
    return VT_OK;
}

// ------------------------------------------------------------------------------------
// Per-block vector segments: block b owns [seg[b], seg[b+1]) of a concatenated vector.
// ------------------------------------------------------------------------------------

// res = r - AX - lam_b X  for every element (elem_job maps element -> block)
__global__ void residual_batched_kernel(double *__restrict__ res, const double *__restrict__ r,
                                        const double *__restrict__ AX, const double *__restrict__ X,
                                        const double *__restrict__ lam, const int *__restrict__ elem_job,
                                        long long n)
{
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (long long)gridDim.x * blockDim.x)
        res[i] = r[i] - AX[i] - lam[2 * elem_job[i]] * X[i];
}

VT_EXPORT int vt_residual_batched(double *res, const double *r, const double *AX, const double *X,
                                  const double *lam, const int *elem_job, long long n, void *stream)
{
    VT_CHECK_ARG(res && r && AX && X && lam && elem_job && n >= 0);
    if (n == 0) return VT_OK;
    unsigned g = vt_blocks(n, 256);
    residual_batched_kernel<<<g < 4096 ? g : 4096, 256, 0, (hipStream_t)stream>>>(res, r, AX, X, lam,
                                                                                 elem_job, n);
    VT_LAUNCH_CHECK();
    return VT_OK;
}

// out[b * out_stride] = sum over segment b of x^2  (one workgroup per block)
__global__ void __launch_bounds__(256)
seg_sumsq_kernel(const double *__restrict__ x, const long long *__restrict__ seg,
                 double *__restrict__ out, int out_stride)
{
    __shared__ double red[256];
    const long long a = seg[blockIdx.x], b = seg[blockIdx.x + 1];
    double s = 0.0;
    for (long long i = a + threadIdx.x; i < b; i += 256) s += x[i] * x[i];
    red[threadIdx.x] = s;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if (threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[(long long)blockIdx.x * out_stride] = red[0];
}

VT_EXPORT int vt_seg_sumsq(const double *x, const long long *seg, int nseg, double *out,
                           int out_stride, void *stream)
{
    VT_CHECK_ARG(x && seg && out && nseg >= 1 && out_stride >= 1);
    seg_sumsq_kernel<<<nseg, 256, 0, (hipStream_t)stream>>>(x, seg, out, out_stride);
    VT_LAUNCH_CHECK();
    return VT_OK;
}

// A_b += lam_b I for a batch of matrices {ptr, lda, D} (dtype 0 fp64, 1 fp32)
struct DiagDesc {
    void *A;
    long long lda;
    int D, pad;
};

template <typename T>
__global__ void diag_add_batched_kernel(const DiagDesc *__restrict__ dd, const double *__restrict__ lam)
{
    const DiagDesc d = dd[blockIdx.y];
    const double l = lam[2 * blockIdx.y];
    T *A = (T *)d.A;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < d.D; i += gridDim.x * blockDim.x)
        A[(long long)i * d.lda + i] += (T)l;
}

VT_EXPORT int vt_diag_add_batched(const void *descs, int batch, int dtype, const double *lam,
                                  int max_D, void *stream)
{
    VT_CHECK_ARG(descs && lam && batch >= 1 && batch <= 65535 && max_D >= 1);
    dim3 g(vt_blocks(max_D, 256) < 16 ? vt_blocks(max_D, 256) : 16, batch);
    hipStream_t s = (hipStream_t)stream;
    if (dtype == 0)
        diag_add_batched_kernel<double><<<g, 256, 0, s>>>((const DiagDesc *)descs, lam);
    else
        diag_add_batched_kernel<float><<<g, 256, 0, s>>>((const DiagDesc *)descs, lam);
    VT_LAUNCH_CHECK();
    return VT_OK;
}

// out = x + y (n doubles): refinement update over the whole batch
__global__ void add_kernel(double *__restrict__ x, const double *__restrict__ y, long long n)
{
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (long long)gridDim.x * blockDim.x)
        x[i] += y[i];
}

VT_EXPORT int vt_add_inplace(double *x, const double *y, long long n, void *stream)
{
    VT_CHECK_ARG(x && y && n >= 0);
    if (n == 0) return VT_OK;
    unsigned g = vt_blocks(n, 256);
    add_kernel<<<g < 4096 ? g : 4096, 256, 0, (hipStream_t)stream>>>(x, y, n);
    VT_LAUNCH_CHECK();
    return VT_OK;
}

// n_b = offsets[base[b+1]] - offsets[base[b]]  (intervals per block, on device)
__global__ void seg_counts_kernel(const int *__restrict__ offsets, const int *__restrict__ base,
                                  int nseg, int *__restrict__ out)
{
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b < nseg) out[b] = offsets[base[b + 1]] - offsets[base[b]];
}

VT_EXPORT int vt_seg_counts(const int *offsets, const int *base, int nseg, int *out, void *stream)
{
    VT_CHECK_ARG(offsets && base && out && nseg >= 1);
    seg_counts_kernel<<<vt_blocks(nseg, 256), 256, 0, (hipStream_t)stream>>>(offsets, base, nseg, out);
    VT_LAUNCH_CHECK();
    return VT_OK;
}

// ------------------------------------------------------------------------------------
// Skinny grouped GEMM (fp64): problem b computes C_b (M x Nc) = A_b (M x K) B_b (K x Nc)
// for its own Nc <= 32 (the launch's max_Nc bounds them all), with A_b(m, k) at A[m*sam + k*sak] (either orientation), B and C
// row-major.  Optionally each output row m is then multiplied by the Nc x Nc matrix
// T_b + m*tstride (C_m <- T_m C_m): with tstride = Nc*Nc this fuses Z_j = T_j (S P)_j
// into the product; with tstride = 0 every row shares one matrix (right factor C T^T).
//
// One workgroup owns 16 or 64 rows of one K split; its 8 waves take 16-wide K chunks in turn
// (wave w: chunks w, w+8, ...).  Operands go from global memory straight into the MFMA
// f64 16x16x4 operand layout: lane (r = l%16, g = l/16) loads the 4 consecutive k values
// c*16 + 4g + s (s = 0..3) of its row, and MFMA step s pairs them with the B rows of the
// same k, so each step sums k in {c*16 + 4g + s} and the 4 steps cover the chunk.  The
// wave partials are added in LDS in a fixed tree order.
//
// Long reductions are split: a problem with nch = ceil(K/16) chunks uses
// ks = clamp(nch/32, 1, 8) splits (a function of its own K only, so a problem's rounding
// does not depend on the batch), each writing a partial tile to its workspace; a second
// pass adds the ks partials in order and applies the row factor.
// ------------------------------------------------------------------------------------

// NT = 1: problems with Nc <= 16, and Nc = 17 with X16; NT = 2: any Nc <= 32, problems
// with Nc <= 17 taking the NT = 1 routes (every problem's route is fixed by its own Nc:
// <= 16 one MFMA tile, 17 one tile + the vector column, more two tiles), so a launch's
// width changes no result.
template <int NT, int RT, bool X16>
__global__ void __launch_bounds__(64 * SK_WAVES) skinny_kernel(const SkDesc *__restrict__ descs)
{
    __shared__ double red[SK_WAVES / 2][16 * RT][33];
    const SkDesc d = descs[blockIdx.y];
    const int m0 = blockIdx.x * 16 * RT;
    if (m0 >= d.M || (int)blockIdx.z >= skinny_splits(d.K)) return;
    if (d.active && *d.active == 0.0) return;
    if constexpr (NT == 1) {
        // the vector column reads B's column 16: only problems that have one take it
        if (X16 && d.Nc == 17)
            skinny_tile<1, 1, true>(d, m0, blockIdx.z, red);
        else
            skinny_tile<1, RT>(d, m0, blockIdx.z, red);
    } else {
        if (d.Nc == 17)
            skinny_tile<1, 1, true>(d, m0, blockIdx.z, red);
        else if (d.Nc <= 16)
            skinny_tile<1, 1>(d, m0, blockIdx.z, red);
        else
            skinny_tile<2, RT>(d, m0, blockIdx.z, red);
    }
}

// second pass of split problems: C = (sum_z partial_z) (row factor applied)
__global__ void __launch_bounds__(256) skinny_reduce_kernel(const SkDesc *__restrict__ descs)
{
    __shared__ double rowbuf[8][33];
    const SkDesc d = descs[blockIdx.y];
    const int ks = skinny_splits(d.K), Nc = d.Nc;
    const int m0 = blockIdx.x * 8;
    if (ks == 1 || m0 >= d.M) return;
    if (d.active && *d.active == 0.0) return;
    const int rr = threadIdx.x >> 5, col = threadIdx.x & 31;
    const int m = m0 + rr;
    double v = 0.0;
    if (m < d.M && col < Nc)
        for (int zz = 0; zz < ks; ++zz) v += d.ws[((long long)zz * d.M + m) * Nc + col];
    rowbuf[rr][col] = v;
    __syncthreads();
    if (m < d.M && col < Nc) {
        const double o = d.T ? skinny_row_factor(d, m, col, Nc, rowbuf[rr]) : v;
        d.C[sk_c_offset(d, m, col)] = o;
    }
}

// the K-split rule of the skinny product: ks = clamp(ceil(K / 16) / chunks_per_split, 1,
// max_splits), for hosts that size many problems' workspaces at once
VT_EXPORT int vt_skinny_split_rule(int *chunks_per_split_host, int *max_splits_host)
{
    VT_CHECK_ARG(chunks_per_split_host && max_splits_host);
    *chunks_per_split_host = VT_SKINNY_CHUNKS_PER_SPLIT;
    *max_splits_host = 8;
    return VT_OK;
}

// workspace (doubles) one problem of the skinny product needs: 0 unless its K is split
VT_EXPORT int vt_gemm_skinny_workspace_size(int M, int K, int Nc, long long *doubles_host)
{
    VT_CHECK_ARG(M >= 0 && K >= 0 && Nc >= 1 && Nc <= 32 && doubles_host);
    const int ks = skinny_splits(K);
    *doubles_host = ks > 1 ? (long long)ks * M * Nc : 0;
    return VT_OK;
}

int vt_skinny_first_pass(const void *descs, int batch, int max_M, int max_K, int max_Nc,
                         hipStream_t s)
{
    const int ks = skinny_splits(max_K);
    const SkDesc *d = (const SkDesc *)descs;
    // 16-row workgroups: measured faster than 64-row ones at every launch size of the
    // operator, the preconditioner and the encoder's coefficient products
    dim3 grid(vt_blocks(max_M, 16), batch, ks);
    if (max_Nc <= 16)
        skinny_kernel<1, 1, false><<<grid, 64 * SK_WAVES, 0, s>>>(d);
    else if (max_Nc == 17)
        skinny_kernel<1, 1, true><<<grid, 64 * SK_WAVES, 0, s>>>(d);
    else
        skinny_kernel<2, 1, false><<<grid, 64 * SK_WAVES, 0, s>>>(d);
    VT_LAUNCH_CHECK();
    return VT_OK;
}

VT_EXPORT int vt_gemm_skinny(const void *descs, int batch, int max_M, int max_K, int max_Nc,
                             void *stream)
{
    VT_CHECK_ARG(descs && batch >= 1 && batch <= 65535 && max_M >= 0 && max_K >= 0 && max_Nc >= 1 &&
                 max_Nc <= 32);
    if (max_M == 0) return VT_OK;
    hipStream_t s = (hipStream_t)stream;
    int r = vt_skinny_first_pass(descs, batch, max_M, max_K, max_Nc, s);
    if (r) return r;
    if (skinny_splits(max_K) > 1)
        skinny_reduce_kernel<<<dim3(vt_blocks(max_M, 8), batch), 256, 0, s>>>((const SkDesc *)descs);
    VT_LAUNCH_CHECK();
    return VT_OK;
}

