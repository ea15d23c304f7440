// Time-parallel IAF encoder.
//
// The reference integrates each neuron serially over all samples, one thread per neuron
// (vtem/ensemble_encode.py:438-552).  With reset to zero, spike k of a neuron is the first
// time the integral F(t) = int_0^t y reaches the level L_k = k delta - v0 (v0 the carried
// voltage): after spike k-1 the voltage is F(t) - L_{k-1}, so the next threshold crossing is
// F = L_{k-1} + delta, and no earlier point can have reached a higher level.  The spike
// count up to t is therefore the number of levels at or below the running maximum of F,
// and the time axis splits into independent segments:
//
//   1. iafp_partial: per (neuron, segment of IAFP_SEG sample intervals) the segment's area
//      A and the maximum M of F over the segment relative to its start (step ends, and the
//      interior peak of a falling step y1 > 0 > y2);
//   2. iafp_scan: per neuron, the exclusive prefix (F_s, M_s) of every segment under
//      (A, M) o (A', M') = (A + A', max(M, A + M')), the totals, and the final state;
//   3. iafp_walk: per segment holding spikes (level count at its end above the count at its
//      start), the segment is integrated again from F_s and every level crossed gets its
//      time from the same quadratic root as the reference (:501, :521).
//
// The samples, the trapezoid areas, the branch conditions and the root formulas are the
// reference's; only the bookkeeping differs (F relative to the call start instead of a
// voltage reset at each spike), so counts agree with the serial encoder and times agree to
// rounding.  Work per neuron drops from `size` dependent steps to IAFP_SEG plus a scan of
// size / IAFP_SEG segments, and the launch holds N x size / IAFP_SEG lanes.
#include "vt_common.hpp"
#include "vt_iaf.hpp"
#include <mutex>

#define IAFP_SEG 64        // sample intervals per segment
#define IAFP_WAVES 4       // segments per workgroup (one wave each, 64 neurons per wave)
#ifndef IAFP_ILP
#define IAFP_ILP 4         // samples whose inputs are evaluated together
#endif
#ifndef IAFP_TRIG_WAVES
#define IAFP_TRIG_WAVES 3  // minimum waves per SIMD of the trigonometric-input kernels
#endif

// One encoder call on one block of neurons.  ws holds (S + 2) x N pairs of doubles,
// S = ceil((size - 1) / IAFP_SEG): segment partials, then prefixes in place; row S the
// totals (F, M); row S + 1 the initial (v0, time_count).
struct IafParDesc {
    const double *coef;     // trig source: coefficient pt of neuron n at
    long long coef_ld;      //   coef[pt * coef_ld + n * coef_ns], and the basis
    long long coef_ns;
    const double *basis;    // size x Dt
    const void *input;      // sampled source: size x N (ld input_ld), FP64 or FP32
    long long input_ld;
    double *spikes;         // max_spike x N (ld spike_ld)
    long long spike_ld;
    double *v0, *time_count;
    const double *kappa, *bias, *delta;
    int *spike_count;
    double2 *ws;
    int *overflow;          // optional: set to 1 when a neuron fires more than max_spike times
    double dt;
    int N, size, Dt, max_spike, root_mode, in_dtype;
    double *ybuf;           // optional: size x N (ld y_ld), receives y = (v + b) / kappa at
    long long y_ld;         //   every sample in pass 1; the walk then reads it back
};

static __host__ __device__ inline int iafp_segments(int size)
{
    return size > 1 ? (size - 1 + IAFP_SEG - 1) / IAFP_SEG : 0;
}

// Number of levels k >= 1 with k delta - v0 <= M, at most cap, evaluated with the same
// level expression the walk compares against.
__device__ __forceinline__ int iafp_levels(double M, double v0, double delta, int cap)
{
#pragma clang fp contract(off)
    if (!(M + v0 >= delta)) return 0;    // also M = -inf (no step yet) and NaN
    double q = floor((M + v0) / delta);
    if (!(q < (double)cap)) return cap;
    int k = (int)q;
    if (k < cap && (double)(k + 1) * delta - v0 <= M)
        ++k;
    else if (k >= 1 && (double)k * delta - v0 > M)
        --k;
    return k;
}

template <class Src>
__device__ __forceinline__ void iafp_partial_dev(const IafParDesc &d, const Src &src, int n, int s,
                                                 int S)
{
#pragma clang fp contract(off)
    const int steps = d.size - 1;
    const int i0 = s * IAFP_SEG, i1 = min(i0 + IAFP_SEG, steps);
    const double kappa = d.kappa[n], bias = d.bias[n], dt = d.dt;
    double *yout = d.ybuf ? d.ybuf + n : nullptr;
    const long long yld = d.y_ld;
    double y1 = (src(i0) + bias) / kappa;
    if (yout && s == 0) yout[0] = y1;
    double P = 0.0, M = -INFINITY;
    auto step = [&](int i, double v2) {
#pragma clang fp contract(off)
        const double y2 = (v2 + bias) / kappa;
        if (yout) yout[(long long)(i + 1) * yld] = y2;
        const double area = (y1 + y2) / 2 * dt;
        double top;
        if (y2 >= 0 || y1 <= 0) {
            top = P + area;
        } else {
            const double a = (y2 - y1) / dt;
            top = P + y1 * y1 / (-2 * a);
        }
        M = fmax(M, top);
        P += area;
        y1 = y2;
    };
    int i = i0;
    for (; i + IAFP_ILP <= i1; i += IAFP_ILP) {
        double v[IAFP_ILP];
        src.template eval<IAFP_ILP>(i + 1, v);
#pragma unroll
        for (int r = 0; r < IAFP_ILP; ++r) step(i + r, v[r]);
    }
    for (; i < i1; ++i) step(i, src(i + 1));
    d.ws[(long long)s * d.N + n] = make_double2(P, M);
}

// The walk's input: y = (v + b) / kappa from a source of v, or y as pass 1 stored it.
template <class Src>
struct ScaledInput {
    const Src &src;
    double bias, kappa;
    __device__ __forceinline__ double operator()(int i) const
    {
#pragma clang fp contract(off)
        return (src(i) + bias) / kappa;
    }
    template <int R>
    __device__ __forceinline__ void eval(int i, double *out) const
    {
#pragma clang fp contract(off)
        src.template eval<R>(i, out);
#pragma unroll
        for (int r = 0; r < R; ++r) out[r] = (out[r] + bias) / kappa;
    }
};

struct StoredInput {
    const double *y;
    long long ld;
    __device__ __forceinline__ double operator()(int i) const { return y[(long long)i * ld]; }
    template <int R>
    __device__ __forceinline__ void eval(int i, double *out) const
    {
#pragma unroll
        for (int r = 0; r < R; ++r) out[r] = y[(long long)(i + r) * ld];
    }
};

template <class Ysrc>
__device__ __forceinline__ void iafp_walk_dev(const IafParDesc &d, const Ysrc &ys, int n, int s,
                                              int S)
{
#pragma clang fp contract(off)
    const long long N = d.N;
    const int cap = d.max_spike + 1;
    const double delta = d.delta[n];
    const double2 init = d.ws[(long long)(S + 1) * N + n];
    const double v0 = init.x, tc0 = init.y;
    const double2 pre = d.ws[(long long)s * N + n];
    const double2 nxt = d.ws[(long long)(s + 1) * N + n];
    int k = iafp_levels(pre.y, v0, delta, cap);
    const int k1 = iafp_levels(nxt.y, v0, delta, cap);
    if (k >= k1) return;
    const int kfinal = iafp_levels(d.ws[(long long)S * N + n].y, v0, delta, cap);
    double *col = d.spikes + n;
    auto emit = [&](int j, double t) {
        // Rows beyond the buffer: the reference keeps overwriting the last row
        // (ensemble_encode.py:503-507), so it ends holding spike max_spike; skip the
        // spike that the later one overwrites.
        if (j == d.max_spike - 1 && kfinal == cap) return;
        const int row = j < d.max_spike - 1 ? j : d.max_spike - 1;
        col[(long long)row * d.spike_ld] = t;
    };
    const int steps = d.size - 1;
    const int i0 = s * IAFP_SEG, i1 = min(i0 + IAFP_SEG, steps);
    const double dt = d.dt;
    const int mode = d.root_mode;
    double y1 = ys(i0);
    double P = 0.0;
    double best = -INFINITY, tbest = tc0 + (double)i0 * dt;
    auto step = [&](int i, double y2) {
#pragma clang fp contract(off)
        const double area = (y1 + y2) / 2 * dt;
        const double F = pre.x + P;
        const double t0 = tc0 + (double)i * dt;
        if (y2 >= 0 || y1 <= 0) {
            // rising / normal step (ensemble_encode.py:493-514); the slope (a division)
            // only where a level is crossed
            while (k < k1) {
                const double L = (double)(k + 1) * delta - v0;
                if (!(F + area >= L)) break;
                const double a = (y2 - y1) / dt;
                const double dv = L - F;
                double remain;
                if (mode == 0)
                    remain = (fabs(a) <= 1e-12) ? (dv / y1) : (sqrt(y1 * y1 + 2 * dv * a) - y1) / (a);
                else
                    remain = 2 * dv / (sqrt(y1 * y1 + 2 * dv * a) + y1);
                emit(k, t0 + remain);
                ++k;
            }
            if (F + area > best) { best = F + area; tbest = t0 + dt; }
        } else {
            // falling step y1 > 0 > y2 (ensemble_encode.py:515-535)
            const double a = (y2 - y1) / dt;
            while (k < k1) {
                const double L = (double)(k + 1) * delta - v0;
                const double dv = L - F;
                const double disc = y1 * y1 + 2 * dv * a;
                if (!(disc >= 0)) break;
                double remain;
                if (mode == 0)
                    remain = (sqrt(disc) - y1) / a;
                else
                    remain = 2 * dv / (sqrt(disc) + y1);
                emit(k, t0 + remain);
                ++k;
            }
            const double top = F + y1 * y1 / (-2 * a);
            if (top > best) { best = top; tbest = t0 + y1 / (-a); }
        }
        P += area;
        y1 = y2;
    };
    // samples in groups of IAFP_ILP (their inputs evaluated together), each step taken
    // only while levels remain, as in a one-sample loop
    int i = i0;
    for (; i + IAFP_ILP <= i1 && k < k1; i += IAFP_ILP) {
        double v[IAFP_ILP];
        ys.template eval<IAFP_ILP>(i + 1, v);
#pragma unroll
        for (int r = 0; r < IAFP_ILP; ++r) {
            if (k >= k1) break;
            step(i + r, v[r]);
        }
    }
    for (; i < i1 && k < k1; ++i) step(i, ys(i + 1));
    // Levels the prefix maximum counts in this segment but the step tests missed by a
    // rounding tie are reached at the segment's highest point.
    for (; k < k1; ++k) emit(k, tbest);
}

// A workgroup's IAFP_WAVES waves cover 2^wl neuron groups (64 lanes each) on each of
// IAFP_WAVES / 2^wl consecutive segments.  With wl = 2 the four waves evaluate the same
// basis rows at about the same time, so three of them find the rows in the scalar cache.
// Workgroup x: neuron-group tile x % GW, segment group x / GW.
__device__ __forceinline__ bool iafp_task(const IafParDesc &d, int wl, int &n, int &s, int &S)
{
    S = iafp_segments(d.size);
    const int G = (d.N + 63) / 64;
    const int GW = (G + (1 << wl) - 1) >> wl;
    const int x = blockIdx.x;
    const int sg = x / GW;
    // wave-uniform segment: the basis row addresses become scalar loads shared by the wave
    const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    n = (((x - sg * GW) << wl) + (w & ((1 << wl) - 1))) * 64 + (threadIdx.x & 63);
    s = sg * (IAFP_WAVES >> wl) + (w >> wl);
    return n < d.N && s < S;
}

// BATCHED selects the descriptor source at compile time (a runtime select between the
// array and the by-value argument would turn every field into a generic-pointer load,
// defeating scalar loads of the uniform fields).
// At least 3 waves per SIMD: the 65-term order needs 130 VGPRs of coefficients, and a few
// more would leave 2 waves to hide the latency of the basis loads.
template <int DT, int PHASE, bool BATCHED>
__global__ void __launch_bounds__(64 * IAFP_WAVES) __attribute__((amdgpu_waves_per_eu(IAFP_TRIG_WAVES)))
iafp_trig_kernel(const IafParDesc *__restrict__ descs, IafParDesc d0, int wl)
{
    const IafParDesc d = BATCHED ? descs[blockIdx.y] : d0;
    int n, s, S;
    if (!iafp_task(d, wl, n, s, S)) return;
    TrigSource<DT> src;
    src.load(d.coef, d.coef_ld, d.basis, d.Dt, n, d.coef_ns);
    if (PHASE == 0)
        iafp_partial_dev(d, src, n, s, S);
    else
        iafp_walk_dev(d, ScaledInput<TrigSource<DT>>{src, d.bias[n], d.kappa[n]}, n, s, S);
}

// Pass 2 over the input that pass 1 stored (every descriptor has ybuf): a stream of
// y = (v + b) / kappa instead of the trigonometric evaluation.
template <bool BATCHED>
__global__ void __launch_bounds__(64 * IAFP_WAVES)
iafp_walk_stored_kernel(const IafParDesc *__restrict__ descs, IafParDesc d0, int wl)
{
    const IafParDesc d = BATCHED ? descs[blockIdx.y] : d0;
    int n, s, S;
    if (!iafp_task(d, wl, n, s, S)) return;
    iafp_walk_dev(d, StoredInput{d.ybuf + n, d.y_ld}, n, s, S);
}

template <typename TIn, int PHASE>
__global__ void __launch_bounds__(64 * IAFP_WAVES)
iafp_sampled_kernel(IafParDesc d, int wl)
{
    int n, s, S;
    if (!iafp_task(d, wl, n, s, S)) return;
    SampledSource<TIn> src{(const TIn *)d.input + n, d.input_ld};
    if (PHASE == 0)
        iafp_partial_dev(d, src, n, s, S);
    else
        iafp_walk_dev(d, ScaledInput<SampledSource<TIn>>{src, d.bias[n], d.kappa[n]}, n, s, S);
}

// Pass 1 of the batched stored-input form on MFMA.  The input of neuron n at sample i is
// v = sum_p C[p][n] Theta[i][p]: a product of the basis rows (samples x Dt) and the
// coefficients (Dt x neurons).  A workgroup takes IAFM_SEGS consecutive segments and 64
// neurons: the basis rows of its IAFM_SEGS * 64 + 1 sample points are staged in LDS, each
// wave forms 16 neurons' inputs as FP64 MFMA 16x16x4 tiles (A = basis rows, B = the
// neurons' coefficients, K = the padded term count), then y = (v + b) / kappa goes to the
// stored input and to LDS, where one lane per (segment, neuron) takes the segment's steps
// in order exactly as iafp_partial_dev does.  A sample shared by two workgroups (the last
// point of one, the first of the next) is the same MFMA element in both, so both see the
// same value.  The sums run in the MFMA's order, not the VALU form's p = 0..Dt-1, so the
// inputs differ from the other forms by rounding (the walk reads the stored y, so the
// pass and the walk agree with each other).
#define IAFM_SEGS 2
#define IAFM_PTS (IAFM_SEGS * IAFP_SEG + 1)
#define IAFM_ROWS (16 * ((IAFM_PTS + 15) / 16))

#define IAFM_YN 65

template <int KS>
__global__ void __launch_bounds__(256) iafp_trig_mfma_kernel(const IafParDesc *__restrict__ descs)
{
    constexpr int DTP = 4 * KS;
    constexpr int LD = DTP + 1;
    constexpr int RT = IAFM_ROWS / 16;
    constexpr int LDS_TH = IAFM_ROWS * LD, LDS_Y = IAFM_PTS * IAFM_YN;
    // the basis rows, then (after the products) y of the workgroup's points x 64 neurons
    __shared__ double sh[LDS_TH > LDS_Y ? LDS_TH : LDS_Y];
    const IafParDesc d = descs[blockIdx.z];
    const int S = iafp_segments(d.size);
    const int s0 = blockIdx.x * IAFM_SEGS, n0 = blockIdx.y * 64;
    if (s0 >= S || n0 >= d.N) return;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int steps = d.size - 1, Dt = d.Dt;
    const int i0 = s0 * IAFP_SEG;
    // basis rows of the sample points i0 .. (rows past the call's samples read its last
    // row and are not used; terms past Dt are zero)
    for (int e = tid; e < IAFM_ROWS * DTP; e += 256) {
        const int row = e / DTP, p = e % DTP;
        const int i = min(i0 + row, d.size - 1);
        sh[row * LD + p] = p < Dt ? d.basis[(long long)i * Dt + p] : 0.0;
    }
    // this wave's 16 neurons' coefficients in the MFMA B layout: slot g of k-step j holds
    // term 4 j + g of neuron nb
    const int nb = n0 + 16 * wave + (lane & 15), g = lane >> 4;
    const bool nin = nb < d.N;
    double c[KS];
#pragma unroll
    for (int j = 0; j < KS; ++j) {
        const int p = 4 * j + g;
        c[j] = (nin && p < Dt) ? d.coef[(long long)p * d.coef_ld + (long long)nb * d.coef_ns] : 0.0;
    }
    __syncthreads();
    v4d acc[RT];
#pragma unroll
    for (int t = 0; t < RT; ++t) acc[t] = (v4d){0, 0, 0, 0};
#pragma unroll
    for (int j = 0; j < KS; ++j)
#pragma unroll
        for (int t = 0; t < RT; ++t)
            acc[t] = __builtin_amdgcn_mfma_f64_16x16x4f64(sh[(16 * t + (lane & 15)) * LD + 4 * j + g], c[j],
                                                          acc[t], 0, 0, 0);
    __syncthreads();                           // the basis rows are no longer read
    const double bias = nin ? d.bias[nb] : 0.0, kappa = nin ? d.kappa[nb] : 1.0;
    const int last = min(i0 + IAFM_SEGS * IAFP_SEG, steps);
    double *yout = d.ybuf + nb;
#pragma unroll
    for (int t = 0; t < RT; ++t)
#pragma unroll
        for (int q = 0; q < 4; ++q) {
#pragma clang fp contract(off)
            const int row = 16 * t + mfma_row_f64(lane, q);
            if (row < IAFM_PTS) {
                const int i = i0 + row;
                const double y = (acc[t][q] + bias) / kappa;
                sh[row * IAFM_YN + 16 * wave + (lane & 15)] = y;
                // every sample once: a workgroup stores its points after i0 (and point 0)
                if (nin && i <= last && (i > i0 || i0 == 0)) yout[(long long)i * d.y_ld] = y;
            }
        }
    __syncthreads();
    if (tid >= 64 * IAFM_SEGS) return;
    const int u = tid >> 6, col = tid & 63, n = n0 + col, s = s0 + u;
    if (n >= d.N || s >= S) return;
    {
#pragma clang fp contract(off)
        const int si0 = s * IAFP_SEG, si1 = min(si0 + IAFP_SEG, steps);
        const double dt = d.dt;
        double y1 = sh[(si0 - i0) * IAFM_YN + col];
        double P = 0.0, M = -INFINITY;
        for (int i = si0; i < si1; ++i) {
            const double y2 = sh[(i + 1 - i0) * IAFM_YN + col];
            const double area = (y1 + y2) / 2 * dt;
            double top;
            if (y2 >= 0 || y1 <= 0) {
                top = P + area;
            } else {
                const double a = (y2 - y1) / dt;
                top = P + y1 * y1 / (-2 * a);
            }
            M = fmax(M, top);
            P += area;
            y1 = y2;
        }
        d.ws[(long long)s * d.N + n] = make_double2(P, M);
    }
}

// Per neuron: exclusive prefixes of the segment partials, totals, and the carried state
// the serial encoder would leave (ensemble_encode.py:541-545): v0 <- F_T - L_K with K the
// spike count (at most max_spike + 1, the reference's overflow stop), time_count advanced
// by the call's size - 1 intervals.
template <bool BATCHED>
__global__ void __launch_bounds__(256)
iafp_scan_kernel(const IafParDesc *__restrict__ descs, IafParDesc d0)
{
#pragma clang fp contract(off)
    const IafParDesc d = BATCHED ? descs[blockIdx.y] : d0;
    const int n = blockIdx.x * 256 + threadIdx.x;
    if (n >= d.N) return;
    const long long N = d.N;
    const int S = iafp_segments(d.size);
    const double v0 = d.v0[n], tc0 = d.time_count[n], delta = d.delta[n];
    double F = 0.0, M = -INFINITY;
    double2 *w = d.ws + n;
    int s = 0;
    for (; s + 8 <= S; s += 8) {
        double2 p[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) p[u] = w[(long long)(s + u) * N];
#pragma unroll
        for (int u = 0; u < 8; ++u) {
            w[(long long)(s + u) * N] = make_double2(F, M);
            M = fmax(M, F + p[u].y);
            F += p[u].x;
        }
    }
    for (; s < S; ++s) {
        const double2 p = w[(long long)s * N];
        w[(long long)s * N] = make_double2(F, M);
        M = fmax(M, F + p.y);
        F += p.x;
    }
    w[(long long)S * N] = make_double2(F, M);
    w[(long long)(S + 1) * N] = make_double2(v0, tc0);
    if (S == 0) {
        d.spike_count[n] = 0;
        return;
    }
    const int K = iafp_levels(M, v0, delta, d.max_spike + 1);
    d.spike_count[n] = K;
    if (K > d.max_spike && d.overflow) *d.overflow = 1;
    d.v0[n] = K > 0 ? F - ((double)K * delta - v0) : v0 + F;
    d.time_count[n] = tc0 + (double)(d.size - 1) * d.dt;
}

static int iafp_check(const IafParDesc &d)
{
    VT_CHECK_ARG(d.N >= 0 && d.size >= 0);
    if (d.N == 0) return VT_OK;
    VT_CHECK_ARG(d.spikes && d.v0 && d.kappa && d.bias && d.delta && d.time_count &&
                 d.spike_count && d.ws);
    VT_CHECK_ARG(d.max_spike >= 1 && d.spike_ld >= d.N && d.dt > 0);
    VT_CHECK_ARG(d.root_mode == 0 || d.root_mode == 1);
    return VT_OK;
}

static long long iafp_ws_bytes(int N, int size)
{
    return (long long)(iafp_segments(size) + 2) * N * (long long)sizeof(double2);
}

VT_EXPORT int vt_iaf_par_workspace_size(int N, int size, long long *bytes)
{
    VT_CHECK_ARG(N >= 0 && size >= 0 && bytes);
    *bytes = iafp_ws_bytes(N, size);
    return VT_OK;
}

// Grid of the segment kernels for blocks of at most maxN neurons and maxSize samples, and
// the number 2^wl of neuron groups a workgroup holds per segment: 4 once there are enough
// groups to fill the workgroups (the waves then share basis rows), fewer for small blocks.
struct IafpGrid {
    dim3 g;
    int wl;
};

static IafpGrid iafp_grid(int maxN, int maxSize, int nd)
{
    const int G = (maxN + 63) / 64;
#ifdef IAFP_WL
    const int wl = IAFP_WL;
#else
    const int wl = G >= 8 ? 2 : (G >= 2 ? 1 : 0);
#endif
    const int GW = (G + (1 << wl) - 1) >> wl;
    const int per = IAFP_WAVES >> wl;
    const int sg = (iafp_segments(maxSize) + per - 1) / per;
    return {dim3((unsigned)(GW * (sg > 0 ? sg : 1)), (unsigned)nd), wl};
}

template <int DT>
static void iafp_trig_launch(const IafParDesc *descs, const IafParDesc &d0, IafpGrid g, dim3 gs,
                             bool stored, hipStream_t s)
{
    if (descs) {
        iafp_trig_kernel<DT, 0, true><<<g.g, 64 * IAFP_WAVES, 0, s>>>(descs, d0, g.wl);
        iafp_scan_kernel<true><<<gs, 256, 0, s>>>(descs, d0);
        if (stored)
            iafp_walk_stored_kernel<true><<<g.g, 64 * IAFP_WAVES, 0, s>>>(descs, d0, g.wl);
        else
            iafp_trig_kernel<DT, 1, true><<<g.g, 64 * IAFP_WAVES, 0, s>>>(descs, d0, g.wl);
    } else {
        iafp_trig_kernel<DT, 0, false><<<g.g, 64 * IAFP_WAVES, 0, s>>>(descs, d0, g.wl);
        iafp_scan_kernel<false><<<gs, 256, 0, s>>>(descs, d0);
        iafp_trig_kernel<DT, 1, false><<<g.g, 64 * IAFP_WAVES, 0, s>>>(descs, d0, g.wl);
    }
}

template <typename TIn>
static void iafp_sampled_launch(const IafParDesc &d, IafpGrid g, dim3 gs, hipStream_t s)
{
    iafp_sampled_kernel<TIn, 0><<<g.g, 64 * IAFP_WAVES, 0, s>>>(d, g.wl);
    iafp_scan_kernel<false><<<gs, 256, 0, s>>>(nullptr, d);
    iafp_sampled_kernel<TIn, 1><<<g.g, 64 * IAFP_WAVES, 0, s>>>(d, g.wl);
}

static int iafp_trig_dispatch(const IafParDesc *descs, const IafParDesc &d0, int maxDt,
                              IafpGrid g, dim3 gs, bool stored, hipStream_t s)
{
    if (maxDt <= 9)
        iafp_trig_launch<9>(descs, d0, g, gs, stored, s);
    else if (maxDt <= 17)
        iafp_trig_launch<17>(descs, d0, g, gs, stored, s);
    else if (maxDt <= 33)
        iafp_trig_launch<33>(descs, d0, g, gs, stored, s);
    else if (maxDt <= 65)
        iafp_trig_launch<65>(descs, d0, g, gs, stored, s);
    else
        return VT_ERANGE;
    VT_LAUNCH_CHECK();
    return VT_OK;
}

VT_EXPORT int vt_iaf_encode_trig_par(const double *coef, long long coef_ld, const double *basis,
                                     int Dt, int N, int size, double *spikes, long long spike_ld,
                                     int max_spike, double *v0, const double *kappa,
                                     const double *bias, const double *delta, double *time_count,
                                     int *spike_count, double dt, int root_mode, void *ws,
                                     long long ws_bytes, void *stream)
{
    IafParDesc d{coef, coef_ld, 1, basis, nullptr, 0, spikes, spike_ld, v0, time_count, kappa,
                 bias, delta, spike_count, (double2 *)ws, nullptr, dt, N, size, Dt, max_spike,
                 root_mode, 0, nullptr, 0};
    int r = iafp_check(d);
    if (r || N == 0) return r;
    VT_CHECK_ARG(coef && basis && Dt >= 1 && coef_ld >= N);
    VT_CHECK_ARG(ws_bytes >= iafp_ws_bytes(N, size));
    return iafp_trig_dispatch(nullptr, d, Dt, iafp_grid(N, size, 1), dim3(vt_blocks(N, 256), 1),
                              false, (hipStream_t)stream);
}

VT_EXPORT int vt_iaf_encode_par(const void *input, int in_dtype, long long input_ld, int N,
                                int size, double *spikes, long long spike_ld, int max_spike,
                                double *v0, const double *kappa, const double *bias,
                                const double *delta, double *time_count, int *spike_count,
                                double dt, int root_mode, void *ws, long long ws_bytes,
                                void *stream)
{
    IafParDesc d{nullptr, 0, 1, nullptr, input, input_ld, spikes, spike_ld, v0, time_count,
                 kappa, bias, delta, spike_count, (double2 *)ws, nullptr, dt, N, size, 0,
                 max_spike, root_mode, in_dtype, nullptr, 0};
    int r = iafp_check(d);
    if (r || N == 0) return r;
    VT_CHECK_ARG((input || size == 0) && input_ld >= N && (in_dtype == 0 || in_dtype == 1));
    VT_CHECK_ARG(ws_bytes >= iafp_ws_bytes(N, size));
    hipStream_t s = (hipStream_t)stream;
    IafpGrid g = iafp_grid(N, size, 1);
    dim3 gs(vt_blocks(N, 256), 1);
    if (in_dtype == 0)
        iafp_sampled_launch<double>(d, g, gs, s);
    else
        iafp_sampled_launch<float>(d, g, gs, s);
    VT_LAUNCH_CHECK();
    return VT_OK;
}

// ------------------------------------------------------------------------------------
// The same encoder for callers that cannot allocate device memory (code under HIP graph
// capture, tight serving loops, harnesses that forbid hipMalloc): the segment workspace
// lives in a file-scope device array of VT_NOALLOC_SCRATCH_BYTES, one per GPU, held until
// the process ends.  Its address is taken once per device with hipGetSymbolAddress (the
// array's name in host code is only a stand-in) and used from then on.  Every call on a
// GPU shares that GPU's array: calls that overlap on one GPU (two streams, two threads)
// race on it.  Pass 1 writes every workspace entry the scan and the walk read, so the
// contents a previous call left need no clearing.
// ------------------------------------------------------------------------------------
#define VT_NOALLOC_SCRATCH_BYTES (64LL << 20)
#define VT_MAX_DEVICES 64

static __device__ __align__(256) unsigned char scratch[VT_NOALLOC_SCRATCH_BYTES];

static std::mutex g_scratch_mutex;
static void *g_scratch_ptr[VT_MAX_DEVICES];

// Device address of the current GPU's scratch array (looked up on the first call per device).
static int scratch_address(void **out)
{
    int dev = 0;
    VT_HIP(hipGetDevice(&dev));
    if (dev < 0 || dev >= VT_MAX_DEVICES) return VT_EUNSUPPORTED;
    std::lock_guard<std::mutex> lock(g_scratch_mutex);
    if (!g_scratch_ptr[dev]) {
        void *ptr = nullptr;
        hipError_t e = hipGetSymbolAddress((void **)&ptr, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return (int)e;
        if (!ptr) return VT_EUNSUPPORTED;
        g_scratch_ptr[dev] = ptr;
    }
    *out = g_scratch_ptr[dev];
    return VT_OK;
}

VT_EXPORT int vt_noalloc_scratch_bytes(long long *bytes)
{
    VT_CHECK_ARG(bytes);
    *bytes = VT_NOALLOC_SCRATCH_BYTES;
    return VT_OK;
}

VT_EXPORT int vt_iaf_encode_par_noalloc(const void *input, int in_dtype, long long input_ld,
                                        int N, int size, double *spikes, long long spike_ld,
                                        int max_spike, double *v0, const double *kappa,
                                        const double *bias, const double *delta,
                                        double *time_count, int *spike_count, double dt,
                                        int root_mode, void *stream)
{
    VT_CHECK_ARG(N >= 0 && size >= 0);
    const long long need = iafp_ws_bytes(N, size);
    if (need > VT_NOALLOC_SCRATCH_BYTES) return VT_ERANGE;
    void *ws = nullptr;
    int r = scratch_address(&ws);
    if (r) return r;
    return vt_iaf_encode_par(input, in_dtype, input_ld, N, size, spikes, spike_ld, max_spike, v0,
                             kappa, bias, delta, time_count, spike_count, dt, root_mode, ws,
                             need, stream);
}

// Batched trig encode: one descriptor per block (device array), launched together.  The
// host passes the maxima of N, size and Dt over the descriptors for the grid; workgroups
// beyond a block's own extent exit at once.  flags bit 0: every descriptor carries ybuf, so
// pass 1 stores the input and pass 2 reads it instead of evaluating it again.
template <int KS>
static void iafp_mfma_launch(const IafParDesc *descs, int nd, int maxN, int maxSize, hipStream_t s)
{
    const dim3 g1((unsigned)((iafp_segments(maxSize) + IAFM_SEGS - 1) / IAFM_SEGS),
                  (unsigned)((maxN + 63) / 64), (unsigned)nd);
    const IafpGrid g = iafp_grid(maxN, maxSize, nd);
    IafParDesc d0{};
    iafp_trig_mfma_kernel<KS><<<g1, 256, 0, s>>>(descs);
    iafp_scan_kernel<true><<<dim3(vt_blocks(maxN, 256), (unsigned)nd), 256, 0, s>>>(descs, d0);
    iafp_walk_stored_kernel<true><<<g.g, 64 * IAFP_WAVES, 0, s>>>(descs, d0, g.wl);
}

// flags: 1 = store the scaled input in pass 1 (every descriptor has ybuf) and walk over
// it; 2 (with 1) = form pass 1's inputs on MFMA (iafp_trig_mfma_kernel).
VT_EXPORT int vt_iaf_encode_trig_batched(const void *descs, int nd, int maxN, int maxSize,
                                         int maxDt, int flags, void *stream)
{
    VT_CHECK_ARG(nd >= 0 && maxN >= 0 && maxSize >= 0 && maxDt >= 1);
    VT_CHECK_ARG(descs || nd == 0);
    VT_CHECK_ARG((flags & ~3) == 0 && flags != 2);
    VT_CHECK_ARG(nd <= 65535 && (maxN + 63) / 64 <= 65535);
    if (nd == 0 || maxN == 0) return VT_OK;
    hipStream_t s = (hipStream_t)stream;
    if (flags & 2) {
        if (maxDt <= 12)
            iafp_mfma_launch<3>((const IafParDesc *)descs, nd, maxN, maxSize, s);
        else if (maxDt <= 20)
            iafp_mfma_launch<5>((const IafParDesc *)descs, nd, maxN, maxSize, s);
        else if (maxDt <= 36)
            iafp_mfma_launch<9>((const IafParDesc *)descs, nd, maxN, maxSize, s);
        else if (maxDt <= 52)
            iafp_mfma_launch<13>((const IafParDesc *)descs, nd, maxN, maxSize, s);
        else if (maxDt <= 68)
            iafp_mfma_launch<17>((const IafParDesc *)descs, nd, maxN, maxSize, s);
        else
            return VT_ERANGE;
        VT_LAUNCH_CHECK();
        return VT_OK;
    }
    IafParDesc d0{};
    return iafp_trig_dispatch((const IafParDesc *)descs, d0, maxDt, iafp_grid(maxN, maxSize, nd),
                              dim3(vt_blocks(maxN, 256), (unsigned)nd), (flags & 1) != 0, s);
}

VT_EXPORT int vt_iaf_par_desc_size(int *bytes)
{
    VT_CHECK_ARG(bytes);
    *bytes = (int)sizeof(IafParDesc);
    return VT_OK;
}
