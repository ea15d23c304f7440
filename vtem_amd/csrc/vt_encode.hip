// Encode side: receptive-field banks, IAF population encoder, spike-interval extraction.
//
// RF generation restates vtem/vrf.py:1197-1291 (Gabor) and :1120-1193 (DoG) for a mixed
// bank (per-neuron kind).  The IAF kernels restate vtem/ensemble_encode.py:438-552 (and
// :555-674 for random thresholds) with one lane per neuron, FP64 state, the reference's
// branch structure and operation order, and FP contraction disabled so each step is the
// same sequence of IEEE operations as the reference's kernel.  Unlike the reference,
// every lane is bounds-guarded (the reference writes state past num_neurons,
// ensemble_encode.py:543-545).
#include "vt_common.hpp"
#include "vt_iaf.hpp"

#define VT_PI 3.141592653589793238462643383279

// ------------------------------------------------------------------------------------
// Receptive fields
// ------------------------------------------------------------------------------------

struct RFParams {
    const int *kind;          // 0 Gabor, 1 DoG (nullptr: all `default_kind`)
    const double *alpha_inv;  // 1/h_alpha  (vrf.py:258)
    const double *theta;      // rotation (Gabor)
    const double *x0, *y0;
    const int *ab;            // 0 even / 1 odd (Gabor)
    double kappa;             // Gabor carrier
    double sc2, ss2;          // DoG sigma_center^2, sigma_surround^2
    int default_kind;
};

template <typename T>
__global__ void __launch_bounds__(256)
rf_kernel(T *__restrict__ out, long long ld, RFParams p, int Px, int Py, double Sx, double Sy,
          double x_start, double y_start)
{
    const int j = blockIdx.y;
    const int pix = blockIdx.x * blockDim.x + threadIdx.x;
    if (pix >= Px * Py) return;
    const int px = pix % Px, py = pix / Px;
    const int kind = p.kind ? p.kind[j] : p.default_kind;
    const double alpha = p.alpha_inv[j];
    double dxdy = Sx / Px;
    double x = (double)px * dxdy + x_start;
    dxdy = Sy / Py;
    double y = -((double)py * dxdy + y_start);
    x = alpha * (x - p.x0[j]);
    y = alpha * (y - p.y0[j]);
    double g;
    if (kind == 0) {
        double st, ct;
        sincos(p.theta[j], &st, &ct);
        double X = x * ct + y * st;
        double Y = -x * st + y * ct;
        double first = alpha * (1 / sqrt(2 * VT_PI)) * exp(-((4 * X * X) + (Y * Y)) / 8);
        sincos(p.kappa * X, &st, &ct);
        g = first * (p.ab[j] == 0 ? ct : st);
    } else {
        double XY = -(x * x + y * y);
        g = alpha * (exp(XY / (2 * p.sc2)) / p.sc2 - 0.9 * exp(XY / (2 * p.ss2)) / p.ss2);
        g = g / (2 * VT_PI);
    }
    out[(long long)j * ld + pix] = (T)g;
}

static int launch_rf(void *out, int dtype, long long ld, const RFParams &p, int N, int Px,
                     int Py, double Sx, double Sy, double x_start, double y_start,
                     hipStream_t s)
{
    VT_CHECK_ARG(out && p.alpha_inv && p.x0 && p.y0 && N >= 0 && Px > 0 && Py > 0);
    VT_CHECK_ARG(ld >= (long long)Px * Py);
    if (N == 0) return VT_OK;
    dim3 grid(vt_blocks((long long)Px * Py, 256), N);
    if (dtype == 0)
        rf_kernel<double><<<grid, 256, 0, s>>>((double *)out, ld, p, Px, Py, Sx, Sy, x_start, y_start);
    else
        rf_kernel<float><<<grid, 256, 0, s>>>((float *)out, ld, p, Px, Py, Sx, Sy, x_start, y_start);
    VT_LAUNCH_CHECK();
    return VT_OK;
}

VT_EXPORT int vt_rf_bank(void *out, int dtype, long long ld, int N, const int *kind,
                         const double *alpha_inv, const double *theta, const double *x0,
                         const double *y0, const int *ab, double kappa, double sigma_c2,
                         double sigma_s2, int Px, int Py, double Sx, double Sy, double x_start,
                         double y_start, void *stream)
{
    VT_CHECK_ARG(kind && theta && ab);
    RFParams p{kind, alpha_inv, theta, x0, y0, ab, kappa, sigma_c2, sigma_s2, 0};
    return launch_rf(out, dtype, ld, p, N, Px, Py, Sx, Sy, x_start, y_start, (hipStream_t)stream);
}

VT_EXPORT int vt_rf_gabor(void *out, int dtype, long long ld, int N, const double *alpha_inv,
                          const double *theta, const double *x0, const double *y0,
                          const int *ab, double kappa, int Px, int Py, double Sx, double Sy,
                          double x_start, double y_start, void *stream)
{
    VT_CHECK_ARG(theta && ab);
    RFParams p{nullptr, alpha_inv, theta, x0, y0, ab, kappa, 1.0, 1.0, 0};
    return launch_rf(out, dtype, ld, p, N, Px, Py, Sx, Sy, x_start, y_start, (hipStream_t)stream);
}

VT_EXPORT int vt_rf_dog(void *out, int dtype, long long ld, int N, const double *alpha_inv,
                        const double *x0, const double *y0, double sigma_c2, double sigma_s2,
                        int Px, int Py, double Sx, double Sy, double x_start, double y_start,
                        void *stream)
{
    VT_CHECK_ARG(sigma_c2 > 0 && sigma_s2 > 0);
    RFParams p{nullptr, alpha_inv, nullptr, x0, y0, nullptr, 0.0, sigma_c2, sigma_s2, 1};
    return launch_rf(out, dtype, ld, p, N, Px, Py, Sx, Sy, x_start, y_start, (hipStream_t)stream);
}

// ------------------------------------------------------------------------------------
// IAF encoder
// ------------------------------------------------------------------------------------

struct IAFState {
    double *spikes; long long spike_ld; int max_spike;
    double *v0; const double *kappa, *bias, *delta; double *time_count; int *spike_count;
    double dt; int root_mode;
};

// One sample interval [y1, y2] of the reference's loop (ensemble_encode.py:490-535).
// The spike loops stop once the buffer has overflowed (count > max_spike): the result is
// then invalid and the host raises, as the reference does (:163-164); this bounds every
// loop even for degenerate parameters.
__device__ __forceinline__ void iaf_step(double y1, double y2, double &v0, double &ddt,
                                         double &time_count, int &spike_count, double delta,
                                         double dt, int root_mode, double *__restrict__ col,
                                         long long spike_ld, int max_spike)
{
#pragma clang fp contract(off)
    if (y2 >= 0 || (y2 < 0 && y1 <= 0)) {
        double area = (y1 + y2) / 2 * ddt;
        while (spike_count <= max_spike && v0 + area >= delta) {
            double a = ((y2 - y1) / ddt);
            double dv = delta - v0;
            double remain;
            if (root_mode == 0)
                remain = (fabs(a) <= 1e-12) ? (dv / y1) : (sqrt(y1 * y1 + 2 * dv * a) - y1) / (a);
            else
                remain = 2 * dv / (sqrt(y1 * y1 + 2 * dv * a) + y1);
            time_count += remain;
            int row = spike_count < max_spike - 1 ? spike_count : max_spike - 1;
            col[(long long)row * spike_ld] = time_count;
            spike_count++;
            area -= (delta - v0);
            v0 = 0;
            ddt -= remain;
            y1 = y1 + remain * a;
        }
        v0 += area;
        time_count += ddt;
        ddt = dt;
    } else {
        double a = ((y2 - y1) / ddt);
        double disc;
        while (spike_count <= max_spike && (disc = y1 * y1 + 2 * (delta - v0) * a) >= 0) {
            double remain;
            if (root_mode == 0)
                remain = (sqrt(disc) - y1) / a;
            else
                remain = 2 * (delta - v0) / (sqrt(disc) + y1);
            time_count += remain;
            int row = spike_count < max_spike - 1 ? spike_count : max_spike - 1;
            col[(long long)row * spike_ld] = time_count;
            spike_count++;
            v0 = 0;
            ddt -= remain;
            y1 = y1 + remain * a;
            a = ((y2 - y1) / ddt);
        }
        v0 += (y1 + y2) / 2 * ddt;
        time_count += ddt;
        ddt = dt;
    }
}

template <typename TIn>
__global__ void __launch_bounds__(64)
iaf_kernel(const TIn *__restrict__ input, long long input_ld, int N, int size, IAFState st)
{
#pragma clang fp contract(off)
    const int n = blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= N) return;
    const double kappa = st.kappa[n], bias = st.bias[n], delta = st.delta[n];
    double v0 = st.v0[n], time_count = st.time_count[n], ddt = st.dt;
    int spike_count = 0;
    double *col = st.spikes + n;
    const TIn *in = input + n;
    double y_next = size > 0 ? ((double)in[0] + bias) / kappa : 0.0;
    for (int i = 0; i + 1 < size; ++i) {
        double y1 = y_next;
        double y2 = ((double)in[(long long)(i + 1) * input_ld] + bias) / kappa;
        y_next = y2;
        iaf_step(y1, y2, v0, ddt, time_count, spike_count, delta, st.dt, st.root_mode, col,
                 st.spike_ld, st.max_spike);
    }
    st.v0[n] = v0;
    st.time_count[n] = time_count;
    st.spike_count[n] = spike_count;
}

// Fused trigonometric upsampling: v_j(t_i) = sum_pt C[pt, j] * basis[i, pt] evaluated in
// registers, so the dt-rate input never touches HBM.  The basis table (size x Dt) is the
// same for all lanes and stays in L2.
template <int DT>
__global__ void __launch_bounds__(64)
iaf_trig_kernel(const double *__restrict__ coef, long long coef_ld, const double *__restrict__ basis,
                int Dt, int N, int size, IAFState st)
{
    const int n = blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= N) return;
    TrigSource<DT> eval;
    eval.load(coef, coef_ld, basis, Dt, n);
    const double kappa = st.kappa[n], bias = st.bias[n], delta = st.delta[n];
    double v0 = st.v0[n], time_count = st.time_count[n], ddt = st.dt;
    int spike_count = 0;
    double *col = st.spikes + n;
    double y_next = size > 0 ? (eval(0) + bias) / kappa : 0.0;
    for (int i = 0; i + 1 < size; ++i) {
        double y1 = y_next;
        double y2 = (eval(i + 1) + bias) / kappa;
        y_next = y2;
        iaf_step(y1, y2, v0, ddt, time_count, spike_count, delta, st.dt, st.root_mode, col,
                 st.spike_ld, st.max_spike);
    }
    st.v0[n] = v0;
    st.time_count[n] = time_count;
    st.spike_count[n] = spike_count;
}

static int check_state(const IAFState &st, int N)
{
    VT_CHECK_ARG(st.spikes && st.v0 && st.kappa && st.bias && st.delta && st.time_count &&
                 st.spike_count);
    VT_CHECK_ARG(st.max_spike >= 1 && st.spike_ld >= N && st.dt > 0);
    VT_CHECK_ARG(st.root_mode == 0 || st.root_mode == 1);
    return VT_OK;
}

VT_EXPORT int vt_iaf_encode(const void *input, int in_dtype, long long input_ld, int N, int size,
                            double *spikes, long long spike_ld, int max_spike, double *v0,
                            const double *kappa, const double *bias, const double *delta,
                            double *time_count, int *spike_count, double dt, int root_mode,
                            void *stream)
{
    IAFState st{spikes, spike_ld, max_spike, v0, kappa, bias, delta, time_count, spike_count, dt,
                root_mode};
    VT_CHECK_ARG(N >= 0 && size >= 0 && input_ld >= N && (input || size == 0));
    if (N == 0) return VT_OK;
    int r = check_state(st, N);
    if (r) return r;
    hipStream_t s = (hipStream_t)stream;
    if (in_dtype == 0)
        iaf_kernel<double><<<vt_blocks(N, 64), 64, 0, s>>>((const double *)input, input_ld, N, size, st);
    else
        iaf_kernel<float><<<vt_blocks(N, 64), 64, 0, s>>>((const float *)input, input_ld, N, size, st);
    VT_LAUNCH_CHECK();
    return VT_OK;
}

VT_EXPORT int vt_iaf_encode_trig(const double *coef, long long coef_ld, const double *basis,
                                 int Dt, int N, int size, double *spikes, long long spike_ld,
                                 int max_spike, double *v0, const double *kappa,
                                 const double *bias, const double *delta, double *time_count,
                                 int *spike_count, double dt, int root_mode, void *stream)
{
    IAFState st{spikes, spike_ld, max_spike, v0, kappa, bias, delta, time_count, spike_count, dt,
                root_mode};
    VT_CHECK_ARG(N >= 0 && size >= 0 && coef && basis && Dt >= 1 && coef_ld >= N);
    if (N == 0) return VT_OK;
    int r = check_state(st, N);
    if (r) return r;
    hipStream_t s = (hipStream_t)stream;
    dim3 g(vt_blocks(N, 64));
    if (Dt <= 9)
        iaf_trig_kernel<9><<<g, 64, 0, s>>>(coef, coef_ld, basis, Dt, N, size, st);
    else if (Dt <= 17)
        iaf_trig_kernel<17><<<g, 64, 0, s>>>(coef, coef_ld, basis, Dt, N, size, st);
    else if (Dt <= 33)
        iaf_trig_kernel<33><<<g, 64, 0, s>>>(coef, coef_ld, basis, Dt, N, size, st);
    else if (Dt <= 65)
        iaf_trig_kernel<65><<<g, 64, 0, s>>>(coef, coef_ld, basis, Dt, N, size, st);
    else
        return VT_ERANGE;
    VT_LAUNCH_CHECK();
    return VT_OK;
}

// Random-threshold IAF (ensemble_encode.py:606-668).  `spikes` holds N(0,1) draws on entry.
__global__ void __launch_bounds__(64)
iaf_rt_kernel(const double *__restrict__ input, long long input_ld, int N, int size, IAFState st,
              double *delta_value, const double *sigma)
{
#pragma clang fp contract(off)
    const int n = blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= N) return;
    const double kappa = st.kappa[n], bias = st.bias[n], delta_mean = st.delta[n], sg = sigma[n];
    double delta = delta_value[n];
    double v0 = st.v0[n], time_count = st.time_count[n], ddt = st.dt;
    const double dt = st.dt;
    const int max_spike = st.max_spike;
    int spike_count = 0;
    double *col = st.spikes + n;
    const long long ld = st.spike_ld;
    for (int i = 0; i + 1 < size; ++i) {
        double y1 = (input[(long long)i * input_ld + n] + bias) / kappa;
        double y2 = (input[(long long)(i + 1) * input_ld + n] + bias) / kappa;
        if (y2 >= 0 || (y2 < 0 && y1 <= 0)) {
            double area = (y1 + y2) / 2 * ddt;
            while (spike_count <= max_spike && v0 + area >= delta) {
                double a = ((y2 - y1) / ddt);
                double dv = delta - v0;
                double remain;
                if (st.root_mode == 0)
                    remain = (fabs(a) <= 1e-12) ? (dv / y1) : (sqrt(y1 * y1 + 2 * dv * a) - y1) / (a);
                else
                    remain = 2 * dv / (sqrt(y1 * y1 + 2 * dv * a) + y1);
                time_count += remain;
                int row = spike_count < max_spike - 1 ? spike_count : max_spike - 1;
                delta = col[(long long)row * ld] * sg + delta_mean;
                col[(long long)row * ld] = time_count;
                spike_count++;
                area -= (delta - v0);
                v0 = 0;
                ddt -= remain;
                y1 = y1 + remain * a;
            }
            v0 += area;
            time_count += ddt;
            ddt = dt;
        } else {
            double a = ((y2 - y1) / ddt);
            double disc;
            while (spike_count <= max_spike && (disc = y1 * y1 + 2 * (delta - v0) * a) >= 0) {
                double remain;
                if (st.root_mode == 0)
                    remain = (sqrt(disc) - y1) / a;
                else
                    remain = 2 * (delta - v0) / (sqrt(disc) + y1);
                time_count += remain;
                int row = spike_count < max_spike - 1 ? spike_count : max_spike - 1;
                delta = col[(long long)row * ld] * sg + delta_mean;
                col[(long long)row * ld] = time_count;
                spike_count++;
                v0 = 0;
                ddt -= remain;
                y1 = y1 + remain * a;
                a = ((y2 - y1) / ddt);
            }
            v0 += (y1 + y2) / 2 * ddt;
            time_count += ddt;
            ddt = dt;
        }
    }
    st.v0[n] = v0;
    st.time_count[n] = time_count;
    st.spike_count[n] = spike_count;
    delta_value[n] = delta;
}

VT_EXPORT int vt_iaf_encode_rt(const double *input, long long input_ld, int N, int size,
                               double *spikes, long long spike_ld, int max_spike, double *v0,
                               const double *kappa, const double *bias, const double *delta_mean,
                               double *time_count, int *spike_count, double dt,
                               double *delta_value, const double *sigma, int root_mode,
                               void *stream)
{
    IAFState st{spikes, spike_ld, max_spike, v0, kappa, bias, delta_mean, time_count,
                spike_count, dt, root_mode};
    VT_CHECK_ARG(N >= 0 && size >= 0 && input_ld >= N && (input || size == 0));
    VT_CHECK_ARG(delta_value && sigma);
    if (N == 0) return VT_OK;
    int r = check_state(st, N);
    if (r) return r;
    iaf_rt_kernel<<<vt_blocks(N, 64), 64, 0, (hipStream_t)stream>>>(input, input_ld, N, size, st,
                                                                   delta_value, sigma);
    VT_LAUNCH_CHECK();
    return VT_OK;
}

// ------------------------------------------------------------------------------------
// Interval extraction (vtem/utils/spike_store.py:279-351), on device.
// Spike buffer as the encoder writes it: (max_spike x N), column j sorted ascending.
// For a window [t0, t0+Tw] a neuron keeps the spikes in (0, Tw) after shifting by t0,
// plus one before and one after; consecutive pairs are its intervals.
// ------------------------------------------------------------------------------------

__global__ void interval_count_kernel(const double *__restrict__ spikes, long long ld,
                                      const int *__restrict__ counts, int N, int max_rows,
                                      double t0, double Tw, int *__restrict__ lo_out,
                                      int *__restrict__ cnt_out)
{
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= N) return;
    const int c = min(counts[j], max_rows);   // overflowed buffers hold max_rows valid rows
    int first = -1, last = -1;
    for (int i = 0; i < c; ++i) {
        double r = spikes[(long long)i * ld + j] - t0;
        if (r > 0 && r < Tw) {
            if (first < 0) first = i;
            last = i;
        }
    }
    int lo = 0, cnt = 0;
    if (first >= 0) {
        lo = first > 0 ? first - 1 : 0;
        int hi = (last + 2 < c) ? last + 2 : c;
        cnt = hi - lo - 1;
        if (cnt < 0) cnt = 0;
    }
    lo_out[j] = lo;
    cnt_out[j] = cnt;
}

// Exclusive scan of n ints with one workgroup (offsets has n+1 entries; offsets[n] = total).
__global__ void __launch_bounds__(1024) scan_kernel(const int *__restrict__ in, int n,
                                                   int *__restrict__ offsets)
{
    __shared__ long long part[1024];
    const int t = threadIdx.x;
    const int chunk = (n + 1023) / 1024;
    const int b = t * chunk, e = min(n, b + chunk);
    long long s = 0;
    for (int i = b; i < e; ++i) s += in[i];
    part[t] = s;
    __syncthreads();
    for (int off = 1; off < 1024; off <<= 1) {
        long long v = t >= off ? part[t - off] : 0;
        __syncthreads();
        part[t] += v;
        __syncthreads();
    }
    long long run = t > 0 ? part[t - 1] : 0;
    for (int i = b; i < e; ++i) {
        offsets[i] = (int)run;
        run += in[i];
    }
    if (t == 1023) offsets[n] = (int)part[1023];
}

__global__ void interval_write_kernel(const double *__restrict__ spikes, long long ld, int N,
                                      double t0, const int *__restrict__ lo,
                                      const int *__restrict__ offsets, double *__restrict__ tk1,
                                      double *__restrict__ tk2, int *__restrict__ nid, int cap)
{
    const int j = blockIdx.x;
    const int base = offsets[j], cnt = offsets[j + 1] - base, l = lo[j];
    for (int k = threadIdx.x; k < cnt; k += blockDim.x) {
        int o = base + k;
        if (o >= cap) break;
        tk1[o] = spikes[(long long)(l + k) * ld + j] - t0;
        tk2[o] = spikes[(long long)(l + k + 1) * ld + j] - t0;
        nid[o] = j;
    }
}

// workspace: 2*N ints.  offsets: N+1 ints (offsets[N] = number of intervals, on device).
VT_EXPORT int vt_intervals(const double *spikes, long long spike_ld, int max_rows,
                           const int *spike_count, int N, double t0, double Tw, int *workspace,
                           int *offsets, double *tk1, double *tk2, int *neuron_ind, int capacity,
                           void *stream)
{
    VT_CHECK_ARG(N >= 1 && spikes && spike_count && workspace && offsets && tk1 && tk2 &&
                 neuron_ind && spike_ld >= N && Tw > 0 && capacity >= 0 && max_rows >= 0);
    hipStream_t s = (hipStream_t)stream;
    int *lo = workspace, *cnt = workspace + N;
    interval_count_kernel<<<vt_blocks(N, 256), 256, 0, s>>>(spikes, spike_ld, spike_count, N,
                                                          max_rows, t0, Tw, lo, cnt);
    scan_kernel<<<1, 1024, 0, s>>>(cnt, N, offsets);
    interval_write_kernel<<<N, 64, 0, s>>>(spikes, spike_ld, N, t0, lo, offsets, tk1, tk2,
                                           neuron_ind, capacity);
    VT_LAUNCH_CHECK();
    return VT_OK;
}

VT_EXPORT int vt_exclusive_scan(const int *in, int n, int *offsets, void *stream)
{
    VT_CHECK_ARG(n >= 0 && offsets && (in || n == 0));
    scan_kernel<<<1, 1024, 0, (hipStream_t)stream>>>(in, n, offsets);
    VT_LAUNCH_CHECK();
    return VT_OK;
}
