/*
 * TEST INFRASTRUCTURE ONLY -- the CPU oracle for the IAF population encoder.
 * Nothing in the product path links or calls this file; tests, smoke() and bench.py's
 * cpu_baseline leg load it through ctypes as the checker.
 *
 * Plain-C restatement of the reference's linear IAF kernel
 * (vtem/ensemble_encode.py:438-552, host wrapper :106-168) and of its random-threshold
 * variant (vtem/ensemble_encode.py:555-674, host wrapper :265-329).  One neuron at a
 * time, samples walked in order, the same branch structure, the same crossing formula
 * and the same operation order as the reference, so that with -ffp-contract=off every
 * operation is a single IEEE-754 double operation.
 *
 * Layout as the reference: input (size x num_neurons), row stride input_ld;
 * spikes (max_spike x num_neurons), row stride spike_ld, spike-major and
 * neuron-contiguous; overflowing spikes are written to row max_spike-1 (the reference's
 * min(max_spike-1, count) clamp) and the caller raises.
 *
 * root_mode 0: the reference's root (sqrt(y1^2 + 2(d-v0)a) - y1)/a with the linear
 *              fallback (d-v0)/y1 when |a| <= 1e-12 (ensemble_encode.py:501, :521);
 * root_mode 1: the algebraically identical cancellation-free form
 *              2(d-v0)/(sqrt(y1^2 + 2(d-v0)a) + y1).
 */
#include <math.h>
#include <stdint.h>

static inline double crossing(double y1, double a, double dv, int root_mode)
{
    if (root_mode == 0)
        return (fabs(a) <= 1e-12) ? (dv / y1) : (sqrt(y1 * y1 + 2 * dv * a) - y1) / (a);
    return 2 * dv / (sqrt(y1 * y1 + 2 * dv * a) + y1);
}

static inline double crossing_falling(double disc, double y1, double a, double dv, int root_mode)
{
    if (root_mode == 0)
        return (sqrt(disc) - y1) / a;
    return 2 * dv / (sqrt(disc) + y1);
}

/* Deterministic-threshold IAF (ensemble_encode.py:486-545). */
int oracle_iaf_encode(const double *input, int64_t input_ld, int num_neurons, int size,
                      double *spikes, int64_t spike_ld, double *g_v0, const double *g_kappa,
                      const double *g_bias, const double *g_delta, double *g_time_count,
                      int *g_spike_count, int max_spike, double dt, int root_mode)
{
    for (int n = 0; n < num_neurons; ++n) {
        double kappa = g_kappa[n], bias = g_bias[n], delta = g_delta[n];
        double ddt = dt, v0 = g_v0[n], time_count = g_time_count[n];
        int spike_count = 0;
        for (int i = 0; i + 1 < size; ++i) {
            double y1 = (input[(int64_t)i * input_ld + n] + bias) / kappa;
            double y2 = (input[(int64_t)(i + 1) * input_ld + n] + bias) / kappa;
            if (y2 >= 0 || (y2 < 0 && y1 <= 0)) {
                double area = (y1 + y2) / 2 * ddt;
                while (v0 + area >= delta) {
                    double a = ((y2 - y1) / ddt);
                    double remain = crossing(y1, a, delta - v0, root_mode);
                    time_count += remain;
                    int row = spike_count < max_spike - 1 ? spike_count : max_spike - 1;
                    spikes[(int64_t)row * spike_ld + n] = time_count;
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
                while ((disc = y1 * y1 + 2 * (delta - v0) * a) >= 0) {
                    double remain = crossing_falling(disc, y1, a, delta - v0, root_mode);
                    time_count += remain;
                    int row = spike_count < max_spike - 1 ? spike_count : max_spike - 1;
                    spikes[(int64_t)row * spike_ld + n] = time_count;
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
        g_v0[n] = v0;
        g_time_count[n] = time_count;
        g_spike_count[n] = spike_count;
    }
    return 0;
}

/*
 * Random-threshold IAF (ensemble_encode.py:606-668).  `spikes` holds N(0,1) draws on
 * entry (the reference uploads host randn into the spike buffer, :304-306); after each
 * spike the next threshold is randn*sigma + delta_mean read from the slot before that
 * slot is overwritten with the spike time (:623-624, :645-646).  The reference subtracts
 * the NEW threshold from the area in the rising branch (:627); that is restated as is.
 */
int oracle_iaf_encode_rt(const double *input, int64_t input_ld, int num_neurons, int size,
                         double *spikes, int64_t spike_ld, double *g_v0,
                         const double *g_kappa, const double *g_bias,
                         const double *g_delta_mean, double *g_time_count,
                         int *g_spike_count, int max_spike, double dt,
                         double *g_delta_value, const double *g_sigma, int root_mode)
{
    for (int n = 0; n < num_neurons; ++n) {
        double kappa = g_kappa[n], bias = g_bias[n];
        double ddt = dt, v0 = g_v0[n], time_count = g_time_count[n];
        double delta_mean = g_delta_mean[n], delta = g_delta_value[n], sigma = g_sigma[n];
        int spike_count = 0;
        for (int i = 0; i + 1 < size; ++i) {
            double y1 = (input[(int64_t)i * input_ld + n] + bias) / kappa;
            double y2 = (input[(int64_t)(i + 1) * input_ld + n] + bias) / kappa;
            if (y2 >= 0 || (y2 < 0 && y1 <= 0)) {
                double area = (y1 + y2) / 2 * ddt;
                while (v0 + area >= delta) {
                    double a = ((y2 - y1) / ddt);
                    double remain = crossing(y1, a, delta - v0, root_mode);
                    time_count += remain;
                    int row = spike_count < max_spike - 1 ? spike_count : max_spike - 1;
                    delta = spikes[(int64_t)row * spike_ld + n] * sigma + delta_mean;
                    spikes[(int64_t)row * spike_ld + n] = time_count;
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
                while ((disc = y1 * y1 + 2 * (delta - v0) * a) >= 0) {
                    double remain = crossing_falling(disc, y1, a, delta - v0, root_mode);
                    time_count += remain;
                    int row = spike_count < max_spike - 1 ? spike_count : max_spike - 1;
                    delta = spikes[(int64_t)row * spike_ld + n] * sigma + delta_mean;
                    spikes[(int64_t)row * spike_ld + n] = time_count;
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
        g_v0[n] = v0;
        g_time_count[n] = time_count;
        g_spike_count[n] = spike_count;
        g_delta_value[n] = delta;
    }
    return 0;
}
