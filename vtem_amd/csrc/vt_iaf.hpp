// Input sources of the IAF kernels: the neuron's filtered input v_j at sample i.
// The sequential and the time-parallel encoders evaluate the same source code, so both
// integrate bit-identical samples.
#pragma once
#include "vt_common.hpp"

#ifndef VT_TRIG_CHUNK
#define VT_TRIG_CHUNK 4    // basis doubles per scalar load (32 bytes)
#endif

// Trigonometric upsampling fused into the encoder: v_j(t_i) = sum_pt C[pt, j] * basis[i, pt]
// with the neuron's Dt coefficients in registers.  The basis table (size x Dt) is shared by
// all lanes of a wave (uniform row address) and stays in L2.  DT bounds Dt.
template <int DT>
struct TrigSource {
    double c[DT];
    const double *basis;
    int Dt;
    // coefficient pt of neuron n at coef[pt * coef_ld + n * coef_ns]
    __device__ __forceinline__ void load(const double *coef, long long coef_ld,
                                         const double *basis_, int Dt_, int n,
                                         long long coef_ns = 1)
    {
        basis = basis_;
        Dt = Dt_;
#pragma unroll
        for (int p = 0; p < DT; ++p)
            c[p] = p < Dt ? coef[(long long)p * coef_ld + (long long)n * coef_ns] : 0.0;
    }
    // v_j at the R samples i .. i+R-1 (rows i .. i+R-1 of the table).  The table is
    // read-only for the kernel's lifetime: read it through the constant address space, so
    // a wave-uniform row becomes scalar loads into SGPRs instead of one vector load per
    // lane.  The rows are read in chunks of VT_TRIG_CHUNK doubles under one uniform branch
    // per chunk (one scalar load per row and one wait per chunk, the R rows' loads in
    // flight together; a branch per term would serialise every load); only the last,
    // partial chunk tests its terms one by one.  Every sample sums its terms in the order
    // p = 0 .. Dt-1, whatever R.
    template <int R>
    __device__ __forceinline__ void eval(int i, double *out) const
    {
        typedef const __attribute__((address_space(4))) double *const_row;
        const_row b[R];
        double s[R];
#pragma unroll
        for (int r = 0; r < R; ++r) {
            b[r] = (const_row)(basis + (long long)(i + r) * Dt);
            s[r] = 0.0;
        }
        constexpr int CH = VT_TRIG_CHUNK;
#pragma unroll
        for (int k = 0; k < (DT + CH - 1) / CH; ++k) {
            if (CH * k + CH <= Dt) {
#pragma unroll
                for (int u = 0; u < CH; ++u)
#pragma unroll
                    for (int r = 0; r < R; ++r)
                        if (CH * k + u < DT) s[r] += c[CH * k + u] * b[r][CH * k + u];
            } else {
#pragma unroll
                for (int u = 0; u < CH; ++u)
#pragma unroll
                    for (int r = 0; r < R; ++r)
                        if (CH * k + u < DT && CH * k + u < Dt) s[r] += c[CH * k + u] * b[r][CH * k + u];
                break;
            }
        }
#pragma unroll
        for (int r = 0; r < R; ++r) out[r] = s[r];
    }
    __device__ __forceinline__ double operator()(int i) const
    {
        double v;
        eval<1>(i, &v);
        return v;
    }
};

// Sampled input (T x N), neuron-contiguous, as the reference's encoder reads it
// (vtem/ensemble_encode.py:146, :484-488).
template <typename TIn>
struct SampledSource {
    const TIn *in;
    long long ld;
    __device__ __forceinline__ double operator()(int i) const
    {
        return (double)in[(long long)i * ld];
    }
    template <int R>
    __device__ __forceinline__ void eval(int i, double *out) const
    {
#pragma unroll
        for (int r = 0; r < R; ++r) out[r] = (double)in[(long long)(i + r) * ld];
    }
};
