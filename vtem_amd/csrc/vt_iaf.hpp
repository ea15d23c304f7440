// Input sources of the IAF kernels: the neuron's filtered input v_j at sample i.
// The sequential and the time-parallel encoders evaluate the same source code, so both
// integrate bit-identical samples.
#pragma once
#include "vt_common.hpp"

// Trigonometric upsampling fused into the encoder: v_j(t_i) = sum_pt C[pt, j] * basis[i, pt]
// with the neuron's Dt coefficients in registers.  The basis table (size x Dt) is shared by
// all lanes of a wave (uniform row address) and stays in L2.
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
    __device__ __forceinline__ double operator()(int i) const
    {
        // The table is read-only for the kernel's lifetime: read it through the constant
        // address space, so a wave-uniform row becomes scalar loads into SGPRs instead of
        // one vector load per lane.
        typedef const __attribute__((address_space(4))) double *const_row;
        const_row b = (const_row)(basis + (long long)i * Dt);
        double s = 0.0;
#pragma unroll
        for (int p = 0; p < DT; ++p)
            if (p < Dt) s += c[p] * b[p];
        return s;
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
};
