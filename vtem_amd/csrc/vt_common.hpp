// Shared definitions for the vtem HIP kernels (gfx950 / CDNA4, wave64).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#define VT_EXPORT extern "C" __attribute__((visibility("default")))

// Status convention of the C-ABI: 0 = OK, > 0 = hipError_t, < 0 = argument error.
enum {
    VT_OK = 0,
    VT_EARG = -1,        // invalid argument (size, null pointer, stride)
    VT_EUNSUPPORTED = -2,
    VT_EOVERFLOW = -3,   // spike buffer overflow (reported, not raised, by the kernel)
    VT_ENOTPD = -4,      // Cholesky met a non-positive pivot
    VT_ERANGE = -5,      // size beyond a documented limit
};

#define VT_CHECK_ARG(cond)          \
    do {                            \
        if (!(cond)) return VT_EARG; \
    } while (0)

#define VT_LAUNCH_CHECK()                          \
    do {                                           \
        hipError_t e__ = hipGetLastError();        \
        if (e__ != hipSuccess) return (int)e__;    \
    } while (0)

#define VT_HIP(call)                               \
    do {                                           \
        hipError_t e__ = (call);                   \
        if (e__ != hipSuccess) return (int)e__;    \
    } while (0)

// return a failing library status from the enclosing entry point
#define VT_TRY(call)                    \
    do {                                \
        int st__ = (call);              \
        if (st__ != VT_OK) return st__; \
    } while (0)

typedef double v4d __attribute__((ext_vector_type(4)));
typedef float v4f __attribute__((ext_vector_type(4)));
typedef float v16f __attribute__((ext_vector_type(16)));
typedef short v8s __attribute__((ext_vector_type(8)));
typedef __bf16 v8bf __attribute__((ext_vector_type(8)));
typedef _Float16 v8h __attribute__((ext_vector_type(8)));

static inline unsigned vt_blocks(long long n, int threads) {
    long long b = (n + threads - 1) / threads;
    return (unsigned)(b < 1 ? 1 : b);
}

// MFMA accumulator layouts for the 16x16 tiles used here, measured on gfx950 with a
// lane-layout probe (vt_mfma_probe):
//   fp32 16x16x4, bf16/f16 16x16x32: lane l, register i holds C[4*(l/16) + i][l % 16];
//   fp64 16x16x4:                    lane l, register i holds C[(l/16) + 4*i][l % 16].
// A operand: lane l supplies row l % 16 (k slice l/16); B operand: column l % 16.
__device__ __forceinline__ int mfma_row(int lane, int i) { return 4 * (lane >> 4) + i; }
__device__ __forceinline__ int mfma_row_f64(int lane, int i) { return (lane >> 4) + 4 * i; }
__device__ __forceinline__ int mfma_col(int lane) { return lane & 15; }
template <typename T> __device__ __forceinline__ int acc_row(int lane, int i);
template <> __device__ __forceinline__ int acc_row<double>(int lane, int i) { return mfma_row_f64(lane, i); }
template <> __device__ __forceinline__ int acc_row<float>(int lane, int i) { return mfma_row(lane, i); }
