// Skinny grouped product: descriptor layout and split rule shared by the product kernels
// and the solver loops that fuse its second pass.
#pragma once
#include "vt_common.hpp"

#ifndef VT_SKINNY_CHUNKS_PER_SPLIT
#define VT_SKINNY_CHUNKS_PER_SPLIT 32
#endif

struct SkDesc {
    const double *A, *B;
    double *C;
    const double *T;
    double *ws;                 // ks * M * Nc doubles when ks > 1
    const double *active;       // optional: the problem is skipped while *active == 0
    long long sam, sak, ldb, ldc, tstride;
    int M, K, Nc, pad;
};

// K splits of a problem: one per VT_SKINNY_CHUNKS_PER_SPLIT 16-wide chunks, at most 8 --
// a function of the problem's own K, so its rounding does not depend on the batch.
__host__ __device__ __forceinline__ int skinny_splits(int K)
{
    const int nch = (K + 15) / 16;
    const int ks = nch / VT_SKINNY_CHUNKS_PER_SPLIT;
    return ks < 1 ? 1 : (ks > 8 ? 8 : ks);
}

// First pass only (library-internal): problems with ks > 1 leave their ks partial tiles in
// ws (partial z of row m, column c at ws[(z * M + m) * Nc + c]) for a fused second pass.
int vt_skinny_first_pass(const void *descs, int batch, int max_M, int max_K, int max_Nc,
                         hipStream_t stream);
