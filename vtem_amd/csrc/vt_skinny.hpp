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

// Output store of a product tile.  COH: a relaxed agent-scope atomic store (written through
// to the device coherence point), for tiles that other workgroups of a persistent kernel
// read back within the same launch; the value stored is the same.
template <bool COH>
__device__ __forceinline__ void sk_store(double *p, double v)
{
    if constexpr (COH)
        __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    else
        *p = v;
}

// One gradient-flow step of an element, x + eta (r - q - lam x): the launch-per-phase and
// the persistent RNN loops share it, so both round alike.
__device__ __forceinline__ double rnn_step(double x, double r, double q, double lam, double eta)
{
    return x + eta * (r - q - lam * x);
}

#ifndef SK_WAVES
#define SK_WAVES 8
#endif
#ifndef SK_INFLIGHT
#define SK_INFLIGHT 2      // chunks each wave has in flight
#endif

// SkDesc.pad bits.  SK_A_TILED: A is stored in MFMA tiles (sk_tile_index), sam = the tiles
// per row of tiles (ceil(K / 16)); rows and columns beyond M and K are zeros.  SK_B_TILED: B
// is stored as the tiles of its transpose (row n of the tiles is column n of B), ldb = the
// tiles per row (ceil(K / 16)), with two rows of tiles when Nc > 16.  SK_C_TILED: C is
// stored the same way (the B layout of a product that takes C as its B), ldc = ceil(M / 16);
// only its entries are written, so the caller zeroes the padding once.
#define SK_A_TILED 1
#define SK_B_TILED 2
#define SK_C_TILED 4

// Position of A[m][k] in the tiled layout: tile (m / 16, k / 16) holds 256 doubles, and lane
// L = r + 16 g of a wave (row r, k group g) finds its four values of the tile (k = 4g + s)
// as two pairs, s = 0, 1 at 2L and s = 2, 3 at 128 + 2L: each of a wave's two 16-byte loads
// reads 1 KB in one piece.
__host__ __device__ __forceinline__ long long sk_tile_index(long long m, long long k, long long tiles_per_row)
{
    const long long r = m & 15, kk = k & 15, g = kk >> 2, s = kk & 3;
    return ((m >> 4) * tiles_per_row + (k >> 4)) * 256 + (s >> 1) * 128 + (16 * g + r) * 2 + (s & 1);
}

// offset of C[m][col] in the problem's C layout
__device__ __forceinline__ long long sk_c_offset(const SkDesc &d, int m, int col)
{
    return (d.pad & SK_C_TILED) ? sk_tile_index(col, m, d.ldc) : (long long)m * d.ldc + col;
}

// operands of one 16-wide K chunk for lane (r, g); chunks at or beyond c_end load zeros
template <int NT, int RT>
__device__ __forceinline__ void skinny_load(const SkDesc &d, int m0, int c, int c_end, int r, int g,
                                            int Nc, double (&a)[RT][4], double (&bv)[NT][4])
{
    const int kb = c * 16 + 4 * g;
    const bool cin = c < c_end;
    typedef double d2a16 __attribute__((ext_vector_type(2)));
    if (d.pad & SK_A_TILED) {
#pragma unroll
        for (int t = 0; t < RT; ++t) {
            const int mt = (m0 >> 4) + t;
            if (cin && 16 * mt < d.M) {
                const double *p = d.A + ((long long)mt * d.sam + c) * 256 + 2 * (16 * g + r);
                const d2a16 x = *(const d2a16 *)p, y = *(const d2a16 *)(p + 128);
                a[t][0] = x[0];
                a[t][1] = x[1];
                a[t][2] = y[0];
                a[t][3] = y[1];
            } else {
                a[t][0] = a[t][1] = a[t][2] = a[t][3] = 0.0;
            }
        }
    } else {
#pragma unroll
        for (int t = 0; t < RT; ++t) {
            const int m = m0 + 16 * t + r;
            if (d.sak == 1 && cin && m < d.M && kb + 3 < d.K) {
                // row-major A: the lane's 4 consecutive k values as two 16-byte loads
                typedef double d2a8 __attribute__((ext_vector_type(2), aligned(8)));
                const double *p = d.A + (long long)m * d.sam + kb;
                const d2a8 x = *(const d2a8 *)p, y = *(const d2a8 *)(p + 2);
                a[t][0] = x[0];
                a[t][1] = x[1];
                a[t][2] = y[0];
                a[t][3] = y[1];
                continue;
            }
#pragma unroll
            for (int s = 0; s < 4; ++s) {
                const int k = kb + s;
                a[t][s] = (cin && m < d.M && k < d.K) ? d.A[(long long)m * d.sam + (long long)k * d.sak] : 0.0;
            }
        }
    }
    if (d.pad & SK_B_TILED) {
#pragma unroll
        for (int j = 0; j < NT; ++j) {
            if (cin && 16 * j < Nc) {
                const double *p = d.B + ((long long)j * d.ldb + c) * 256 + 2 * (16 * g + r);
                const d2a16 x = *(const d2a16 *)p, y = *(const d2a16 *)(p + 128);
                bv[j][0] = x[0];
                bv[j][1] = x[1];
                bv[j][2] = y[0];
                bv[j][3] = y[1];
            } else {
                bv[j][0] = bv[j][1] = bv[j][2] = bv[j][3] = 0.0;
            }
        }
        return;
    }
#pragma unroll
    for (int j = 0; j < NT; ++j) {
        const int n = 16 * j + r;
#pragma unroll
        for (int s = 0; s < 4; ++s) {
            const int k = kb + s;
            bv[j][s] = (cin && n < Nc && k < d.K) ? d.B[(long long)k * d.ldb + n] : 0.0;
        }
    }
}

template <int NT, int RT>
__device__ __forceinline__ void skinny_mma(v4d (&acc)[RT][NT], const double (&a)[RT][4],
                                           const double (&bv)[NT][4])
{
#pragma unroll
    for (int s = 0; s < 4; ++s)
#pragma unroll
        for (int t = 0; t < RT; ++t)
#pragma unroll
            for (int j = 0; j < NT; ++j)
                acc[t][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[t][s], bv[j][s], acc[t][j], 0, 0, 0);
}

__device__ __forceinline__ double skinny_row_factor(const SkDesc &d, int m, int col, int Nc,
                                                    const double *row)
{
    const double *Tr = d.T + (long long)m * d.tstride + (long long)col * Nc;
    double v = 0.0;
    for (int k = 0; k < Nc; ++k) v += Tr[k] * row[k];
    return v;
}

// Column 16 of a problem with Nc = 17 (the temporal order of Lt = 8 blocks) is not given a
// second 16-column MFMA tile, of which it would use one column: each lane adds its four
// products a[s] * B[k][16] of the chunk with FMAs (k = c*16 + 4g + s in order s = 0..3,
// chunks in the wave's order), the four lane groups g are added as (g0 + g1) + (g2 + g3),
// and the waves' sums join the MFMA columns in the tree reduction.
__device__ __forceinline__ void skinny_load_x16(const SkDesc &d, int c, int c_end, int g, double (&x)[4])
{
    const int kb = c * 16 + 4 * g;
    const bool cin = c < c_end;
    if (d.pad & SK_B_TILED) {
        // column 16 is row 0 of the second row of tiles: lane group g's pairs at 2 (16 g)
        typedef double d2a16 __attribute__((ext_vector_type(2)));
        if (cin && d.Nc > 16) {
            const double *p = d.B + ((long long)d.ldb + c) * 256 + 32 * g;
            const d2a16 u = *(const d2a16 *)p, v = *(const d2a16 *)(p + 128);
            x[0] = u[0];
            x[1] = u[1];
            x[2] = v[0];
            x[3] = v[1];
        } else {
            x[0] = x[1] = x[2] = x[3] = 0.0;
        }
        return;
    }
#pragma unroll
    for (int s = 0; s < 4; ++s) {
        const int k = kb + s;
        x[s] = (cin && k < d.K && d.Nc > 16) ? d.B[(long long)k * d.ldb + 16] : 0.0;
    }
}

// One workgroup of SK_WAVES waves computes rows m0 .. m0 + 16 RT - 1 of K split z of
// problem d (RT row tiles of 16: the row tiling does not change any sum's order, so a
// launch may pick it from the batch size).  The waves split K and are reduced in LDS in a
// fixed tree order; red is the workgroup's LDS scratch.  X16: the problem has Nc = 17 and
// its column 16 is summed on the vector ALU (NT = 1, RT = 1); the route depends only on the
// problem's own Nc, so a problem's sums do not depend on the batch.  COH: the outputs are
// stored agent-coherent (sk_store).
template <int NT, int RT, bool X16 = false, bool COH = false>
__device__ __forceinline__ void skinny_tile(const SkDesc &d, int m0, int z,
                                            double (&red)[SK_WAVES / 2][16 * RT][33])
{
    static_assert(!X16 || (NT == 1 && RT == 1), "the vector column extends one 16-column tile");
    constexpr int R = 16 * RT;
    const int ks = skinny_splits(d.K), Nc = d.Nc;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int r = lane & 15, g = lane >> 4;
    v4d acc[RT][NT];
#pragma unroll
    for (int t = 0; t < RT; ++t)
#pragma unroll
        for (int j = 0; j < NT; ++j) acc[t][j] = (v4d){0, 0, 0, 0};
    const int nch = (d.K + 15) / 16;
    const int per = (nch + ks - 1) / ks;
    const int c_beg = z * per, c_end = min(nch, c_beg + per);
    double p16 = 0.0;                      // X16: this lane's part of column 16
    // SK_INFLIGHT chunks in flight per wave: c, c + SK_WAVES, ... (added in that order)
    for (int c = c_beg + wave; c < c_end; c += SK_INFLIGHT * SK_WAVES) {
        double a[SK_INFLIGHT][RT][4], b[SK_INFLIGHT][NT][4], x[SK_INFLIGHT][4];
#pragma unroll
        for (int u = 0; u < SK_INFLIGHT; ++u) {
            skinny_load<NT, RT>(d, m0, c + u * SK_WAVES, c_end, r, g, Nc, a[u], b[u]);
            if constexpr (X16) skinny_load_x16(d, c + u * SK_WAVES, c_end, g, x[u]);
        }
#pragma unroll
        for (int u = 0; u < SK_INFLIGHT; ++u) {
            skinny_mma<NT, RT>(acc, a[u], b[u]);
            if constexpr (X16) {
#pragma unroll
                for (int s = 0; s < 4; ++s) p16 = __builtin_fma(a[u][0][s], x[u][s], p16);
            }
        }
    }
    if constexpr (X16) {
        p16 += __shfl_xor(p16, 16);        // (g0 + g1), (g2 + g3) in every lane
        p16 += __shfl_xor(p16, 32);        // (g0 + g1) + (g2 + g3)
    }
    // cross-wave tree reduction: at each level the upper half adds into the lower half
#pragma unroll
    for (int half = SK_WAVES / 2; half >= 1; half >>= 1) {
        if (wave >= half && wave < 2 * half) {
#pragma unroll
            for (int t = 0; t < RT; ++t)
#pragma unroll
                for (int j = 0; j < NT; ++j)
#pragma unroll
                    for (int i = 0; i < 4; ++i)
                        red[wave - half][16 * t + mfma_row_f64(lane, i)][16 * j + mfma_col(lane)] = acc[t][j][i];
            if constexpr (X16) {
                if (g == 0) red[wave - half][r][16] = p16;
            }
        }
        __syncthreads();
        if (wave < half) {
#pragma unroll
            for (int t = 0; t < RT; ++t)
#pragma unroll
                for (int j = 0; j < NT; ++j)
#pragma unroll
                    for (int i = 0; i < 4; ++i)
                        acc[t][j][i] += red[wave][16 * t + mfma_row_f64(lane, i)][16 * j + mfma_col(lane)];
            if constexpr (X16) p16 += red[wave][r][16];
        }
        __syncthreads();
    }
    if (wave == 0) {
#pragma unroll
        for (int t = 0; t < RT; ++t)
#pragma unroll
            for (int j = 0; j < NT; ++j)
#pragma unroll
                for (int i = 0; i < 4; ++i)
                    red[0][16 * t + mfma_row_f64(lane, i)][16 * j + mfma_col(lane)] = acc[t][j][i];
        if constexpr (X16) {
            if (g == 0) red[0][r][16] = p16;
        }
    }
    __syncthreads();
    const int rows = min(R, d.M - m0);
    if (ks > 1) {
        double *W = d.ws + (long long)z * d.M * Nc;
        for (int e = tid; e < rows * Nc; e += 64 * SK_WAVES) {
            const int row = e / Nc, col = e % Nc;
            sk_store<COH>(W + (long long)(m0 + row) * Nc + col, red[0][row][col]);
        }
        return;
    }
    for (int e = tid; e < rows * Nc; e += 64 * SK_WAVES) {
        const int row = e / Nc, col = e % Nc;
        const double v = d.T ? skinny_row_factor(d, m0 + row, col, Nc, red[0][row]) : red[0][row][col];
        sk_store<COH>(d.C + sk_c_offset(d, m0 + row, col), v);
    }
}


// First pass only (library-internal): problems with ks > 1 leave their ks partial tiles in
// ws (partial z of row m, column c at ws[(z * M + m) * Nc + c]) for a fused second pass.
int vt_skinny_first_pass(const void *descs, int batch, int max_M, int max_K, int max_Nc,
                         hipStream_t stream);
