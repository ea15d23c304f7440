// Coefficient-space Gram  A = Phi^T Phi = sum_j (S_j S_j^T) kron T_j  on MFMA.
//
// Phi (n x D) is the row-wise Khatri-Rao product of the spatial coefficients S[j(k), :]
// and the interval integrals alpha[k, :], so its Gram needs only the per-neuron temporal
// moments T_j = sum_{k in j} alpha_k alpha_k^T:
//
//   A[(ps,pt), (ps',pt')] = sum_j S[j,ps] S[j,ps'] T_j[pt,pt'],   D index = ps*Dt + pt.
//
// Viewed as a GEMM with K = neurons, rows = (ps, ps') pairs and columns = (pt, pt')
// pairs: the left operand W[(ps,ps'), j] = S[j,ps] S[j,ps'] is formed in registers from
// two LDS-staged S tiles (never materialised: it would be Ds^2 x N), the right operand
// is T (N x Dt^2) staged in LDS.  This is the reference's G (vtem/dirichlet.py:272-330,
// n x n with n ~ 2.6e5 at 64x64 px) mapped through the push-through identity to D x D,
// with 2*N*D^2 instead of 2*n*D^2 flops.
//
// A workgroup owns a pair of spatial blocks (b1 >= b2) of BP = 8 spatial indices each,
// i.e. the 64 rows (ps in b1, ps' in b2) x the Dt (Dt + 1) / 2 temporal columns; that
// covers the D-space block rows [b1*8*Dt, ...) x cols [b2*8*Dt, ...), so only the lower
// block triangle of A is computed (diagonal blocks in full).  4 waves, wave w owns row
// tile w.
#include "vt_common.hpp"

// T_j is symmetric, so the columns (pt, pt') and (pt', pt) of every row are equal.  The
// kernels compute only the Dt (Dt + 1) / 2 columns pt <= pt' (numbered row by row) and
// write each result at both positions: 153 of 289 columns at Dt = 17, the same products in
// the same order as the mirrored column, so the matrix is unchanged bit for bit.
__host__ __device__ __forceinline__ int gram_tcols(int Dt) { return Dt * (Dt + 1) / 2; }

// column c < gram_tcols(Dt) of the upper triangle -> (pt, pt')
__device__ __forceinline__ void gram_tpair(int c, int Dt, int &pt, int &ptp)
{
    int p = 0;
    while (c >= Dt - p) {
        c -= Dt - p;
        ++p;
    }
    pt = p;
    ptp = p + c;
}

// column tiles of 16 for a temporal order Dt (the kernels' template buckets)
__host__ __device__ __forceinline__ int gram_nct(int Dt)
{
    const int nct = (gram_tcols(Dt) + 15) / 16;
    return nct <= 1 ? 1 : (nct <= 3 ? 3 : (nct <= 6 ? 6 : (nct <= 10 ? 10 : 0)));
}

#define BP 8
#define KC 16

struct Gram64Desc {
    const double *S;      // N x Ds (row stride lds)
    const double *Tm;     // N x Dt^2
    double *A;            // D x D (row stride lda)
    long long lds, lda;
    int N, Ds, Dt, pad;
};

template <int NCT>
__global__ void __launch_bounds__(256)
gram_f64_kernel(const Gram64Desc *__restrict__ descs, Gram64Desc d0)
{
    constexpr int NC = NCT * 16;
    constexpr int TLD = NC + 2;
    __shared__ double S1[KC][BP];
    __shared__ double S2[KC][BP];
    __shared__ double Tt[KC][TLD];
    __shared__ int toff[NC];              // column -> offset pt * Dt + pt' in T_j, -1 past the end

    const Gram64Desc d = descs ? descs[blockIdx.y] : d0;
    const double *__restrict__ S = d.S;
    const double *__restrict__ Tm = d.Tm;
    double *__restrict__ A = d.A;
    const long long lds = d.lds, lda = d.lda;
    const int N = d.N, Ds = d.Ds, Dt = d.Dt;
    const int nbp = (Ds + BP - 1) / BP;
    // block pair from the linear index: b1 >= b2
    const int pid = blockIdx.x;
    if (pid >= nbp * (nbp + 1) / 2) return;
    int b1 = (int)((sqrt(8.0 * pid + 1.0) - 1.0) * 0.5);
    while ((b1 + 1) * (b1 + 2) / 2 <= pid) ++b1;
    while (b1 * (b1 + 1) / 2 > pid) --b1;
    const int b2 = pid - b1 * (b1 + 1) / 2;

    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int DT2 = Dt * Dt, DTU = gram_tcols(Dt);
    for (int c = tid; c < NC; c += 256) {
        int pt = 0, ptp = 0;
        if (c < DTU) gram_tpair(c, Dt, pt, ptp);
        toff[c] = c < DTU ? pt * Dt + ptp : -1;
    }
    v4d acc[NCT];
#pragma unroll
    for (int c = 0; c < NCT; ++c) acc[c] = (v4d){0, 0, 0, 0};

    const int r = wave * 16 + (lane & 15);   // row of this lane's A operand
    const int i1 = r >> 3, i2 = r & 7;
    __syncthreads();

    for (int j0 = 0; j0 < N; j0 += KC) {
        // stage S tiles
        if (tid < KC * BP) {
            int kk = tid / BP, i = tid % BP;
            int j = j0 + kk, p1 = b1 * BP + i, p2 = b2 * BP + i;
            S1[kk][i] = (j < N && p1 < Ds) ? S[(long long)j * lds + p1] : 0.0;
            S2[kk][i] = (j < N && p2 < Ds) ? S[(long long)j * lds + p2] : 0.0;
        }
        // stage the upper-triangle columns of the T tile
        const double *Tg = Tm + (long long)j0 * DT2;
        for (int e = tid; e < KC * NC; e += 256) {
            int kk = e / NC, c = e % NC;
            const int o = toff[c];
            Tt[kk][c] = (o >= 0 && j0 + kk < N) ? Tg[(long long)kk * DT2 + o] : 0.0;
        }
        __syncthreads();
#pragma unroll
        for (int ks = 0; ks < KC; ks += 4) {
            const int k = ks + (lane >> 4);
            const double a = S1[k][i1] * S2[k][i2];
#pragma unroll
            for (int c = 0; c < NCT; ++c) {
                const double b = Tt[k][c * 16 + (lane & 15)];
                acc[c] = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, acc[c], 0, 0, 0);
            }
        }
        __syncthreads();
    }
#pragma unroll
    for (int c = 0; c < NCT; ++c) {
        const int col = c * 16 + mfma_col(lane);
        if (col >= DTU) continue;
        const int pt = toff[col] / Dt, ptp = toff[col] % Dt;
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const int rr = wave * 16 + mfma_row_f64(lane, q);
            const int ps = b1 * BP + (rr >> 3), psp = b2 * BP + (rr & 7);
            if (ps < Ds && psp < Ds) {
                A[(long long)(ps * Dt + pt) * lda + psp * Dt + ptp] = acc[c][q];
                if (ptp != pt) A[(long long)(ps * Dt + ptp) * lda + psp * Dt + pt] = acc[c][q];
            }
        }
    }
}

static int launch_gram_f64(const Gram64Desc *descs, const Gram64Desc &d0, int batch, int max_Ds,
                           int Dt, hipStream_t s)
{
    const int nbp = (max_Ds + BP - 1) / BP;
    dim3 grid(nbp * (nbp + 1) / 2, batch);
    const int nct = gram_nct(Dt);
    if (nct == 1) gram_f64_kernel<1><<<grid, 256, 0, s>>>(descs, d0);
    else if (nct == 3) gram_f64_kernel<3><<<grid, 256, 0, s>>>(descs, d0);
    else if (nct == 6) gram_f64_kernel<6><<<grid, 256, 0, s>>>(descs, d0);
    else if (nct == 10) gram_f64_kernel<10><<<grid, 256, 0, s>>>(descs, d0);
    else return VT_ERANGE;
    VT_LAUNCH_CHECK();
    return VT_OK;
}

// A (D x D, row stride lda, D = Ds*Dt): lower block triangle written (see above).
VT_EXPORT int vt_gram_f64(const double *S, long long lds, const double *Tm, int N, int Ds, int Dt,
                          double *A, long long lda, void *stream)
{
    VT_CHECK_ARG(S && Tm && A && N >= 1 && Ds >= 1 && Dt >= 1 && lds >= Ds &&
                 lda >= (long long)Ds * Dt);
    Gram64Desc d0{S, Tm, A, lds, lda, N, Ds, Dt, 0};
    return launch_gram_f64(nullptr, d0, 1, Ds, Dt, (hipStream_t)stream);
}

// Batched FP64 Gram: descs = device array of {S, Tm, A, lds, lda, N, Ds, Dt, pad}
// (56 bytes); all problems share Dt.
VT_EXPORT int vt_gram_f64_batched(const void *descs, int batch, int max_Ds, int Dt, void *stream)
{
    VT_CHECK_ARG(descs && batch >= 1 && batch <= 65535 && max_Ds >= 1 && Dt >= 1);
    Gram64Desc d0{};
    return launch_gram_f64((const Gram64Desc *)descs, d0, batch, max_Ds, Dt, (hipStream_t)stream);
}

// Mirror the lower triangle into the upper one (for callers that want the full matrix).
template <typename T>
__global__ void symmetrize_kernel(T *A, long long lda, int D)
{
    const int i = blockIdx.y * 16 + threadIdx.y;
    const int j = blockIdx.x * 16 + threadIdx.x;
    if (i < D && j < D && j > i) A[(long long)i * lda + j] = A[(long long)j * lda + i];
}

VT_EXPORT int vt_symmetrize(void *A, int dtype, long long lda, int D, void *stream)
{
    VT_CHECK_ARG(A && D >= 0 && lda >= D);
    if (D == 0) return VT_OK;
    dim3 g(vt_blocks(D, 16), vt_blocks(D, 16));
    dim3 b(16, 16);
    if (dtype == 0)
        symmetrize_kernel<double><<<g, b, 0, (hipStream_t)stream>>>((double *)A, lda, D);
    else
        symmetrize_kernel<float><<<g, b, 0, (hipStream_t)stream>>>((float *)A, lda, D);
    VT_LAUNCH_CHECK();
    return VT_OK;
}

// ------------------------------------------------------------------------------------
// Low-precision Gram on the gfx950 16x16x32 MFMAs, FP32 accumulation, FP32 output
// (lower block triangle), batched over independent blocks by descriptor.
//
//   mode 0 (split-BF16, 3 products): x = hi + lo with hi = bf16(x), lo = bf16(x - hi);
//          A += Whi Thi + Whi Tlo + Wlo Thi   (about 16 significant bits per operand);
//   mode 1 (FP16, 1 product): W and T pre-scaled by powers of two (scales on device) so
//          they sit in FP16's normal range; the product is unscaled in the epilogue.
//
// T_j is split once per decode (vt_gram_split_T) into 16-bit arrays tiled per 32-neuron
// chunk as [column][neuron], so a chunk is staged into LDS with 16-byte copies and each
// lane's 8-neuron B operand is one 16-byte LDS read; W = S[j,ps] S[j,ps'] is formed in
// FP32 registers and split per k-step.  The LDS stage is double-buffered: chunk k+1 is
// copied while chunk k feeds the MFMAs, one barrier per chunk.
// A workgroup owns a 16 x 16 block of spatial pairs (256 pair rows, so T_j is streamed
// once per 256 rows) and the Dt (Dt + 1) / 2 temporal columns, split over its 8 waves by
// rows (split-BF16: 8 x 32 rows) or by rows and column halves (FP16: 4 x 64 rows x 2).
// ------------------------------------------------------------------------------------
struct GramDesc {
    const double *S;      // N x Ds (row stride lds)
    const unsigned short *Th, *Tl;   // N x Dt^2 16-bit (Tl unused in mode 1)
    float *A;             // D x D (row stride lda)
    long long lds, lda;
    int N, Ds, Dt, pad;
    const float *scales;  // mode 1: [sT, sW, 1/(sT sW)]
};

#define KL 32
#define KP (KL + 8)           // padded k stride (16-bit elements) of a T tile row, in HBM and LDS
#define LP_BP 16          // spatial indices per block: a workgroup owns 16 x 16 = 256 pair rows

#define LP_WAVES 8        // waves per workgroup


// MODE 0: split-BF16 3-product; MODE 1: FP16.  RT: row tiles of 16 per wave, so the
// workgroup's 256 pair rows form 16/RT row groups and its 8 waves RT/2 column groups.
// Every tiling issues the same MFMAs in the same order for each output element (the
// result is bit-identical); they differ in how often an operand is formed or read:
// RT = 2 forms each W element once per workgroup (every wave takes all columns; split-BF16
// pays most for W), RT = 4 and 8 let each B fragment read from LDS feed more row tiles.
// Tail split of a single-problem launch.  A problem of P pair blocks at one workgroup per
// CU fills the chip for P / CUs rounds; a thin last round would leave most CUs idle for a
// whole round.  Its ntail blocks are instead split over f neuron ranges each: workgroup
// nfull + t f + part computes block nfull + t over chunk range `part` and writes its FP32
// partial tile (before unscaling) to ws[((t f + part) * 256 + row) * NC + col];
// gram_lp_fixup_kernel adds the f partials in part order and writes the block.
struct GramSplit {
    int nfull, ntail, f, pad;
    float *ws;
};

template <int NCT, int MODE, int RT>
__global__ void __launch_bounds__(64 * LP_WAVES)
gram_lp_kernel(const GramDesc *__restrict__ descs, const GramSplit sp)
{
    constexpr int NC = NCT * 16;
    static_assert(RT == 2 || RT == 4 || RT == 8, "row tiles per wave");
    constexpr int NRG = 16 / RT;             // row groups
    constexpr int NCG = LP_WAVES / NRG;      // column groups
    constexpr int NCT2 = (NCT + NCG - 1) / NCG;  // column tiles per column group
    // rows padded to LP_BP + 2 floats: the four 16-lane groups of a wave read rows 8 apart,
    // which a 16-float stride maps onto the same banks
    __shared__ float S1[2][KL][LP_BP + 2];
    __shared__ float S2[2][KL][LP_BP + 2];
    __shared__ __attribute__((aligned(16))) unsigned short Th[2][NC][KP];
    __shared__ __attribute__((aligned(16))) unsigned short Tl[MODE == 0 ? 2 : 1][MODE == 0 ? NC : 1][KP];

    const GramDesc d = descs[blockIdx.y];
    const int nbp = (d.Ds + LP_BP - 1) / LP_BP;
    const int nsteps = (d.N + KL - 1) / KL;
    int pid = blockIdx.x, kb = 0, ke = nsteps, part = -1, tb = 0;
    if (sp.f > 1 && pid >= sp.nfull) {
        const int it = pid - sp.nfull;
        tb = it / sp.f;
        part = it % sp.f;
        if (tb >= sp.ntail) return;
        pid = sp.nfull + tb;
        const int per = (nsteps + sp.f - 1) / sp.f;
        kb = min(nsteps, part * per);
        ke = min(nsteps, kb + per);
    }
    if (pid >= nbp * (nbp + 1) / 2) return;
    int b1 = (int)((sqrt(8.0 * pid + 1.0) - 1.0) * 0.5);
    while ((b1 + 1) * (b1 + 2) / 2 <= pid) ++b1;
    while (b1 * (b1 + 1) / 2 > pid) --b1;
    const int b2 = pid - b1 * (b1 + 1) / 2;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int rg = wave % NRG;                           // row group (16 RT rows)
    const int cg = wave / NRG;                           // column group
    const int c0 = cg * NCT2;                            // first column tile of the group
    const int Dt = d.Dt, N = d.N;
    const float sW = MODE == 1 ? d.scales[1] : 1.0f;
    const int g = lane >> 4;

    // Chunk k (neurons 32k..32k+31).  The pre-tiled T rows go from HBM straight into LDS
    // (global_load_lds, 16 bytes per lane at the wave's LDS base + 16 * lane; the tile's
    // HBM layout already carries the LDS row padding), and the S slices through one
    // register per thread: the next chunk's loads are issued before this chunk's MFMAs and
    // waited for at the step's barrier, so their latency overlaps the math.
    typedef __attribute__((address_space(1))) const void *gptr_t;
    typedef __attribute__((address_space(3))) void *lptr_t;
    constexpr int NTH = 64 * LP_WAVES;
    constexpr int TV = NC * KP / 8;                      // 16-byte vectors per T tile
    float rs1 = 0.f, rs2 = 0.f;
    const int se_kk = tid / LP_BP, se_i = tid % LP_BP;  // S element of this thread (tid < KL * LP_BP)
    auto load_S = [&](int k) {
        if (tid < KL * LP_BP) {
            const int j = k * KL + se_kk, p1 = b1 * LP_BP + se_i, p2 = b2 * LP_BP + se_i;
            rs1 = (j < N && p1 < d.Ds) ? (float)d.S[(long long)j * d.lds + p1] : 0.f;
            // FP16: S2 carries the power-of-two scale sW (exact), so W = S1 S2 sW is one product
            rs2 = (j < N && p2 < d.Ds) ? (float)d.S[(long long)j * d.lds + p2] * sW : 0.f;
        }
    };
    auto store_S = [&](int bf) {
        if (tid < KL * LP_BP) {
            S1[bf][se_kk][se_i] = rs1;
            S2[bf][se_kk][se_i] = rs2;
        }
    };
    auto dma_T = [&](int k, int bf) {
        const unsigned short *th = d.Th + (long long)k * NC * KP;
        const unsigned short *tl = d.Tl + (long long)k * NC * KP;
#pragma unroll
        for (int q = 0; q < (TV + NTH - 1) / NTH; ++q) {
            const int v0 = q * NTH + wave * 64;          // this wave's first vector (uniform)
            if (v0 >= TV) break;
            const int v = v0 + lane;
            if (v < TV) {
                __builtin_amdgcn_global_load_lds((gptr_t)(th + 8 * v), (lptr_t)(&Th[bf][0][0] + 8 * v0),
                                                 16, 0, 0);
                if (MODE == 0)
                    __builtin_amdgcn_global_load_lds((gptr_t)(tl + 8 * v), (lptr_t)(&Tl[bf][0][0] + 8 * v0),
                                                     16, 0, 0);
            }
        }
    };

    v4f acc[RT][NCT2];
#pragma unroll
    for (int t = 0; t < RT; ++t)
#pragma unroll
        for (int c = 0; c < NCT2; ++c) acc[t][c] = (v4f){0, 0, 0, 0};

    if (kb < ke) {                                      // uniform (an empty split range adds zeros)
        load_S(kb);
        store_S(0);
        dma_T(kb, 0);
        __builtin_amdgcn_s_waitcnt(0);                  // this wave's tile loads have landed
        __syncthreads();
    }
    for (int k = kb; k < ke; ++k) {
        const int bf = (k - kb) & 1;
        if (k + 1 < ke) {                               // the other buffer: free since the last barrier
            load_S(k + 1);
            dma_T(k + 1, bf ^ 1);
        }
        // A operands: W = S1 S2 for this wave's 4 row tiles, neurons 8g..8g+7 of the chunk
        if (MODE == 0) {
            v8bf wh[RT], wl[RT];
#pragma unroll
            for (int t = 0; t < RT; ++t) {
                const int r = rg * 16 * RT + t * 16 + (lane & 15);
                const int i1 = r >> 4, i2 = r & 15;
#pragma unroll
                for (int i = 0; i < 8; ++i) {
                    const float w = S1[bf][8 * g + i][i1] * S2[bf][8 * g + i][i2];
                    const __bf16 h = (__bf16)w;
                    wh[t][i] = h;
                    wl[t][i] = (__bf16)(w - (float)h);
                }
            }
#pragma unroll
            for (int c = 0; c < NCT2; ++c) {
                if (c0 + c >= NCT) break;
                const int cc = (c0 + c) * 16 + (lane & 15);
                const v8bf bh = *(const v8bf *)&Th[bf][cc][8 * g];
                const v8bf bl = *(const v8bf *)&Tl[bf][cc][8 * g];
#pragma unroll
                for (int t = 0; t < RT; ++t) {
                    acc[t][c] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(wh[t], bh, acc[t][c], 0, 0, 0);
                    acc[t][c] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(wh[t], bl, acc[t][c], 0, 0, 0);
                    acc[t][c] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(wl[t], bh, acc[t][c], 0, 0, 0);
                }
            }
        } else {
            v8h wh[RT];
#pragma unroll
            for (int t = 0; t < RT; ++t) {
                const int r = rg * 16 * RT + t * 16 + (lane & 15);
                const int i1 = r >> 4, i2 = r & 15;
#pragma unroll
                for (int i = 0; i < 8; ++i)
                    wh[t][i] = (_Float16)(S1[bf][8 * g + i][i1] * S2[bf][8 * g + i][i2]);
            }
#pragma unroll
            for (int c = 0; c < NCT2; ++c) {
                if (c0 + c >= NCT) break;
                const int cc = (c0 + c) * 16 + (lane & 15);
                const v8h bh = *(const v8h *)&Th[bf][cc][8 * g];
#pragma unroll
                for (int t = 0; t < RT; ++t)
                    acc[t][c] = __builtin_amdgcn_mfma_f32_16x16x32_f16(wh[t], bh, acc[t][c], 0, 0, 0);
            }
        }
        if (k + 1 < ke) store_S(bf ^ 1);
        __builtin_amdgcn_s_waitcnt(0);                  // the next chunk's tile loads have landed
        __syncthreads();
    }
    if (part >= 0) {
        float *W = sp.ws + (long long)(tb * sp.f + part) * 256 * NC;
#pragma unroll
        for (int c = 0; c < NCT2; ++c) {
            if (c0 + c >= NCT) break;
            const int col = (c0 + c) * 16 + mfma_col(lane);
#pragma unroll
            for (int t = 0; t < RT; ++t)
#pragma unroll
                for (int q = 0; q < 4; ++q)
                    W[(rg * 16 * RT + t * 16 + mfma_row(lane, q)) * NC + col] = acc[t][c][q];
        }
        return;
    }
    const float unscale = MODE == 1 ? d.scales[2] : 1.0f;
    const int DTU = gram_tcols(Dt);
#pragma unroll
    for (int c = 0; c < NCT2; ++c) {
        if (c0 + c >= NCT) break;
        const int col = (c0 + c) * 16 + mfma_col(lane);
        if (col >= DTU) continue;
        int pt, ptp;
        gram_tpair(col, Dt, pt, ptp);
#pragma unroll
        for (int t = 0; t < RT; ++t)
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const int rr = rg * 16 * RT + t * 16 + mfma_row(lane, q);
                const int ps = b1 * LP_BP + (rr >> 4), psp = b2 * LP_BP + (rr & 15);
                if (ps < d.Ds && psp < d.Ds) {
                    const float v = acc[t][c][q] * unscale;
                    d.A[(long long)(ps * Dt + pt) * d.lda + psp * Dt + ptp] = v;
                    if (ptp != pt) d.A[(long long)(ps * Dt + ptp) * d.lda + psp * Dt + pt] = v;
                }
            }
    }
}

// Tail blocks of a split launch: sum the f partial tiles in part order, unscale, write the
// block's pairs (pt <= pt') at both mirrored places, as the Gram kernel's epilogue does.
__global__ void __launch_bounds__(256)
gram_lp_fixup_kernel(const GramDesc *__restrict__ descs, const GramSplit sp, int NC, int mode)
{
    const GramDesc d = descs[0];
    const int tb = blockIdx.x;
    const int pid = sp.nfull + tb;
    int b1 = (int)((sqrt(8.0 * pid + 1.0) - 1.0) * 0.5);
    while ((b1 + 1) * (b1 + 2) / 2 <= pid) ++b1;
    while (b1 * (b1 + 1) / 2 > pid) --b1;
    const int b2 = pid - b1 * (b1 + 1) / 2;
    const int Dt = d.Dt, DTU = gram_tcols(Dt);
    const float unscale = mode == 1 ? d.scales[2] : 1.0f;
    const float *W = sp.ws + (long long)tb * sp.f * 256 * NC;
    for (int e = blockIdx.y * 256 + threadIdx.x; e < 256 * DTU; e += 256 * gridDim.y) {
        const int rr = e / DTU, col = e % DTU;
        float v = 0.f;
        for (int p = 0; p < sp.f; ++p) v += W[((long long)p * 256 + rr) * NC + col];
        v *= unscale;
        int pt, ptp;
        gram_tpair(col, Dt, pt, ptp);
        const int ps = b1 * LP_BP + (rr >> 4), psp = b2 * LP_BP + (rr & 15);
        if (ps < d.Ds && psp < d.Ds) {
            d.A[(long long)(ps * Dt + pt) * d.lda + psp * Dt + ptp] = v;
            if (ptp != pt) d.A[(long long)(ps * Dt + ptp) * d.lda + psp * Dt + pt] = v;
        }
    }
}

// Split T_j (N x Dt^2 doubles) into the tiled 16-bit layout of the Gram kernels: chunk k of
// 32 neurons is a contiguous [NC][KP] tile (column-major within the chunk, rows padded from
// 32 to KP elements with zeros, as the kernels keep it in LDS), so the kernels copy it to
// LDS with 16-byte direct loads.  Tile column c holds T_j[pt][pt'] for the c-th upper-
// triangle pair.  mode 0: hi / lo bf16; mode 1: fp16 of T*scales[0].
__global__ void __launch_bounds__(256)
split_T_kernel(const double *__restrict__ T, int N, int Dt, int NC, long long total,
               unsigned short *hi, unsigned short *lo, int mode, const float *scales)
{
    __shared__ int toff[256];
    const int DTU = gram_tcols(Dt);
    for (int c = threadIdx.x; c < NC; c += 256) {
        int pt = 0, ptp = 0;
        if (c < DTU) gram_tpair(c, Dt, pt, ptp);
        toff[c] = c < DTU ? pt * Dt + ptp : -1;
    }
    __syncthreads();
    const float s = mode == 1 ? scales[0] : 1.0f;
    const int DT2 = Dt * Dt;
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < total;
         i += (long long)gridDim.x * blockDim.x) {
        const int kk = (int)(i % KP);
        const long long r = i / KP;
        const int c = (int)(r % NC);
        const int j = (int)(r / NC) * KL + kk;
        const int o = toff[c];
        const float t = (kk < KL && j < N && o >= 0) ? (float)T[(long long)j * DT2 + o] : 0.f;
        if (mode == 0) {
            const __bf16 h = (__bf16)t;
            const __bf16 l = (__bf16)(t - (float)h);
            hi[i] = __builtin_bit_cast(unsigned short, h);
            lo[i] = __builtin_bit_cast(unsigned short, l);
        } else {
            hi[i] = __builtin_bit_cast(unsigned short, (_Float16)(t * s));
        }
    }
}

// FP16 scales from max|T| and max|S|: powers of two putting max|T| * sT and
// max|S|^2 * sW near 2^14 (FP16 max 65504), stored as [sT, sW, 1/(sT sW)].  The maxima
// are reduced over a grid with atomicMax on the bit patterns of non-negative floats (which
// order like the values), so the result does not depend on the reduction order.
__global__ void __launch_bounds__(256)
absmax_kernel(const double *__restrict__ T, long long nT, const double *__restrict__ S, long long nS,
              unsigned int *bits)
{
    __shared__ float red[2][256];
    float mt = 0.f, ms = 0.f;
    for (long long i = (long long)blockIdx.x * 256 + threadIdx.x; i < nT; i += (long long)gridDim.x * 256)
        mt = fmaxf(mt, (float)fabs(T[i]));
    for (long long i = (long long)blockIdx.x * 256 + threadIdx.x; i < nS; i += (long long)gridDim.x * 256)
        ms = fmaxf(ms, (float)fabs(S[i]));
    red[0][threadIdx.x] = mt;
    red[1][threadIdx.x] = ms;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if (threadIdx.x < o) {
            red[0][threadIdx.x] = fmaxf(red[0][threadIdx.x], red[0][threadIdx.x + o]);
            red[1][threadIdx.x] = fmaxf(red[1][threadIdx.x], red[1][threadIdx.x + o]);
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        atomicMax(bits, __float_as_uint(red[0][0]));
        atomicMax(bits + 1, __float_as_uint(red[1][0]));
    }
}

__global__ void fp16_scales_kernel(const unsigned int *bits, float *scales)
{
    const double tm = __uint_as_float(bits[0]), sm = __uint_as_float(bits[1]);
    const double tmax = tm > 0 ? tm : 1.0, smax = sm > 0 ? sm : 1.0;
    const double sT = exp2(floor(log2(16384.0 / tmax)));
    const double sW = exp2(floor(log2(16384.0 / (smax * smax))));
    scales[0] = (float)sT;
    scales[1] = (float)sW;
    scales[2] = (float)(1.0 / (sT * sW));
}

__global__ void zero_bits_kernel(unsigned int *bits) { bits[0] = 0u; bits[1] = 0u; }

// 16-bit elements of one block's tiled T (per array: hi, and lo in mode 0)
VT_EXPORT int vt_gram_split_T_size(int N, int Dt, long long *elems_host)
{
    VT_CHECK_ARG(N >= 0 && Dt >= 1 && elems_host);
    const int nct = gram_nct(Dt);
    if (!nct) return VT_ERANGE;
    *elems_host = (long long)((N + KL - 1) / KL) * nct * 16 * KP;
    return VT_OK;
}

VT_EXPORT int vt_gram_split_T(const double *Tm, int N, int Dt, void *hi, void *lo, int mode,
                              const double *S, long long nS, float *scales, void *stream)
{
    VT_CHECK_ARG(Tm && hi && N >= 1 && Dt >= 1 && (mode == 0 || mode == 1));
    VT_CHECK_ARG(mode == 1 ? (S && scales && nS > 0) : (lo != nullptr));
    const int nct = gram_nct(Dt);
    if (!nct) return VT_ERANGE;
    hipStream_t s = (hipStream_t)stream;
    if (mode == 1) {
        // scales[3..4] hold the maxima's bit patterns while they are reduced
        unsigned int *bits = (unsigned int *)(scales + 3);
        zero_bits_kernel<<<1, 1, 0, s>>>(bits);
        absmax_kernel<<<256, 256, 0, s>>>(Tm, (long long)N * Dt * Dt, S, nS, bits);
        fp16_scales_kernel<<<1, 1, 0, s>>>(bits, scales);
    }
    const long long total = (long long)((N + KL - 1) / KL) * nct * 16 * KP;
    unsigned g = vt_blocks(total, 256);
    split_T_kernel<<<g < 4096 ? g : 4096, 256, 0, s>>>(Tm, N, Dt, nct * 16, total, (unsigned short *)hi,
                                                      (unsigned short *)lo, mode, scales);
    VT_LAUNCH_CHECK();
    return VT_OK;
}

// The kernel instance of a launch: nct column tiles, mode, tile (row tiles per wave).
static const void *gram_lp_fn(int nct, int mode, int tile)
{
#define LP_FN_T(NCTV, M)                                                                       \
    return tile == 2 ? (const void *)gram_lp_kernel<NCTV, M, 2>                                \
           : tile == 4 ? (const void *)gram_lp_kernel<NCTV, M, 4>                              \
                       : (const void *)gram_lp_kernel<NCTV, M, 8>
#define LP_FN(NCTV)                                                                            \
    do {                                                                                       \
        if (mode == 0) { LP_FN_T(NCTV, 0); }                                                   \
        else { LP_FN_T(NCTV, 1); }                                                             \
    } while (0)
    if (nct == 1) LP_FN(1);
    if (nct == 3) LP_FN(3);
    if (nct == 6) LP_FN(6);
    if (nct == 10) LP_FN(10);
#undef LP_FN
#undef LP_FN_T
    return nullptr;
}

// Tail-split plan of a launch (f = 1: none).  Only single-problem launches split: in a
// batch, a block's result must not depend on the other blocks.  ws_floats: workspace the
// plan needs.
static void gram_lp_plan(int batch, int max_Ds, int nct, const void *fn, GramSplit *sp,
                         long long *ws_floats)
{
    const int nbp = (max_Ds + LP_BP - 1) / LP_BP;
    const int P = nbp * (nbp + 1) / 2;
    *sp = GramSplit{P, 0, 1, 0, nullptr};
    *ws_floats = 0;
    if (batch != 1 || !fn) return;
    int dev = 0, cus = 0, occ = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus < 1)
        return;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, fn, 64 * LP_WAVES, 0) != hipSuccess || occ < 1)
        occ = 1;
    const int W = cus * occ;                 // workgroups in flight
    const int rounds = P / W, tail = P % W;
    if (rounds < 1 || tail == 0 || 2 * tail > W) return;
    const int f = W / tail > 8 ? 8 : W / tail;
    if (f < 2) return;
    sp->nfull = P - tail;
    sp->ntail = tail;
    sp->f = f;
    *ws_floats = (long long)tail * f * 256 * nct * 16;
}

VT_EXPORT int vt_gram_lp_ws_size(int batch, int max_Ds, int Dt, int mode, int tile, long long *bytes_host)
{
    VT_CHECK_ARG(bytes_host && batch >= 1 && max_Ds >= 1 && Dt >= 1 && (mode == 0 || mode == 1) &&
                 (tile == 0 || tile == 2 || tile == 4 || tile == 8));
    if (tile == 0) tile = mode == 0 ? 2 : 4;
    const int nct = gram_nct(Dt);
    if (!nct) return VT_ERANGE;
    GramSplit sp;
    long long wf;
    gram_lp_plan(batch, max_Ds, nct, gram_lp_fn(nct, mode, tile), &sp, &wf);
    *bytes_host = wf * (long long)sizeof(float);
    return VT_OK;
}

// Batched low-precision Gram.  descs: device array of GramDesc; every problem must have
// the same Dt (column tiling), max_Ds bounds the grid.  ws (ws_bytes): workspace of the
// tail split (single-problem launches; vt_gram_lp_ws_size), or null: no split.
VT_EXPORT int vt_gram_lp(const void *descs, int batch, int max_Ds, int Dt, int mode, int tile,
                         void *ws, long long ws_bytes, void *stream)
{
    VT_CHECK_ARG(descs && batch >= 1 && batch <= 65535 && max_Ds >= 1 && Dt >= 1 &&
                 (mode == 0 || mode == 1) && (tile == 0 || tile == 2 || tile == 4 || tile == 8) &&
                 ws_bytes >= 0);
    // default tiling per form (measured): split-BF16 2 row tiles per wave, FP16 4
    if (tile == 0) tile = mode == 0 ? 2 : 4;
    const int nct = gram_nct(Dt);
    if (!nct) return VT_ERANGE;
    GramSplit sp;
    long long wf = 0;
    if (ws) gram_lp_plan(batch, max_Ds, nct, gram_lp_fn(nct, mode, tile), &sp, &wf);
    if (!ws || sp.f < 2 || ws_bytes < wf * (long long)sizeof(float)) {
        const int nbp = (max_Ds + LP_BP - 1) / LP_BP;
        sp = GramSplit{nbp * (nbp + 1) / 2, 0, 1, 0, nullptr};
    } else {
        sp.ws = (float *)ws;
    }
    dim3 grid(sp.nfull + sp.ntail * sp.f, batch);
    hipStream_t s = (hipStream_t)stream;
    const GramDesc *dd = (const GramDesc *)descs;
#define LAUNCH_LP_T(NCTV, M)                                                                  \
    do {                                                                                      \
        if (tile == 2) gram_lp_kernel<NCTV, M, 2><<<grid, 64 * LP_WAVES, 0, s>>>(dd, sp);     \
        else if (tile == 4) gram_lp_kernel<NCTV, M, 4><<<grid, 64 * LP_WAVES, 0, s>>>(dd, sp); \
        else gram_lp_kernel<NCTV, M, 8><<<grid, 64 * LP_WAVES, 0, s>>>(dd, sp);               \
    } while (0)
#define LAUNCH_LP(NCTV)                                                                       \
    do {                                                                                      \
        if (mode == 0) LAUNCH_LP_T(NCTV, 0);                                                  \
        else LAUNCH_LP_T(NCTV, 1);                                                            \
    } while (0)
    if (nct == 1) LAUNCH_LP(1);
    else if (nct == 3) LAUNCH_LP(3);
    else if (nct == 6) LAUNCH_LP(6);
    else LAUNCH_LP(10);
#undef LAUNCH_LP
#undef LAUNCH_LP_T
    VT_LAUNCH_CHECK();
    if (sp.f > 1) {
        gram_lp_fixup_kernel<<<dim3(sp.ntail, 8), 256, 0, s>>>(dd, sp, nct * 16, mode);
        VT_LAUNCH_CHECK();
    }
    return VT_OK;
}
