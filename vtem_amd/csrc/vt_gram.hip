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
// i.e. the 64 rows (ps in b1, ps' in b2) x all Dt^2 columns; that covers the D-space
// block rows [b1*8*Dt, ...) x cols [b2*8*Dt, ...), so only the lower block triangle of A
// is computed (diagonal blocks in full).  4 waves, wave w owns row tile w.
#include "vt_common.hpp"

#define BP 8
#define KC 16

template <int NCT>
__global__ void __launch_bounds__(256)
gram_f64_kernel(const double *__restrict__ S, long long lds, const double *__restrict__ Tm, int N,
                int Ds, int Dt, double *__restrict__ A, long long lda, int nbp)
{
    constexpr int NC = NCT * 16;
    constexpr int TLD = NC + 2;
    __shared__ double S1[KC][BP];
    __shared__ double S2[KC][BP];
    __shared__ double Tt[KC][TLD];

    // block pair from the linear index: b1 >= b2
    const int pid = blockIdx.x;
    int b1 = (int)((sqrt(8.0 * pid + 1.0) - 1.0) * 0.5);
    while ((b1 + 1) * (b1 + 2) / 2 <= pid) ++b1;
    while (b1 * (b1 + 1) / 2 > pid) --b1;
    const int b2 = pid - b1 * (b1 + 1) / 2;

    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int DT2 = Dt * Dt;
    v4d acc[NCT];
#pragma unroll
    for (int c = 0; c < NCT; ++c) acc[c] = (v4d){0, 0, 0, 0};

    const int r = wave * 16 + (lane & 15);   // row of this lane's A operand
    const int i1 = r >> 3, i2 = r & 7;

    for (int j0 = 0; j0 < N; j0 += KC) {
        // stage S tiles
        if (tid < KC * BP) {
            int kk = tid / BP, i = tid % BP;
            int j = j0 + kk, p1 = b1 * BP + i, p2 = b2 * BP + i;
            S1[kk][i] = (j < N && p1 < Ds) ? S[(long long)j * lds + p1] : 0.0;
            S2[kk][i] = (j < N && p2 < Ds) ? S[(long long)j * lds + p2] : 0.0;
        }
        // stage T tile (contiguous KC*Dt^2 doubles in global)
        const double *Tg = Tm + (long long)j0 * DT2;
        for (int e = tid; e < KC * NC; e += 256) {
            int kk = e / NC, c = e % NC;
            Tt[kk][c] = (c < DT2 && j0 + kk < N) ? Tg[(long long)kk * DT2 + c] : 0.0;
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
        if (col >= DT2) continue;
        const int pt = col / Dt, ptp = col % Dt;
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const int rr = wave * 16 + mfma_row_f64(lane, q);
            const int ps = b1 * BP + (rr >> 3), psp = b2 * BP + (rr & 7);
            if (ps < Ds && psp < Ds)
                A[(long long)(ps * Dt + pt) * lda + psp * Dt + ptp] = acc[c][q];
        }
    }
}

template <int NCT>
static void launch_gram_f64(const double *S, long long lds, const double *Tm, int N, int Ds,
                            int Dt, double *A, long long lda, hipStream_t s)
{
    const int nbp = (Ds + BP - 1) / BP;
    const int npairs = nbp * (nbp + 1) / 2;
    gram_f64_kernel<NCT><<<dim3(npairs, 1), 256, 0, s>>>(S, lds, Tm, N, Ds, Dt, A, lda, nbp);
}

// A (D x D, row stride lda, D = Ds*Dt): lower block triangle written (see above).
VT_EXPORT int vt_gram_f64(const double *S, long long lds, const double *Tm, int N, int Ds, int Dt,
                          double *A, long long lda, void *stream)
{
    VT_CHECK_ARG(S && Tm && A && N >= 1 && Ds >= 1 && Dt >= 1 && lds >= Ds &&
                 lda >= (long long)Ds * Dt);
    hipStream_t s = (hipStream_t)stream;
    const int nct = (Dt * Dt + 15) / 16;
    if (nct <= 2) launch_gram_f64<2>(S, lds, Tm, N, Ds, Dt, A, lda, s);
    else if (nct <= 6) launch_gram_f64<6>(S, lds, Tm, N, Ds, Dt, A, lda, s);
    else if (nct <= 19) launch_gram_f64<19>(S, lds, Tm, N, Ds, Dt, A, lda, s);
    else return VT_ERANGE;
    VT_LAUNCH_CHECK();
    return VT_OK;
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
