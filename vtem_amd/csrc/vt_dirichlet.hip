// Spatial Dirichlet coefficients, the decode-space functions psi and block stitching: the
// spike-space side of the reference's decoder and its window / tile drivers.
//
// Pixel grid (vtem/vrf.py:1254-1257, origin :323-327): x_p = x_start + p dx,
// y_q = -(y_start + q dy).  Mode m = (mx, my), |mx| <= Mx, |my| <= My, is column
// m = (my + My)(2Mx + 1) + (mx + Mx) of the table; theta = mx Ox x_p + my Oy y_q.
#include "vt_common.hpp"

// table (P x 2M, row stride ld): [cos theta | -sin theta], row q*Px + p.  RF (N x P) times
// it gives sum RF e^{-i theta} split in real / imaginary halves; [Re Ds | Im Ds] times its
// transpose gives Re sum Ds e^{i theta} (psi).
__global__ void dirichlet_table_kernel(double *__restrict__ out, long long ld, int Px, int Py,
                                       double x_start, double dx, double y_start, double dy, int Mx,
                                       int My, double Ox, double Oy)
{
    const int M = (2 * Mx + 1) * (2 * My + 1);
    const long long P = (long long)Px * Py;
    const long long total = P * M;
    for (long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x; e < total;
         e += (long long)gridDim.x * blockDim.x) {
        const long long pix = e / M;
        const int m = (int)(e % M);
        const int q = (int)(pix / Px), p = (int)(pix % Px);
        const int mx = m % (2 * Mx + 1) - Mx, my = m / (2 * Mx + 1) - My;
        const double x = x_start + p * dx, y = -(y_start + q * dy);
        double s, c;
        sincos(mx * Ox * x + my * Oy * y, &s, &c);
        out[pix * ld + m] = c;
        out[pix * ld + M + m] = -s;
    }
}

VT_EXPORT int vt_dirichlet_table(double *out, long long ld, int Px, int Py, double x_start, double dx,
                                 double y_start, double dy, int Mx, int My, double Ox, double Oy,
                                 void *stream)
{
    const int M = (2 * Mx + 1) * (2 * My + 1);
    VT_CHECK_ARG(out && Px >= 1 && Py >= 1 && Mx >= 0 && My >= 0 && ld >= 2 * M);
    const long long total = (long long)Px * Py * M;
    unsigned g = vt_blocks(total, 256);
    dirichlet_table_kernel<<<g < 8192 ? g : 8192, 256, 0, (hipStream_t)stream>>>(
        out, ld, Px, Py, x_start, dx, y_start, dy, Mx, My, Ox, Oy);
    VT_LAUNCH_CHECK();
    return VT_OK;
}

// Ds (N x M complex, interleaved re/im, the reference's d_Ds[j, my+My, mx+Mx] layout) from
// C = RF table (N x 2M, row stride ldc): Ds = scale (C[:, m] + i C[:, M+m]), zero outside
// the mask (1 = ellipse mx^2 My^2 + my^2 Mx^2 <= Mx^2 My^2, vtem/vrf.py:1331; 0 = rect).
__global__ void ds_assemble_kernel(double *__restrict__ Ds, const double *__restrict__ C, long long ldc,
                                   int N, int Mx, int My, int mask, double scale)
{
    const int M = (2 * Mx + 1) * (2 * My + 1);
    const long long total = (long long)N * M;
    for (long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x; e < total;
         e += (long long)gridDim.x * blockDim.x) {
        const long long j = e / M;
        const int m = (int)(e % M);
        const long long mx = m % (2 * Mx + 1) - Mx, my = m / (2 * Mx + 1) - My;
        const bool in = !mask || mx * mx * My * My + my * my * Mx * Mx <= (long long)Mx * Mx * My * My;
        Ds[2 * e] = in ? scale * C[j * ldc + m] : 0.0;
        Ds[2 * e + 1] = in ? scale * C[j * ldc + M + m] : 0.0;
    }
}

VT_EXPORT int vt_ds_assemble(double *Ds, const double *C, long long ldc, int N, int Mx, int My, int mask,
                             double scale, void *stream)
{
    VT_CHECK_ARG(Ds && C && N >= 1 && Mx >= 0 && My >= 0 && (mask == 0 || mask == 1));
    const long long total = (long long)N * (2 * Mx + 1) * (2 * My + 1);
    unsigned g = vt_blocks(total, 256);
    ds_assemble_kernel<<<g < 8192 ? g : 8192, 256, 0, (hipStream_t)stream>>>(Ds, C, ldc, N, Mx, My,
                                                                             mask, scale);
    VT_LAUNCH_CHECK();
    return VT_OK;
}

// W[i, j] *= norm[i] norm[j]  (the put_norm step of compute_Dsw, vtem/vrf.py:529-533, :540-559)
__global__ void scale_outer_kernel(double *__restrict__ W, long long ldw, int n, const double *__restrict__ norm)
{
    const int i = blockIdx.y;
    const double ni = norm[i];
    for (int j = blockIdx.x * blockDim.x + threadIdx.x; j < n; j += gridDim.x * blockDim.x)
        W[(long long)i * ldw + j] *= ni * norm[j];
}

VT_EXPORT int vt_scale_outer(double *W, long long ldw, int n, const double *norm, void *stream)
{
    VT_CHECK_ARG(W && norm && n >= 1 && n <= 65535 && ldw >= n);
    dim3 g(vt_blocks(n, 256) < 16 ? vt_blocks(n, 256) : 16, n);
    scale_outer_kernel<<<g, 256, 0, (hipStream_t)stream>>>(W, ldw, n, norm);
    VT_LAUNCH_CHECK();
    return VT_OK;
}

// ------------------------------------------------------------------------------------
// Stitching: out[t0+t, y0+y, x0+x] += wt[t] wy[y] wx[x] blk[t, y, x] for a set of blocks
// that do not overlap each other (the caller launches overlapping blocks in separate
// calls, so every output element has one writer per call).  Window products form a
// partition of unity over the overlaps (raised-cosine ramps, vtem/vtdm.py:250-268).
// ------------------------------------------------------------------------------------
struct StitchDesc {
    const double *blk;       // T x Py x Px, contiguous
    const double *wt, *wy, *wx;
    int T, Py, Px, t0, y0, x0;
};

__global__ void __launch_bounds__(256)
stitch_kernel(double *__restrict__ out, long long ld_t, long long ld_y, const StitchDesc *__restrict__ descs)
{
    const StitchDesc d = descs[blockIdx.y];
    const long long total = (long long)d.T * d.Py * d.Px;
    for (long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x; e < total;
         e += (long long)gridDim.x * blockDim.x) {
        const int x = (int)(e % d.Px);
        const long long r = e / d.Px;
        const int y = (int)(r % d.Py), t = (int)(r / d.Py);
        out[(long long)(d.t0 + t) * ld_t + (long long)(d.y0 + y) * ld_y + d.x0 + x] +=
            d.wt[t] * d.wy[y] * d.wx[x] * d.blk[e];
    }
}

VT_EXPORT int vt_stitch_accumulate(double *out, long long ld_t, long long ld_y, const void *descs,
                                   int nblocks, long long max_elems, void *stream)
{
    VT_CHECK_ARG(out && descs && nblocks >= 1 && nblocks <= 65535 && max_elems >= 0 && ld_y >= 1 &&
                 ld_t >= ld_y);
    if (max_elems == 0) return VT_OK;
    unsigned g = vt_blocks(max_elems, 256);
    dim3 grid(g < 512 ? g : 512, nblocks);
    stitch_kernel<<<grid, 256, 0, (hipStream_t)stream>>>(out, ld_t, ld_y, (const StitchDesc *)descs);
    VT_LAUNCH_CHECK();
    return VT_OK;
}
