/*
 * vtem_hip.h -- C-ABI of libvtem_hip.so, the MI355X (gfx950) implementation of the video
 * time-encoding / time-decoding path of bionet/vtem.
 *
 * Conventions (all entries):
 *   - pointers are device addresses (hipMalloc / torch tensors) unless stated;
 *   - matrices are row-major with an explicit leading dimension in elements (the
 *     reference's PitchArray `ld`, vtem/utils/parray.py:136-157);
 *   - dtype selectors: 0 = float64, 1 = float32;
 *   - `stream` is a hipStream_t (void*); 0 is the legacy default stream;
 *   - every entry only enqueues work on `stream` (no host synchronisation) and returns
 *     an int status: 0 OK, > 0 hipError_t, < 0 argument error (-1 invalid argument,
 *     -2 unsupported, -3 spike buffer overflow, -4 not positive definite, -5 beyond a
 *     documented size limit);
 *   - callers own every buffer; workspaces are passed in explicitly.
 *
 * Coefficient-space conventions: real spatial basis index ps (0: constant; 1+2i, 2+2i:
 * sqrt2 cos / sqrt2 sin of half-set mode i), real temporal basis index pt (0: constant;
 * 2m-1, 2m: sqrt2 cos / sqrt2 sin of m*W*t); D-space index d = ps*Dt + pt.
 */
#ifndef VTEM_HIP_H
#define VTEM_HIP_H

#ifdef __cplusplus
extern "C" {
#endif

/* ---------------------------------------------------------------- receptive fields -- */

/* Mixed Gabor/DoG bank on a Px x Py grid -> out (N x Px*Py, row stride ld).
 * kind[j]: 0 Gabor, 1 DoG.  alpha_inv = 1/h_alpha.
 * Replaces vrf_gabor/vrf_cs.generate_visual_receptive_fields (vtem/vrf.py:296-331,
 * :876-913) and the gabor_Kernel / cs_Kernel they launch (vtem/vrf.py:1197-1291,
 * :1120-1193). */
int vt_rf_bank(void *out, int dtype, long long ld, int N, const int *kind,
               const double *alpha_inv, const double *theta, const double *x0,
               const double *y0, const int *ab, double kappa, double sigma_c2,
               double sigma_s2, int Px, int Py, double Sx, double Sy, double x_start,
               double y_start, void *stream);
/* Gabor-only bank: gabor_Kernel (vtem/vrf.py:1197-1291). */
int vt_rf_gabor(void *out, int dtype, long long ld, int N, const double *alpha_inv,
                const double *theta, const double *x0, const double *y0, const int *ab,
                double kappa, int Px, int Py, double Sx, double Sy, double x_start,
                double y_start, void *stream);
/* DoG-only bank: cs_Kernel (vtem/vrf.py:1120-1193). */
int vt_rf_dog(void *out, int dtype, long long ld, int N, const double *alpha_inv,
              const double *x0, const double *y0, double sigma_c2, double sigma_s2, int Px,
              int Py, double Sx, double Sy, double x_start, double y_start, void *stream);

/* ---------------------------------------------------------------- dense algebra ----- */

/* C[b] = alpha * A[b](m,k) B[b](k,n) + beta * C[b]; A(m,k) = A[m*sam + k*sak],
 * B(k,n) = B[k*sbk + n*sbn], C[m*ldc + n]; batch strides bsa/bsb/bsc.
 * Replaces the cuBLAS GEMMs of the path: filtering (vtem/vrf.py:291, :871), the
 * spatial Gram Zgemm (vtem/vrf.py:529, :1112; vtem/dirichlet.py:109-112) and
 * vtem/utils/linalg.py dot (:32-338). */
int vt_gemm_f64(int M, int N, int K, double alpha, const double *A, long long sam,
                long long sak, long long bsa, const double *B, long long sbk, long long sbn,
                long long bsb, double beta, double *C, long long ldc, long long bsc, int batch,
                void *stream);
int vt_gemm_f32(int M, int N, int K, float alpha, const float *A, long long sam, long long sak,
                long long bsa, const float *B, long long sbk, long long sbn, long long bsb,
                float beta, float *C, long long ldc, long long bsc, int batch, void *stream);
/* C[m,n] = alpha sum_k A[k*lda+m] B[k*ldb+n] + beta C[m,n] for narrow outputs (Nc <= 32)
 * and long K: split over 16 K-chunks into workspace (16*M*Nc doubles), then summed in a
 * fixed order (deterministic).  Used for S^T t (the D-space right
 * hand side, replacing the reference's per-interval q vector use in
 * vtem/utils/linalg.py:623). */
int vt_gemm_tn_small(int M, int Nc, int K, double alpha, const double *A, long long lda,
                     const double *B, long long ldb, double beta, double *C, long long ldc,
                     double *workspace, void *stream);
/* One MFMA of a given kind (0 fp64 16x16x4, 1 fp32 16x16x4, 2 bf16 16x16x32,
 * 3 f16 16x16x32) on 16xK / Kx16 inputs -> 16x16 out.  Test hook for the lane layouts;
 * no reference counterpart. */
int vt_mfma_probe(int kind, const void *A, const void *B, void *out, void *stream);

/* ---------------------------------------------------------------- IAF encoder ------- */

/* Linear IAF population encoder, one lane per neuron.  input (size x N, row stride
 * input_ld, dtype in_dtype); spikes (max_spike x N, row stride spike_ld) fp64; v0 and
 * time_count carried in/out; spike_count out (a count > max_spike - 1 means overflow,
 * as the reference's clamp, and the host raises).  root_mode 0 = the reference's
 * crossing formula, 1 = the cancellation-free equivalent.
 * Replaces IAF_encode.encode and its ensemble_encode kernel
 * (vtem/ensemble_encode.py:106-168, :438-552). */
int vt_iaf_encode(const void *input, int in_dtype, long long input_ld, int N, int size,
                  double *spikes, long long spike_ld, int max_spike, double *v0,
                  const double *kappa, const double *bias, const double *delta,
                  double *time_count, int *spike_count, double dt, int root_mode,
                  void *stream);
/* Same encoder with the input given as temporal trigonometric coefficients coef
 * (Dt x N, row stride coef_ld) and the basis table basis (size x Dt): the dt-rate input
 * is evaluated in registers.  Replaces the frame-rate input of VTEM_IAF
 * (vtem/vtem.py:176-200) with exact interpolation to dt; no reference counterpart for
 * the interpolation itself. */
int vt_iaf_encode_trig(const double *coef, long long coef_ld, const double *basis, int Dt,
                       int N, int size, double *spikes, long long spike_ld, int max_spike,
                       double *v0, const double *kappa, const double *bias,
                       const double *delta, double *time_count, int *spike_count, double dt,
                       int root_mode, void *stream);
/* Time-parallel forms of vt_iaf_encode and vt_iaf_encode_trig: same arguments, state
 * semantics and outputs (counts equal, spike times equal to rounding; the reference's
 * IAF_encode.encode, vtem/ensemble_encode.py:106-168, kernel :438-552).  Spike k is the
 * first passage of the integrated input to the level k*delta - v0, so the samples are cut
 * into segments integrated in parallel, joined by a per-neuron scan of (area, maximum),
 * and only segments holding spikes are walked again for the crossing times.  ws holds
 * vt_iaf_par_workspace_size(N, size) bytes of device memory. */
int vt_iaf_encode_trig_par(const double *coef, long long coef_ld, const double *basis, int Dt,
                           int N, int size, double *spikes, long long spike_ld, int max_spike,
                           double *v0, const double *kappa, const double *bias,
                           const double *delta, double *time_count, int *spike_count,
                           double dt, int root_mode, void *ws, long long ws_bytes,
                           void *stream);
int vt_iaf_encode_par(const void *input, int in_dtype, long long input_ld, int N, int size,
                      double *spikes, long long spike_ld, int max_spike, double *v0,
                      const double *kappa, const double *bias, const double *delta,
                      double *time_count, int *spike_count, double dt, int root_mode, void *ws,
                      long long ws_bytes, void *stream);
/* Host query: workspace bytes of one time-parallel encoder call. */
int vt_iaf_par_workspace_size(int N, int size, long long *bytes_host);
/* vt_iaf_encode_par without the workspace argument: the segment workspace is a static
 * device array of vt_noalloc_scratch_bytes() bytes per GPU (64 MiB), so the call allocates
 * nothing.  No counterpart in the reference: it is for callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  Returns -5 when vt_iaf_par_workspace_size(N, size) exceeds the array.  All
 * calls on one GPU share its array: calls that overlap on one GPU race. */
int vt_iaf_encode_par_noalloc(const void *input, int in_dtype, long long input_ld, int N,
                              int size, double *spikes, long long spike_ld, int max_spike,
                              double *v0, const double *kappa, const double *bias,
                              const double *delta, double *time_count, int *spike_count,
                              double dt, int root_mode, void *stream);
/* Host query: bytes of the per-GPU array of vt_iaf_encode_par_noalloc. */
int vt_noalloc_scratch_bytes(long long *bytes_host);
/* Batched time-parallel trig encoder: nd descriptors (device array) of the layout
 *   struct { const double *coef; long long coef_ld, coef_ns; const double *basis;
 *            const void *input; long long input_ld;          (unused here)
 *            double *spikes; long long spike_ld; double *v0, *time_count;
 *            const double *kappa, *bias, *delta; int *spike_count; void *ws;
 *            int *overflow; double dt; int N, size, Dt, max_spike, root_mode, in_dtype;
 *            double *ybuf; long long y_ld; }
 * (176 bytes), each one an independent vt_iaf_encode_trig_par call with coefficient pt of
 * neuron n at coef[pt*coef_ld + n*coef_ns]; *overflow (if not null) is set to 1 when a
 * neuron fires more than max_spike times.  maxN, maxSize and maxDt bound the descriptors'
 * N, size and Dt.  flags bit 0: every descriptor's ybuf (size x N, ld y_ld >= N) receives
 * the scaled input y = (v + b)/kappa at every sample in the first pass, and the crossing
 * walk reads it back instead of evaluating the input again (same results; 8 bytes of
 * device memory per neuron and sample); with flags 0 ybuf is unused.  Encodes many blocks
 * (served requests, stitching tiles) in one launch sequence; the reference encodes one
 * video at a time (vtem/vtem.py:165-200). */
int vt_iaf_encode_trig_batched(const void *descs, int nd, int maxN, int maxSize, int maxDt,
                               int flags, void *stream);
/* Host query: the descriptor size above, for layout checks. */
int vt_iaf_par_desc_size(int *bytes_host);
/* Random-threshold IAF: spikes holds N(0,1) draws on entry; delta_value carried in/out.
 * Replaces IAF_encode_rt.encode (vtem/ensemble_encode.py:265-329, kernel :555-674). */
int vt_iaf_encode_rt(const double *input, long long input_ld, int N, int size, double *spikes,
                     long long spike_ld, int max_spike, double *v0, const double *kappa,
                     const double *bias, const double *delta_mean, double *time_count,
                     int *spike_count, double dt, double *delta_value, const double *sigma,
                     int root_mode, void *stream);

/* ---------------------------------------------------------------- intervals, q ------- */

/* Spike buffer -> intervals of a window [t0, t0+Tw] (times shifted by t0; per neuron
 * the spikes in (0, Tw) plus one before and one after), CSR by neuron: offsets (N+1,
 * offsets[N] = n on device), tk1, tk2, neuron_ind (capacity entries).  workspace: 2N
 * ints.  Replaces vtem_storage_read.read_select_spikes (vtem/utils/spike_store.py:279-351). */
int vt_intervals(const double *spikes, long long spike_ld, int max_rows,
                 const int *spike_count, int N, double t0, double Tw, int *workspace,
                 int *offsets, double *tk1, double *tk2, int *neuron_ind, int capacity,
                 void *stream);
/* offsets[0..n] = exclusive scan of in[0..n) (one workgroup). */
int vt_exclusive_scan(const int *in, int n, int *offsets, void *stream);
/* q_k = (kappa_j delta_j - b_j (tk2-tk1)) norm_j; n from n_dev when given.
 * Replaces dirichlet.compute_q / compute_q_Kernel (vtem/dirichlet.py:55-60, :204-222). */
int vt_q(double *q, const double *tk1, const double *tk2, const int *neuron_ind,
         const double *kappa, const double *delta, const double *bias, const double *norm,
         const int *n_dev, int n, void *stream);
/* alpha (n x Dt): integrals of the temporal basis over each interval (Lt <= 32). */
int vt_temporal_factors(double *alpha, const double *tk1, const double *tk2, int Lt, double W,
                        const int *n_dev, int n, void *stream);
/* Per-neuron temporal moments T_j = sum norm^2 alpha alpha^T (N x Dt x Dt) and
 * t_j = sum norm alpha q (N x Dt), Lt <= 8.  Together with vt_gram_f64 this replaces
 * dirichlet.compute_G (vtem/dirichlet.py:62-85, kernel :272-330). */
int vt_neuron_T(double *Tm, double *tv, const double *tk1, const double *tk2, const double *q,
                const double *norm, const int *offsets, int N, int Lt, double W, void *stream);
/* Batched over independent problems described by a device descriptor array. */
int vt_neuron_T_batched(double *Tm, double *tv, const double *tk1, const double *tk2,
                        const double *q, const double *norm, const int *offsets,
                        const void *descs_dev, int n_desc, int max_neurons, void *stream);

/* ---------------------------------------------------------------- Gram / operator --- */

/* A = sum_j (S_j S_j^T) kron T_j on FP64 MFMA, lower block triangle of A (D x D, D =
 * Ds*Dt).  Replaces compute_Dsw + compute_G (vtem/vrf.py:502-536,
 * vtem/dirichlet.py:62-85) through the push-through identity. */
int vt_gram_f64(const double *S, long long lds, const double *Tm, int N, int Ds, int Dt,
                double *A, long long lda, void *stream);
/* Split one block's T (N x Dt^2 doubles) for the low-precision Gram into the tiled 16-bit
 * layout the Gram kernel stages with 16-byte copies: per 32-neuron chunk a contiguous
 * [NC][32] tile, NC = 16 * column tiles of Dt^2 (zero padded); the element count per array
 * comes from vt_gram_split_T_size.  mode 0 -> hi/lo bf16 arrays; mode 1 -> fp16(T*sT) in
 * hi, with power-of-two scales [sT, sW, 1/(sT sW)] computed on device from max|T| and
 * max|S| (S: nS doubles) into `scales` (5 floats: the last two are reduction scratch).
 * Dt <= 17 (else -5).  No reference
 * counterpart (the reference has only fp32/fp64 paths). */
int vt_gram_split_T(const double *Tm, int N, int Dt, void *hi, void *lo, int mode, const double *S,
                    long long nS, float *scales, void *stream);
/* 16-bit elements of one block's tiled T (per array), written to the HOST pointer. */
int vt_gram_split_T_size(int N, int Dt, long long *elems_host);
/* Batched low-precision Gram (mode 0 split-BF16 3-product, mode 1 FP16) on the 16x16x32
 * MFMAs with fp32 accumulation into fp32 A (lower block triangle).  descs: device array of
 * {const double *S; const u16 *Th, *Tl; float *A; long long lds, lda; int N, Ds, Dt, pad;
 * const float *scales;} (72 bytes); all problems share Dt.  Same role as vt_gram_f64.
 * tile: row tiles of 16 per wave (2, 4 or 8; 0 = 2 for split-BF16, 4 for FP16).  Every
 * tiling gives a bit-identical A; vtem_amd/autotune.py picks the fastest per shape.
 * ws (ws_bytes >= vt_gram_lp_ws_size): a single-problem launch whose last round of pair
 * blocks would leave most CUs idle splits those blocks over up to 8 neuron ranges, FP32
 * partials in ws summed in order by a second kernel (A then differs from the unsplit one
 * by FP32 reassociation only, and not between tilings); null: no split. */
int vt_gram_lp(const void *descs, int batch, int max_Ds, int Dt, int mode, int tile,
               void *ws, long long ws_bytes, void *stream);
/* Bytes of workspace the tail split of such a launch needs (0: it does not split),
 * written to the host pointer. */
int vt_gram_lp_ws_size(int batch, int max_Ds, int Dt, int mode, int tile, long long *bytes_host);
/* s2[j] = |S_j|^2 (once per neuron set). */
int vt_row_sumsq(const double *S, long long lds, int N, int Ds, double *s2, void *stream);
/* dst (cols x rows, row stride ldd) = src^T (rows x cols, row stride lds).  Makes the
 * decoder's resident S^T (Ds x N), so the operator's second product Q = S^T Z streams rows
 * as the first does.  No reference counterpart (the reference forms no such operator). */
int vt_transpose_f64(const double *src, long long lds, int rows, int cols, double *dst,
                     long long ldd, void *stream);
/* dst = A in the skinny product's MFMA tile layout, A = src (rows x cols, row stride lds)
 * or, with trans = 1, its transpose: for A of M x K, ceil(M/16) * ceil(K/16) tiles of 256
 * doubles, tile (m/16, k/16) holding A[m][k] at (k%16 %4 / 2) * 128 + 2 (16 ((k%16)/4) + m%16)
 * + k%2, zeros beyond M and K.  Each lane of a wave then reads its operands of a 16 x 16
 * tile as two 16-byte loads that cover 1 KB in one piece.  Makes the decoder's resident S
 * and S^T operands.  No reference counterpart (the reference forms no such operator). */
int vt_tile_f64(const double *src, long long lds, int rows, int cols, int trans, double *dst,
                void *stream);
/* Exact tr(A) = sum_j s2_j tr(T_j) in fp64; lam_out = {lam_rel tr(A)/n, tr(A)}.  Replaces
 * the trace needed for lambda = 1e-6 tr(G)/n without forming G or A. */
int vt_kron_trace(const double *s2, const double *Tm, int N, int Dt, const int *n_dev,
                  double lam_rel, double *lam_out, void *stream);
/* Batched over problems b: N[b], s2 + s2_off[b], Tm + T_off[b]*Dt*Dt, n_dev[b*n_stride];
 * lam_out[2b], lam_out[2b+1]. */
int vt_kron_trace_batched(const double *s2, const double *Tm, const int *Nb,
                          const long long *s2_off, const long long *T_off, int Dt,
                          const int *n_dev, int n_stride, double lam_rel, double *lam_out,
                          int batch, void *stream);
/* Batched FP64 Gram: descs = {const double *S, *Tm; double *A; long long lds, lda; int N,
 * Ds, Dt, pad;} per problem (56 bytes), common Dt. */
int vt_gram_f64_batched(const void *descs, int batch, int max_Ds, int Dt, void *stream);
/* Copy the lower triangle into the upper one. */
int vt_symmetrize(void *A, int dtype, long long lda, int D, void *stream);
/* Z[j,:] = T_j Y[j,:] (structured operator step). */
int vt_tj_apply(double *Z, const double *Tm, const double *Y, int N, int Dt, void *stream);
/* out = a x + b y + c z (y, z may be null). */
int vt_axpbypcz(double *out, double a, const double *x, double b, const double *y, double c,
                const double *z, long long n, void *stream);
/* res = r - AX - lam_dev[0] X. */
int vt_residual(double *res, const double *r, const double *AX, const double *X,
                const double *lam_dev, long long n, void *stream);
/* out[0] = sum x^2. */
int vt_sumsq(const double *x, long long n, double *out, void *stream);
/* lam_out[0] = lam_rel * trace(A) / n_dev[0], lam_out[1] = trace(A). */
int vt_trace_lambda(const void *A, int dtype, long long lda, int D, const int *n_dev,
                    double lam_rel, double *lam_out, void *stream);
/* A += lam I (lam from lam_dev[0] when given).  Replaces diag_add_Kernel
 * (vtem/dirichlet.py:83-85, :334-350). */
int vt_diag_add(void *A, int dtype, long long lda, int D, double lam, const double *lam_dev,
                void *stream);

/* ---------------------------------------------------------------- solves ------------- */

/* Blocked Cholesky (block 64) of a batch of matrices described by a device array of
 * {T *A; long long lda; int D; int pad;} records; inv_ws receives the diagonal-block
 * inverses (per matrix ceil(maxD/64)*4096 elements, inv_stride apart); info[b] = first
 * failing pivot + 1 (caller zeroes it).  Replaces linalg.solve_eq_sym's factorisation
 * step (CULA gesvd, vtem/utils/linalg.py:591-631) with the Tikhonov-regularised form. */
int vt_potrf_f64(const void *mats_dev, int batch, int maxD, double *inv_ws,
                 long long inv_stride, int *info, void *stream);
int vt_potrf_f32(const void *mats_dev, int batch, int maxD, float *inv_ws,
                 long long inv_stride, int *info, void *stream);
/* x <- (L L^T)^-1 x for each matrix of the batch (x fp64, x_stride apart), one workgroup
 * per 64-row block with flag-synchronised forward/backward sweeps.  sync: 2 +
 * 2*batch*ceil(maxD/64) ints zeroed once by the caller and reused (sync[0] is an epoch
 * bumped on device, sync[1] != 0 reports an expired bounded wait).  Replaces the U^H q /
 * U qq products of solve_eq_sym (vtem/utils/linalg.py:623-630). */
int vt_potrs_f64(const void *mats_dev, int batch, int maxD, const double *inv_ws,
                 long long inv_stride, double *x, long long x_stride, int *sync, void *stream);
int vt_potrs_f32(const void *mats_dev, int batch, int maxD, const float *inv_ws,
                 long long inv_stride, double *x, long long x_stride, int *sync, void *stream);
/* fp32 copy of the lower triangle of an fp64 matrix, lam_dev[0] added to the diagonal. */
int vt_lower_to_f32(float *dst, long long ldd, const double *src, long long lds, int D,
                    const double *lam_dev, void *stream);

/* ---------------------------------------------------------------- spike-space path --- */

/* G (n x n) = Gt(k,l) Dsw[j(k), j(l)], Gt the double integral of the temporal Dirichlet
 * kernel over intervals k and l, formed as the dot product of the two intervals' integrals
 * of the real temporal basis (Mt <= 1024, else -5).  Replaces compute_G_Kernel
 * (vtem/dirichlet.py:272-330). */
int vt_G_spike(double *G, long long ldg, const double *tk1, const double *tk2, double Wt,
               int Mt, const double *Dsw, long long ldw, const int *neuron_ind, int n,
               void *stream);
/* y = G x (one wave per row). */
int vt_gemv(double *y, const double *G, long long ldg, const double *x, int n, void *stream);
/* steps+1 iterations of c += eta (q - G c).  Replaces neural_network.rnn3
 * (vtem/neural_network.py:15-99, update kernel :102-128). */
int vt_rnn_spike(double *c, const double *G, long long ldg, const double *q, int n,
                 double eta, int steps, double *work_Gc, double *work_y, void *stream);
/* u(t,p) for T output times.  Replaces dirichlet.reconstruct / reconstruct_Kernel
 * (vtem/dirichlet.py:165-189, :354-420). */
int vt_reconstruct_spike(double *u, long long ldu, const double *psi, long long ldpsi,
                         const double *tk1, const double *tk2, const double *c,
                         const double *times, int T, const int *neuron_ind,
                         const double *norm, int Mt, double WM, int n, int P, void *stream);

/* ---------------------------------------------------------------- batched stages ----- */
/* Many independent blocks per launch (tiles/windows of a stitched video, served
 * requests); the reference runs them one after another (vtem/vtdm.py:236-302, :630-692).
 * Neurons are numbered globally across the batch. */

/* Intervals of all blocks: spike_descs = {const double *spikes; const int *counts;
 * long long ld; int max_rows, neuron_base; double t0, Tw;} per block (48 bytes);
 * job_of[g] = block of global neuron g; offsets (Ntot+1) global CSR; neuron_ind global. */
int vt_intervals_batched(const void *spike_descs, const int *job_of, int Ntot, int *workspace,
                         int *offsets, double *tk1, double *tk2, int *neuron_ind, int capacity,
                         void *stream);
/* Grouped fp64 GEMM: descs = {const double *A, *B; double *C; long long sam, sak, sbk,
 * sbn, ldc; int M, N, K, pad; double alpha, beta;} per problem (96 bytes). */
int vt_gemm_grouped(const void *descs, int batch, int max_M, int max_N, void *stream);
/* Batched skinny A^T B: descs = {const double *A, *B; double *C, *ws; long long lda, ldb,
 * ldc; int M, K;} (64 bytes), ws: 16*M*Nc doubles per problem. */
int vt_gemm_tn_batched(const void *descs, int batch, int max_M, int Nc, void *stream);
/* res = r - AX - lam[2*elem_job[i]] X over a concatenated vector. */
int vt_residual_batched(double *res, const double *r, const double *AX, const double *X,
                        const double *lam, const int *elem_job, long long n, void *stream);
/* out[b*out_stride] = sum_{seg[b] <= i < seg[b+1]} x_i^2. */
int vt_seg_sumsq(const double *x, const long long *seg, int nseg, double *out, int out_stride,
                 void *stream);
/* A_b += lam[2b] I; descs = {void *A; long long lda; int D, pad;} (24 bytes). */
int vt_diag_add_batched(const void *descs, int batch, int dtype, const double *lam, int max_D,
                        void *stream);
/* Skinny grouped product: C_b (M x Nc) = A_b B_b, Nc <= max_Nc <= 32 per problem, with
 * A_b(m,k) at A[m*sam + k*sak], B and C row-major; if T != 0 each row is then multiplied
 * by T + m*tstride (C_m <- T_m C_m); a problem whose `active` flag (optional double) reads
 * 0 is skipped.  descs = {const double *A, *B; double *C; const double *T; double *ws;
 * const double *active; long long sam, sak, ldb, ldc, tstride; int M, K, Nc, pad;} (104
 * bytes).  A problem with K > 1008 splits its reduction (one part per 512 terms, at most
 * 8, a function of its own K) and needs the ws that vt_gemm_skinny_workspace_size
 * reports; max_K bounds every problem's K.  The structured operator S^T (T_j (S X)) of the coefficient-space
 * Gram (vtem/dirichlet.py:62-85 builds that Gram in spike space) is two such products. */
int vt_gemm_skinny(const void *descs, int batch, int max_M, int max_K, int max_Nc, void *stream);
/* The skinny product's K-split rule, ks = clamp(ceil(K/16) / chunks_per_split, 1,
 * max_splits), for a host that sizes the split-K workspaces of many problems at once
 * (vt_gemm_skinny_workspace_size gives one problem's).  Host only.  No reference
 * counterpart (the reference forms no such operator). */
int vt_skinny_split_rule(int *chunks_per_split_host, int *max_splits_host);
/* Doubles of ws one skinny problem (M x K, Nc columns) needs, written to the HOST pointer
 * doubles_host (0 when its reduction is not split). */
int vt_gemm_skinny_workspace_size(int M, int K, int Nc, long long *doubles_host);
/* x += y. */
int vt_add_inplace(double *x, const double *y, long long n, void *stream);
/* out[b] = offsets[base[b+1]] - offsets[base[b]] (intervals per block). */
int vt_seg_counts(const int *offsets, const int *base, int nseg, int *out, void *stream);

/* ------------------------------------------- Kronecker-preconditioned CG (batched) -- */
/* Matrix-free solve of (A + lam I) X = r with A = sum_j S_j S_j^T (x) T_j and the
 * preconditioner (S^T C S + lam I) (x) Tbar, c_j = tr(T_j)/Dt, Tbar = sum_j T_j / sum_j c_j.
 * It solves the same Tikhonov system as the Cholesky path and replaces, for that system,
 * the reference's pseudoinverse solve (vtem/utils/linalg.py:591-631, called from
 * vtem/vtdm.py:651) and its Gram build (vtem/dirichlet.py:62-85), which it never forms. */
/* Tbar inverse per block, stored transposed (TiT, batch x Dt x Dt), and c_j (global neuron
 * numbering, nbase has batch+1 entries); ws: 16*batch*Dt*Dt doubles.  Dt <= 33 (else -5). */
int vt_kpcg_tbar(const double *Tm, const int *nbase, int Dt, int batch, double *ws, double *c,
                 double *Ti, void *stream);
/* M_b = S_b^T diag(c_b) S_b; descs = {const double *S, *c; double *M; long long lds, ldm;
 * int N, Ds;} per block (48 bytes). */
int vt_kpcg_mgram(const void *descs, int batch, int max_Ds, void *stream);
/* Diagonal 64x64 blocks of L_b^-1 from the Cholesky block inverses (inv_ws +
 * b*inv_stride + i*4096) into X + xoff[b] (dim[b] x dim[b], leading dimension dim[b]). */
int vt_kpcg_place_diag(double *X, const long long *xoff, const int *dim, int batch, int max_dim,
                       const double *inv_ws, long long inv_stride, void *stream);
/* CG recurrences over concatenated per-block vectors (block b owns [seg[b], seg[b+1]));
 * st: 4 doubles per block {rz, |r0|^2, active, iterations}; hist[b*hist_stride + k] =
 * |r_k|^2.  init: x = 0, p = z, rz = <r,z>.  step_a: q += lam p, alpha = rz/<p,q>,
 * x += alpha p, r -= alpha q, stop when |r|^2 <= stop_rel2 |r0|^2.  step_b: beta =
 * <r,z>/rz, p = z + beta p.  Stopped blocks are left unchanged by later steps.  pt
 * (optional, with ptab): p also kept in the forward product's tiled B layout (SK_B_TILED,
 * the tiles of block b's P^T, Dt x Ds) at pt + ptab[2b], ptab[2b+1] = block b's Dt; init
 * and step_b write each element there too (the padding is the caller's zeros). */
int vt_kpcg_init(const double *r, const double *z, double *p, double *x, const long long *seg,
                 int batch, double *st, double *hist, int hist_stride, double *pt,
                 const long long *ptab, void *stream);
/* it = -1: each block takes its iteration index from its own count st[4b+3] (equal to
 * the loop index while it is active), so every iteration has the same arguments.
 * prune > 0 (>= 10): from iteration `prune` on, every 10 iterations, a block whose last
 * 10 iterations' residual decay, continued geometrically, would not reach the stop by
 * iteration `cap` is stopped too (its result then misses the caller's guard); 0: off. */
int vt_kpcg_step_a(double *x, double *r, const double *p, double *q, const double *lam,
                   const long long *seg, int batch, double *st, double *hist, int hist_stride,
                   int it, double stop_rel2, int prune, int cap, void *stream);
int vt_kpcg_step_b(const double *r, const double *z, double *p, const long long *seg, int batch,
                   double *st, double *pt, const long long *ptab, void *stream);

/* `iters` CG iterations after vt_kpcg_init, enqueued by one call: per iteration the
 * operator (skinny products fwd then bwd, q = A p) with step_a fused into the backward
 * launch (the last row tile of a block to finish runs it), the preconditioner (skinny
 * products pre, z = Mi r Ti) and step_b.  Each product is given as its descriptors and
 * its max_M / max_K; max_Nc bounds every problem's column count; bwd's descriptors are
 * in block order.  cnt: batch * (ceil(bwd_M / 16) + 1) words of counters (zeroed here).
 * graph_chunk >= 2 replays the iterations as a HIP graph of up to graph_chunk (at most
 * iters/2) iterations, as vt_kron_rnn does; 0 launches each.  Same results either way.
 * pt, ptab: p's tiled copy as in vt_kpcg_init, for fwd descriptors that read it. */
int vt_kpcg_iterate(const void *fwd, int fwd_M, int fwd_K, const void *bwd, int bwd_M, int bwd_K,
                    const void *pre, int pre_n, int pre_M, int pre_K, int pre_Nc, const void *pre2,
                    int pre2_n, int pre2_M, int pre2_K, int pre2_Nc, const void *pre3, int pre3_n,
                    int pre3_M, int pre3_K, int pre3_Nc, int batch, int max_Nc, double *x,
                    double *r, double *p, double *q, double *z, const double *lam,
                    const long long *seg, double *st, double *hist, int hist_stride, int iters,
                    double stop_rel2, int prune, int graph_chunk, unsigned *cnt, double *pt,
                    const long long *ptab, void *stream);
/* Kronecker-sum preconditioner (S^T C S) (x) Tbar + lam I for blocks whose decode period
 * is longer than their window: c_j as vt_kpcg_tbar, and per block Tbar's eigenvalues theta
 * (batch x Dt) and eigenvectors U (eigenvector k in column k of the Dt x Dt block; UT its
 * transpose), by two-sided cyclic Jacobi.  Then P^-1 R = sum_k (theta_k K + lam I)^-1
 * (R u_k) u_k^T; vt_kpcg_shift_pool forms the Dt matrices theta_k K_b + lam_b I of every
 * block (K_b at kpool + kbase[b], matrix k of block b at opool + obase[b] + k Ds_b^2). */
int vt_kpcg_tbar_eig(const double *Tm, const int *nbase, int Dt, int batch, double *ws,
                     double *c, double *theta, double *U, double *UT, void *stream);
int vt_kpcg_shift_pool(const double *kpool, const long long *kbase, const long long *obase,
                       const int *ds, const double *theta, const double *lam, int Dt, int batch,
                       int max_Ds, double *opool, void *stream);

/* ----------------------------------------- gradient-flow (RNN) decoder, batched -- */
/* Coefficient-space image of neural_network.rnn3 (vtem/neural_network.py:15-99, update
 * kernel :102-128): X <- X + eta_b (r - (A + lam_b I) X), steps+1 iterations from X = 0,
 * A applied through the skinny operator launches fwd (reads X) / bwd (writes q, or its
 * split partials), the backward launch doing the update too: the last of a row tile's K
 * splits to finish adds the partials in split order and updates those rows.  bwd's
 * descriptors are in block order; seg (batch + 1) gives each block's range of the
 * concatenated X; cnt is a workspace of batch * ceil(bwd_M / 16) words (zeroed here).
 * graph_chunk >= 2 replays the iterations as a HIP graph of graph_chunk iterations,
 * captured once per call on a private stream of the current device (one host call per
 * replay instead of two launches per iteration); 0 launches them one by one.  Results
 * are identical.  xt (optional; n_xt doubles, zeroed here): X also kept as the forward
 * product's tiled B operand (descriptor flag SK_B_TILED: the tiles of X_b^T, block b's at
 * xt + xtoff[b]), written by each update; null when the forward reads X row-major.  The
 * layout changes no value or sum order: iterates are identical either way. */
int vt_kron_rnn(const void *fwd, int fwd_M, int fwd_K, const void *bwd, int bwd_M, int bwd_K,
                int batch, int max_Nc, double *x, const double *r, double *q, const double *lam,
                const double *eta, const long long *seg, long long n, int steps, int graph_chunk,
                unsigned *cnt, double *xt, const long long *xtoff, long long n_xt, void *stream);
/* The same steps+1 iterations in ONE persistent launch (vtem/neural_network.py:78-92, update
 * kernel :102-128: the loop the reference drives from the host, one launch per step).  Each
 * block owns the workgroups [gstart[b], gstart[b+1]) (device array, batch + 1 ints, grid =
 * gstart[batch] within vt_kron_rnn_persistent_capacity, groups of 1..max_group); a
 * block's group runs its forward tiles, backward tiles and update with barriers of that
 * group only.  fwd / bwd: the operator descriptors in block order (forward K <= 1008, so
 * unsplit); max_Nc their widest column count.  x (n doubles) is zeroed here and holds
 * X_{steps+1} on return, bit-identical to vt_kron_rnn.  X and Z pass through ring slots
 * (ring_x: ring slots of slot_x >= n doubles, ring_z: ring slots of slot_z >= the joint Z,
 * slot sizes multiples of 32); ring >= steps + 1 needs no cache invalidation, a shorter
 * ring (even, >= 4) invalidates every ring/2 iterations.  bar: barrier workspace of
 * vt_kron_rnn_persistent_bar_words words, zeroed here; bar[0] != 0 afterwards means a
 * barrier wait exceeded limit_s (the grid was not co-resident) and x is invalid.  No
 * counterpart call in the reference (its loop is host-driven); this replaces vt_kron_rnn's
 * launches, not its arithmetic. */
int vt_kron_rnn_persistent(const void *fwd, const void *bwd, int batch, int max_Nc, const int *gstart,
                           int grid, double *x, long long n, const double *r, const double *zbase,
                           const double *lam, const double *eta, const long long *seg, int steps,
                           double *ring_x, long long slot_x, double *ring_z, long long slot_z, int ring,
                           unsigned *bar, double limit_s, void *stream);
/* Workgroups of the persistent RNN launch that fit the current device at once:
 * max_wgs[0] for batches of at most 17 columns, max_wgs[1] for wider ones; *max_group the
 * largest group of one block. */
int vt_kron_rnn_persistent_capacity(int *max_wgs, int *max_group);
/* 32-bit words of the persistent launch's barrier workspace for a batch. */
int vt_kron_rnn_persistent_bar_words(int batch, long long *words);

/* Step sizes eta_b = 1/theta_b, theta_b the largest Ritz value of A_b + lam_b I after
 * `steps` Lanczos steps from the normalised all-ones vector (v, vp: two work vectors of
 * the joint length; q: the backward's output): u = (A + lam I) v_j, alpha_j = <u, v_j>,
 * w = u - alpha_j v_j - beta_{j-1} v_{j-1}, beta_j = |w|, v_{j+1} = w / beta_j, fused into
 * the backward launch (two launches per step), then theta_b by bisection on the Sturm
 * count of the tridiagonal.  The reference steps its ODE by dt*alpha (vtem/vtdm.py:21,
 * vtem/neural_network.py:78-92); 1/lambda_max is the configurations' rule.  bwd's
 * descriptors in block order; cnt: batch * ceil(bwd_M / 16) + 2 batch words (zeroed
 * here); part: batch * ceil(bwd_M / 16) doubles; tri: 2 * steps * batch doubles.  * vt, vtab (optional): v also kept as the forward product's tiled B (as p in
 * vt_kpcg_init: block b's tiles at vt + vtab[2b], vtab[2b+1] its Dt), written with v. */
int vt_kron_lanczos(const void *fwd, int fwd_M, int fwd_K, const void *bwd, int bwd_M, int bwd_K,
                    int batch, int max_Nc, double *v, double *vp, double *q, const double *lam,
                    const long long *seg, int steps, double *eta, int graph_chunk, unsigned *cnt,
                    double *part, double *tri, double *vt, const long long *vtab, void *stream);


/* ------------------------------------------ spatial Dirichlet side, pseudoinverse -- */
/* Mode table (P x 2M, row stride ld >= 2M, M = (2Mx+1)(2My+1)): row q*Px + p holds
 * [cos theta_m | -sin theta_m], theta_m = mx Ox x_p + my Oy y_q, x_p = x_start + p dx,
 * y_q = -(y_start + q dy), column m = (my+My)(2Mx+1) + (mx+Mx).  RF times this table is
 * the direct DFT that replaces the cuFFT path of compute_Ds (vtem/vrf.py:333-377,
 * fft2Ds kernel :1293-1350); [Re Ds | Im Ds] times its transpose is compute_dirich_space
 * (vtem/vrf.py:380-499, kernels :1353-1433). */
int vt_dirichlet_table(double *out, long long ld, int Px, int Py, double x_start, double dx,
                       double y_start, double dy, int Mx, int My, double Ox, double Oy,
                       void *stream);
/* d_Ds (N x (2My+1) x (2Mx+1) complex128, interleaved) = scale (C[:, m] + i C[:, M+m]) from
 * C = RF table (row stride ldc); mask 1 = the reference's ellipse (vtem/vrf.py:1331), 0 = all
 * modes. */
int vt_ds_assemble(double *Ds, const double *C, long long ldc, int N, int Mx, int My, int mask,
                   double scale, void *stream);
/* W[i,j] *= norm[i] norm[j] (put_norm of compute_Dsw, vtem/vrf.py:529-533, :540-559). */
int vt_scale_outer(double *W, long long ldw, int n, const double *norm, void *stream);
/* One-sided Jacobi (FP64) on the rows of A (n x n; initially the symmetric G, whose rows
 * are its columns): one sweep of n'-1 rounds (n' = n rounded up to even) of disjoint row
 * pairs, a pair rotated when |a_p.a_q| > tol |a_p| |a_q| and both |a|^2 > small2 (rows at
 * or below it are numerically zero and left alone); V (optional, null to skip)
 * accumulates the rotations.  *rotations (device int) counts the rotations applied; the
 * caller zeroes it (vt_zero_int) and sweeps until a sweep leaves it at 0.  At convergence
 * A = (G V)^T with rows a_j = s_j u_j: the singular values and left singular vectors of
 * G, which the reference gets from CULA's SVD (vtem/utils/linalg.py:366-523). */
int vt_svd_jacobi_sweep(double *A, long long lda, double *V, long long ldv, int n, double tol,
                        double small2, int *rotations, void *stream);
/* *x = 0 (device int). */
int vt_zero_int(int *x, void *stream);
/* c = sum_{s_j > rcond s_max} u_j (u_j.q) / s_j from the converged rows of A: the
 * truncated-SVD solve of solve_eq_sym (vtem/utils/linalg.py:591-631, sq_Kernel :790-831).
 * ws: 2n doubles. */
int vt_svd_pinv_apply(const double *A, long long lda, int n, const double *q, double rcond,
                      double *c, double *ws, void *stream);
/* Eigendecomposition from the converged rows of A = G + shift I (shift > |G|, so A is
 * positive definite) and the accumulated V: eigenvalues w_j = a_j.v_j - shift ascending in
 * w, eigenvector k in row k of E (row stride lde).  Replaces eig_sym
 * (vtem/utils/linalg.py:634-718).  ws: n doubles. */
int vt_svd_eig_finish(const double *A, long long lda, const double *V, long long ldv, int n,
                      double shift, double *w, double *E, long long lde, double *ws, void *stream);
/* Stitching: for each block descriptor {const double *blk, *wt, *wy, *wx; int T, Py, Px,
 * t0, y0, x0;} (56 bytes), out[t0+t, y0+y, x0+x] += wt[t] wy[y] wx[x] blk[t, y, x].  Blocks
 * of one call must not overlap.  Replaces the tile / window accumulation of decode_video
 * and VTDM (vtem/vtdm.py:236-302, :604-683). */
int vt_stitch_accumulate(double *out, long long ld_t, long long ld_y, const void *descs,
                         int nblocks, long long max_elems, void *stream);

#ifdef __cplusplus
}
#endif
#endif /* VTEM_HIP_H */
