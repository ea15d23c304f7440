// Persistent gradient-flow (RNN) decoder in coefficient space: every iteration of
//   X <- X + eta_b (r - (A_b + lam_b I) X),   A_b X = S_b^T (T_j (S_b X)_j)_j,
// the D-space image of the reference's c <- c + eta (q - G c) (vtem/neural_network.py:78-92,
// :102-128), in ONE launch instead of three launches per iteration.
//
// Each block of the batch owns a group of workgroups.  An iteration is three phases of the
// group, separated by barriers of that group only (blocks never wait for each other):
//   F: Z = T_j (S X)_j, the forward product's row tiles,
//   B: the backward product S^T Z, its row tiles and K splits (partials in the workspace),
//   U: X' = X + eta (r - sum of partials - lam X), elementwise.
// The tiles are the launch form's own (skinny_tile, same routes by the block's own Nc), the
// partials are added in split order and the step is rnn_step, so the iterates are those of
// vt_kron_rnn bit for bit; which workgroup computes a tile changes no sum.
//
// Coherence without cache flushes.  The eight XCDs of the device have separate L2 caches.
// Data one workgroup writes and another reads inside the launch is stored with relaxed
// agent-scope atomic stores (written through to the device coherence point).  X and Z live
// in ring slots: slot k % ring holds iteration k's X / Z and is written once before any
// workgroup reads it, so plain (cached) loads cannot meet a stale line when the ring holds
// every iteration (ring >= steps + 1).  A shorter ring (even, >= 4) invalidates each
// workgroup's caches (agent-scope acquire) at every iteration that is a multiple of
// ring/2, between a slot's last read and its next write.  The partials and the unsplit
// outputs, rewritten every iteration, are read with agent-scope atomic loads.  S, S^T, T_j,
// r and the descriptors are read-only here and stay cached.
//
// Barriers: counters in the caller's workspace, zeroed by this entry.  The group's ranks
// count in sub-groups of 16 (one 256-byte line each); the last of a sub-group counts on the
// group's line, the last of those releases the epoch.  Every wait is bounded by a wall-clock
// limit: a workgroup whose wait expires sets the abort word (line 0 of the workspace) and
// every workgroup then leaves; the iterate is then invalid and the caller runs the launch
// form.  The grid must be co-resident: the caller sizes it from
// vt_kron_rnn_persistent_capacity.
#include "vt_common.hpp"
#include "vt_skinny.hpp"

#define PR_LINE 64             // 32-bit words per 256-byte line
#define PR_BLOCK_LINES 20      // per block: counter, release flag, 16 sub-group counters, pad
#define PR_MAX_GROUP 256       // 16 sub-groups of 16

struct PrnnArgs {
    const SkDesc *fwd, *bwd;   // per block, in block order
    const int *gstart;         // batch + 1: block b owns workgroups [gstart[b], gstart[b+1])
    double *x;                 // joint iterate: zeroed here, X_{steps+1} on return
    const double *r, *lam, *eta;
    const long long *seg;      // block b owns [seg[b], seg[b+1]) of x, r
    const double *zbase;       // the joint Z buffer the forward descriptors point into
    double *ring_x, *ring_z;
    long long slot_x, slot_z;  // doubles per ring slot
    int ring, steps, batch, pad;
    unsigned *bar;             // line 0: abort word; PR_BLOCK_LINES lines per block after it
    unsigned long long limit;  // wall-clock ticks a barrier wait may last
};

__device__ __forceinline__ unsigned pr_load(const unsigned *p)
{
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ __forceinline__ double pr_loadd(const double *p)
{
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Barrier of the g workgroups of one block (bl: the block's lines); epoch e completes when
// every rank has arrived e + 1 times.  Returns false when the launch is aborting.
__device__ bool pr_barrier(unsigned *bl, unsigned *abortw, int rank, int g, unsigned epoch,
                           unsigned long long limit, int *sflag)
{
    __builtin_amdgcn_s_waitcnt(0);           // this wave's stores have completed
    __syncthreads();
    if (threadIdx.x == 0) {
        const int sg = rank >> 4, nsg = (g + 15) >> 4;
        const unsigned ss = (unsigned)min(16, g - 16 * sg);
        const unsigned o1 = __hip_atomic_fetch_add(bl + PR_LINE * (2 + sg), 1u, __ATOMIC_RELAXED,
                                                   __HIP_MEMORY_SCOPE_AGENT);
        if (o1 == epoch * ss + ss - 1) {
            const unsigned o2 = __hip_atomic_fetch_add(bl, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            if (o2 == epoch * (unsigned)nsg + (unsigned)nsg - 1)
                __hip_atomic_store(bl + PR_LINE, epoch + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
        int ok = 1;
        const unsigned long long t0 = wall_clock64();
        while (pr_load(bl + PR_LINE) < epoch + 1) {
            if (pr_load(abortw)) {
                ok = 0;
                break;
            }
            if (wall_clock64() - t0 > limit) {
                __hip_atomic_store(abortw, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                ok = 0;
                break;
            }
            __builtin_amdgcn_s_sleep(1);
        }
        *sflag = ok;
    }
    __syncthreads();
    return *sflag != 0;
}

// One 16-row tile of a product, on the route the problem's own Nc fixes (as skinny_kernel).
// WIDE: the batch has problems of more than 17 columns (two MFMA tiles); without them the
// kernel is compiled for the narrow routes only, which needs fewer registers.  The split
// index passes an empty asm so that a constant split 0 does not specialise the tile code
// (that specialisation takes more registers).
template <bool WIDE>
__device__ __forceinline__ void pr_tile(const SkDesc &d, int m0, int z, double (&red)[SK_WAVES / 2][16][33])
{
    asm volatile("" : "+s"(z));
    if (d.Nc == 17)
        skinny_tile<1, 1, true, true>(d, m0, z, red);
    else if (!WIDE || d.Nc <= 16)
        skinny_tile<1, 1, false, true>(d, m0, z, red);
    else if constexpr (WIDE)
        skinny_tile<2, 1, false, true>(d, m0, z, red);
}

template <bool RING_INV, bool WIDE>
__global__ void __launch_bounds__(64 * SK_WAVES) rnn_persistent_kernel(PrnnArgs a)
{
    __shared__ double red[SK_WAVES / 2][16][33];
    __shared__ int sflag;
    if ((int)blockIdx.x >= a.gstart[a.batch]) return;
    int b = 0;
    while (a.gstart[b + 1] <= (int)blockIdx.x) ++b;
    const int rank = (int)blockIdx.x - a.gstart[b], g = a.gstart[b + 1] - a.gstart[b];
    unsigned *abortw = a.bar, *bl = a.bar + PR_LINE * (1 + PR_BLOCK_LINES * b);
    if (g > PR_MAX_GROUP) {                 // a group the barrier lines cannot hold
        if (threadIdx.x == 0) __hip_atomic_store(abortw, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        return;
    }
    const SkDesc F0 = a.fwd[b], B0 = a.bwd[b];
    const long long s0 = a.seg[b], s1 = a.seg[b + 1];
    const long long zoff = F0.C - a.zbase;
    const double lam = a.lam[2 * b], eta = a.eta[b];
    const int ftiles = (F0.M + 15) / 16;
    const int ksb = skinny_splits(B0.K), btr = (B0.M + 15) / 16, bitems = btr * ksb;
    const long long mn = (long long)B0.M * B0.Nc;
    unsigned epoch = 0;
    for (int k = 0; k <= a.steps; ++k) {
        if (RING_INV && k > 0 && k % (a.ring / 2) == 0) __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
        const double *xin = k == 0 ? a.x : a.ring_x + (long long)(k % a.ring) * a.slot_x;
        double *zk = a.ring_z + (long long)(k % a.ring) * a.slot_z;
        // F: Z_k = T_j (S x_k)_j
        SkDesc f = F0;
        f.B = xin + s0;
        f.C = zk + zoff;
        for (int t = rank; t < ftiles; t += g) {
            pr_tile<WIDE>(f, 16 * t, 0, red);
            __syncthreads();
        }
        if (!pr_barrier(bl, abortw, rank, g, epoch++, a.limit, &sflag)) return;
        // B: S^T Z_k, its K splits into the workspace
        SkDesc bd = B0;
        bd.B = zk + zoff;
        for (int t = rank; t < bitems; t += g) {
            pr_tile<WIDE>(bd, 16 * (t % btr), t / btr, red);
            __syncthreads();
        }
        if (!pr_barrier(bl, abortw, rank, g, epoch++, a.limit, &sflag)) return;
        // U: x_{k+1} = x_k + eta (r - q - lam x_k), q summed in split order
        double *xout = k == a.steps ? a.x : a.ring_x + (long long)((k + 1) % a.ring) * a.slot_x;
        for (long long i = s0 + (long long)rank * blockDim.x + threadIdx.x; i < s1;
             i += (long long)g * blockDim.x) {
            const long long loc = i - s0;
            double qi;
            if (ksb > 1) {
                qi = 0.0;
                for (int zz = 0; zz < ksb; ++zz) qi += pr_loadd(B0.ws + zz * mn + loc);
            } else {
                qi = pr_loadd(B0.C + loc);
            }
            sk_store<true>(xout + i, rnn_step(xin[i], a.r[i], qi, lam, eta));
        }
        if (k < a.steps && !pr_barrier(bl, abortw, rank, g, epoch++, a.limit, &sflag)) return;
    }
}

__global__ void pr_zero_words(unsigned *w, long long n)
{
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (long long)gridDim.x * blockDim.x)
        w[i] = 0u;
}

__global__ void pr_zero_doubles(double *x, long long n)
{
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (long long)gridDim.x * blockDim.x)
        x[i] = 0.0;
}

// 32-bit words of barrier workspace for a batch (line 0 + PR_BLOCK_LINES lines per block)
VT_EXPORT int vt_kron_rnn_persistent_bar_words(int batch, long long *words_host)
{
    VT_CHECK_ARG(batch >= 1 && words_host);
    *words_host = (long long)PR_LINE * (1 + (long long)PR_BLOCK_LINES * batch);
    return VT_OK;
}

// workgroups of the persistent kernel that fit on the current device at once: [0] for
// batches of at most 17 columns, [1] for wider ones; and the largest group a block may have
VT_EXPORT int vt_kron_rnn_persistent_capacity(int *max_wgs_host, int *max_group_host)
{
    VT_CHECK_ARG(max_wgs_host && max_group_host);
    int dev = 0, cus = 0, o[4] = {0, 0, 0, 0};
    VT_HIP(hipGetDevice(&dev));
    VT_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
    const int th = 64 * SK_WAVES;
    VT_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&o[0], rnn_persistent_kernel<false, false>, th, 0));
    VT_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&o[1], rnn_persistent_kernel<true, false>, th, 0));
    VT_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&o[2], rnn_persistent_kernel<false, true>, th, 0));
    VT_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&o[3], rnn_persistent_kernel<true, true>, th, 0));
    max_wgs_host[0] = cus * (o[0] < o[1] ? o[0] : o[1]);
    max_wgs_host[1] = cus * (o[2] < o[3] ? o[2] : o[3]);
    *max_group_host = PR_MAX_GROUP;
    return VT_OK;
}

// steps+1 gradient-flow iterations from x = 0 in one persistent launch (the iterates of
// vt_kron_rnn).  fwd / bwd: the operator's descriptors in block order, max_Nc their widest
// column count (<= 32) (forward: A = S,
// B into x, C into zbase, row factor T_j, K = Ds <= 1008 so unsplit; backward: A = S^T,
// B into zbase, C = q, its split workspace).  gstart (device, batch + 1 ints, grid =
// gstart[batch] <= the capacity for max_Nc, every group 1 .. PR_MAX_GROUP).  Rings: ring_x holds `ring`
// slots of slot_x doubles (>= n), ring_z `ring` slots of slot_z doubles (>= the joint Z),
// slot sizes multiples of 32 doubles; ring >= steps + 1, or even and >= 4.  bar: the
// barrier workspace (vt_kron_rnn_persistent_bar_words), zeroed here; afterwards bar[0] != 0
// means a barrier wait exceeded limit_s and x is invalid.
VT_EXPORT int vt_kron_rnn_persistent(const void *fwd, const void *bwd, int batch, int max_Nc, const int *gstart,
                                     int grid,
                                     double *x, long long n, const double *r, const double *zbase,
                                     const double *lam, const double *eta, const long long *seg, int steps,
                                     double *ring_x, long long slot_x, double *ring_z, long long slot_z,
                                     int ring, unsigned *bar, double limit_s, void *stream)
{
    VT_CHECK_ARG(fwd && bwd && gstart && x && r && zbase && lam && eta && seg && ring_x && ring_z && bar);
    VT_CHECK_ARG(batch >= 1 && grid >= batch && n >= 0 && steps >= 0 && limit_s > 0);
    VT_CHECK_ARG(slot_x >= n && slot_x % 32 == 0 && slot_z >= 1 && slot_z % 32 == 0);
    const bool full = ring >= steps + 1;
    VT_CHECK_ARG(full || (ring >= 4 && ring % 2 == 0));
    VT_CHECK_ARG(max_Nc >= 1 && max_Nc <= 32);
    const bool wide = max_Nc > 17;
    int dev = 0, khz = 0, cap[2] = {0, 0}, mg = 0;
    VT_HIP(hipGetDevice(&dev));
    VT_HIP(hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, dev));
    VT_TRY(vt_kron_rnn_persistent_capacity(cap, &mg));
    if (grid > cap[wide ? 1 : 0]) return VT_ERANGE;
    hipStream_t s = (hipStream_t)stream;
    long long words = 0;
    VT_TRY(vt_kron_rnn_persistent_bar_words(batch, &words));
    pr_zero_words<<<vt_blocks(words, 256) < 64 ? vt_blocks(words, 256) : 64, 256, 0, s>>>(bar, words);
    if (n > 0) {
        const unsigned g = vt_blocks(n, 256);
        pr_zero_doubles<<<g < 4096 ? g : 4096, 256, 0, s>>>(x, n);
    }
    VT_LAUNCH_CHECK();
    PrnnArgs a;
    a.fwd = (const SkDesc *)fwd;
    a.bwd = (const SkDesc *)bwd;
    a.gstart = gstart;
    a.x = x;
    a.r = r;
    a.lam = lam;
    a.eta = eta;
    a.seg = seg;
    a.zbase = zbase;
    a.ring_x = ring_x;
    a.ring_z = ring_z;
    a.slot_x = slot_x;
    a.slot_z = slot_z;
    a.ring = ring;
    a.steps = steps;
    a.batch = batch;
    a.pad = 0;
    a.bar = bar;
    a.limit = (unsigned long long)(limit_s * 1000.0 * (double)khz);
    if (full && !wide)
        rnn_persistent_kernel<false, false><<<grid, 64 * SK_WAVES, 0, s>>>(a);
    else if (!wide)
        rnn_persistent_kernel<true, false><<<grid, 64 * SK_WAVES, 0, s>>>(a);
    else if (full)
        rnn_persistent_kernel<false, true><<<grid, 64 * SK_WAVES, 0, s>>>(a);
    else
        rnn_persistent_kernel<true, true><<<grid, 64 * SK_WAVES, 0, s>>>(a);
    VT_LAUNCH_CHECK();
    return VT_OK;
}
