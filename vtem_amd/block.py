"""GPU pipeline for one block (a video window decoded in its own Dirichlet space).

Everything stays resident in HBM: RF bank, spatial coefficients S, spike buffer,
intervals, per-neuron temporal moments T_j, the D x D Gram and its factor.  Host code
only sizes buffers and launches the C-ABI entries of libvtem_hip.so.

Encode  (vtem/vtem.py:176-200 -> vrf.filter :264-292 -> IAF_encode.encode
        vtem/ensemble_encode.py:106-168):
    RF (N x P)  ->  v = dx dy V RF^T (frames x N)  ->  temporal coefficients
    C = Theta_f^T v / frames (Dt x N)  ->  IAF with on-the-fly trigonometric
    interpolation to dt  ->  spike buffer (max_spike x N) + counts.
Decode  (vtem/vtdm.py:630-652 -> dirichlet.compute_q/compute_G, linalg.solve_eq_sym,
        dirichlet.reconstruct), in coefficient space:
    intervals (CSR by neuron) -> q -> T_j, t_j -> A = sum_j S_j S_j^T kron T_j,
    r = sum_j S_j kron t_j -> lam = lam_rel tr(A)/n -> Cholesky (A + lam I) X = r
    -> frames = B_s X Theta_out^T.
"""

from __future__ import annotations

import copy
import ctypes
import dataclasses
import math
from typing import Optional

import numpy as np
import torch

from ._lib import VtemError, lib, ptr, stream_ptr
from .autotune import gram_tuner
from .basis import Geometry
from .synthetic import Bank, IAFParams

LANCZOS_STEPS = 16    # Lanczos steps for the RNN step size 1/lambda_max

F64 = torch.float64
I32 = torch.int32

MAT_DTYPE = np.dtype([("A", np.uint64), ("lda", np.int64), ("D", np.int32), ("pad", np.int32)])


def mat_descs(items, device) -> torch.Tensor:
    """Device array of {A, lda, D} descriptors for the batched Cholesky entries."""
    arr = np.zeros(len(items), MAT_DTYPE)
    for i, (A, lda, D) in enumerate(items):
        arr[i] = (int(A), int(lda), int(D), 0)
    return torch.from_numpy(arr.view(np.uint8).copy()).to(device)


GRAM_DTYPE = np.dtype([("S", np.uint64), ("Th", np.uint64), ("Tl", np.uint64), ("A", np.uint64),
                       ("lds", np.int64), ("lda", np.int64), ("N", np.int32), ("Ds", np.int32),
                       ("Dt", np.int32), ("pad", np.int32), ("scales", np.uint64)])


def gram_descs(items, device) -> torch.Tensor:
    """Device array of low-precision Gram problem descriptors (72 bytes each)."""
    arr = np.zeros(len(items), GRAM_DTYPE)
    for i, (S, Th, Tl, A, N, Ds, Dt, scales) in enumerate(items):
        arr[i] = (S.data_ptr(), Th.data_ptr(), Tl.data_ptr(), A.data_ptr(), S.stride(0),
                  A.stride(0), N, Ds, Dt, 0, scales.data_ptr())
    return torch.from_numpy(arr.view(np.uint8).copy()).to(device)


def gram_lp_ws_bytes(Ds: int, Dt: int, mode: int, tile: int) -> int:
    """Workspace bytes of a single-problem low-precision Gram's tail split (0: no split)."""
    out = ctypes.c_longlong(0)
    lib.vt_gram_lp_ws_size(1, int(Ds), int(Dt), int(mode), int(tile), ctypes.addressof(out))
    return int(out.value)


def split_T_size(N: int, Dt: int) -> int:
    """16-bit elements of one block's tiled low-precision T (queried from the library)."""
    out = ctypes.c_longlong(0)
    lib.vt_gram_split_T_size(int(N), int(Dt), ctypes.addressof(out))
    return int(out.value)


def dev(a, dtype=F64, device="cuda"):
    return torch.as_tensor(np.ascontiguousarray(a), dtype=dtype).to(device)


# --------------------------------------------------------------------------------------
# Kernels wrapped with torch buffers
# --------------------------------------------------------------------------------------

def gemm(M, N, K, alpha, A, sam, sak, B, sbk, sbn, beta, C, ldc, batch=1, bsa=0, bsb=0,
         bsc=0, stream=None):
    f = lib.vt_gemm_f64 if C.dtype == F64 else lib.vt_gemm_f32
    f(M, N, K, alpha, ptr(A), sam, sak, bsa, ptr(B), sbk, sbn, bsb, beta, ptr(C), ldc, bsc,
      batch, stream_ptr(stream))


def rf_bank(bank: Bank, Px: int, Py: int, dx: float = 1.0, dy: float = 1.0, dtype=F64,
            device="cuda", x_start=None, y_start=None, out=None) -> torch.Tensor:
    """(N, Py*Px) receptive fields of a mixed Gabor/DoG bank (vtem/vrf.py:1197-1291)."""
    N = bank.num_neurons
    Sx, Sy = Px * dx, Py * dy
    x_start = -Sx / 2 if x_start is None else x_start
    y_start = -Sy / 2 if y_start is None else y_start
    if out is None:
        out = torch.empty((N, Px * Py), dtype=dtype, device=device)
    kind = dev(bank.kind, I32, device)
    ainv = dev(1.0 / bank.h_alpha, F64, device)
    th = dev(bank.h_l, F64, device)
    x0 = dev(bank.h_x0, F64, device)
    y0 = dev(bank.h_y0, F64, device)
    ab = dev(bank.h_ab, I32, device)
    lib.vt_rf_bank(ptr(out), 0 if dtype == F64 else 1, out.stride(0), N, ptr(kind), ptr(ainv),
                   ptr(th), ptr(x0), ptr(y0), ptr(ab), float(bank.KAPPA),
                   float(bank.sigma_center) ** 2, float(bank.sigma_surround) ** 2, Px, Py, Sx, Sy,
                   x_start, y_start, stream_ptr())
    return out


def spatial_coeffs(RF: torch.Tensor, geom: Geometry, Bs: torch.Tensor = None) -> torch.Tensor:
    """S = dx dy RF B_s  (N x Ds): the real form of the reference's Ds (vtem/vrf.py:333-377)."""
    N, P = RF.shape
    if Bs is None:
        Bs = dev(geom.spatial(), F64, RF.device)
    Ds = Bs.shape[1]
    S = torch.empty((N, Ds), dtype=F64, device=RF.device)
    gemm(N, Ds, P, geom.dx * geom.dy, RF, RF.stride(0), 1, Bs, Ds, 1, 0.0, S, Ds)
    return S


@dataclasses.dataclass
class SpikeTrains:
    """Device spike buffer as the encoder writes it (vtem/ensemble_encode.py:146):
    spikes (max_spike x N), neuron-contiguous; counts (N) int32."""
    spikes: torch.Tensor
    counts: torch.Tensor

    @property
    def num_neurons(self):
        return self.spikes.shape[1]

    def to_lists(self):
        """Host list of per-neuron arrays (vtem/ensemble_encode.py:335-340)."""
        c = self.counts.cpu().numpy()
        s = self.spikes.cpu().numpy()
        return [s[:c[j], j].copy() for j in range(c.size)]


class Encoder:
    """Filter + IAF encoder of one block on the GPU."""

    def __init__(self, geom: Geometry, bank: Bank, device="cuda", filter_dtype=F64):
        self.geom = geom
        self.bank = bank
        self.device = device
        self.filter_dtype = filter_dtype
        self.RF = rf_bank(bank, geom.Px, geom.Py, geom.dx, geom.dy, filter_dtype, device)
        self.theta_f = dev(geom.temporal(geom.frame_times()), filter_dtype, device)  # (F, Dt)
        self.theta_iaf = dev(geom.temporal(geom.iaf_times()), F64, device)          # (T, Dt)

    def filter(self, V: torch.Tensor) -> torch.Tensor:
        """v = dx dy V RF^T, V (frames x P) -> (frames x N) (vtem/vrf.py:264-292)."""
        g = self.geom
        T, P = V.shape
        N = self.RF.shape[0]
        out = torch.empty((T, N), dtype=self.filter_dtype, device=self.device)
        gemm(T, N, P, g.dx * g.dy, V, V.stride(0), 1, self.RF, 1, self.RF.stride(0), 0.0, out, N)
        return out

    def trig_coeffs(self, v: torch.Tensor) -> torch.Tensor:
        """C = Theta_f^T v / frames  (Dt x N), exact for trig polynomials of order Lt."""
        g = self.geom
        F_, N = v.shape
        Dt = g.Dt
        C = torch.empty((Dt, N), dtype=self.filter_dtype, device=self.device)
        gemm(Dt, N, F_, 1.0 / F_, self.theta_f, 1, Dt, v, N, 1, 0.0, C, N)
        return C.to(F64) if C.dtype != F64 else C

    def max_spikes(self, C: torch.Tensor, iaf: IAFParams) -> int:
        """Spike buffer rows: rate bound (sum|C| sqrt2 + b)/(kappa delta) over the window."""
        bound = (C.abs().sum(0) * math.sqrt(2.0)).cpu().numpy()
        rate = (bound + iaf.bias) / (iaf.kappa * iaf.delta)
        return int(math.ceil(self.geom.duration * float(rate.max()) * 1.05)) + 8

    def encode(self, V: torch.Tensor, iaf: IAFParams, root_mode: int = 0,
               max_spike: Optional[int] = None, parallel: bool = True):
        v = self.filter(V)
        C = self.trig_coeffs(v)
        return self.encode_coeffs(C, iaf, root_mode, max_spike, parallel)

    def encode_coeffs(self, C: torch.Tensor, iaf: IAFParams, root_mode: int = 0,
                      max_spike: Optional[int] = None, parallel: bool = True):
        """IAF over the dt grid with the input evaluated from its temporal coefficients.
        parallel=True runs the time-parallel encoder (segments joined by a scan of the
        integrated input); parallel=False the one-lane-per-neuron serial loop of the
        reference (vtem/ensemble_encode.py:438-552).  Counts agree, times to rounding."""
        g = self.geom
        N = C.shape[1]
        if max_spike is None:
            max_spike = self.max_spikes(C, iaf)
        spikes = torch.empty((max_spike, N), dtype=F64, device=self.device)
        counts = torch.empty(N, dtype=I32, device=self.device)
        v0 = dev(iaf.v0, F64, self.device)
        tc = torch.zeros(N, dtype=F64, device=self.device)
        kap, bia, dlt = (dev(a, F64, self.device) for a in (iaf.kappa, iaf.bias, iaf.delta))
        args = (ptr(C), C.stride(0), ptr(self.theta_iaf), g.Dt, N, g.iaf_samples, ptr(spikes), N,
                max_spike, ptr(v0), ptr(kap), ptr(bia), ptr(dlt), ptr(tc), ptr(counts), g.dt,
                root_mode)
        if parallel:
            ws = iaf_par_workspace(N, g.iaf_samples, self.device)
            lib.vt_iaf_encode_trig_par(*args, ptr(ws), ws.numel(), stream_ptr())
        else:
            lib.vt_iaf_encode_trig(*args, stream_ptr())
        self.last_state = (v0, tc)
        return SpikeTrains(spikes, counts)


def iaf_par_workspace_bytes(N: int, size: int) -> int:
    """Device bytes of one time-parallel IAF call (segment partials and prefixes)."""
    b = ctypes.c_longlong()
    lib.vt_iaf_par_workspace_size(int(N), int(size), ctypes.addressof(b))
    return int(b.value)


def iaf_par_workspace(N: int, size: int, device="cuda") -> torch.Tensor:
    return torch.empty(iaf_par_workspace_bytes(N, size), dtype=torch.uint8, device=device)


# --------------------------------------------------------------------------------------
# Decoder
# --------------------------------------------------------------------------------------

@dataclasses.dataclass
class DecodeResult:
    frames: torch.Tensor          # (T_out, Py*Px) float64
    X: torch.Tensor               # (D,) coefficients, index ps*Dt + pt
    lam: torch.Tensor             # (2,) [lambda, trace(A)]
    n: torch.Tensor               # number of intervals (device int)
    info: Optional[torch.Tensor] = None
    resid: Optional[torch.Tensor] = None
    fallback: bool = False
    state: Optional[torch.Tensor] = None   # iterative solver state (CG: rz, |r0|^2, active, iterations)


def tile_doubles(M: int, K: int) -> int:
    """Doubles of an M x K matrix in the skinny product's MFMA tile layout (vt_tile_f64):
    ceil(M/16) * ceil(K/16) tiles of 256."""
    return -(-M // 16) * -(-K // 16) * 256


class Decoder:
    """Coefficient-space decoder for one block geometry and neuron set.

    S (N x Ds) is the decoder's prepared spatial data (the reference's VTDM_prep output,
    vtem/vtdm.py:311-447, kept resident instead of written to dsw.h5/dirich.h5).
    """

    def __init__(self, geom: Geometry, S: torch.Tensor, kappa, delta, bias, norm=None,
                 device="cuda"):
        self.geom = geom
        self.device = device
        self.S = S.contiguous()
        N = S.shape[0]
        self.N = N
        self.kappa = dev(kappa, F64, device)
        self.delta = dev(delta, F64, device)
        self.bias = dev(bias, F64, device)
        self.norm = dev(np.ones(N) if norm is None else norm, F64, device)
        self.Bs = dev(geom.spatial(), F64, device)
        self.theta_out = dev(geom.temporal(geom.frame_times()), F64, device)
        self._tn_ws = torch.empty(16 * geom.Ds * geom.Dt, dtype=F64, device=device)
        self.s2 = torch.empty(N, dtype=F64, device=device)
        lib.vt_row_sumsq(ptr(self.S), self.S.stride(0), N, geom.Ds, ptr(self.s2), stream_ptr())
        # S and S^T resident in the skinny product's MFMA tile layout (vt_tile_f64): the
        # operator's products Z = T_j (S X)_j and Q = S^T Z take them as their A operands
        self.S_tiled = torch.empty(tile_doubles(N, geom.Ds), dtype=F64, device=device)
        lib.vt_tile_f64(ptr(self.S), self.S.stride(0), N, geom.Ds, 0, ptr(self.S_tiled), stream_ptr())
        self.St_tiled = torch.empty(tile_doubles(geom.Ds, N), dtype=F64, device=device)
        lib.vt_tile_f64(ptr(self.S), self.S.stride(0), N, geom.Ds, 1, ptr(self.St_tiled), stream_ptr())

    def with_iaf(self, kappa, delta, bias, gain: float = 1.0) -> "Decoder":
        """The same decoder (shared S, its tiled copies and tables) for a population with other IAF
        parameters.  gain c: the neurons integrated c u + b, which decodes as u with
        kappa / c and b / c (q_k = kappa delta - b (t_k+1 - t_k) scales by 1 / c)."""
        d = copy.copy(self)
        d.kappa = dev(np.asarray(kappa, np.float64) / gain, F64, self.device)
        d.delta = dev(delta, F64, self.device)
        d.bias = dev(np.asarray(bias, np.float64) / gain, F64, self.device)
        d._tn_ws = torch.empty_like(self._tn_ws)
        return d

    # -- spikes -> intervals -> moments ------------------------------------------------
    def intervals(self, st: SpikeTrains, t0: float = 0.0, Tw: float = None):
        N = self.N
        Tw = self.geom.duration if Tw is None else Tw
        cap = st.spikes.shape[0] * N
        ws = torch.empty(2 * N, dtype=I32, device=self.device)
        offsets = torch.empty(N + 1, dtype=I32, device=self.device)
        tk1 = torch.empty(cap, dtype=F64, device=self.device)
        tk2 = torch.empty(cap, dtype=F64, device=self.device)
        nid = torch.empty(cap, dtype=I32, device=self.device)
        lib.vt_intervals(ptr(st.spikes), st.spikes.stride(0), st.spikes.shape[0], ptr(st.counts),
                         N, t0, Tw, ptr(ws),
                         ptr(offsets), ptr(tk1), ptr(tk2), ptr(nid), cap, stream_ptr())
        return tk1, tk2, nid, offsets

    def moments(self, tk1, tk2, nid, offsets):
        g = self.geom
        N = self.N
        cap = tk1.shape[0]
        n_dev = offsets[N:]
        q = torch.empty(cap, dtype=F64, device=self.device)
        lib.vt_q(ptr(q), ptr(tk1), ptr(tk2), ptr(nid), ptr(self.kappa), ptr(self.delta),
                 ptr(self.bias), ptr(self.norm), ptr(n_dev), cap, stream_ptr())
        Tm = torch.empty((N, g.Dt, g.Dt), dtype=F64, device=self.device)
        tv = torch.empty((N, g.Dt), dtype=F64, device=self.device)
        lib.vt_neuron_T(ptr(Tm), ptr(tv), ptr(tk1), ptr(tk2), ptr(q), ptr(self.norm), ptr(offsets),
                        N, g.Lt, g.W, stream_ptr())
        return q, Tm, tv

    # -- Gram, rhs, structured operator --------------------------------------------------
    def gram(self, Tm, out=None):
        g = self.geom
        D = g.D
        A = torch.empty((D, D), dtype=F64, device=self.device) if out is None else out
        lib.vt_gram_f64(ptr(self.S), self.S.stride(0), ptr(Tm), self.N, g.Ds, g.Dt, ptr(A),
                        A.stride(0), stream_ptr())
        return A

    def rhs(self, tv):
        g = self.geom
        r = torch.empty(g.D, dtype=F64, device=self.device)
        lib.vt_gemm_tn_small(g.Ds, g.Dt, self.N, 1.0, ptr(self.S), self.S.stride(0), ptr(tv), g.Dt,
                             0.0, ptr(r), g.Dt, ptr(self._tn_ws), stream_ptr())
        return r

    def apply_A(self, Tm, X, out=None):
        """out = A X through the Kronecker structure: S^T (T_j (S X)), 4 N D flops."""
        g = self.geom
        N, Ds, Dt = self.N, g.Ds, g.Dt
        Y = torch.empty((N, Dt), dtype=F64, device=self.device)
        gemm(N, Dt, Ds, 1.0, self.S, self.S.stride(0), 1, X, Dt, 1, 0.0, Y, Dt)
        Z = torch.empty_like(Y)
        lib.vt_tj_apply(ptr(Z), ptr(Tm), ptr(Y), N, Dt, stream_ptr())
        out = torch.empty(g.D, dtype=F64, device=self.device) if out is None else out
        lib.vt_gemm_tn_small(Ds, Dt, N, 1.0, ptr(self.S), self.S.stride(0), ptr(Z), Dt, 0.0,
                             ptr(out), Dt, ptr(self._tn_ws), stream_ptr())
        return out

    def synthesize(self, X, out=None):
        """frames (T_out x P) = Theta_out Xm^T B_s^T."""
        g = self.geom
        T = self.theta_out.shape[0]
        Y = torch.empty((T, g.Ds), dtype=F64, device=self.device)
        gemm(T, g.Ds, g.Dt, 1.0, self.theta_out, g.Dt, 1, X, 1, g.Dt, 0.0, Y, g.Ds)
        out = torch.empty((T, g.P), dtype=F64, device=self.device) if out is None else out
        gemm(T, g.P, g.Ds, 1.0, Y, g.Ds, 1, self.Bs, 1, g.Ds, 0.0, out, g.P)
        return out

    # -- solves ---------------------------------------------------------------------------
    def _sync_ws(self, batch, maxD):
        """Flag workspace of the triangular solves (zeroed once, reused across calls)."""
        need = 2 + 2 * batch * ((maxD + 63) // 64)
        ws = getattr(self, "_sync", None)
        if ws is None or ws.numel() < need:
            ws = torch.zeros(need, dtype=I32, device=self.device)
            self._sync = ws
        return ws
    def solve_expired(self) -> bool:
        """True when a bounded wait of the triangular solves expired since the flag
        workspace was zeroed (word 1 of the workspace, sticky; a host synchronisation)."""
        ws = getattr(self, "_sync", None)
        return ws is not None and int(ws[1].item()) != 0

    def lam_exact(self, Tm, n_dev, lam_rel):
        """[lam, tr(A)] with tr(A) = sum_j |S_j|^2 tr(T_j) in FP64 (no Gram needed)."""
        g = self.geom
        lam = torch.empty(2, dtype=F64, device=self.device)
        lib.vt_kron_trace(ptr(self.s2), ptr(Tm), self.N, g.Dt, ptr(n_dev), lam_rel, ptr(lam),
                          stream_ptr())
        return lam

    def gram_lp(self, Tm, mode: int, out=None):
        """Low-precision Gram into FP32 (mode 0 split-BF16 x3, mode 1 FP16)."""
        g = self.geom
        D, N = g.D, self.N
        n = split_T_size(N, g.Dt)
        hi = torch.empty(n, dtype=torch.int16, device=self.device)
        lo = torch.empty(n if mode == 0 else 1, dtype=torch.int16, device=self.device)
        scales = torch.ones(5, dtype=torch.float32, device=self.device)
        lib.vt_gram_split_T(ptr(Tm), N, g.Dt, ptr(hi), ptr(lo), mode, ptr(self.S), self.S.numel(),
                            ptr(scales), stream_ptr())
        A = torch.empty((D, D), dtype=torch.float32, device=self.device) if out is None else out
        desc = gram_descs([(self.S, hi, lo, A, N, g.Ds, g.Dt, scales)], self.device)
        s = stream_ptr()
        wss = {}

        def run(tile):
            # workspace of the tail split (a thin last round of pair blocks spread over
            # the chip), per tiling: the plan depends on the kernel's occupancy
            ws = wss.get(tile)
            if ws is None:
                ws = wss[tile] = torch.empty(max(1, gram_lp_ws_bytes(g.Ds, g.Dt, mode, tile) // 4),
                                             dtype=torch.float32, device=self.device)
            lib.vt_gram_lp(ptr(desc), 1, g.Ds, g.Dt, mode, tile, ptr(ws), 4 * ws.numel(), s)

        gram_tuner().launch(run, "bf16x3" if mode == 0 else "fp16", g.Dt, g.Ds, 1, N)
        self._keep = (hi, lo, scales, desc, wss)   # alive until the next decode
        return A

    def solve_tikhonov(self, Tm, r, n_dev, lam_rel=1e-6, precision="fp64", refine_iters=4,
                       A=None):
        """(A + lam I) X = r, lam = lam_rel tr(A)/n, A = sum_j S_j S_j^T kron T_j.

        "fp64"   : FP64 Gram, FP64 Cholesky (the plain path).
        "mixed"  : FP64 Gram rounded to FP32, FP32 Cholesky, FP64 refinement.
        "bf16x3" : split-BF16 3-product Gram (FP32 accumulation), FP32 Cholesky, FP64
                   refinement.
        "fp16"   : FP16 Gram (power-of-two scaled, FP32 accumulation), FP32 Cholesky, FP64
                   refinement.
        Refinement residuals use the exact FP64 operator through the Kronecker structure,
        so the refined X solves the FP64 system; the residual history is returned.
        """
        g = self.geom
        D = g.D
        nblk = (D + 63) // 64
        info = torch.zeros(1, dtype=I32, device=self.device)
        sync = self._sync_ws(1, D)
        X = r.clone()
        if precision == "fp64":
            A = self.gram(Tm) if A is None else A
            lam = torch.empty(2, dtype=F64, device=self.device)
            lib.vt_trace_lambda(ptr(A), 0, A.stride(0), D, ptr(n_dev), lam_rel, ptr(lam),
                                stream_ptr())
            lib.vt_diag_add(ptr(A), 0, A.stride(0), D, 0.0, ptr(lam), stream_ptr())
            inv = torch.empty(nblk * 64 * 64, dtype=F64, device=self.device)
            mats = mat_descs([(A.data_ptr(), A.stride(0), D)], self.device)
            lib.vt_potrf_f64(ptr(mats), 1, D, ptr(inv), inv.numel(), ptr(info), stream_ptr())
            lib.vt_potrs_f64(ptr(mats), 1, D, ptr(inv), inv.numel(), ptr(X), D, ptr(sync),
                             stream_ptr())
            self._keep_solve = (A, inv, mats)
            return X, lam, info, None
        lam = self.lam_exact(Tm, n_dev, lam_rel)
        if precision == "mixed":
            A64 = self.gram(Tm) if A is None else A
            A32 = torch.empty((D, D), dtype=torch.float32, device=self.device)
            lib.vt_lower_to_f32(ptr(A32), D, ptr(A64), A64.stride(0), D, ptr(lam), stream_ptr())
        elif precision in ("bf16x3", "fp16"):
            A32 = self.gram_lp(Tm, 0 if precision == "bf16x3" else 1)
            lib.vt_diag_add(ptr(A32), 1, D, D, 0.0, ptr(lam), stream_ptr())
        else:
            raise ValueError(f"unknown precision {precision}")
        inv = torch.empty(nblk * 64 * 64, dtype=torch.float32, device=self.device)
        mats = mat_descs([(A32.data_ptr(), D, D)], self.device)
        lib.vt_potrf_f32(ptr(mats), 1, D, ptr(inv), inv.numel(), ptr(info), stream_ptr())
        lib.vt_potrs_f32(ptr(mats), 1, D, ptr(inv), inv.numel(), ptr(X), D, ptr(sync), stream_ptr())
        res = torch.empty(D, dtype=F64, device=self.device)
        AX = torch.empty(D, dtype=F64, device=self.device)
        nrm = torch.zeros(refine_iters + 2, dtype=F64, device=self.device)
        lib.vt_sumsq(ptr(r), D, ptr(nrm[0:]), stream_ptr())
        for it in range(refine_iters + 1):
            self.apply_A(Tm, X, out=AX)
            lib.vt_residual(ptr(res), ptr(r), ptr(AX), ptr(X), ptr(lam), D, stream_ptr())
            lib.vt_sumsq(ptr(res), D, ptr(nrm[it + 1:]), stream_ptr())
            if it == refine_iters:
                break
            lib.vt_potrs_f32(ptr(mats), 1, D, ptr(inv), inv.numel(), ptr(res), D, ptr(sync),
                             stream_ptr())
            lib.vt_axpbypcz(ptr(X), 1.0, ptr(X), 1.0, ptr(res), 0.0, 0, D, stream_ptr())
        self._keep_solve = (A32, inv, mats, res, AX)
        return X, lam, info, nrm

    def power_eta(self, Tm, lam, steps: int = LANCZOS_STEPS) -> float:
        """The RNN step 1/theta, theta the largest Ritz value of A + lam I after `steps`
        Lanczos steps from the normalised all-ones vector (u = (A + lam I) v_j,
        alpha_j = <u, v_j>, w = u - alpha_j v_j - beta_{j-1} v_{j-1}, beta_j = |w|,
        v_{j+1} = w / beta_j; a zero beta ends the recurrence), the batched decoder's rule."""
        g = self.geom
        v = torch.full((g.D,), 1.0 / math.sqrt(g.D), dtype=F64, device=self.device)
        vp = torch.zeros_like(v)
        q = torch.empty_like(v)
        ab = torch.zeros((steps, 2), dtype=F64, device=self.device)
        bprev = torch.zeros((), dtype=F64, device=self.device)
        for j in range(steps):
            self.apply_A(Tm, v, out=q)
            u = q + lam[0] * v
            a = torch.dot(u, v)
            w = u - a * v - bprev * vp
            b = torch.linalg.vector_norm(w)
            b = torch.where(torch.isfinite(b) & (b > 0), b, torch.zeros_like(b))
            ab[j, 0], ab[j, 1] = a, b
            vp = v
            v = torch.where(b > 0, w / torch.where(b > 0, b, torch.ones_like(b)), torch.zeros_like(w))
            bprev = b
        h = ab.cpu().numpy()
        T = np.diag(h[:, 0]) + np.diag(h[:-1, 1], 1) + np.diag(h[:-1, 1], -1)
        th = float(np.linalg.eigvalsh(T)[-1])
        return 1.0 / th if th > 0.0 else 0.0

    def solve_rnn(self, Tm, r, lam, steps: int, eta: float):
        """X <- X + eta (r - (A + lam I) X), steps+1 iterations from 0 (the D-space image of
        vtem/neural_network.py:78-82), with A applied through its Kronecker structure."""
        g = self.geom
        X = torch.zeros(g.D, dtype=F64, device=self.device)
        AX = torch.empty_like(X)
        res = torch.empty_like(X)
        for _ in range(steps + 1):
            self.apply_A(Tm, X, out=AX)
            lib.vt_residual(ptr(res), ptr(r), ptr(AX), ptr(X), ptr(lam), g.D, stream_ptr())
            lib.vt_axpbypcz(ptr(X), 1.0, ptr(X), eta, ptr(res), 0.0, 0, g.D, stream_ptr())
        return X

    # -- full decode ----------------------------------------------------------------------
    def decode(self, st: SpikeTrains, t0: float = 0.0, method: str = "tikhonov",
               precision: str = "fp64", lam_rel: float = 1e-6, rnn_steps: int = 500,
               rnn_eta: Optional[float] = None, refine_iters: int = 4,
               verify: bool = True, tol: float = 1e-10, pinv_rcond: float = 1e-8) -> DecodeResult:
        """Decode one window.  For the low-precision forms with `verify`, a factorisation
        that lost positive definiteness or a refinement that did not reach `tol` relative
        residual is redone on the FP64 path (result.fallback = True); this costs one host
        synchronisation."""
        tk1, tk2, nid, offsets = self.intervals(st, t0)
        q, Tm, tv = self.moments(tk1, tk2, nid, offsets)
        n_dev = offsets[self.N:]
        r = self.rhs(tv)
        fallback = False
        if method == "tikhonov":
            X, lam, info, nrm = self.solve_tikhonov(Tm, r, n_dev, lam_rel, precision, refine_iters)
            if verify and precision != "fp64":
                h = nrm.cpu().numpy()
                ok = int(info.item()) == 0 and np.isfinite(h).all() and \
                    math.sqrt(h[-1]) <= tol * math.sqrt(h[0]) and not self.solve_expired()
                if not ok:
                    self._sync.zero_()
                    X, lam, info, nrm = self.solve_tikhonov(Tm, r, n_dev, lam_rel, "fp64")
                    fallback = True
            if verify and self.solve_expired():
                raise VtemError("a triangular solve's bounded wait expired: the solution is invalid")
        elif method == "pinv":
            # the reference's decode: truncated-SVD pseudoinverse (rcond 1e-8, no lambda) of
            # the Gram (vtem/vtdm.py:451, vtem/utils/linalg.py:591-631).  In D-space: the
            # nonzero spectra of Phi Phi^T and Phi^T Phi coincide, so the truncation and
            # X = pinv(A) r equal Phi^T pinv(G) q
            from .utils.linalg import solve_eq_sym
            A = self.gram(Tm)
            lib.vt_symmetrize(ptr(A), 0, A.stride(0), self.geom.D, stream_ptr())
            X = solve_eq_sym(A, r, rcond=pinv_rcond)
            lam, info, nrm = None, None, None
        elif method == "rnn":
            lam = self.lam_exact(Tm, n_dev, lam_rel)
            eta = self.power_eta(Tm, lam) if rnn_eta is None else rnn_eta
            X = self.solve_rnn(Tm, r, lam, rnn_steps, eta)
            info, nrm = None, None
        else:
            raise ValueError(f"unknown method {method}")
        frames = self.synthesize(X)
        return DecodeResult(frames, X, lam, n_dev, info, nrm, fallback)
