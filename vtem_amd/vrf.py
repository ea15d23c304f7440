"""Visual receptive fields with the reference's API (vtem/vrf.py): vrf_gabor and vrf_cs.

Same class, method and keyword names as the reference; arrays are torch tensors on the
GPU (numpy inputs are copied in, as the reference's to_gpu did).  Every computation runs
in the HIP library:

* generate_visual_receptive_fields -> vt_rf_gabor / vt_rf_dog (vtem/vrf.py:296-331,
  :876-913; kernels :1120-1291);
* filter -> RF generation + vt_gemm_f64 (vtem/vrf.py:264-292, :845-872);
* compute_Ds -> vt_dirichlet_table + GEMM + vt_ds_assemble, a direct DFT of the RFs
  sampled over one period (vtem/vrf.py:333-377, :915-968);
* compute_Dsw -> two GEMMs + vt_scale_outer (vtem/vrf.py:502-536, :1087-1117);
* compute_dirich_space[_fft] -> GEMM against the mode table (vtem/vrf.py:380-499).

The default parameter laws (get_gabor_parameters / get_cs_parameters) follow
vtem/vrf.py:136-205 and :726-790: dilations alpha0^m, rotations l (R-1) pi / R,
translations on a (2t+1)^2 grid of spacing b0 alpha0^m b0scalebase^-m.
"""

from __future__ import annotations

import math

import numpy as np
import torch

from ._lib import lib, ptr, stream_ptr

F64 = torch.float64


def _dev(a, dtype=F64, device="cuda"):
    if isinstance(a, torch.Tensor):
        return a.to(device=device, dtype=dtype).contiguous()
    return torch.as_tensor(np.ascontiguousarray(a), dtype=dtype, device=device)


def _gemm(M, N, K, alpha, A, sam, sak, B, sbk, sbn, beta, C, ldc):
    lib.vt_gemm_f64(M, N, K, alpha, ptr(A), sam, sak, 0, ptr(B), sbk, sbn, 0, beta, ptr(C), ldc, 0, 1,
                    stream_ptr())


class vrf:
    """Visual receptive field base: grid geometry (vtem/vrf.py:17-82).

    shape (Py, Px) pixels; dx, dy pixel pitch; scale multiplies the pixel count and
    divides the pitch; domain (Sx, Sy) restricts the extent the parameter laws cover."""

    def __init__(self, shape, dtype=np.float64, dx=1.0 / 16, dy=1.0 / 16, scale=1, domain=None,
                 device="cuda"):
        self.shape = shape
        self.dtype = np.dtype(dtype)
        self.device = device
        self.Pyall = int(shape[-2] * scale)
        self.Pxall = int(shape[-1] * scale)
        self.dx = float(dx) / scale
        self.dy = float(dy) / scale
        self.Sxall = self.Pxall * self.dx
        self.Syall = self.Pyall * self.dy
        if domain is None:
            self.Sx, self.Sy = self.Sxall, self.Syall
        else:
            self.Sx, self.Sy = float(domain[0]), float(domain[1])
        self.Px = int(round(self.Sx / self.dx))
        self.Py = int(round(self.Sy / self.dy))
        self.num_neurons = 0

    # -- RFs ---------------------------------------------------------------------------
    def _generate(self, out, startbias, N, x_start, y_start):
        raise NotImplementedError

    def generate_visual_receptive_fields(self, startbias=0, N_filters=None, x_start=None,
                                         y_start=None):
        """RFs startbias .. startbias+N_filters-1 on the full grid -> self.rfs
        (N_filters, Pyall, Pxall) float64 (also available as .gabors / .filters)."""
        if N_filters is None:
            N_filters = self.num_neurons - startbias
        x_start = -self.Sxall / 2 if x_start is None else float(x_start)
        y_start = -self.Syall / 2 if y_start is None else float(y_start)
        out = torch.empty((N_filters, self.Pyall, self.Pxall), dtype=F64, device=self.device)
        self._generate(out, startbias, N_filters, x_start, y_start)
        self.rfs = out
        return out

    @property
    def gabors(self):
        return self.rfs

    def filter(self, V, max_bytes: int = 8 << 30):
        """v[t, j] = dx dy sum_p V[t, p] RF_j[p] for a video V (T, Py, Px) -> (T, N) float64
        on the GPU (time-major, neuron-contiguous: the encoder's input layout).  RFs are
        generated in batches bounded by max_bytes (the reference's ONE_TIME_FILTERS)."""
        Vd = _dev(V).reshape(V.shape[0], -1)
        T, P = Vd.shape
        if P != self.Pxall * self.Pyall:
            raise ValueError("video frame size does not match the receptive fields")
        out = torch.empty((T, self.num_neurons), dtype=F64, device=self.device)
        step = max(1, min(self.num_neurons, max_bytes // (8 * P)))
        for i in range(0, self.num_neurons, step):
            n = min(step, self.num_neurons - i)
            rf = self.generate_visual_receptive_fields(startbias=i, N_filters=n).reshape(n, P)
            # out[:, i:i+n] = dxdy V RF^T
            _gemm(T, n, P, self.dx * self.dy, Vd, P, 1, rf, 1, P, 0.0, out[:, i:], self.num_neurons)
        return out

    # -- Dirichlet space ---------------------------------------------------------------
    def _table(self, Px, Py, x_start, dx, y_start, dy, Mx, My, Ox, Oy):
        M = (2 * Mx + 1) * (2 * My + 1)
        tab = torch.empty((Px * Py, 2 * M), dtype=F64, device=self.device)
        lib.vt_dirichlet_table(ptr(tab), 2 * M, Px, Py, x_start, dx, y_start, dy, Mx, My, Ox, Oy,
                               stream_ptr())
        return tab

    def compute_Ds(self, Mx, My, mask: str = "ellipse"):
        """Dirichlet coefficients of every RF, d_Ds (N, 2My+1, 2Mx+1) complex128 with row
        r <-> m_y = r - My and column c <-> m_x = c - Mx:
        Ds_j(m) = dx dy sum_p RF_j(x_p, y_q) exp(-i (mx Ox x_p + my Oy y_q)), Ox = 2 pi/Sx,
        over the object's grid, which is taken to span exactly one period.  mask
        "ellipse" keeps mx^2 My^2 + my^2 Mx^2 <= Mx^2 My^2 (vtem/vrf.py:1331), "rect"
        every mode."""
        Mx, My = int(Mx), int(My)
        M = (2 * Mx + 1) * (2 * My + 1)
        P = self.Pxall * self.Pyall
        tab = self._table(self.Pxall, self.Pyall, -self.Sxall / 2, self.dx, -self.Syall / 2, self.dy,
                          Mx, My, 2 * math.pi / self.Sxall, 2 * math.pi / self.Syall)
        N = self.num_neurons
        C = torch.empty((N, 2 * M), dtype=F64, device=self.device)
        step = max(1, min(N, (4 << 30) // (8 * P)))
        for i in range(0, N, step):
            n = min(step, N - i)
            rf = self.generate_visual_receptive_fields(startbias=i, N_filters=n).reshape(n, P)
            _gemm(n, 2 * M, P, 1.0, rf, P, 1, tab, 2 * M, 1, 0.0, C[i:], 2 * M)
        Ds = torch.empty((N, 2 * My + 1, 2 * Mx + 1), dtype=torch.complex128, device=self.device)
        lib.vt_ds_assemble(ptr(Ds), ptr(C), 2 * M, N, Mx, My, 1 if mask == "ellipse" else 0,
                           self.dx * self.dy, stream_ptr())
        return Ds

    def compute_Dsw(self, d_Ds, Mx=None, My=None, h_norm=None):
        """Dsw = Re(Ds Ds^H) * norm norm^T (N x N float64)."""
        return weighted_gram(d_Ds, h_norm, self.device)

    def compute_dirich_space(self, d_Ds, Mx, My, Px, Py, Sx, Sy, Wx, Wy, x_start=None, y_start=None):
        """psi_j(x, y) = Re sum_m Ds_j(m) e^{i(mx Wx/Mx x + my Wy/My y)} on a Px x Py grid
        of extent Sx x Sy (origin -S/2 unless given) -> (N, Py, Px) float64."""
        Mx, My = int(Mx), int(My)
        Ds = _dev(d_Ds, torch.complex128, self.device)
        N = Ds.shape[0]
        M = (2 * Mx + 1) * (2 * My + 1)
        if Ds[0].numel() != M:
            raise ValueError("d_Ds does not have (2My+1) x (2Mx+1) modes")
        x_start = -Sx / 2 if x_start is None else float(x_start)
        y_start = -Sy / 2 if y_start is None else float(y_start)
        tab = self._table(int(Px), int(Py), x_start, Sx / Px, y_start, Sy / Py, Mx, My, Wx / Mx if Mx else 0.0,
                          Wy / My if My else 0.0)
        R = torch.view_as_real(Ds).reshape(N, 2 * M)
        psi = torch.empty((N, int(Py), int(Px)), dtype=F64, device=self.device)
        P = int(Px) * int(Py)
        # psi = sum_m Re Ds cos - Im Ds sin: interleaved A (re at 2m, im at 2m+1) against the
        # table's cos column m and -sin column M+m
        _gemm(N, P, M, 1.0, R, 2 * M, 2, tab, 1, 2 * M, 0.0, psi, P)
        _gemm(N, P, M, 1.0, R[:, 1:], 2 * M, 2, tab[:, M:], 1, 2 * M, 1.0, psi, P)
        return psi

    def compute_dirich_space_fft(self, d_Ds, Mx, My, Px, Py, Sx, Sy, Wx, Wy):
        """The same functions as compute_dirich_space (the reference computes them by an
        inverse FFT; both forms agree)."""
        return self.compute_dirich_space(d_Ds, Mx, My, Px, Py, Sx, Sy, Wx, Wy)


def weighted_gram(d_Ds, h_norm=None, device="cuda"):
    """Re(Ds Ds^H) * norm norm^T (N x N float64) for Dirichlet coefficients Ds (N x
    modes, complex): the RFs' inner products in the Dirichlet space, weighted
    (vtem/vtdm.py:440-446 with put_norm; vtem/dirichlet.py:101-115)."""
    Ds = _dev(d_Ds, torch.complex128, device)
    N = Ds.shape[0]
    M = Ds[0].numel()
    R = torch.view_as_real(Ds).reshape(N, 2 * M)      # interleaved re, im
    W = torch.empty((N, N), dtype=F64, device=device)
    # Re Re^T + Im Im^T = sum over the interleaved pairs
    _gemm(N, N, 2 * M, 1.0, R, 2 * M, 1, R, 1, 2 * M, 0.0, W, N)
    if h_norm is not None:
        nrm = _dev(h_norm, F64, device)
        lib.vt_scale_outer(ptr(W), N, N, ptr(nrm), stream_ptr())
    return W


def _grid_law(S, dilations, alpha0, b0, b0scalebase):
    """Per dilation: spacing b_i and half-counts of the translation grid in x and y."""
    out = []
    for m in dilations:
        b = float(b0) * float(alpha0) ** m * float(b0scalebase) ** (-m)
        out.append((m, b, int(math.floor(S[0] / 2 / b)), int(math.floor(S[1] / 2 / b))))
    return out


class vrf_gabor(vrf):
    """Gabor bank (vtem/vrf.py:88-536): RF_j = exp(-X^2/2 - Y^2/8) (cos|sin)(KAPPA X) in
    coordinates rotated by h_l and scaled by 1/h_alpha about (h_x0, h_y0)."""

    def get_gabor_parameters(self, dilations=(1, 0, -1, -2, -3), num_rotations=8, alpha0=2.0,
                             b0=1.5, b0scalebase=math.sqrt(2), KAPPA=2.5):
        self.KAPPA = float(KAPPA)
        l0 = (num_rotations - 1) * math.pi / num_rotations
        alpha, theta, x0, y0, ab = [], [], [], [], []
        for m, b, tx, ty in _grid_law((self.Sx, self.Sy), dilations, alpha0, b0, b0scalebase):
            ll = np.arange(num_rotations) * l0
            k = np.arange(-ty, ty + 1)
            n = np.arange(-tx, tx + 1)
            L, K, Nn = np.meshgrid(ll, k, n, indexing="ij")
            cnt = L.size
            # each (rotation, y, x) gives an even (ab 0) and an odd (ab 1) filter
            alpha.append(np.full(2 * cnt, float(alpha0) ** m))
            theta.append(np.repeat(L.ravel(), 2))
            y0.append(np.repeat(K.ravel() * b, 2))
            x0.append(np.repeat(Nn.ravel() * b, 2))
            ab.append(np.tile(np.array([0, 1], np.int32), cnt))
        self.h_alpha = np.concatenate(alpha)
        self.h_l = np.concatenate(theta)
        self.h_x0 = np.concatenate(x0)
        self.h_y0 = np.concatenate(y0)
        self.h_ab = np.concatenate(ab).astype(np.int32)
        self.num_neurons = int(self.h_alpha.size)

    def load_parameters(self, num_neurons=None, h_alpha=None, h_l=None, h_x0=None, h_y0=None,
                        h_ab=None, KAPPA=2.5, set=0):
        if num_neurons is not None:
            if h_alpha is None or h_l is None or h_x0 is None or h_y0 is None or h_ab is None:
                raise ValueError("must specify the gabor parameters")
            self.num_neurons = int(num_neurons)
            self.h_alpha, self.h_l = np.asarray(h_alpha, np.float64), np.asarray(h_l, np.float64)
            self.h_x0, self.h_y0 = np.asarray(h_x0, np.float64), np.asarray(h_y0, np.float64)
            self.h_ab = np.asarray(h_ab, np.int32)
            self.KAPPA = float(KAPPA)
        elif not set:
            self.get_gabor_parameters(KAPPA=KAPPA)
        self.d_alpha = _dev(1.0 / self.h_alpha, F64, self.device)
        self.d_l = _dev(self.h_l, F64, self.device)
        self.d_x0 = _dev(self.h_x0, F64, self.device)
        self.d_y0 = _dev(self.h_y0, F64, self.device)
        self.d_ab = _dev(self.h_ab, torch.int32, self.device)

    def _generate(self, out, startbias, N, x_start, y_start):
        o = startbias
        lib.vt_rf_gabor(ptr(out), 0, self.Pxall * self.Pyall, N, ptr(self.d_alpha) + 8 * o,
                        ptr(self.d_l) + 8 * o, ptr(self.d_x0) + 8 * o, ptr(self.d_y0) + 8 * o,
                        ptr(self.d_ab) + 4 * o, self.KAPPA, self.Pxall, self.Pyall, self.Sxall,
                        self.Syall, x_start, y_start, stream_ptr())


class vrf_cs(vrf):
    """Centre-surround (difference of Gaussians) bank (vtem/vrf.py:681-1117)."""

    def get_cs_parameters(self, dilations=(1, 0, -1, -2, -3), alpha0=2.0, b0=1.0, b0scalebase=1.0,
                          sigma_center=0.5, sigma_surround=0.8):
        self.sigma_center, self.sigma_surround = float(sigma_center), float(sigma_surround)
        alpha, x0, y0 = [], [], []
        for m, b, tx, ty in _grid_law((self.Sx, self.Sy), dilations, alpha0, b0, b0scalebase):
            K, Nn = np.meshgrid(np.arange(-ty, ty + 1), np.arange(-tx, tx + 1), indexing="ij")
            alpha.append(np.full(K.size, float(alpha0) ** m))
            y0.append(K.ravel() * b)
            x0.append(Nn.ravel() * b)
        self.h_alpha = np.concatenate(alpha)
        self.h_x0 = np.concatenate(x0)
        self.h_y0 = np.concatenate(y0)
        self.num_neurons = int(self.h_alpha.size)

    def load_parameters(self, num_neurons=None, h_alpha=None, h_x0=None, h_y0=None, set=0,
                        sigma_center=0.5, sigma_surround=0.8):
        if num_neurons is not None:
            if h_alpha is None or h_x0 is None or h_y0 is None:
                raise ValueError("must specify the centre-surround parameters")
            self.num_neurons = int(num_neurons)
            self.h_alpha = np.asarray(h_alpha, np.float64)
            self.h_x0, self.h_y0 = np.asarray(h_x0, np.float64), np.asarray(h_y0, np.float64)
            self.sigma_center, self.sigma_surround = float(sigma_center), float(sigma_surround)
        elif not set:
            self.get_cs_parameters(sigma_center=sigma_center, sigma_surround=sigma_surround)
        self.d_alpha = _dev(1.0 / self.h_alpha, F64, self.device)
        self.d_x0 = _dev(self.h_x0, F64, self.device)
        self.d_y0 = _dev(self.h_y0, F64, self.device)

    def _generate(self, out, startbias, N, x_start, y_start):
        o = startbias
        lib.vt_rf_dog(ptr(out), 0, self.Pxall * self.Pyall, N, ptr(self.d_alpha) + 8 * o,
                      ptr(self.d_x0) + 8 * o, ptr(self.d_y0) + 8 * o, self.sigma_center ** 2,
                      self.sigma_surround ** 2, self.Pxall, self.Pyall, self.Sxall, self.Syall,
                      x_start, y_start, stream_ptr())
