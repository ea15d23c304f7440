"""TEST INFRASTRUCTURE ONLY: float64 CPU oracle for the video TEM/TDM path.

This module restates the reference algorithm (bionet/vtem) in NumPy/SciPy float64 (and,
for the IAF encoder, in plain C: iaf_ref.c).  Only tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg import it, and only as the checker or the timed CPU
baseline; the product path (vtem_amd) never imports anything from here and fails loudly
when its HIP library is missing.

Each function cites the reference file:line it follows.  Functions marked "(new)" have
no reference counterpart; they implement the coefficient-space (D-space) form of the
reference's decoder that the configurations need, and are pinned to the reference's
spike-space form by the push-through identity in the tests.

Conventions shared with the HIP path (stated here so the oracle stands alone):

* pixel coordinates x_p = p*dx - Sx/2, y_q = -(q*dy - Sy/2) (vtem/vrf.py:1254-1257);
* real spatial basis, index ps: ps = 0 is the constant; mode i of the half set
  H = {(mx, my): my > 0} U {(mx, 0): mx > 0} (ordered my = 0..Ly, mx = -Lx..Lx, filtered
  by the mask) contributes ps = 1+2i -> sqrt2*cos(mx*Ox*x + my*Oy*y) and ps = 2+2i ->
  sqrt2*sin(...);
* real temporal basis, index pt: pt = 0 is the constant, pt = 2m-1 -> sqrt2*cos(m*W*t),
  pt = 2m -> sqrt2*sin(m*W*t), m = 1..Lt;
* D-space index d = ps*Dt + pt.
"""

from __future__ import annotations

import ctypes
import math
import os
from typing import Sequence, Tuple

import numpy as np
import scipy.linalg as sla

SQRT2 = math.sqrt(2.0)
_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None


# --------------------------------------------------------------------------------------
# C part (IAF)
# --------------------------------------------------------------------------------------

def lib_path() -> str:
    return os.path.join(_HERE, "_build", "liboracle_iaf.so")


def build(force: bool = False) -> str:
    """Compile iaf_ref.c with gcc (no FP contraction, no fast-math)."""
    out = lib_path()
    src = os.path.join(_HERE, "iaf_ref.c")
    if force or not os.path.exists(out) or os.path.getmtime(out) < os.path.getmtime(src):
        os.makedirs(os.path.dirname(out), exist_ok=True)
        import subprocess
        subprocess.check_call(["gcc", "-O2", "-fPIC", "-shared", "-ffp-contract=off",
                               "-fno-fast-math", "-o", out, src, "-lm"])
    return out


def _lib():
    global _LIB
    if _LIB is None:
        path = lib_path()
        if not os.path.exists(path):
            build()
        L = ctypes.CDLL(path)
        dp = ctypes.POINTER(ctypes.c_double)
        ip = ctypes.POINTER(ctypes.c_int)
        L.oracle_iaf_encode.argtypes = [dp, ctypes.c_int64, ctypes.c_int, ctypes.c_int, dp,
                                        ctypes.c_int64, dp, dp, dp, dp, dp, ip, ctypes.c_int,
                                        ctypes.c_double, ctypes.c_int]
        L.oracle_iaf_encode_rt.argtypes = [dp, ctypes.c_int64, ctypes.c_int, ctypes.c_int, dp,
                                           ctypes.c_int64, dp, dp, dp, dp, dp, ip,
                                           ctypes.c_int, ctypes.c_double, dp, dp, ctypes.c_int]
        _LIB = L
    return _LIB


def _dp(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def iaf_encode(inputs: np.ndarray, dt: float, kappa, bias, delta, v0, time_count=None,
               avg_rate: float = 0.1, root_mode: int = 0, max_spike: int = None):
    """IAF population encode (vtem/ensemble_encode.py:106-168, kernel :486-545).

    inputs: (T, N) float64, time-major and neuron-contiguous.  Returns
    (spike_lists, counts, v0_out, time_count_out, buffer) where spike_lists[j] holds
    neuron j's spike times relative to the call start plus time_count, like the
    reference's rearrange_spikes (:335-340).  Raises ValueError on buffer overflow
    exactly as the reference (:163-164).
    """
    x = np.ascontiguousarray(inputs, dtype=np.float64)
    T, N = x.shape
    if max_spike is None:
        max_spike = int(np.ceil(T / avg_rate))
    spikes = np.zeros((max_spike, N), np.float64)
    v0_ = np.array(v0, np.float64, copy=True).reshape(N)
    tc = np.zeros(N) if time_count is None else np.array(time_count, np.float64).reshape(N)
    kap = np.ascontiguousarray(kappa, np.float64).reshape(N)
    bia = np.ascontiguousarray(bias, np.float64).reshape(N)
    dlt = np.ascontiguousarray(delta, np.float64).reshape(N)
    cnt = np.zeros(N, np.int32)
    _lib().oracle_iaf_encode(_dp(x), N, N, T, _dp(spikes), N, _dp(v0_), _dp(kap), _dp(bia),
                             _dp(dlt), _dp(tc), cnt.ctypes.data_as(ctypes.POINTER(ctypes.c_int)),
                             max_spike, float(dt), int(root_mode))
    if cnt.size and cnt.max() >= max_spike:
        raise ValueError("number of spikes exceeded the limit of buffer")
    lists = [spikes[:cnt[j], j].copy() for j in range(N)]
    return lists, cnt, v0_, tc, spikes


def iaf_encode_rt(inputs: np.ndarray, dt: float, kappa, bias, delta_mean, sigma,
                  delta_value, v0, randn: np.ndarray, time_count=None, root_mode: int = 0):
    """Random-threshold IAF (vtem/ensemble_encode.py:265-329, kernel :606-668).

    randn: (max_spike, N) N(0,1) draws (the reference draws them on the host, :304).
    """
    x = np.ascontiguousarray(inputs, dtype=np.float64)
    T, N = x.shape
    spikes = np.array(randn, np.float64, copy=True, order="C")
    max_spike = spikes.shape[0]
    v0_ = np.array(v0, np.float64, copy=True).reshape(N)
    tc = np.zeros(N) if time_count is None else np.array(time_count, np.float64).reshape(N)
    dv = np.array(delta_value, np.float64, copy=True).reshape(N)
    cnt = np.zeros(N, np.int32)
    f = lambda a: np.ascontiguousarray(a, np.float64).reshape(N)
    kap, bia, dm, sg = f(kappa), f(bias), f(delta_mean), f(sigma)
    _lib().oracle_iaf_encode_rt(_dp(x), N, N, T, _dp(spikes), N, _dp(v0_), _dp(kap), _dp(bia),
                                _dp(dm), _dp(tc), cnt.ctypes.data_as(ctypes.POINTER(ctypes.c_int)),
                                max_spike, float(dt), _dp(dv), _dp(sg), int(root_mode))
    if cnt.size and cnt.max() >= max_spike:
        raise ValueError("number of spikes exceeded the limit of buffer")
    lists = [spikes[:cnt[j], j].copy() for j in range(N)]
    return lists, cnt, v0_, tc, dv


# --------------------------------------------------------------------------------------
# Receptive fields (vtem/vrf.py)
# --------------------------------------------------------------------------------------

def _grid(Px, Py, Sx, Sy, x_start=None, y_start=None):
    """Physical (x, y) of every pixel (vtem/vrf.py:1254-1257, origin :323-327)."""
    if x_start is None:
        x_start = -Sx / 2
    if y_start is None:
        y_start = -Sy / 2
    x = np.arange(Px, dtype=np.float64) * (Sx / Px) + x_start
    y = -(np.arange(Py, dtype=np.float64) * (Sy / Py) + y_start)
    return x, y


def rf_gabor(h_alpha, h_l, h_x0, h_y0, h_ab, KAPPA, Px, Py, Sx, Sy, x_start=None,
             y_start=None) -> np.ndarray:
    """Gabor RFs on the pixel grid (vtem/vrf.py:1246-1286; 1/h_alpha from :258).

    RF = a/sqrt(2pi) exp(-(4X^2+Y^2)/8) * (cos(KAPPA X) | sin(KAPPA X)), with
    (x, y) <- a(x - x0), a(y - y0), a = 1/h_alpha, X = x cos t + y sin t,
    Y = -x sin t + y cos t.
    """
    x, y = _grid(Px, Py, Sx, Sy, x_start, y_start)
    a = (1.0 / np.asarray(h_alpha, np.float64))[:, None, None]
    xx = a * (x[None, None, :] - np.asarray(h_x0)[:, None, None])
    yy = a * (y[None, :, None] - np.asarray(h_y0)[:, None, None])
    th = np.asarray(h_l, np.float64)[:, None, None]
    c, s = np.cos(th), np.sin(th)
    X = xx * c + yy * s
    Y = -xx * s + yy * c
    first = a * (1 / np.sqrt(2 * np.pi)) * np.exp(-((4 * X * X) + (Y * Y)) / 8)
    odd = (np.asarray(h_ab) != 0)[:, None, None]
    return first * np.where(odd, np.sin(KAPPA * X), np.cos(KAPPA * X))


def rf_dog(h_alpha, h_x0, h_y0, sigma_center, sigma_surround, Px, Py, Sx, Sy,
           x_start=None, y_start=None) -> np.ndarray:
    """Centre-surround RFs (vtem/vrf.py:1161-1188; sigma^2 passed at :913).

    RF = a (exp(-r^2/(2 sc^2))/sc^2 - 0.9 exp(-r^2/(2 ss^2))/ss^2) / (2pi), r in a-scaled
    coordinates.
    """
    x, y = _grid(Px, Py, Sx, Sy, x_start, y_start)
    a = (1.0 / np.asarray(h_alpha, np.float64))[:, None, None]
    xx = a * (x[None, None, :] - np.asarray(h_x0)[:, None, None])
    yy = a * (y[None, :, None] - np.asarray(h_y0)[:, None, None])
    XY = -(xx * xx + yy * yy)
    sc2, ss2 = sigma_center ** 2, sigma_surround ** 2
    gb = a * (np.exp(XY / (2 * sc2)) / sc2 - 0.9 * np.exp(XY / (2 * ss2)) / ss2)
    return gb / (2 * np.pi)


def rf_bank(bank, Px, Py, dx=1.0, dy=1.0, x_start=None, y_start=None) -> np.ndarray:
    """Mixed bank (kind 0 Gabor, 1 DoG) -> (N, Py, Px)."""
    Sx, Sy = Px * dx, Py * dy
    out = np.empty((bank.num_neurons, Py, Px), np.float64)
    g = bank.kind == 0
    if g.any():
        out[g] = rf_gabor(bank.h_alpha[g], bank.h_l[g], bank.h_x0[g], bank.h_y0[g],
                          bank.h_ab[g], bank.KAPPA, Px, Py, Sx, Sy, x_start, y_start)
    d = ~g
    if d.any():
        out[d] = rf_dog(bank.h_alpha[d], bank.h_x0[d], bank.h_y0[d], bank.sigma_center,
                        bank.sigma_surround, Px, Py, Sx, Sy, x_start, y_start)
    return out


def filter_video(V: np.ndarray, RF: np.ndarray, dx=1.0, dy=1.0) -> np.ndarray:
    """v[t, j] = dx*dy*sum_p V[t,p] RF_j[p]  (vtem/vrf.py:264-292: DGEMM then transpose)."""
    T = V.shape[0]
    N = RF.shape[0]
    return (dx * dy) * (V.reshape(T, -1) @ RF.reshape(N, -1).T)


# --------------------------------------------------------------------------------------
# Bases (new): real trigonometric bases of the Dirichlet space
# --------------------------------------------------------------------------------------

def spatial_modes(Lx: int, Ly: int, mask: str = "rect") -> np.ndarray:
    """Half-set modes (mx, my) in basis order; the ellipse mask is the reference's
    mx^2 My^2 + my^2 Mx^2 <= Mx^2 My^2 (vtem/vrf.py:1331, :1419)."""
    out = []
    for my in range(0, Ly + 1):
        for mx in range(-Lx, Lx + 1):
            if my == 0 and mx <= 0:
                continue
            if mask == "ellipse" and mx * mx * Ly * Ly + my * my * Lx * Lx > Lx * Lx * Ly * Ly:
                continue
            out.append((mx, my))
    return np.array(out, np.int64).reshape(-1, 2)


def spatial_basis(x: np.ndarray, y: np.ndarray, modes: np.ndarray, Ox: float, Oy: float):
    """B[q, p, ps] for the grid x (Px,), y (Py,) -> (Py, Px, Ds)."""
    Ds = 1 + 2 * len(modes)
    B = np.empty((len(y), len(x), Ds), np.float64)
    B[..., 0] = 1.0
    for i, (mx, my) in enumerate(modes):
        ph = mx * Ox * x[None, :] + my * Oy * y[:, None]
        B[..., 1 + 2 * i] = SQRT2 * np.cos(ph)
        B[..., 2 + 2 * i] = SQRT2 * np.sin(ph)
    return B


def temporal_basis(t: np.ndarray, Lt: int, W: float) -> np.ndarray:
    """Theta[i, pt] at times t -> (len(t), 2Lt+1)."""
    t = np.asarray(t, np.float64)
    out = np.empty((t.size, 2 * Lt + 1), np.float64)
    out[:, 0] = 1.0
    for m in range(1, Lt + 1):
        out[:, 2 * m - 1] = SQRT2 * np.cos(m * W * t)
        out[:, 2 * m] = SQRT2 * np.sin(m * W * t)
    return out


def temporal_factors(tk1: np.ndarray, tk2: np.ndarray, Lt: int, W: float) -> np.ndarray:
    """alpha[k, pt] = integral over [tk1, tk2] of Theta_pt (n, 2Lt+1).  (new)

    Product-to-sum form: sin b - sin a = 2 cos((a+b)/2) sin((b-a)/2),
    cos a - cos b = 2 sin((a+b)/2) sin((b-a)/2).
    """
    tk1 = np.asarray(tk1, np.float64)
    tk2 = np.asarray(tk2, np.float64)
    out = np.empty((tk1.size, 2 * Lt + 1), np.float64)
    out[:, 0] = tk2 - tk1
    mid = 0.5 * (tk1 + tk2)
    half = 0.5 * (tk2 - tk1)
    for m in range(1, Lt + 1):
        w = m * W
        s = 2.0 * SQRT2 * np.sin(w * half) / w
        out[:, 2 * m - 1] = s * np.cos(w * mid)
        out[:, 2 * m] = s * np.sin(w * mid)
    return out


def spatial_coeffs(RF: np.ndarray, B: np.ndarray, dx=1.0, dy=1.0) -> np.ndarray:
    """S[j, ps] = dx*dy*sum_pixels RF_j * B_ps  (N, Ds).  (new; real form of Ds)"""
    N = RF.shape[0]
    return (dx * dy) * (RF.reshape(N, -1) @ B.reshape(-1, B.shape[-1]))


# --------------------------------------------------------------------------------------
# Temporal upsampling (new; R4')
# --------------------------------------------------------------------------------------

def trig_coeffs(v: np.ndarray, Lt: int, period: float, fps: float) -> np.ndarray:
    """Real temporal coefficients C (Dt, N) of frame-rate samples v (T, N) taken at
    t_f = f/fps over one period: exact for trig polynomials of order Lt < T/2."""
    T = v.shape[0]
    assert abs(T / fps - period) < 1e-9 * max(1.0, period), "frames must span one period"
    Th = temporal_basis(np.arange(T) / fps, Lt, 2 * np.pi / period)
    return (Th.T @ v) / T


def upsample(v: np.ndarray, Lt: int, period: float, fps: float, t_out: np.ndarray):
    C = trig_coeffs(v, Lt, period, fps)
    return temporal_basis(t_out, Lt, 2 * np.pi / period) @ C


# --------------------------------------------------------------------------------------
# Interval selection and t-transform (vtem/utils/spike_store.py, vtem/dirichlet.py)
# --------------------------------------------------------------------------------------

def select_intervals(spike_lists: Sequence[np.ndarray], window: Tuple[float, float]):
    """Intervals per neuron for a decode window (vtem/utils/spike_store.py:279-351).

    Spike times are shifted to the window start (:319); per neuron the spikes in
    (0, T_win) plus the one before and the one after are kept (:327-334); consecutive
    pairs give (tk1, tk2), flattened neuron-major with neuron_ind (:340-349).
    Returns tk1, tk2 (float64), neuron_ind (int32), counts (int32 per neuron).
    """
    t0, t1 = window
    Tw = t1 - t0
    tk1s, tk2s, inds, counts = [], [], [], []
    for j, s in enumerate(spike_lists):
        s = np.asarray(s, np.float64) - t0
        a = np.nonzero((s > 0) & (s < Tw))[0]
        if a.size > 0:
            tk = s[max(0, a[0] - 1):min(a[-1] + 2, s.size)]
        else:
            tk = np.zeros(0)
        c = max(0, tk.size - 1)
        counts.append(c)
        if c:
            tk1s.append(tk[:-1])
            tk2s.append(tk[1:])
            inds.append(np.full(c, j, np.int32))
    if tk1s:
        return (np.concatenate(tk1s), np.concatenate(tk2s), np.concatenate(inds),
                np.array(counts, np.int32))
    return np.zeros(0), np.zeros(0), np.zeros(0, np.int32), np.array(counts, np.int32)


def select_neurons(h_x0, h_y0, domain, tol: float = 0.01):
    """Neurons to decode for a domain [x_start, x_end, y_start, y_end]
    (vtem/utils/spike_store.py:243-277): indices whose RF centre satisfies
    x_start - 0.01 < x0 < x_end + 0.01 and the same in y (strict, :266), with the centres
    shifted by the domain's centre x_start + Sx/2, y_start + Sy/2 (:260-271).
    Returns (choose, x0 shifted, y0 shifted)."""
    h_x0 = np.asarray(h_x0, np.float64)
    h_y0 = np.asarray(h_y0, np.float64)
    Sx = domain[1] - domain[0]
    Sy = domain[3] - domain[2]
    shiftx = domain[0] + Sx / 2
    shifty = domain[2] + Sy / 2
    choose = []
    for j in range(h_x0.size):
        if domain[0] - tol < h_x0[j] < domain[1] + tol and domain[2] - tol < h_y0[j] < domain[3] + tol:
            choose.append(j)
    choose = np.array(choose, np.int64)
    return choose, h_x0[choose] - shiftx, h_y0[choose] - shifty


def decode_norm(kappa, sigma) -> np.ndarray:
    """norm_j = 1/(kappa sigma) for sigma != 0, else 1, divided by the max
    (vtem/vtdm.py:634-638)."""
    kappa = np.asarray(kappa, np.float64)
    sigma = np.asarray(sigma, np.float64)
    h = np.ones(kappa.size)
    nz = sigma != 0
    h[nz] = 1.0 / (kappa[nz] * sigma[nz])
    return h / h.max()


def compute_q(tk1, tk2, neuron_ind, kappa, delta, bias, norm) -> np.ndarray:
    """q_k = (kappa_j delta_j - b_j (tk2 - tk1)) norm_j  (vtem/dirichlet.py:204-222)."""
    j = np.asarray(neuron_ind)
    return (kappa[j] * delta[j] - bias[j] * (tk2 - tk1)) * norm[j]


# --------------------------------------------------------------------------------------
# Reference spike-space decoder (vtem/vrf.py, vtem/dirichlet.py, vtem/utils/linalg.py)
# --------------------------------------------------------------------------------------

def compute_Ds_fft(RF: np.ndarray, dx: float, dy: float, Mx: int, My: int,
                   mask: str = "ellipse") -> np.ndarray:
    """Spatial Dirichlet coefficients by 2-D FFT (vtem/vrf.py:333-377, :1293-1350).

    RF (N, Py, Px) sampled over exactly one period.  Returns d_Ds (N, 2My+1, 2Mx+1)
    complex with the reference's storage: FFT row-frequency my goes to row My-my
    (:1337), sign (-1)^(mx+my) dx dy (:1321-1327), zero outside the mask (:1331).
    """
    N, Py, Px = RF.shape
    F = np.fft.fft2(RF, axes=(1, 2))
    out = np.zeros((N, 2 * My + 1, 2 * Mx + 1), np.complex128)
    for my in range(-My, My + 1):
        for mx in range(-Mx, Mx + 1):
            if mask == "ellipse" and mx * mx * My * My + my * my * Mx * Mx > Mx * Mx * My * My:
                continue
            sgn = dx * dy if (mx + my) % 2 == 0 else -dx * dy
            out[:, My - my, mx + Mx] = sgn * F[:, my % Py, mx % Px]
    return out


def compute_Dsw(d_Ds: np.ndarray, norm: np.ndarray) -> np.ndarray:
    """Dsw = Re(Ds Ds^H) * norm norm^T  (vtem/vrf.py:502-536, put_norm :540-559)."""
    N = d_Ds.shape[0]
    D = d_Ds.reshape(N, -1)
    W = (D @ D.conj().T).real
    return W * np.outer(norm, norm)


def compute_psi(d_Ds: np.ndarray, Px: int, Py: int, Sx: float, Sy: float, WMx: float,
                WMy: float, mask: str = "ellipse", x_start=None, y_start=None) -> np.ndarray:
    """psi_j(x,y) = Re sum_mask Ds_j(m) e^{i(mx WMx x + my WMy y)} on the decode grid
    (vtem/vrf.py:1373-1433; row my+My read at :1412)."""
    N, Dy, Dx = d_Ds.shape
    My, Mx = (Dy - 1) // 2, (Dx - 1) // 2
    x, y = _grid(Px, Py, Sx, Sy, x_start, y_start)
    out = np.zeros((N, Py, Px), np.complex128)
    for my in range(-My, My + 1):
        for mx in range(-Mx, Mx + 1):
            if mask == "ellipse" and mx * mx * My * My + my * my * Mx * Mx > Mx * Mx * My * My:
                continue
            e = np.exp(1j * (mx * WMx * x[None, :] + my * WMy * y[:, None]))
            out += d_Ds[:, my + My, mx + Mx][:, None, None] * e[None]
    return out.real


def compute_G_spike(tk1, tk2, neuron_ind, Dsw, Wt: float, Mt: int) -> np.ndarray:
    """Spike-space Gram, separable closed form (vtem/dirichlet.py:272-330):

    G_kl = [2/WM^2 sum_{m=1}^{M} (cos mWM(tk2-tl2) - cos mWM(tk2-tl1) - cos mWM(tk1-tl2)
            + cos mWM(tk1-tl1))/m^2 + (tk2-tk1)(tl2-tl1)] * Dsw[j(k), j(l)],  WM = Wt/Mt.
    """
    WM = Wt / Mt
    a2 = tk2[:, None]
    a1 = tk1[:, None]
    b2 = tk2[None, :]
    b1 = tk1[None, :]
    s = np.zeros((tk1.size, tk1.size))
    for m in range(1, Mt + 1):
        s += (np.cos(m * WM * (a2 - b2)) - np.cos(m * WM * (a2 - b1))
              - np.cos(m * WM * (a1 - b2)) + np.cos(m * WM * (a1 - b1))) / (m * m)
    s = s * 2 / (WM * WM) + (a2 - a1) * (b2 - b1)
    j = np.asarray(neuron_ind)
    return s * Dsw[np.ix_(j, j)]


def add_diag(G: np.ndarray, lamb: float) -> np.ndarray:
    """G += lamb I  (vtem/dirichlet.py:334-350)."""
    G = G.copy()
    G[np.diag_indices_from(G)] += lamb
    return G


def solve_eq_sym(G: np.ndarray, q: np.ndarray, rcond: float = 1e-8) -> np.ndarray:
    """c = U diag(1/s_i if s_i > rcond s_0 else 0) U^T q with U the LEFT singular vectors
    and s the singular values of G, used on both sides (vtem/utils/linalg.py:622-630:
    svd(G, compute_v=0), then dot(U, q, opa='c'), sq_Kernel :790-831 divides by s, then
    dot(U, qq)).  For symmetric G = sum w_i u_i u_i^T the left singular vectors are the
    eigenvectors and s_i = |w_i|, so this is the truncated pinv of |G| = sum |w_i| u_i u_i^T:
    the pinv of G itself when G is positive semidefinite (every Gram), and a different
    matrix for an indefinite G."""
    w, U = np.linalg.eigh(G)
    s = np.abs(w)
    qq = U.T @ q
    keep = s > rcond * s.max()
    qq = np.where(keep, qq / np.where(keep, s, 1.0), 0.0)
    return U @ qq


def rnn3(G: np.ndarray, q: np.ndarray, dt: float = 1e-6, alpha: float = 5000,
         steps: int = 4000) -> np.ndarray:
    """Gradient-flow decoder: c <- c + dt*alpha*(q - G c), steps+1 iterations from 0
    (vtem/neural_network.py:60-92, update kernel :102-128)."""
    c = np.zeros_like(q)
    eta = float(dt) * float(alpha)
    for _ in range(steps + 1):
        y = q - G @ c
        c = c + eta * y
    return c


def reconstruct_spike(c, tk1, tk2, neuron_ind, norm, psi: np.ndarray, t: np.ndarray,
                      Mt: int, WM: float) -> np.ndarray:
    """u(t,p) = sum_k c_k norm_j psi_j(p) [2/WM sum_{m=1}^{M} (sin mWM(t-tk1) -
    sin mWM(t-tk2))/m + tk2 - tk1]  (vtem/dirichlet.py:165-189, kernel :380-413)."""
    t = np.asarray(t, np.float64)[:, None]
    phi = np.zeros((t.size, tk1.size))
    for m in range(1, Mt + 1):
        phi += (np.sin(m * WM * (t - tk1[None])) - np.sin(m * WM * (t - tk2[None]))) / m
    temporal = phi * 2 / WM + (tk2 - tk1)[None]                    # (T, n)
    j = np.asarray(neuron_ind)
    w = c * norm[j]
    N, Py, Px = psi.shape
    # sum over k: temporal[t,k] w[k] psi[j(k), p]
    acc = np.zeros((N, t.size))
    np.add.at(acc.T, (slice(None), j), temporal * w[None])  # acc[j, t] += ...
    return np.einsum("jt,jp->tp", acc, psi.reshape(N, -1)).reshape(t.size, Py, Px)


# --------------------------------------------------------------------------------------
# Coefficient-space (D-space) decoder (new; exactly equivalent by push-through)
# --------------------------------------------------------------------------------------

def phi_matrix(S: np.ndarray, alpha: np.ndarray, neuron_ind: np.ndarray, norm=None):
    """Phi[k, ps*Dt + pt] = norm_j S[j, ps] alpha[k, pt] (row-wise Khatri-Rao)."""
    j = np.asarray(neuron_ind)
    nk = 1.0 if norm is None else np.asarray(norm)[j][:, None, None]
    P = S[j][:, :, None] * alpha[:, None, :] * nk
    return P.reshape(alpha.shape[0], -1)


def neuron_T(alpha: np.ndarray, q: np.ndarray, neuron_ind: np.ndarray, N: int, norm=None):
    """T_j = sum_{k in j} w_k^2 alpha_k alpha_k^T and t_j = sum_{k in j} w_k alpha_k q_k,
    with w_k = norm_j (segmented reductions over each neuron's intervals)."""
    j = np.asarray(neuron_ind)
    w = np.ones(j.size) if norm is None else np.asarray(norm)[j]
    Dt = alpha.shape[1]
    Tm = np.zeros((N, Dt, Dt))
    tv = np.zeros((N, Dt))
    aw = alpha * w[:, None]
    np.add.at(Tm, j, aw[:, :, None] * aw[:, None, :])
    np.add.at(tv, j, aw * q[:, None])
    return Tm, tv


def gram_kron(S: np.ndarray, Tm: np.ndarray, chunk: int = 16) -> np.ndarray:
    """A = sum_j (S_j S_j^T) kron T_j  (D x D, index ps*Dt + pt), built `chunk` spatial
    rows at a time so the N x (chunk*Ds) product stays small."""
    N, Ds = S.shape
    Dt = Tm.shape[1]
    T2 = Tm.reshape(N, Dt * Dt)
    A = np.empty((Ds * Dt, Ds * Dt))
    for p0 in range(0, Ds, chunk):
        p1 = min(Ds, p0 + chunk)
        P = (S[:, p0:p1, None] * S[:, None, :]).reshape(N, (p1 - p0) * Ds)
        A4 = (P.T @ T2).reshape(p1 - p0, Ds, Dt, Dt)
        A[p0 * Dt:p1 * Dt] = A4.transpose(0, 2, 1, 3).reshape((p1 - p0) * Dt, Ds * Dt)
    return A


def rhs_kron(S: np.ndarray, tv: np.ndarray) -> np.ndarray:
    """r = sum_j S_j kron t_j (D,)."""
    return (S.T @ tv).reshape(-1)


def tikhonov_lambda(A: np.ndarray, n_rows: int, lam_rel: float = 1e-6) -> float:
    """lambda = lam_rel * tr(G) / n with tr(G) = tr(A) and n the number of intervals
    (rows of the spike-space G)."""
    return lam_rel * float(np.trace(A)) / float(n_rows)


def solve_tikhonov(A: np.ndarray, r: np.ndarray, lam: float) -> np.ndarray:
    """(A + lam I) X = r by Cholesky."""
    M = A.copy()
    M[np.diag_indices_from(M)] += lam
    c, low = sla.cho_factor(M, lower=True, check_finite=False)
    return sla.cho_solve((c, low), r, check_finite=False)


def rnn_dspace(A: np.ndarray, r: np.ndarray, eta: float, steps: int, lam: float = 0.0):
    """X <- X + eta (r - (A + lam I) X), steps+1 iterations from 0: the image under
    Phi^T of the reference's rnn3 iterates (vtem/neural_network.py:78-82)."""
    X = np.zeros_like(r)
    for _ in range(steps + 1):
        X = X + eta * (r - A @ X - lam * X)
    return X


def lanczos_eta(A: np.ndarray, lam: float, steps: int = 16) -> float:
    """The RNN step eta = 1/theta, theta the largest Ritz value of A + lam I after `steps`
    Lanczos steps from the normalised all-ones vector (no reorthogonalisation; a zero
    beta ends the recurrence).  The reference steps its ODE by dt*alpha
    (vtem/neural_network.py:78-92); the step 1/lambda_max is this build's rule for the
    configurations, and the GPU path computes it the same way."""
    n = A.shape[0]
    v = np.ones(n) / np.sqrt(n)
    vp = np.zeros(n)
    bprev = 0.0
    al, be = [], []
    for _ in range(steps):
        u = A @ v + lam * v
        a = float(u @ v)
        w = u - a * v - bprev * vp
        b = float(np.sqrt(w @ w))
        if not (b > 0.0 and np.isfinite(b)):
            b = 0.0
        al.append(a)
        be.append(b)
        vp, v = v, (w / b if b > 0.0 else np.zeros(n))
        bprev = b
    T = np.diag(al) + np.diag(be[:-1], 1) + np.diag(be[:-1], -1)
    th = float(np.linalg.eigvalsh(T)[-1])
    return 1.0 / th if th > 0.0 else 0.0


def synthesize(X: np.ndarray, Bs: np.ndarray, Theta: np.ndarray) -> np.ndarray:
    """u[t, q, p] = sum_{ps, pt} X[ps, pt] Bs[q, p, ps] Theta[t, pt]."""
    Py, Px, Ds = Bs.shape
    Dt = Theta.shape[1]
    Xm = X.reshape(Ds, Dt)
    return (Theta @ (Bs.reshape(-1, Ds) @ Xm).T).reshape(Theta.shape[0], Py, Px)


# --------------------------------------------------------------------------------------
# Stitching windows and metrics
# --------------------------------------------------------------------------------------

def hann_ramp(L: int) -> np.ndarray:
    """Rising half of a length-2L Hann window (vtem/vtdm.py:254, :617)."""
    if L <= 0:
        return np.zeros(0)
    n = np.arange(2 * L)
    return (0.5 - 0.5 * np.cos(2 * np.pi * n / (2 * L - 1)))[:L]


def tile_window(length: int, left: int, right: int) -> np.ndarray:
    """[rising ramp, ones, 1 - rising ramp] (vtem/vtdm.py:252-258)."""
    return np.concatenate([hann_ramp(left), np.ones(length - left - right),
                           1.0 - hann_ramp(right)])


def snr_db(u: np.ndarray, urec: np.ndarray) -> float:
    u = np.asarray(u, np.float64)
    e = u - np.asarray(urec, np.float64)
    return 10.0 * math.log10(float(np.sum(u * u)) / float(np.sum(e * e)))


def vtdm_windows(decode_window, start_time: float, end_time: float, dt: float,
                 stitching: bool = False, stitch_interval=None) -> np.ndarray:
    """Temporal window loop of VTDM (vtem/vtdm.py:570-692): windows of stitch_interval
    overlapping by 20 % (4 frames when that is under 3 frames and the window has at least
    8), each decoded by decode_window(start, end) -> (frames, Py, Px) sampled at dt from
    the window start, then written with a Hann crossfade over the overlaps."""
    end_time = float(end_time)
    if stitching:
        stitch_interval = 20 * dt if stitch_interval is None else float(stitch_interval)
        stitch_interval = min(stitch_interval, end_time)
        end = round(float(min(start_time + stitch_interval, end_time)) / dt) * dt
        overlap = 0.2 * stitch_interval
    else:
        stitch_interval = end_time
        end = end_time
        overlap = 0.0
    total_l = int(round((end_time - start_time) / dt))
    overlap_l = int(round(overlap / dt))
    if overlap_l < 3:
        if round(stitch_interval / dt) >= 8:
            overlap_l, overlap = 4, 4 * dt
        else:
            overlap_l, overlap = 0, 0.0
    interval_l = int(round(stitch_interval / dt))
    lw = hann_ramp(overlap_l)[:, None, None]
    rw = 1.0 - lw
    start = float(start_time)
    u = None
    i = 0
    prev = None
    while True:
        u_rec = decode_window(start, end)
        if u is None:
            u = np.empty((total_l,) + u_rec.shape[1:])
        if overlap_l:
            if i == 0:
                if end_time - end < dt:
                    return u_rec
                u[0:interval_l - overlap_l] = u_rec[0:-overlap_l]
                prev = u_rec[-overlap_l:]
            else:
                l_ind = i * (interval_l - overlap_l)
                r_ind = (i + 1) * (interval_l - overlap_l)
                u[l_ind:l_ind + overlap_l] = u_rec[0:overlap_l] * lw + prev * rw
                if end_time - end < dt:
                    u[l_ind + overlap_l:] = u_rec[-u.shape[0] + l_ind + overlap_l:]
                    return u
                u[l_ind + overlap_l:r_ind] = u_rec[overlap_l:interval_l - overlap_l]
                prev = u_rec[-overlap_l:]
        else:
            u[i * interval_l:i * interval_l + u_rec.shape[0]] = u_rec
            if end_time - end < dt:
                return u
        start = start + round(stitch_interval / dt) * dt - round(overlap / dt) * dt
        end = min(start + round(stitch_interval / dt) * dt, end_time)
        if end_time - end < overlap:
            end = end_time
        i += 1


def decode_tiles(decode_tile, domain, dx: float, dy: float, spatial_interval, total_l: int) -> np.ndarray:
    """Spatial tile loop of decode_video (vtem/vtdm.py:179-309): tiles of spatial_interval
    pixels overlapping by 20 %, placed from the right / top edges (the last tile clamped to
    the domain), each decoded by decode_tile(domain) -> (total_l, py, px) and accumulated
    with [Hann ramp, ones, 1 - Hann ramp] windows on the sides shared with other tiles."""
    xs, xe, ys, ye = domain
    decPx, decPy = int(round((xe - xs) / dx)), int(round((ye - ys) / dy))
    sx, sy = min(spatial_interval[0], decPx), min(spatial_interval[1], decPy)
    ox, oy = int(round(0.2 * sx)), int(round(0.2 * sy))
    u = np.zeros((total_l, decPy, decPx))
    start_y, end_y = max(ye - sy * dy, ys), ye
    y0, y1 = 0, sy
    while True:
        start_x, end_x = max(xe - sx * dx, xs), xe
        x0, x1 = decPx - sx, decPx
        while True:
            left = 0 if x0 == 0 else ox
            right = 0 if x1 == decPx else ox
            top = 0 if y0 == 0 else oy
            bottom = 0 if y1 == decPy else oy
            w = np.outer(tile_window(y1 - y0, top, bottom), tile_window(x1 - x0, left, right))
            u[:, y0:y1, x0:x1] += decode_tile([start_x, end_x, start_y, end_y]) * w[None]
            if start_x == xs:
                break
            start_x = max(start_x - sx * dx + ox * dx, xs)
            end_x = end_x - sx * dx + ox * dx
            x0 = max(x0 - sx + ox, 0)
            x1 = x1 - sx + ox
        if start_y == ys:
            break
        start_y = max(start_y - sy * dy + oy * dy, ys)
        end_y = end_y - sy * dy + oy * dy
        y0 = y1 - oy
        y1 = min(y0 + sy, decPy)
    return u
