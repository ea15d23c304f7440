"""Seeded synthetic workloads: bandlimited video, receptive-field banks, IAF parameters.

Everything here is input generation shared by the HIP path and the float64 oracle,
so both see bit-identical inputs.  Nothing in this module is on the hot path.

Coordinate conventions follow the reference's RF kernels (vtem/vrf.py:1254-1257):
pixel (row q, column p) of a frame of Px x Py pixels with spacing dx, dy sits at

    x_p = p*dx - Sx/2,     y_q = -(q*dy - Sy/2),     Sx = Px*dx, Sy = Py*dy,

i.e. x grows to the right and y grows upwards while the row index grows downwards.

Video model (Dirichlet / trigonometric-polynomial space, vtem/dirichlet.py:15-35):

    u(x, y, t) = sum_m  c[m] exp(i(mx*Ox*x + my*Oy*y + mt*Ot*t)),
    |mx| <= Lx, |my| <= Ly, |mt| <= Lt,  Ox = 2pi/Sx, Oy = 2pi/Sy, Ot = 2pi/T,

with c[m] complex N(0,1) (real and imaginary parts each N(0,1/2)), Hermitian
(c[-m] = conj(c[m]), c[0] real N(0,1)), scaled so that max|u| = 1 over the samples.
"""

from __future__ import annotations

import dataclasses
import math
from typing import Optional, Sequence

import numpy as np

TWO_PI = 2.0 * math.pi


# --------------------------------------------------------------------------------------
# Video
# --------------------------------------------------------------------------------------

def video_coefficients(Lx: int, Ly: int, Lt: int, rng: np.random.Generator) -> np.ndarray:
    """Hermitian complex N(0,1) coefficients c[mt+Lt, my+Ly, mx+Lx]."""
    shape = (2 * Lt + 1, 2 * Ly + 1, 2 * Lx + 1)
    z = (rng.standard_normal(shape) + 1j * rng.standard_normal(shape)) / math.sqrt(2.0)
    # (z[m] + conj z[-m]) / sqrt(2): E|c|^2 = 1 for m != 0; c[0] = sqrt(2) Re z[0] ~ N(0,1).
    zr = np.conj(z[::-1, ::-1, ::-1])
    return (z + zr) / math.sqrt(2.0)


def pixel_coords(P: int, d: float, axis: str) -> np.ndarray:
    """Physical coordinate of each pixel along one axis (vtem/vrf.py:1254-1257)."""
    S = P * d
    idx = np.arange(P, dtype=np.float64)
    if axis == "x":
        return idx * d - S / 2.0
    return -(idx * d - S / 2.0)


def evaluate_video(c: np.ndarray, Px: int, Py: int, times: np.ndarray, dx: float = 1.0,
                   dy: float = 1.0, Sx: Optional[float] = None, Sy: Optional[float] = None,
                   T: float = 1.0, x0: float = 0.0, y0: float = 0.0,
                   out: Optional[np.ndarray] = None) -> np.ndarray:
    """Sample u at frame times `times` on a Px x Py pixel grid -> (len(times), Py, Px).

    (Sx, Sy, T) are the spatial/temporal periods (default: the frame extent).  (x0, y0)
    shifts the grid origin so that sub-blocks of a larger video can be sampled with the
    larger video's periods.
    """
    Dt, Dy, Dx = c.shape
    Lt, Ly, Lx = (Dt - 1) // 2, (Dy - 1) // 2, (Dx - 1) // 2
    Sx = Px * dx if Sx is None else Sx
    Sy = Py * dy if Sy is None else Sy
    xs = pixel_coords(Px, dx, "x") + x0
    ys = pixel_coords(Py, dy, "y") + y0
    Ex = np.exp(1j * np.outer(xs, np.arange(-Lx, Lx + 1)) * (TWO_PI / Sx))      # (Px, Dx)
    Ey = np.exp(1j * np.outer(ys, np.arange(-Ly, Ly + 1)) * (TWO_PI / Sy))      # (Py, Dy)
    Et = np.exp(1j * np.outer(np.asarray(times, np.float64), np.arange(-Lt, Lt + 1))
                * (TWO_PI / T))                                                   # (Nt, Dt)
    a = np.einsum("tyx,px->typ", c, Ex)            # (Dt, Dy, Px)
    b = np.einsum("typ,qy->tqp", a, Ey)            # (Dt, Py, Px)
    nt = len(times)
    if out is None:
        out = np.empty((nt, Py, Px), np.float64)
    br = b.reshape(Dt, -1)
    step = max(1, int(2e7 // max(1, Py * Px)))
    for s in range(0, nt, step):
        e = min(nt, s + step)
        out[s:e] = (Et[s:e] @ br).real.reshape(e - s, Py, Px)
    return out


@dataclasses.dataclass
class Video:
    frames: np.ndarray          # (T_frames, Py, Px) float64, max|u| = 1 (times contrast)
    coeffs: np.ndarray          # scaled complex coefficients
    Lx: int
    Ly: int
    Lt: int
    dx: float
    dy: float
    fps: float
    duration: float
    seed: int

    @property
    def Px(self):
        return self.frames.shape[2]

    @property
    def Py(self):
        return self.frames.shape[1]


def make_video(Px: int, Py: int, duration: float, fps: float, Lx: int, Ly: int, Lt: int,
               seed: int = 0, dx: float = 1.0, dy: float = 1.0) -> Video:
    """Seeded bandlimited video whose periods equal the frame extent and the duration."""
    rng = np.random.default_rng(seed)
    c = video_coefficients(Lx, Ly, Lt, rng)
    nf = int(round(duration * fps))
    times = np.arange(nf, dtype=np.float64) / fps
    V = evaluate_video(c, Px, Py, times, dx, dy, T=duration)
    s = 1.0 / np.abs(V).max()
    V *= s
    return Video(V, c * s, Lx, Ly, Lt, dx, dy, fps, duration, seed)


# --------------------------------------------------------------------------------------
# Receptive-field banks (explicit parameter arrays, vtem/vrf.py:209-262 and :794-843)
# --------------------------------------------------------------------------------------

GABOR = 0
DOG = 1


@dataclasses.dataclass
class Bank:
    """A mixed Gabor/DoG bank as per-neuron parameter arrays.

    kind[j] = 0 (Gabor) or 1 (centre-surround DoG).  h_alpha is the dilation (the RF
    kernels use 1/h_alpha, vtem/vrf.py:258), h_l the rotation angle, (h_x0, h_y0) the
    centre, h_ab 0/1 for even/odd Gabor phase.  KAPPA is the Gabor carrier constant and
    (sigma_center, sigma_surround) the DoG widths, shared by all neurons of a kind.
    """
    kind: np.ndarray
    h_alpha: np.ndarray
    h_l: np.ndarray
    h_x0: np.ndarray
    h_y0: np.ndarray
    h_ab: np.ndarray
    KAPPA: float = 2.0 * math.pi
    sigma_center: float = 2.0
    sigma_surround: float = 3.2

    @property
    def num_neurons(self) -> int:
        return int(self.kind.size)

    def subset(self, idx) -> "Bank":
        idx = np.asarray(idx)
        return Bank(self.kind[idx], self.h_alpha[idx], self.h_l[idx], self.h_x0[idx],
                    self.h_y0[idx], self.h_ab[idx], self.KAPPA, self.sigma_center,
                    self.sigma_surround)

    def shifted(self, dx0: float, dy0: float) -> "Bank":
        return Bank(self.kind, self.h_alpha, self.h_l, self.h_x0 - dx0, self.h_y0 - dy0,
                    self.h_ab, self.KAPPA, self.sigma_center, self.sigma_surround)

    @staticmethod
    def concat(banks: Sequence["Bank"]) -> "Bank":
        b0 = banks[0]
        cat = lambda f: np.concatenate([getattr(b, f) for b in banks])
        return Bank(cat("kind"), cat("h_alpha"), cat("h_l"), cat("h_x0"), cat("h_y0"),
                    cat("h_ab"), b0.KAPPA, b0.sigma_center, b0.sigma_surround)


def _grid_centres(S: float, n: int) -> np.ndarray:
    """n centres evenly spaced over an extent S centred at 0 (cell centres)."""
    step = S / n
    return (np.arange(n) + 0.5) * step - S / 2.0


def gabor_grid_bank(Sx: float, Sy: float, sigma: float, grid: int, n_orient: int = 8,
                    KAPPA: float = 2.0 * math.pi) -> Bank:
    """8 orientations k*pi/8 x 2 phases x grid^2 centres at one scale.

    sigma (pixels) maps to h_alpha = sigma; the carrier frequency is KAPPA/(2pi*sigma)
    cycles/px, so KAPPA = 2pi gives 0.25 cyc/px at sigma = 4 px.
    """
    xs = _grid_centres(Sx, grid)
    ys = _grid_centres(Sy, grid)
    rows = []
    for k in range(n_orient):
        th = k * math.pi / n_orient
        for y0 in ys:
            for x0 in xs:
                rows.append((sigma, th, x0, y0, 0))
                rows.append((sigma, th, x0, y0, 1))
    r = np.array(rows, np.float64)
    n = r.shape[0]
    return Bank(np.zeros(n, np.int32), r[:, 0].copy(), r[:, 1].copy(), r[:, 2].copy(),
                r[:, 3].copy(), r[:, 4].astype(np.int32), KAPPA)


def dog_grid_bank(Sx: float, Sy: float, grid: int, sigma_c: float = 2.0,
                  sigma_s: float = 3.2) -> Bank:
    xs = _grid_centres(Sx, grid)
    ys = _grid_centres(Sy, grid)
    X, Y = np.meshgrid(xs, ys)
    n = X.size
    return Bank(np.ones(n, np.int32), np.ones(n), np.zeros(n), X.ravel().copy(),
                Y.ravel().copy(), np.zeros(n, np.int32), 2.0 * math.pi, sigma_c, sigma_s)


def bank_config0(Sx: float = 32.0, Sy: float = 32.0) -> Bank:
    """256 Gabor neurons: 8 orientations x 2 phases x 4x4 centres, sigma 4 px, 0.25 cyc/px."""
    return gabor_grid_bank(Sx, Sy, 4.0, 4)


def bank_dyadic(Sx: float, Sy: float) -> Bank:
    """Config-1 law: Gabor at 4 dyadic scales + DoG, Gabor:DoG = 7:1.

    Per 64x64 px: sigma=2 on an 8x8 grid, sigma=4, 8, 16 on 4x4 grids (112 centres x 16
    = 1792 Gabor) plus 256 DoG on a 16x16 grid (sigma_c = 2, sigma_s = 3.2 px).  Grids
    scale with the frame so 128x128 gives 7168 + 1024 = 8192 neurons.
    """
    f = int(round(Sx / 64.0))
    fy = int(round(Sy / 64.0))
    assert f >= 1 and f == fy, "dyadic bank law is defined for square frames of 64*k px"
    parts = [gabor_grid_bank(Sx, Sy, 2.0, 8 * f)]
    for s in (4.0, 8.0, 16.0):
        parts.append(gabor_grid_bank(Sx, Sy, s, 4 * f))
    parts.append(dog_grid_bank(Sx, Sy, 16 * f))
    return Bank.concat(parts)


def bank_random(Sx: float, Sy: float, N: int, rng: np.random.Generator) -> Bank:
    """N neurons drawn from the config-1 law for arbitrary block sizes (serving form).

    Gabor:DoG = 7:1; Gabor scales sigma = 2, 4, 8, 16 px with weights 4:1:1:1 (the
    centre counts of the dyadic law); orientation k*pi/8 and phase uniform; centres
    uniform over the block.
    """
    n_dog = int(round(N / 8.0))
    n_g = N - n_dog
    sig = rng.choice(np.array([2.0, 4.0, 8.0, 16.0]), size=n_g, p=[4 / 7, 1 / 7, 1 / 7, 1 / 7])
    th = rng.integers(0, 8, size=n_g) * (math.pi / 8)
    ab = rng.integers(0, 2, size=n_g).astype(np.int32)
    gx = rng.uniform(-Sx / 2, Sx / 2, n_g)
    gy = rng.uniform(-Sy / 2, Sy / 2, n_g)
    g = Bank(np.zeros(n_g, np.int32), sig, th, gx, gy, ab)
    dx_ = rng.uniform(-Sx / 2, Sx / 2, n_dog)
    dy_ = rng.uniform(-Sy / 2, Sy / 2, n_dog)
    d = Bank(np.ones(n_dog, np.int32), np.ones(n_dog), np.zeros(n_dog), dx_, dy_,
             np.zeros(n_dog, np.int32))
    return Bank.concat([g, d])


# --------------------------------------------------------------------------------------
# IAF parameters
# --------------------------------------------------------------------------------------

@dataclasses.dataclass
class IAFParams:
    kappa: np.ndarray
    bias: np.ndarray
    delta: np.ndarray
    v0: np.ndarray
    sigma: np.ndarray           # per-spike threshold std (0 = deterministic threshold)

    @property
    def num_neurons(self):
        return self.kappa.size

    def subset(self, idx):
        idx = np.asarray(idx)
        return IAFParams(self.kappa[idx], self.bias[idx], self.delta[idx], self.v0[idx],
                         self.sigma[idx])


def iaf_params(N: int, b: float, kappa: float, delta_mean: float, delta_std: float,
               rng: np.random.Generator) -> IAFParams:
    """Seeded IAF parameters; v0 = U(0,1)*delta as the reference default
    (vtem/ensemble_encode.py:64-65) but drawn from a seeded generator."""
    if delta_std > 0:
        delta = rng.normal(delta_mean, delta_std, N)
    else:
        delta = np.full(N, float(delta_mean))
    v0 = rng.uniform(0.0, 1.0, N) * delta
    return IAFParams(np.full(N, float(kappa)), np.full(N, float(b)), delta, v0, np.zeros(N))


# --------------------------------------------------------------------------------------
# Serving-form blocks: side in {32, 48, 64, 96} px, duration in {0.5, 1, 2} s
# --------------------------------------------------------------------------------------

SERVE_SIDES = (32, 48, 64, 96)
SERVE_DURATIONS = (0.5, 1.0, 2.0)


@dataclasses.dataclass
class ServeBlock:
    side: int
    duration: float
    Lx: int
    Lt: int
    video: Video
    bank: Bank
    iaf: IAFParams

    @property
    def D(self):
        return (2 * self.Lx + 1) ** 2 * (2 * self.Lt + 1)


def serve_block(side: int, duration: float, seed: int) -> ServeBlock:
    """One request block: Lx = Ly = round(side/8), Lt = round(4*duration), N =
    round(2*D/9) neurons from the config-1 law (Gabor:DoG = 7:1), IAF b=2.5, kappa=1,
    delta ~ N(0.04, 0.004^2), video as the other configurations (seeded)."""
    rng = np.random.default_rng(seed)
    L = int(round(side / 8.0))
    Lt = int(round(4.0 * duration))
    D = (2 * L + 1) ** 2 * (2 * Lt + 1)
    N = int(round(2.0 * D / 9.0))
    vid = make_video(side, side, duration, 100.0, L, L, Lt, seed=int(rng.integers(1 << 30)))
    bank = bank_random(float(side), float(side), N, rng)
    iaf = iaf_params(N, 2.5, 1.0, 0.04, 0.004, rng)
    return ServeBlock(side, duration, L, Lt, vid, bank, iaf)
