"""Real trigonometric bases of the Dirichlet space (host-side constant tables).

The reference expands videos in complex exponentials e_m (vtem/dirichlet.py:15-35;
spatial coefficients vtem/vrf.py:333-377, synthesis :1373-1433).  For real videos and
real receptive fields the same space has an orthonormal real basis, which keeps every
kernel in real arithmetic:

* spatial index ps: 0 -> 1; mode i of the half set H = {my > 0} U {my = 0, mx > 0}
  (ordered my = 0..Ly, mx = -Lx..Lx, optionally restricted to the reference's ellipse
  mx^2 Ly^2 + my^2 Lx^2 <= Lx^2 Ly^2, vtem/vrf.py:1331) gives ps = 1+2i -> sqrt2 cos(phase)
  and ps = 2+2i -> sqrt2 sin(phase), phase = mx*Ox*x + my*Oy*y;
* temporal index pt: 0 -> 1, 2m-1 -> sqrt2 cos(m W t), 2m -> sqrt2 sin(m W t);
* coefficient-space index d = ps*Dt + pt.

Pixel coordinates follow vtem/vrf.py:1254-1257 (x_p = p*dx - Sx/2, y_q = -(q*dy - Sy/2)).
"""

from __future__ import annotations

import dataclasses
import math
from functools import lru_cache

import numpy as np

SQRT2 = math.sqrt(2.0)


def modes(Lx: int, Ly: int, mask: str = "rect") -> np.ndarray:
    out = []
    for my in range(Ly + 1):
        for mx in range(-Lx, Lx + 1):
            if my == 0 and mx <= 0:
                continue
            if mask == "ellipse" and mx * mx * Ly * Ly + my * my * Lx * Lx > Lx * Lx * Ly * Ly:
                continue
            out.append((mx, my))
    return np.asarray(out, np.int64).reshape(-1, 2)


def grid_x(Px: int, dx: float, Sx: float = None, x_start: float = None) -> np.ndarray:
    Sx = Px * dx if Sx is None else Sx
    x_start = -Sx / 2 if x_start is None else x_start
    return np.arange(Px, dtype=np.float64) * (Sx / Px) + x_start


def grid_y(Py: int, dy: float, Sy: float = None, y_start: float = None) -> np.ndarray:
    Sy = Py * dy if Sy is None else Sy
    y_start = -Sy / 2 if y_start is None else y_start
    return -(np.arange(Py, dtype=np.float64) * (Sy / Py) + y_start)


def spatial_table(x: np.ndarray, y: np.ndarray, md: np.ndarray, Ox: float, Oy: float) -> np.ndarray:
    """(Py*Px, Ds) table of the real spatial basis on the grid (row-major pixels)."""
    Ds = 1 + 2 * len(md)
    ph = (md[:, 0][None, None, :] * Ox * x[None, :, None]
          + md[:, 1][None, None, :] * Oy * y[:, None, None])          # (Py, Px, H)
    B = np.empty((len(y), len(x), Ds), np.float64)
    B[..., 0] = 1.0
    B[..., 1::2] = SQRT2 * np.cos(ph)
    B[..., 2::2] = SQRT2 * np.sin(ph)
    return B.reshape(-1, Ds)


def temporal_table(t: np.ndarray, Lt: int, W: float) -> np.ndarray:
    t = np.asarray(t, np.float64)
    m = np.arange(1, Lt + 1, dtype=np.float64)
    out = np.empty((t.size, 2 * Lt + 1), np.float64)
    out[:, 0] = 1.0
    ph = np.outer(t, m) * W
    out[:, 1::2] = SQRT2 * np.cos(ph)
    out[:, 2::2] = SQRT2 * np.sin(ph)
    return out


@dataclasses.dataclass(frozen=True)
class Geometry:
    """One decode block: pixel grid, Dirichlet orders, periods and time grids.

    Px, Py, dx, dy    pixel grid of the block (RFs are evaluated on it);
    Lx, Ly, Lt        Dirichlet orders; the spatial periods are `period_x`, `period_y`
                      (default: the block extent) and the temporal period is `period`
                      (seconds; default: the duration);
    fps               frame rate of the video and of the reconstruction;
    dt                IAF integration step;
    duration          length of the encoded / decoded window (seconds).

    A block decoded inside a longer period (the reference's rule: period 2 pi M / W from
    the orders and the bandwidth, independent of the block's extent, vtem/vtdm.py:385-388,
    :964-967) keeps its pixel grid centred at 0 and only changes the basis frequencies
    2 pi / period.
    """
    Px: int
    Py: int
    Lx: int
    Ly: int
    Lt: int
    duration: float
    fps: float = 100.0
    dt: float = 1e-4
    dx: float = 1.0
    dy: float = 1.0
    period: float = None
    mask: str = "rect"
    period_x: float = None
    period_y: float = None

    @property
    def T_period(self) -> float:
        return self.duration if self.period is None else self.period

    @property
    def P(self) -> int:
        return self.Px * self.Py

    @property
    def Sx(self) -> float:
        return self.Px * self.dx

    @property
    def Sy(self) -> float:
        return self.Py * self.dy

    @property
    def Ox(self) -> float:
        return 2 * math.pi / (self.Sx if self.period_x is None else self.period_x)

    @property
    def Oy(self) -> float:
        return 2 * math.pi / (self.Sy if self.period_y is None else self.period_y)

    @property
    def W(self) -> float:
        return 2 * math.pi / self.T_period

    @property
    def Ds(self) -> int:
        return 1 + 2 * len(modes(self.Lx, self.Ly, self.mask))

    @property
    def Dt(self) -> int:
        return 2 * self.Lt + 1

    @property
    def D(self) -> int:
        return self.Ds * self.Dt

    @property
    def frames(self) -> int:
        return int(round(self.duration * self.fps))

    @property
    def iaf_samples(self) -> int:
        return int(round(self.duration / self.dt)) + 1

    def frame_times(self) -> np.ndarray:
        return np.arange(self.frames, dtype=np.float64) / self.fps

    def iaf_times(self) -> np.ndarray:
        return np.arange(self.iaf_samples, dtype=np.float64) * self.dt

    @lru_cache(maxsize=None)
    def spatial(self) -> np.ndarray:
        return spatial_table(grid_x(self.Px, self.dx), grid_y(self.Py, self.dy),
                             modes(self.Lx, self.Ly, self.mask), self.Ox, self.Oy)

    def temporal(self, t) -> np.ndarray:
        return temporal_table(t, self.Lt, self.W)
