"""Host logic of the reference-shaped API (no GPU): parameter laws of vrf_gabor / vrf_cs
and the window / tile schedules of VTDM and decode_video, pinned by invariants."""

import math

import numpy as np
import pytest

from oracle import vtem_oracle as O


def test_gabor_parameter_law_counts_and_order():
    """vtem/vrf.py:136-205: per dilation m the grid spacing is b0 alpha0^m b0scalebase^-m,
    (2t+1)^2 centres x R rotations x 2 phases; rotations l (R-1) pi / R; entries come in
    even/odd pairs with equal parameters."""
    from vtem_amd.vrf import vrf_gabor
    g = vrf_gabor((64, 64), dx=1.0, dy=1.0, device="cpu")
    g.get_gabor_parameters(dilations=[1, 0], num_rotations=4, alpha0=2.0, b0=8.0, b0scalebase=1.0,
                           KAPPA=2 * math.pi)
    # m = 1: b = 16 -> t = 2 -> 25 centres; m = 0: b = 8 -> t = 4 -> 81 centres
    assert g.num_neurons == (25 + 81) * 4 * 2
    assert np.all(g.h_ab[0::2] == 0) and np.all(g.h_ab[1::2] == 1)
    for a in (g.h_alpha, g.h_l, g.h_x0, g.h_y0):
        np.testing.assert_array_equal(a[0::2], a[1::2])
    assert np.all(g.h_alpha[:200] == 2.0) and np.all(g.h_alpha[200:] == 1.0)
    l0 = 3 * math.pi / 4
    assert set(np.round(g.h_l / l0, 9)) == {0.0, 1.0, 2.0, 3.0}
    # first block: rotation 0, y from -t*b, x sweeping fastest
    np.testing.assert_allclose(g.h_x0[0:10:2], [-32, -16, 0, 16, 32])
    assert np.all(g.h_y0[0:10:2] == -32)


def test_cs_parameter_law():
    from vtem_amd.vrf import vrf_cs
    c = vrf_cs((32, 32), dx=1.0, dy=1.0, device="cpu")
    c.get_cs_parameters(dilations=[0, -1], alpha0=2.0, b0=4.0, b0scalebase=1.0)
    # m = 0: b = 4 -> t = 4 -> 81; m = -1: b = 2 -> t = 8 -> 289
    assert c.num_neurons == 81 + 289
    assert np.all(c.h_alpha[:81] == 1.0) and np.all(c.h_alpha[81:] == 0.5)


@pytest.mark.parametrize("stitching,interval", [(False, None), (True, 0.2), (True, 0.33), (True, 0.05)])
def test_vtdm_window_schedule_reproduces_a_common_signal(stitching, interval):
    """If every window decodes the same underlying signal f(t_abs), the crossfaded result
    is f exactly (ramps sum to one, offsets consistent) -- for any window length."""
    dt = 0.01
    f = lambda t: np.sin(7.3 * t) + t ** 2

    def window(start, end):
        n = int(round((end - start) / dt))
        t = start + np.arange(n) * dt
        return np.broadcast_to(f(t)[:, None, None], (n, 2, 3)).copy()

    u = O.vtdm_windows(window, 0.0, 1.0, dt, stitching, interval)
    t = np.arange(100) * dt
    assert u.shape == (100, 2, 3)
    np.testing.assert_allclose(u[:, 0, 0], f(t), atol=1e-12)


@pytest.mark.parametrize("size,interval", [((64, 64), (32, 32)), ((70, 50), (24, 30)), ((40, 40), (96, 96))])
def test_tile_schedule_reproduces_a_common_field(size, interval):
    """Tiles that all return the same field f(x, y) stitch back to f (Hann windows form a
    partition of unity over every overlap, including the clamped last tiles)."""
    Px, Py = size
    dx = dy = 1.0
    dom = [-Px / 2, Px / 2, -Py / 2, Py / 2]
    f = lambda x, y: np.cos(0.3 * x) + 0.01 * x * y

    def tile(d):
        xs, xe, ys, ye = d
        px, py = int(round((xe - xs) / dx)), int(round((ye - ys) / dy))
        # column p <-> x = xs + p dx, row q <-> y = ye - q dy (top first, like the frames)
        x0 = int(round((xs - dom[0]) / dx))
        y0 = int(round((dom[3] - ye) / dy))
        X, Y = np.meshgrid(np.arange(x0, x0 + px), np.arange(y0, y0 + py))
        return f(X, Y)[None].repeat(2, 0)

    u = O.decode_tiles(tile, dom, dx, dy, interval, 2)
    X, Y = np.meshgrid(np.arange(Px), np.arange(Py))
    np.testing.assert_allclose(u[1], f(X, Y), atol=1e-12)


def test_select_neurons_matches_oracle():
    """vtdm.select_neurons (host) equals the oracle's restatement of
    vtem/utils/spike_store.py:243-277, bit for bit, on random centres plus centres placed
    exactly on the widened bounds (excluded: the comparisons are strict)."""
    from types import SimpleNamespace
    from vtem_amd import vtdm
    rng = np.random.default_rng(2)
    dom = [-20.0, 12.5, -7.25, 30.0]
    tol = 0.01
    x0 = np.concatenate([rng.uniform(-40, 40, 300), [dom[0] - tol, dom[1] + tol, 0.0, 0.0]])
    y0 = np.concatenate([rng.uniform(-40, 40, 300), [0.0, 0.0, dom[2] - tol, dom[3] + tol]])
    sel, xs, ys = vtdm.select_neurons(SimpleNamespace(h_x0=x0, h_y0=y0), dom)
    ch, xo, yo = O.select_neurons(x0, y0, dom)
    np.testing.assert_array_equal(sel, ch)
    np.testing.assert_array_equal(xs, xo)
    np.testing.assert_array_equal(ys, yo)
    assert not set(range(300, 304)) & set(sel.tolist())


def test_decode_norm_is_over_all_neurons():
    """The decoder's weights 1/(kappa sigma) are normalised by their maximum over every
    neuron before the selection (vtem/vtdm.py:408-413), so a subset's weights do not
    change with the subset."""
    from types import SimpleNamespace
    from vtem_amd import vtdm
    rng = np.random.default_rng(3)
    N = 50
    kappa = rng.uniform(0.5, 2.0, N)
    sigma = np.where(rng.random(N) < 0.8, rng.uniform(1e-3, 5e-3, N), 0.0)
    iaf = SimpleNamespace(h_kappa=kappa, h_sigma=sigma)
    sel = np.arange(5, 40, 3)
    got = vtdm.decode_norm(iaf, sel)
    np.testing.assert_array_equal(got, O.decode_norm(kappa, sigma)[sel])
    assert got.max() < 1.0 or np.argmax(O.decode_norm(kappa, sigma)) in sel
