"""Video time decoding drivers with the reference's API (vtem/vtdm.py).

VTDM_prep, VTDM and decode_video follow the reference's decoding procedure (spike-space
Dirichlet decoder: Dsw and psi of the selected neurons, intervals of each temporal
window, q, G, truncated-eigen pseudoinverse or RNN, reconstruction, Hann crossfade of the
windows and Hann-windowed accumulation of spatial tiles), with every numerical step in
the HIP library (vrf.compute_Ds / compute_Dsw / compute_dirich_space, vt_intervals,
dirichlet.*, utils.linalg.solve_eq_sym, neural_network.rnn3, vt_stitch_accumulate).

The reference passes spikes and prepared data through HDF5 files (spike file, dsw.h5,
dirich.h5).  Here `spikefile` is a spike file name (utils.spike_store) or the
EncodedVideo that vtem.VTEM_IAF returns; VTDM_prep / VTDM_prepb write Dsw (or Ds) and
dirich to the given file names (utils.simpleio) and also return them, and VTDM / VTDMb
take either the file names or the arrays.  With precompute=False, decode_video runs
VTDM_prepb / VTDMb: the prepared file holds each RF's Dirichlet coefficients Ds and the
decoder forms Dsw = Re(Ds Ds^H) with the neurons' weights itself (vtem/vtdm.py:699-1071,
vtem/dirichlet.py:87-127).

Spatial Dirichlet coefficients are computed over one period on the encoder's pixel grid
(scale = 1 by default: Ds is then exactly the DFT the encoder's filtering applied); the
reference samples the RFs at scale 4 (vtem/vtdm.py:424-425), available through `scale`.
"""

from __future__ import annotations

import math
import os
from typing import Optional, Sequence

import numpy as np
import torch

from . import dirichlet as dr
from . import neural_network as nn
from . import vrf as vrf_mod
from ._lib import lib, ptr, stream_ptr
from .utils import linalg as ls
from .utils import simpleio, videoio

F64 = torch.float64
STITCH_DTYPE = np.dtype([("blk", np.uint64), ("wt", np.uint64), ("wy", np.uint64), ("wx", np.uint64),
                         ("T", np.int32), ("Py", np.int32), ("Px", np.int32), ("t0", np.int32),
                         ("y0", np.int32), ("x0", np.int32)])
assert STITCH_DTYPE.itemsize == 56


def accumulate(out: torch.Tensor, blocks):
    """out[t0+t, y0+y, x0+x] += wt[t] wy[y] wx[x] blk[t, y, x] on the GPU for blocks given
    as (blk (T, Py, Px) tensor, wt, wy, wx, t0, y0, x0); the blocks of one call must not
    overlap (overlapping blocks go in separate calls)."""
    d = np.zeros(len(blocks), STITCH_DTYPE)
    keep = []
    for i, (blk, wt, wy, wx, t0, y0, x0) in enumerate(blocks):
        T, Py, Px = blk.shape
        ws = [torch.as_tensor(np.ascontiguousarray(w, np.float64), device=out.device) for w in (wt, wy, wx)]
        keep += ws
        d[i] = (blk.data_ptr(), ws[0].data_ptr(), ws[1].data_ptr(), ws[2].data_ptr(), T, Py, Px, t0, y0, x0)
    dd = torch.from_numpy(d.view(np.uint8).copy()).to(out.device)
    mx = max(b[0].numel() for b in blocks)
    lib.vt_stitch_accumulate(ptr(out), out.stride(0), out.stride(1), ptr(dd), len(blocks), mx, stream_ptr())
    torch.cuda.current_stream().synchronize()      # the window tables are freed on return


def _hann_ramp(n: int) -> np.ndarray:
    """First half of a 2n-point Hann window (np.hanning(2n)[:n], vtem/vtdm.py:252-268)."""
    if n <= 0:
        return np.zeros(0)
    k = np.arange(n)
    return 0.5 - 0.5 * np.cos(2 * np.pi * k / (2 * n - 1))


def select_neurons(rf, domain, tol: float = 0.01):
    """Neurons whose RF centre lies strictly inside the domain [xs, xe, ys, ye] widened by
    tol on every side, with centres shifted to the domain's centre
    (vtem/utils/spike_store.py:243-277; strict comparisons as at :266, shift as at
    :260-271).  Returns (indices, x0 shifted, y0 shifted)."""
    xs, xe, ys, ye = domain
    x0, y0 = np.asarray(rf.h_x0), np.asarray(rf.h_y0)
    sel = np.nonzero((x0 > xs - tol) & (x0 < xe + tol) & (y0 > ys - tol) & (y0 < ye + tol))[0]
    cx, cy = xs + (xe - xs) / 2, ys + (ye - ys) / 2
    return sel, x0[sel] - cx, y0[sel] - cy


def decode_norm(iaf, neurons) -> np.ndarray:
    """Per selected neuron the decoder's weight: 1/(kappa sigma) for a random-threshold
    neuron (sigma != 0), else 1, divided by its maximum over ALL neurons, then taken at
    `neurons` (vtem/vtdm.py:408-413).  VTDM uses the same array for q and the
    reconstruction (vtem/vtdm.py:634-638 means the same rule; there the reference indexes
    the selected kappa with full-population indices, which only agrees when every neuron
    is selected or kappa is uniform; the full kappa is used here)."""
    kappa = np.asarray(iaf.h_kappa, np.float64)
    sigma = np.asarray(getattr(iaf, "h_sigma", np.zeros_like(kappa)), np.float64)
    h = np.ones(kappa.size)
    nz = sigma != 0
    h[nz] = 1.0 / (kappa[nz] * sigma[nz])
    h = h / h.max()
    return h[np.asarray(neurons, np.int64)]


def _sub_rf(rf, sel, x0, y0, shape, dx, dy, scale):
    """A receptive-field object of rf's type for the selected neurons on a new grid."""
    if isinstance(rf, vrf_mod.vrf_gabor):
        sub = vrf_mod.vrf_gabor(shape, dx=dx, dy=dy, scale=scale, device=rf.device)
        sub.load_parameters(len(sel), np.asarray(rf.h_alpha)[sel], np.asarray(rf.h_l)[sel], x0, y0,
                            np.asarray(rf.h_ab)[sel], KAPPA=rf.KAPPA)
    elif isinstance(rf, vrf_mod.vrf_cs):
        sub = vrf_mod.vrf_cs(shape, dx=dx, dy=dy, scale=scale, device=rf.device)
        sub.load_parameters(len(sel), np.asarray(rf.h_alpha)[sel], x0, y0,
                            sigma_center=rf.sigma_center, sigma_surround=rf.sigma_surround)
    else:
        raise TypeError("Receptive field type not recognized")
    return sub


def _full_domain(enc, domain):
    Sx, Sy = enc.Px * enc.dx, enc.Py * enc.dy
    if domain is None:
        return [-Sx / 2, Sx / 2, -Sy / 2, Sy / 2]
    return [max(float(domain[0]), -Sx / 2), min(float(domain[1]), Sx / 2),
            max(float(domain[2]), -Sy / 2), min(float(domain[3]), Sy / 2)]


def _encoded(spikefile):
    """The EncodedVideo behind a spike-file name or an EncodedVideo."""
    if isinstance(spikefile, (str, bytes, os.PathLike)):
        from .vtem import load_encoded
        return load_encoded(spikefile)
    return spikefile


def _is_name(x) -> bool:
    return isinstance(x, (str, bytes, os.PathLike))


def _period_rf(enc, Mx, My, Wx, Wy, domain, scale):
    """(rf over one spatial period for the selected neurons, selection, domain)."""
    Mx = int(Mx)
    My = Mx if My is None else int(My)
    Wx = enc.Wx if Wx is None else float(Wx)
    Wy = enc.Wy if Wy is None else float(Wy)
    dom = _full_domain(enc, domain)
    sel, x0, y0 = select_neurons(enc.rf, dom)
    if sel.size == 0:
        raise ValueError("no neuron has its receptive field centre in the domain")
    # one spatial period 2 pi M / W on the encoder's grid
    rfPx = int(round(2 * math.pi * Mx / Wx / enc.dx))
    rfPy = int(round(2 * math.pi * My / Wy / enc.dy))
    return _sub_rf(enc.rf, sel, x0, y0, (rfPy, rfPx), enc.dx, enc.dy, scale), sel, dom, Mx, My, Wx, Wy


def VTDM_prep(spikefile, Dswfilename=None, dirichfilename=None, Mx=None, My=None, domain=None, Wx=None,
              Wy=None, dtype=np.complex128, scale: int = 1, mask: str = "ellipse"):
    """Dsw (N_sel x N_sel) and dirich (N_sel x decPy x decPx) of the neurons selected by
    `domain` (vtem/vtdm.py:311-447), written to Dswfilename / dirichfilename when given.
    Returns (Dsw, dirich, selected neuron indices)."""
    enc = _encoded(spikefile)
    period_rf, sel, dom, Mx, My, Wx, Wy = _period_rf(enc, Mx, My, Wx, Wy, domain, scale)
    Ds = period_rf.compute_Ds(Mx, My, mask=mask)
    Dsw = period_rf.compute_Dsw(Ds, Mx, My, decode_norm(enc.iaf, sel))
    decSx, decSy = dom[1] - dom[0], dom[3] - dom[2]
    decPx, decPy = int(round(decSx / enc.dx)), int(round(decSy / enc.dy))
    dirich = period_rf.compute_dirich_space(Ds, Mx, My, decPx, decPy, decSx, decSy, Wx, Wy)
    if _is_name(Dswfilename):
        simpleio.write_memory_to_file(Dsw, Dswfilename)
    if _is_name(dirichfilename):
        simpleio.write_memory_to_file(dirich, dirichfilename)
    return Dsw, dirich, sel


def VTDM_prepb(spikefile, Dsfilename=None, dirichfilename=None, Mx=None, My=None, domain=None, Wx=None,
               Wy=None, dtype=np.complex128, scale: int = 1, mask: str = "ellipse"):
    """The Dirichlet coefficients Ds (N_sel x (2My+1) x (2Mx+1), complex) and dirich of the
    selected neurons (vtem/vtdm.py:699-813), written to Dsfilename / dirichfilename when
    given; VTDMb forms the weighted Gram of the RFs from Ds.  Returns (Ds, dirich, selected
    neuron indices)."""
    enc = _encoded(spikefile)
    period_rf, sel, dom, Mx, My, Wx, Wy = _period_rf(enc, Mx, My, Wx, Wy, domain, scale)
    Ds = period_rf.compute_Ds(Mx, My, mask=mask)
    decSx, decSy = dom[1] - dom[0], dom[3] - dom[2]
    decPx, decPy = int(round(decSx / enc.dx)), int(round(decSy / enc.dy))
    dirich = period_rf.compute_dirich_space(Ds, Mx, My, decPx, decPy, decSx, decSy, Wx, Wy)
    if _is_name(Dsfilename):
        simpleio.write_memory_to_file(Ds, Dsfilename)
    if _is_name(dirichfilename):
        simpleio.write_memory_to_file(dirich, dirichfilename)
    return Ds, dirich, sel


class _SpikeSource:
    """Per-neuron spike arrays resident on the GPU, with the interval selection of
    read_select_spikes (vtem/utils/spike_store.py:279-351) done by vt_intervals."""

    def __init__(self, spikes: Sequence[np.ndarray], device="cuda"):
        self.N = len(spikes)
        rows = max(1, max((s.size for s in spikes), default=0))
        buf = np.zeros((rows, self.N))
        cnt = np.zeros(self.N, np.int32)
        for j, s in enumerate(spikes):
            buf[:s.size, j] = s
            cnt[j] = s.size
        self.spikes = torch.as_tensor(buf, device=device)
        self.counts = torch.as_tensor(cnt, device=device)
        self.rows = rows
        self.device = device

    def select(self, neurons: np.ndarray, t0: float, t1: float):
        """Intervals of the chosen neurons in window (t0, t1), times relative to t0:
        (tk1, tk2, local neuron index) as host arrays plus the count."""
        idx = torch.as_tensor(np.asarray(neurons, np.int64), device=self.device)
        sp = self.spikes.index_select(1, idx).contiguous()
        ct = self.counts.index_select(0, idx).contiguous()
        N = idx.numel()
        cap = self.rows * N
        ws = torch.empty(2 * N, dtype=torch.int32, device=self.device)
        off = torch.empty(N + 1, dtype=torch.int32, device=self.device)
        tk1 = torch.empty(max(cap, 1), dtype=F64, device=self.device)
        tk2 = torch.empty(max(cap, 1), dtype=F64, device=self.device)
        nid = torch.empty(max(cap, 1), dtype=torch.int32, device=self.device)
        lib.vt_intervals(ptr(sp), sp.stride(0), self.rows, ptr(ct), N, float(t0), float(t1 - t0), ptr(ws),
                         ptr(off), ptr(tk1), ptr(tk2), ptr(nid), cap, stream_ptr())
        n = int(off[N].item())
        return tk1[:n], tk2[:n], nid[:n], n


def _write_output(u, output_format, name):
    """output_format 0: an AVI file, 1: an array file, else nothing."""
    if output_format == 0:
        videoio.write_video(u, name)
    elif output_format == 1:
        simpleio.write_memory_to_file(u, name)


def VTDM(spikefile, Dswfilename, dirichfilename, start_time, end_time, dt, Mx, My=None, Mt=None,
         domain=None, Wx=None, Wy=None, Wt=None, lamb=0.0, dtype=np.float64, rnn=False, alpha=50000,
         steps=4000, stitching=False, stitch_interval=None, output_format=0, output="rec", *,
         neurons=None, source: Optional[_SpikeSource] = None, keep_device: bool = False, _gram_from_ds=False):
    """Spike-space decode of [start_time, end_time) at frame step dt, in temporal windows
    of stitch_interval overlapping by 20 % (at least 4 frames when the window has 8 or
    more) and crossfaded with Hann ramps (vtem/vtdm.py:450-692).  Dswfilename /
    dirichfilename are VTDM_prep's file names or arrays; `neurons` the indices it selected
    (selected again from `domain` when not given).  output_format 0 writes output + ".avi",
    1 the array file output + ".h5", anything else nothing.  Returns the video (T, decPy,
    decPx) as a float64 ndarray (a device tensor with keep_device)."""
    enc = _encoded(spikefile)
    Dsw = simpleio.read_file(Dswfilename) if _is_name(Dswfilename) else Dswfilename
    dirich = simpleio.read_file(dirichfilename) if _is_name(dirichfilename) else dirichfilename
    if neurons is None:
        neurons = select_neurons(enc.rf, _full_domain(enc, domain))[0]
    Wt = enc.Wt if Wt is None else float(Wt)
    Mt = int(round(Wt / math.pi)) if Mt is None else int(Mt)      # vtem/vtdm.py:569
    src = source or _SpikeSource(enc.spikes)
    kappa = np.asarray(enc.iaf.h_kappa)[neurons]
    delta = np.asarray(enc.iaf.h_delta)[neurons]
    bias = np.asarray(enc.iaf.h_bias)[neurons]
    norm = decode_norm(enc.iaf, neurons)
    end_time = float(end_time)
    if stitching:
        stitch_interval = 20 * dt if stitch_interval is None else float(stitch_interval)
        stitch_interval = min(stitch_interval, end_time)
        end = round(min(start_time + stitch_interval, end_time) / dt) * dt
        overlap = 0.2 * stitch_interval
    else:
        stitch_interval = end_time
        end = end_time
        overlap = 0.0
    total_l = int(round((end_time - start_time) / dt))
    overlap_l = int(round(overlap / dt))
    interval_l = int(round(stitch_interval / dt))
    if overlap_l < 3:
        if interval_l >= 8:
            overlap_l, overlap = 4, 4 * dt
        else:
            overlap_l, overlap = 0, 0.0
    lw = _hann_ramp(overlap_l)
    psi = dirich
    Py, Px = psi.shape[1], psi.shape[2]
    u = torch.zeros((total_l, Py, Px), dtype=F64, device=src.device)
    start = float(start_time)
    i = 0
    while True:
        tk1, tk2, nid, n = src.select(neurons, start, end)
        cp = dr.dirichlet(len(neurons), tk1, tk2, nid, kappa, delta, bias, norm, Wt, Mt)
        cp.compute_q()
        if _gram_from_ds:
            cp.compute_Gb(Dsw, float(lamb))
        else:
            cp.compute_G(Dsw, float(lamb))
        if rnn:
            cp.d_q = nn.rnn3(cp.d_G, cp.d_q, alpha=alpha, steps=steps)
        else:
            cp.d_q = ls.solve_eq_sym(cp.d_G, cp.d_q)
        cp.freeG()
        u_rec = cp.reconstruct(psi, [0, round((end - start) / dt) * dt], dt)
        last = end_time - end < dt
        L = u_rec.shape[0]
        if overlap_l:
            off = i * (interval_l - overlap_l)
            # Hann crossfade: this window ramps in over the previous window's last frames
            # and (unless last) out over the next window's first frames
            wt = np.ones(L)
            if i > 0:
                wt[:overlap_l] = lw
            if not last:
                wt[interval_l - overlap_l:] = 1.0 - lw
                wt[interval_l:] = 0.0
        else:
            off = i * interval_l
            wt = np.ones(L)
        L = min(L, total_l - off)
        accumulate(u, [(u_rec[:L].contiguous(), wt[:L], np.ones(Py), np.ones(Px), off, 0, 0)])
        if last:
            break
        start = start + round(stitch_interval / dt) * dt - round(overlap / dt) * dt
        end = min(start + round(stitch_interval / dt) * dt, end_time)
        if end_time - end < overlap:
            end = end_time
        i += 1
    if output_format in (0, 1):
        _write_output(u, output_format, output + (".avi" if output_format == 0 else ".h5"))
    return u if keep_device else u.cpu().numpy()


def VTDMb(spikefile, Dsfilename, dirichfilename, start_time, end_time, dt, Mx, My=None, Mt=None,
          domain=None, Wx=None, Wy=None, Wt=None, lamb=0.0, dtype=np.float64, rnn=False, alpha=50000,
          steps=4000, stitching=False, stitch_interval=None, output_format=0, output="rec", *,
          neurons=None, source: Optional[_SpikeSource] = None, keep_device: bool = False):
    """VTDM from VTDM_prepb's Dirichlet coefficients (vtem/vtdm.py:816-1071): each
    window's G comes from Dsw = Re(Ds Ds^H) * norm norm^T formed on the device
    (dirichlet.compute_Gb), the rest as VTDM.  output_format 0 / 1 write the file named
    `output` as given (no extension added, as the reference's VTDMb)."""
    u = VTDM(spikefile, Dsfilename, dirichfilename, start_time, end_time, dt, Mx, My, Mt, domain, Wx, Wy, Wt,
             lamb, dtype, rnn, alpha, steps, stitching, stitch_interval, 2, output, neurons=neurons,
             source=source, keep_device=True, _gram_from_ds=True)
    _write_output(u, output_format, output)
    return u if keep_device else u.cpu().numpy()


def decode_video(spikefile, Dswfilename=None, dirichfilename=None, start_time=0.0, end_time=None,
                 dt=0.01, Mx=None, My=None, Mt=None, domain=None, Wx=None, Wy=None, Wt=None, lamb=0.0,
                 dtype=np.float64, rnn=False, alpha=50000, steps=4000, stitching=False,
                 stitch_interval=None, spatial_stitching=False, spatial_interval=(96, 96),
                 precompute=True, write_blocks=False, output_format=0, output="rec", scale: int = 1,
                 mask: str = "ellipse"):
    """Decode a VTEM_IAF result, a spike file name or an EncodedVideo (vtem/vtdm.py:18-309):
    one VTDM over the domain, or, with spatial_stitching, tiles of spatial_interval pixels
    overlapping by 20 % placed from the right / top edges, each decoded with its own
    VTDM_prep and accumulated with Hann-ramp windows.  precompute=False prepares Ds
    instead of Dsw (VTDM_prepb / VTDMb).  Dswfilename / dirichfilename, when given, receive
    the prepared arrays (of the last tile).  write_blocks writes each tile's video as
    output + its number (".h5" array files with output_format 1, else ".avi"); output_format
    0 / 1 writes the whole video as output + ".avi" / ".h5".  Returns the video (T, Py, Px)
    float64."""
    enc = _encoded(spikefile)
    end_time = enc.duration if end_time is None else end_time
    src = _SpikeSource(enc.spikes)
    kw = dict(Mt=Mt, Wx=Wx, Wy=Wy, Wt=Wt, lamb=lamb, dtype=dtype, rnn=rnn, alpha=alpha, steps=steps,
              stitching=stitching, stitch_interval=stitch_interval, source=src)
    prep, dec = (VTDM_prep, VTDM) if precompute else (VTDM_prepb, VTDMb)
    dom = _full_domain(enc, domain)
    if not spatial_stitching:
        Dsw, dirich, sel = prep(enc, Dswfilename, dirichfilename, Mx, My, dom, Wx, Wy, scale=scale, mask=mask)
        u = dec(enc, Dsw, dirich, start_time, end_time, dt, Mx, My, domain=dom, neurons=sel, output_format=2,
                keep_device=True, **kw)
        if output_format in (0, 1):
            _write_output(u, output_format, output + (".avi" if output_format == 0 else ".h5"))
        return u.cpu().numpy()
    dx, dy = enc.dx, enc.dy
    decPx, decPy = int(round((dom[1] - dom[0]) / dx)), int(round((dom[3] - dom[2]) / dy))
    sx, sy = min(spatial_interval[0], decPx), min(spatial_interval[1], decPy)
    ox, oy = int(round(0.2 * sx)), int(round(0.2 * sy))
    total_l = int(round((end_time - start_time) / dt))
    u = torch.zeros((total_l, decPy, decPx), dtype=F64, device=src.device)
    start_y = max(dom[3] - sy * dy, dom[2])
    end_y = dom[3]
    y0i, y1i = 0, sy
    tile = 0
    while True:
        start_x = max(dom[1] - sx * dx, dom[0])
        end_x = dom[1]
        x0i, x1i = decPx - sx, decPx
        while True:
            cur = [start_x, end_x, start_y, end_y]
            Dsw, dirich, sel = prep(enc, Dswfilename, dirichfilename, Mx, My, cur, Wx, Wy, scale=scale,
                                    mask=mask)
            left = 0 if x0i == 0 else ox
            right = 0 if x1i == decPx else ox
            top = 0 if y0i == 0 else oy
            bottom = 0 if y1i == decPy else oy
            xw = np.concatenate([_hann_ramp(left), np.ones(x1i - x0i - left - right), 1 - _hann_ramp(right)])
            yw = np.concatenate([_hann_ramp(top), np.ones(y1i - y0i - top - bottom), 1 - _hann_ramp(bottom)])
            blk = dec(enc, Dsw, dirich, start_time, end_time, dt, Mx, My, domain=cur, neurons=sel,
                      output_format=2, keep_device=True, **kw)
            if write_blocks:
                fmt = 1 if output_format == 1 else 0
                _write_output(blk, fmt, f"{output}{tile}" + (".h5" if fmt == 1 else ".avi"))
            tile += 1
            accumulate(u, [(blk, np.ones(blk.shape[0]), yw, xw, 0, y0i, x0i)])
            if start_x == dom[0]:
                break
            start_x = max(start_x - sx * dx + ox * dx, dom[0])
            end_x = end_x - sx * dx + ox * dx
            x0i = max(x0i - sx + ox, 0)
            x1i = x1i - sx + ox
        if start_y == dom[2]:
            break
        start_y = max(start_y - sy * dy + oy * dy, dom[2])
        end_y = end_y - sy * dy + oy * dy
        y0i = y1i - oy
        y1i = min(y0i + sy, decPy)
    if output_format in (0, 1):
        _write_output(u, output_format, output + (".avi" if output_format == 0 else ".h5"))
    return u.cpu().numpy()
