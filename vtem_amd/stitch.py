"""Stitched encode/decode of a large video (the batching form).

The video is cut into square spatial tiles that overlap and into temporal windows that
overlap.  Every tile is encoded by its own neuron population; every (tile, window) pair
is one decode block of the batched coefficient-space decoder; the decoded blocks are
blended with raised-cosine ramps that sum to one over every overlap.  This is the
reference's tiling of decode_video and windowing of VTDM (vtem/vtdm.py:179-309,
:575-692) with the blocks decoded together instead of one after another.

* All tiles share one receptive-field bank, laid out on the tile's own grid, so one
  Encoder (and one spatial-coefficient matrix S) serves every tile; the IAF parameters
  (thresholds, initial voltages) are drawn per tile.
* A per-tile contrast multiplier c is the tile's input gain: its neurons integrate
  (c v + b) / kappa, i.e. v with kappa / c and b / c.  The encoder and the tile's decoders
  use those scaled parameters, so each block decodes the video itself.
* The encoder integrates the whole duration per tile (its input is evaluated from the
  tile's temporal coefficients over the video's period); each window's block selects its
  intervals from the tile's spike trains (BlockJob.t0).
"""

from __future__ import annotations

import dataclasses
from typing import List, Optional, Sequence

import numpy as np
import torch

from ._lib import lib, ptr, stream_ptr
from .basis import Geometry
from .batch import BatchDecoder, BlockJob
from .batch_encode import BatchEncoder, EncodeJob, IAFDevice, max_spikes_rate
from .block import Decoder, Encoder, F64, spatial_coeffs
from .synthetic import Bank, IAFParams

STITCH_DTYPE = np.dtype([("blk", np.uint64), ("wt", np.uint64), ("wy", np.uint64), ("wx", np.uint64),
                         ("T", np.int32), ("Py", np.int32), ("Px", np.int32), ("t0", np.int32),
                         ("y0", np.int32), ("x0", np.int32)])
assert STITCH_DTYPE.itemsize == 56


def tile_origins(P: int, B: int, overlap: int) -> List[int]:
    """Origins of length-B tiles covering [0, P) with the given overlap; the last tile is
    moved back to end at P (its overlap with the previous one is then larger)."""
    if B >= P:
        return [0]
    step = B - overlap
    if step <= 0:
        raise ValueError("overlap must be smaller than the tile")
    out = list(range(0, P - B + 1, step))
    if out[-1] + B < P:
        out.append(P - B)
    return out


def coverage(origins: Sequence[int], B: int) -> int:
    """The largest number of length-B tiles that cover one pixel.  With k = coverage,
    tiles i and i + k never overlap (otherwise the start of tile i + k would be covered
    k + 1 times), so k interleaved passes write disjoint tiles."""
    cnt = np.zeros(max(origins) + B, np.int64)
    for a in origins:
        cnt[a:a + B] += 1
    return int(cnt.max())


def blend_weights(origins: Sequence[int], B: int) -> List[np.ndarray]:
    """Per tile a length-B weight: 1 inside, a raised-cosine ramp over the overlap with each
    neighbour.  Where at most two tiles cover a pixel (every plan whose last tile is not
    clamped into a third tile) the weights are cos^2 and sin^2 of the same angle and sum to
    one.  Where a clamped last tile makes three tiles cover a pixel, each tile's ramps are
    divided by the sum of the ramps of all tiles covering it, so they sum to one there too."""
    n = len(origins)
    ws = []
    for i, a in enumerate(origins):
        w = np.ones(B)
        if i > 0:                                   # overlap [a, origins[i-1] + B)
            L = origins[i - 1] + B - a
            k = np.arange(L)
            w[:L] *= np.sin(0.5 * np.pi * (k + 0.5) / L) ** 2
        if i + 1 < n:                               # overlap [origins[i+1], a + B)
            L = a + B - origins[i + 1]
            k = np.arange(L)
            w[B - L:] *= np.cos(0.5 * np.pi * (k + 0.5) / L) ** 2
        ws.append(w)
    if n > 2 and coverage(origins, B) > 2:
        tot = np.zeros(max(origins) + B)
        for a, w in zip(origins, ws):
            tot[a:a + B] += w
        ws = [w / tot[a:a + B] for a, w in zip(origins, ws)]
    return ws


@dataclasses.dataclass
class StitchPlan:
    Px: int
    Py: int
    frames: int
    tile: int
    overlap: int
    window: int          # frames per window
    window_overlap: int  # frames

    def __post_init__(self):
        self.x0 = tile_origins(self.Px, self.tile, self.overlap)
        self.y0 = tile_origins(self.Py, self.tile, self.overlap)
        self.f0 = tile_origins(self.frames, self.window, self.window_overlap)
        self.wx = blend_weights(self.x0, self.tile)
        self.wy = blend_weights(self.y0, self.tile)
        self.wt = blend_weights(self.f0, self.window)
        # interleaved passes per axis: tiles i and i + k never overlap
        self.kx = coverage(self.x0, self.tile)
        self.ky = coverage(self.y0, self.tile)
        self.kt = coverage(self.f0, self.window)

    def passes(self):
        """Groups of blocks (iy, ix, it) of which no two overlap: blocks whose indices agree
        modulo each axis' coverage, one group per residue triple."""
        out = []
        for ry in range(self.ky):
            for rx in range(self.kx):
                for rt in range(self.kt):
                    out.append([(iy, ix, it) for (iy, ix, it) in self.blocks
                                if iy % self.ky == ry and ix % self.kx == rx and it % self.kt == rt])
        return [g for g in out if g]

    @property
    def tiles(self):
        return [(iy, ix) for iy in range(len(self.y0)) for ix in range(len(self.x0))]

    @property
    def blocks(self):
        return [(iy, ix, it) for (iy, ix) in self.tiles for it in range(len(self.f0))]


def stitch_accumulate(out: torch.Tensor, items) -> None:
    """out[t0+t, y0+y, x0+x] += wt[t] wy[y] wx[x] blk[t, y, x] for items
    (blk (T, Py, Px), wt, wy, wx device tensors, t0, y0, x0) that do not overlap."""
    d = np.zeros(len(items), STITCH_DTYPE)
    for i, (blk, wt, wy, wx, t0, y0, x0) in enumerate(items):
        T, Py, Px = blk.shape
        d[i] = (ptr(blk), ptr(wt), ptr(wy), ptr(wx), T, Py, Px, t0, y0, x0)
    dd = torch.from_numpy(d.view(np.uint8).copy()).pin_memory().to(out.device, non_blocking=True)
    mx = max(b[0].numel() for b in items)
    lib.vt_stitch_accumulate(ptr(out), out.stride(0), out.stride(1), ptr(dd), len(items), mx,
                             stream_ptr())
    return dd


class TiledCodec:
    """Encode a (frames, Py, Px) video tile by tile and decode it in (tile, window)
    blocks with the batched decoder, then blend the blocks."""

    def __init__(self, Px: int, Py: int, duration: float, Lt_video: int, bank: Bank,
                 tile: int = 64, overlap: int = 16, window: float = 1.0,
                 window_overlap: float = 0.25, L: int = 8, Lt: int = 8, fps: float = 100.0,
                 Wx: Optional[float] = None, Wt: Optional[float] = None, device="cuda"):
        """Wx (rad per pixel) and Wt (rad/s) are the video's bandwidths.  Each block is
        decoded in the Dirichlet space of orders (L, L, Lt) whose periods follow from the
        orders and the bandwidths, 2 pi L / Wx pixels and 2 pi Lt / Wt seconds, as the
        reference sets a decoder's periods (vtem/vtdm.py:385-388, :964-967) independently of
        the tile it decodes (:212-217); a block then sits inside one period with margin.
        Without a bandwidth the period is the tile's own extent (or window)."""
        self.device = device
        frames = int(round(duration * fps))
        self.plan = StitchPlan(Px, Py, frames, tile, overlap, int(round(window * fps)),
                               int(round(window_overlap * fps)))
        self.fps = fps
        # the encoder integrates a tile over the whole duration (the video's period)
        self.enc_geom = Geometry(tile, tile, L, L, Lt_video, duration, fps=fps)
        self.bank = bank
        self.encoder = Encoder(self.enc_geom, bank, device=device)
        # every window is a block of the decoder's periodic model, periods from M / W
        per_x = None if Wx is None else 2 * np.pi * L / float(Wx)
        per_t = None if Wt is None else 2 * np.pi * Lt / float(Wt)
        if per_x is not None and per_x < tile:
            raise ValueError("the decoder's spatial period 2 pi L / Wx is shorter than the tile")
        if per_t is not None and per_t < window:
            raise ValueError("the decoder's temporal period 2 pi Lt / Wt is shorter than the window")
        self.dec_geom = Geometry(tile, tile, L, L, Lt, window, fps=fps, period=per_t,
                                 period_x=per_x, period_y=per_x)
        self.S = spatial_coeffs(self.encoder.RF, self.dec_geom)
        dev = lambda w: torch.as_tensor(np.ascontiguousarray(w), dtype=F64, device=device)
        self.wx = [dev(w) for w in self.plan.wx]
        self.wy = [dev(w) for w in self.plan.wy]
        self.wt = [dev(w) for w in self.plan.wt]

    @staticmethod
    def tile_params(iaf: Sequence[IAFParams], contrast: Sequence[float]) -> List[IAFParams]:
        """Each tile's IAF parameters with its input gain c folded in: the neuron integrates
        (c v + b) / kappa = (v + b/c) / (kappa/c)."""
        return [dataclasses.replace(p_, kappa=p_.kappa / float(c), bias=p_.bias / float(c))
                for p_, c in zip(iaf, contrast)]

    def iaf_devices(self, iaf: Sequence[IAFParams], contrast: Sequence[float]) -> List[IAFDevice]:
        """The tiles' IAF parameters and state resident on the device (gain folded in), for
        encode(devices=...): built once, like the RF bank and the decoders."""
        return [IAFDevice.from_params(p_, self.device) for p_ in self.tile_params(iaf, contrast)]

    def encode(self, V: torch.Tensor, iaf: Sequence[IAFParams], contrast: Sequence[float],
               group: int = 32, devices: Optional[Sequence[IAFDevice]] = None):
        """Spike trains per tile (plan.tiles order).  V: (frames, Py, Px) float64 on the GPU.
        Tiles are encoded `group` at a time; a tile whose buffer overflows is encoded again
        with a larger one.  `devices` (from iaf_devices) are the resident IAF parameters;
        without them they are copied to the device here.  The state is not advanced."""
        p = self.plan
        T = p.tile
        be = BatchEncoder(self.device)
        pks = self.tile_params(iaf, contrast)
        if devices is None:
            devices = [IAFDevice.from_params(pk, self.device) for pk in pks]
        out = []
        tiles = p.tiles
        for g0 in range(0, len(tiles), group):
            jobs, margins = [], []
            for k in range(g0, min(len(tiles), g0 + group)):
                iy, ix = tiles[k]
                y0, x0 = p.y0[iy], p.x0[ix]
                # the tile's frames (a copy: the product's operand has one pixel stride)
                tv = V[:, y0:y0 + T, x0:x0 + T].reshape(V.shape[0], T * T).contiguous()
                # rows for the bias-driven rate, widened for the input's swing c |v| (|v| of
                # order 1 for unit-peak videos); an overflow is redone below
                m = 1.5 * (1.0 + 0.6 * float(contrast[k]))
                jobs.append(EncodeJob(self.encoder, tv, devices[k],
                                      max_spikes_rate(self.enc_geom, pks[k], margin=m)))
                margins.append(m)
            flags = torch.zeros(len(jobs), dtype=torch.int32, device=self.device)
            sts = be.encode(jobs, flags, carry=False)
            f = flags.cpu().numpy()
            for i in np.nonzero(f)[0]:                   # rare: redo with a bigger buffer
                m = margins[i]
                while True:
                    m *= 2.0
                    j2 = dataclasses.replace(jobs[i], max_spike=max_spikes_rate(self.enc_geom, pks[g0 + i],
                                                                                margin=m))
                    fl = torch.zeros(1, dtype=torch.int32, device=self.device)
                    st = be.encode([j2], fl, carry=False)[0]
                    if not int(fl.item()):
                        sts[i] = st
                        break
            out += sts
        return out

    def decoders(self, iaf: Sequence[IAFParams], contrast: Sequence[float]) -> List[Decoder]:
        """One decoder per tile: the shared S (and S^T), the tile's IAF parameters with the
        input gain c folded in (kappa / c, b / c), as the encoder used them."""
        base = Decoder(self.dec_geom, self.S, iaf[0].kappa, iaf[0].delta, iaf[0].bias,
                       device=self.device)
        out = []
        for p_, c in zip(iaf, contrast):
            out.append(base.with_iaf(p_.kappa, p_.delta, p_.bias, gain=float(c)))
        return out

    def decode(self, spikes, decs: Sequence[Decoder], precision: str = "kpcg",
               max_batch: int = 1024, cg_iters: int = 300, retry_iters: int = 0, prune: int = 40,
               **solver) -> torch.Tensor:
        """The blended video (frames, Py, Px) from the tiles' spike trains.  cg_iters,
        retry_iters, prune and `solver` go to the BatchDecoder (tol, preconditioner, ...).
        A window of a tile whose neurons fall silent for part of it (contrast gain above
        about 2) is far from any separable model of its operator: CG needs hundreds to
        thousands of iterations there, while the batched FP64 Cholesky that the guard falls
        back to costs about as much as 400 iterations.  So the defaults cap CG at 300
        iterations, stop a block from iteration 40 on once its residual decay predicts it
        will not finish by then (prune), retry nothing, and let the guard redo those blocks
        in FP64.
        The blocks go through the decoder max_batch at a time: one joint solve of many blocks
        shares each iteration's launches (about 10 MB of device memory per 64 x 64 x 1 s
        block).  last_stats counts the blocks the guard solved again (kpcg_retries) and
        those it redid in FP64."""
        p = self.plan
        T, W = p.tile, p.window
        jobs, where = [], []
        for k, (iy, ix) in enumerate(p.tiles):
            for it, f0 in enumerate(p.f0):
                jobs.append(BlockJob(decs[k], spikes[k], t0=f0 / self.fps, tag=(iy, ix, it)))
                where.append((iy, ix, it))
        bd = BatchDecoder(precision=precision, cg_iters=cg_iters, retry_iters=retry_iters, prune=prune,
                          **solver)
        frames = []
        fallbacks = retries = 0
        for b0 in range(0, len(jobs), max_batch):
            res = bd.decode(jobs[b0:b0 + max_batch])
            fallbacks += bd.last_stats.get("fallbacks", 0)
            retries += bd.last_stats.get("kpcg_retries", 0)
            frames += [r.frames.view(W, T, T) for r in res]
        self.last_stats = {"blocks": len(jobs), "fallbacks": fallbacks, "kpcg_retries": retries,
                           "preconditioner": bd.last_stats.get("preconditioner")}
        out = torch.zeros((p.frames, p.Py, p.Px), dtype=F64, device=self.device)
        keep = []
        # blocks whose indices agree modulo each axis' coverage never overlap: one
        # accumulation per group (8 for plans covered at most twice), in a fixed order
        fr_of = dict(zip(where, frames))
        for grp in p.passes():
            items = [(fr_of[(iy, ix, it)], self.wt[it], self.wy[iy], self.wx[ix], p.f0[it], p.y0[iy], p.x0[ix])
                     for (iy, ix, it) in grp]
            keep.append(stitch_accumulate(out, items))
        self._keep = keep
        return out
