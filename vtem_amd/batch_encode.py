"""Batched encoder: many video blocks filtered and IAF-encoded together on the GPU.

Each block runs the reference's encode chain (vtem/vtem.py:165-200): filtering by its
receptive-field bank (vtem/vrf.py:264-292), then integrate-and-fire
(vtem/ensemble_encode.py:106-168).  The filter and the temporal-coefficient projection
are one GEMM each per block; the IAF of all blocks is one time-parallel launch sequence
(vt_iaf_encode_trig_batched), so blocks of different sizes and durations share the GPU
instead of each occupying it with one serial lane per neuron.

The reference encodes one video at a time and has no batched form.
"""

from __future__ import annotations

import dataclasses
import math
from typing import Optional, List

import numpy as np
import torch

from ._lib import lib, ptr, stream_ptr
from .batch import SK_DTYPE, _to_dev, skinny_workspace
from .block import Encoder, SpikeTrains, F64, I32, dev, iaf_par_workspace_bytes
from .synthetic import IAFParams

# Descriptor of one time-parallel IAF call (layout of the C struct in vt_iaf_par.hip).
IAFP_DTYPE = np.dtype([
    ("coef", "<u8"), ("coef_ld", "<i8"), ("coef_ns", "<i8"), ("basis", "<u8"), ("input", "<u8"),
    ("input_ld", "<i8"), ("spikes", "<u8"), ("spike_ld", "<i8"), ("v0", "<u8"),
    ("time_count", "<u8"), ("kappa", "<u8"), ("bias", "<u8"), ("delta", "<u8"),
    ("spike_count", "<u8"), ("ws", "<u8"), ("overflow", "<u8"), ("dt", "<f8"), ("N", "<i4"),
    ("size", "<i4"), ("Dt", "<i4"), ("max_spike", "<i4"), ("root_mode", "<i4"),
    ("in_dtype", "<i4"), ("ybuf", "<u8"), ("y_ld", "<i8")])
assert IAFP_DTYPE.itemsize == 176


@dataclasses.dataclass
class IAFDevice:
    """An encoder population's IAF parameters and carried state, resident on the GPU."""
    kappa: torch.Tensor
    bias: torch.Tensor
    delta: torch.Tensor
    v0: torch.Tensor
    time_count: torch.Tensor

    @classmethod
    def from_params(cls, p: IAFParams, device="cuda"):
        N = p.kappa.size
        return cls(dev(p.kappa, F64, device), dev(p.bias, F64, device), dev(p.delta, F64, device),
                   dev(p.v0, F64, device), torch.zeros(N, dtype=F64, device=device))


@dataclasses.dataclass
class EncodeJob:
    """One block to encode: its Encoder (geometry, RF bank), frames (F x Py*Px, device),
    IAF state (updated in place) and the spike buffer's rows (max_spikes_bound gives a
    rate bound that needs no device data)."""
    encoder: Encoder
    frames: torch.Tensor
    iaf: IAFDevice
    max_spike: int


def rf_l1(enc: Encoder) -> torch.Tensor:
    """dx dy ||RF_j||_1 per neuron (device), cached on the encoder."""
    r = getattr(enc, "_rf_l1", None)
    if r is None:
        g = enc.geom
        r = enc.RF.abs().sum(1).to(F64) * (g.dx * g.dy)
        enc._rf_l1 = r
        enc._rf_l1_max = float(r.max().item()) if r.numel() else 0.0
    return r


def max_spikes_bound(enc: Encoder, iaf_host: IAFParams, umax: float) -> int:
    """Spike buffer rows from a rate bound that needs no device data: neuron j's input is
    at most dx dy ||RF_j||_1 umax, so it fires at most (that + b_j) / (kappa_j delta_j) times
    per second.  Safe but loose (the L1 bound is far above typical inputs)."""
    rf_l1(enc)
    rate = (enc._rf_l1_max * umax + np.max(np.abs(iaf_host.bias))) / \
        np.min(iaf_host.kappa * iaf_host.delta)
    return int(math.ceil(enc.geom.duration * rate * 1.05)) + 8


def max_spikes_rate(geom, iaf_host: IAFParams, margin: float = 2.0) -> int:
    """Spike buffer rows from the bias-driven rate b / (kappa delta) of the fastest neuron
    times `margin` (the input's own swing adds at most |v| / (kappa delta); margin 2 covers
    |v| up to b).  Not a bound: a block that exceeds it is reported through the encoder's
    overflow flag and encoded again with an exact buffer."""
    rate = np.max(np.abs(iaf_host.bias) / (iaf_host.kappa * iaf_host.delta))
    return int(math.ceil(geom.duration * rate * margin)) + 16


def encode_footprint(job: EncodeJob) -> int:
    """In-flight HBM of one encode (the RF bank and the frames are resident before
    admission): spike buffer and counts, W^T and C^T, the split-K workspaces, the IAF
    workspace, the stored input y (N x samples) and the carried state."""
    g = job.encoder.geom
    N = job.encoder.RF.shape[0]
    return int(8 * job.max_spike * N + 4 * N + 8 * (g.P + N) * g.Dt
               + 8 * (skinny_workspace(g.P, g.frames, g.Dt) + skinny_workspace(N, g.P, g.Dt))
               + iaf_par_workspace_bytes(N, g.iaf_samples) + 8 * N * g.iaf_samples + 16 * N)


def _theta_scaled(enc: Encoder) -> torch.Tensor:
    """Theta_f dx dy / frames (frames x Dt), cached on the encoder."""
    t = getattr(enc, "_theta_fs", None)
    if t is None:
        g = enc.geom
        t = (enc.theta_f.to(F64) * (g.dx * g.dy / g.frames)).contiguous()
        enc._theta_fs = t
    return t


class BatchEncoder:
    """Filter + IAF of many blocks.  The temporal coefficients of each neuron's filtered
    input, C_j = (1/F) Theta_f^T (dx dy V RF_j^T), are formed as RF (V^T Theta_f dx dy / F):
    the same product reassociated, so the F x P frames are contracted to Dt x P before the
    receptive fields are applied (Dt / F of the filtering flops).  Both products are grouped
    skinny GEMMs over all blocks; the IAF is the batched time-parallel encoder.

    root_mode 1 (default) takes each crossing time from the cancellation-free form
    2 dv / (sqrt(y1^2 + 2 dv a) + y1) of the reference's root (sqrt(y1^2 + 2 dv a) - y1) / a
    (vtem/ensemble_encode.py:501, :521); the two are algebraically identical.  The
    reference's form subtracts two nearly equal numbers when |a| is just above its 1e-12
    switch, so input values 1e-15 apart (the reassociated coefficient products here
    against frame-by-frame filtering) can move a spike by ~1e-8 s; with root mode 1 the
    times of both agree to ~1e-12 s.  root_mode 0 keeps the reference's formula."""

    def __init__(self, device="cuda", root_mode: int = 1):
        self.device = device
        self.root_mode = int(root_mode)

    def coefficients(self, jobs: List[EncodeJob], timers=None) -> List[torch.Tensor]:
        """C^T (N x Dt) per job: coefficient pt of neuron j at [j, pt].  Temporal orders with
        Dt > 32 are computed in column chunks of at most 32 (the skinny product's width).
        Both products' descriptors and workspaces are made first, then the two launches go
        back to back; with `timers`, stage "encode_coeffs" brackets the launches only
        (algorithmic HBM bytes: frames, RF banks and the two products' outputs)."""
        dvc = self.device
        Wt, Ct, rows1, rows2 = [], [], [], []
        for job in jobs:
            g = job.encoder.geom
            N = job.encoder.RF.shape[0]
            if job.encoder.RF.dtype != F64:
                raise ValueError("the batched encoder filters with an FP64 RF bank")
            Wt.append(torch.empty((g.P, g.Dt), dtype=F64, device=dvc))
            Ct.append(torch.empty((N, g.Dt), dtype=F64, device=dvc))
        for job, W, C in zip(jobs, Wt, Ct):
            enc, V = job.encoder, job.frames
            g = enc.geom
            th = _theta_scaled(enc)
            RF = enc.RF
            N = RF.shape[0]
            for c0 in range(0, g.Dt, 32):
                nc = min(32, g.Dt - c0)
                # W^T (P x Dt) = V^T Theta_f dx dy / F: A(m=pixel, k=frame) = V[k, m]
                rows1.append((ptr(V), ptr(th) + 8 * c0, ptr(W) + 8 * c0, 1, V.stride(0), g.Dt, g.Dt,
                              g.P, g.frames, nc))
                # C^T (N x Dt) = RF W^T
                rows2.append((ptr(RF), ptr(W) + 8 * c0, ptr(C) + 8 * c0, RF.stride(0), 1, g.Dt, g.Dt,
                              N, g.P, nc))
        launches, keep = [], []
        for rows in (rows1, rows2):
            if not rows:
                continue
            wsz = [skinny_workspace(r[7], r[8], r[9]) for r in rows]
            ws = torch.empty(max(1, sum(wsz)), dtype=F64, device=dvc)
            d = np.zeros(len(rows), SK_DTYPE)
            off = 0
            for i, (A, B, Cp, sam, sak, ldb, ldc, M, K, nc) in enumerate(rows):
                d[i] = (A, B, Cp, 0, ptr(ws) + 8 * off if wsz[i] else 0, 0, sam, sak, ldb, ldc, 0,
                        M, K, nc, 0)
                off += wsz[i]
            dd = _to_dev(d, dvc)
            launches.append((dd, len(rows), max(r[7] for r in rows), max(r[8] for r in rows),
                             max(r[9] for r in rows)))
            keep += [ws, dd]
        e0 = timers._tic() if timers is not None else None
        s = stream_ptr()
        for dd, nb, mM, mK, mN in launches:
            lib.vt_gemm_skinny(ptr(dd), nb, mM, mK, mN, s)
        if timers is not None:
            work = 0.0
            for j in jobs:
                g, N = j.encoder.geom, j.encoder.RF.shape[0]
                work += 8.0 * (g.frames * g.P + N * g.P + 2 * (g.P + N) * g.Dt)
            timers._toc("encode_coeffs", e0, work)
        self._ws = keep                       # alive until the next call (stream-ordered reuse)
        return Ct

    def encode(self, jobs: List[EncodeJob], overflow: torch.Tensor = None,
               carry: bool = True, timers=None, store_input: bool = True,
               mfma: Optional[bool] = None) -> List[SpikeTrains]:
        """Encode every job; returns resident spike trains (no host synchronisation).
        With `timers`, stages "encode_coeffs" and "iaf" bracket their launches (the host
        work of making descriptors and buffers stays outside the events).
        carry=True advances each job's IAF state (v0, time_count) in place, as the
        reference's encoder state (vtem/ensemble_encode.py:541-545); carry=False leaves it
        and puts the advanced state in self.last_states.  overflow (int32, len(jobs),
        zeroed by the caller) gets 1 where a job's buffer was too small.  store_input keeps
        the scaled input y (8 bytes per neuron and sample) between the two passes, so the
        crossing walk reads it instead of evaluating the trigonometric input again; with it,
        mfma forms the first pass's inputs as FP64 MFMA products of the basis rows and the
        coefficients (the same sums; tested equal to the other forms' spikes).  None
        (default) takes MFMA for term counts above 17, where it is the faster form (the
        stitched tiles' Dt = 49: 15.6 against 21.3 ms per batch of 32 tiles), and the
        per-lane form below (the serving blocks' Dt <= 17: the MFMA form was slower,
        6.7 against 5.1 ms per wave)."""
        dvc = self.device
        coefs = self.coefficients(jobs, timers)
        if carry:
            states = [(j.iaf.v0, j.iaf.time_count) for j in jobs]
        else:
            v0 = torch.cat([j.iaf.v0 for j in jobs]) if jobs else None
            tc = torch.cat([j.iaf.time_count for j in jobs]) if jobs else None
            states, o = [], 0
            for j in jobs:
                n = j.iaf.v0.numel()
                states.append((v0[o:o + n], tc[o:o + n]))
                o += n
        self.last_states = states
        out, ws_sizes = [], []
        for job, C in zip(jobs, coefs):
            g = job.encoder.geom
            N = C.shape[0]
            spikes = torch.empty((job.max_spike, N), dtype=F64, device=dvc)
            counts = torch.empty(N, dtype=I32, device=dvc)
            out.append(SpikeTrains(spikes, counts))
            ws_sizes.append((iaf_par_workspace_bytes(N, g.iaf_samples) + 255) // 256 * 256)
        ws = torch.empty(max(1, sum(ws_sizes)), dtype=torch.uint8, device=dvc)
        ybuf = None
        if store_input and jobs:
            ybuf = torch.empty(sum(C.shape[0] * j.encoder.geom.iaf_samples for j, C in zip(jobs, coefs)),
                               dtype=F64, device=dvc)
        d = np.zeros(len(jobs), IAFP_DTYPE)
        off = yoff = 0
        maxN = maxS = maxDt = 0
        for i, (job, C, st, (v0, tc)) in enumerate(zip(jobs, coefs, out, states)):
            g = job.encoder.geom
            N = C.shape[0]
            d[i] = (ptr(C), 1, C.stride(0), ptr(job.encoder.theta_iaf), 0, 0, ptr(st.spikes), N,
                    ptr(v0), ptr(tc), ptr(job.iaf.kappa),
                    ptr(job.iaf.bias), ptr(job.iaf.delta), ptr(st.counts), ptr(ws) + off,
                    0 if overflow is None else ptr(overflow) + 4 * i, g.dt, N, g.iaf_samples,
                    g.Dt, job.max_spike, self.root_mode, 0,
                    0 if ybuf is None else ptr(ybuf) + 8 * yoff, N)
            off += ws_sizes[i]
            yoff += N * g.iaf_samples
            maxN, maxS, maxDt = max(maxN, N), max(maxS, g.iaf_samples), max(maxDt, g.Dt)
        e0 = None
        if jobs:
            dd = _to_dev(d, dvc)
            e0 = timers._tic() if timers is not None else None
            use_mfma = maxDt > 17 if mfma is None else bool(mfma)
            flags = 0 if ybuf is None else (3 if use_mfma else 1)
            lib.vt_iaf_encode_trig_batched(ptr(dd), len(jobs), maxN, maxS, maxDt, flags, stream_ptr())
        if e0 is not None:
            # algorithmic work: the input's evaluation, 2 Dt flops per sample interval and neuron
            timers._toc("iaf", e0, float(sum(2.0 * j.encoder.geom.Dt * j.encoder.RF.shape[0] *
                                              (j.encoder.geom.iaf_samples - 1) for j in jobs)))
        return out
