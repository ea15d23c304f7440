"""Decoded frames/s of the served, batched, mixed-precision vtem decoder on one MI355X.

Workload (serving form): requests are video blocks drawn uniformly from
{32, 48, 64, 96}^2 px x {0.5, 1, 2} s, with Lx = Ly = round(side/8), Lt = round(4*dur)
(D from 405 to 10625), N = round(2D/9) neurons from the 64x64-px bank law (Gabor:DoG =
7:1), IAF b=2.5, kappa=1, delta ~ N(0.04, 0.004^2), dt = 1e-4, Tikhonov
lam = 1e-6 tr(G)/n.  A request is encode+decode: its video frames, receptive-field bank,
IAF parameters and decoder data (spatial coefficients) are resident in HBM when timing
starts; serving it filters and IAF-encodes the frames into spike trains and decodes them
into 100 fps frames over its duration.

One step = one wave of R requests with seeded Poisson arrivals at `--overload` (default 2)
times the server's capacity, the capacity being measured on a closed batch of the same R
requests just before; served by
  chosen:     the served path (continuous batching on 2 streams, HBM-budget admission):
              batched encode (reassociated filter products + time-parallel IAF), then the
              batched Kronecker-preconditioned CG solver: FP64 conjugate gradients
              through each block's exact Kronecker operator, preconditioner
              (S^T C S + lam I) kron Tbar, no Gram; FP64 guard on the true residual;
  gram_bf16x3: the same server with batched split-BF16 Gram + FP32 Cholesky + FP64
              refinement (FP64 for request classes whose FP32 factor fails);
  plain:      the per-block FP64 path, one request at a time: the reference's encoder
              (every frame filtered, serial one-lane-per-neuron IAF) and FP64 Gram + FP64
              Cholesky.
`value` = chosen decoded frames/s (frames of the wave / time from first arrival to last
result), encode + decode.  The closed-batch figure, a decode-only chosen-vs-plain figure
(spikes resident to frames resident) and a latency sweep at 0.25/0.5/0.75/1.0 of capacity
(p50/p99, arrival -> result in HBM) are reported beside it.  SNR of every request is
checked against its plain FP64 decode (|dSNR| < 0.1 dB).

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--requests R] [--no-cpu]
"""

from __future__ import annotations

import argparse
import dataclasses
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

# stage -> (peak, unit, bound); the stage's recorded work is flops for "mfma" stages and
# HBM bytes for "hbm" stages
RNN_SHARE = 0.2          # decoder mix: 80 % pseudoinverse (Tikhonov), 20 % RNN
RNN_STEPS = 500
LAT_REQUESTS = 1024      # requests per latency-sweep rate point

PEAKS = {"gram_lp": (2500.0, "TFLOP/s", "mfma"), "gram_f64": (78.6, "TFLOP/s", "mfma"),
         "potrf_f32": (157.3, "TFLOP/s", "mfma"), "potrf_f64": (78.6, "TFLOP/s", "mfma"),
         "solve_refine": (157.3, "TFLOP/s", "mfma"), "kpcg_setup": (78.6, "TFLOP/s", "mfma"),
         "kpcg_mgram": (78.6, "TFLOP/s", "mfma"), "kpcg_potrf": (78.6, "TFLOP/s", "mfma"),
         "kpcg_inverse": (78.6, "TFLOP/s", "mfma"), "prologue": (8000.0, "GB/s", "hbm"),
         "synthesis": (78.6, "TFLOP/s", "mfma"), "rnn_iterate": (8000.0, "GB/s", "hbm"),
         "rnn_lanczos": (8000.0, "GB/s", "hbm"),
         "kpcg_iterate": (8000.0, "GB/s", "hbm"), "encode_coeffs": (8000.0, "GB/s", "hbm"),
         "iaf": (78.6, "TFLOP/s", "fp64")}
WORK_NOTES = {"encode_coeffs": "algorithmic HBM bytes 8*(F*P + N*P + 2*(P+N)*Dt) per block (frames, "
                               "RF bank, the two reassociated coefficient products)",
              "iaf": "algorithmic FP64 flops 2*Dt*N*(T_f-1) per block (the trigonometric input's "
                     "evaluation; the walk's second pass and the step arithmetic are extra)",
              "kpcg_iterate": "algorithmic HBM bytes per CG iteration: 8*(2*N*Ds + N*Dt^2 + 2*N*Dt "
                              "+ Ds^2 + 12*D) per block (S twice, T_j, Y/Z, Mi, vectors), times "
                              "the block's own iteration count",
              "synthesis": "algorithmic FP64 flops 2*T_out*Ds*(Dt + P) per block (Theta X^T, then "
                           "times B_s^T)",
              "gram_lp": "algorithmic flops N*D*(D+1) per block",
              "potrf_f32": "algorithmic flops D^3/3 per block"}


def parse():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=10)
    p.add_argument("--warmup", type=int, default=1)
    p.add_argument("--requests", type=int, default=64)
    p.add_argument("--seed", type=int, default=0)
    p.add_argument("--max-batch", type=int, default=32)
    p.add_argument("--streams", type=int, default=2)
    p.add_argument("--rnn-lanes", type=int, default=1, help="streams of the RNN lane")
    p.add_argument("--rnn-lane-delays-ms", default="0,16",
                   help="staggered RNN lanes: comma-separated hold per lane (ms; default: a second "
                        "lane that gathers the RNN requests arriving during the first lane's chain "
                        "for 16 ms); '' = --rnn-lanes lanes splitting the waiting requests, with "
                        "the preparation stage")
    p.add_argument("--graph-chunk", type=int, default=25,
                   help="RNN / power iterations per replayed HIP graph (0 = launch each)")
    p.add_argument("--form", default="kpcg", help="solver form of the chosen served path")
    p.add_argument("--no-rnn-prep", action="store_true",
                   help="no preparation stream: RNN requests behind a running chain wait whole")
    p.add_argument("--rnn-persistent", action="store_true",
                   help="RNN iterations in one persistent launch instead of two launches each")
    p.add_argument("--compare", default="bf16x3", help="second served form measured ('' = none)")
    p.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    p.add_argument("--no-sweep", action="store_true", help="skip the latency sweep")
    p.add_argument("--cpu-budget", type=float, default=15.0, help="seconds of CPU oracle work")
    p.add_argument("--max-delay-ms", type=float, default=0.0,
                   help="dynamic batching: a pseudoinverse lane holds a partial batch this long")
    p.add_argument("--rnn-max-delay-ms", type=float, default=0.0,
                   help="dynamic batching: the RNN lane holds a partial batch this long")
    p.add_argument("--overload", type=float, default=2.0,
                   help="Poisson arrival rate of the timed steps, in multiples of the closed-batch capacity")
    p.add_argument("--no-precision", action="store_true", help="skip the 128x128 precision report")
    p.add_argument("--precision-only", action="store_true", help="only the 128x128 precision report")
    p.add_argument("--no-c3", action="store_true", help="skip the 512x512 stitched batching form")
    p.add_argument("--c3-only", action="store_true", help="only the 512x512 stitched batching form")
    p.add_argument("--roofline-only", action="store_true",
                   help="only time the dominant kernel at the workload's shapes (for rocprofv3 runs)")
    p.add_argument("--cholesky-only", action="store_true",
                   help="only time the batched FP64 Cholesky of 16 SPD matrices of D = 4913")
    return p.parse_args()


class Item:
    """One request's resident data: video frames, RF bank (in the Encoder), IAF state,
    prepared decoder; `job` holds spike trains encoded once before timing (for the CPU
    baseline and the dominant-kernel measurement), `encode` the timed encode."""

    def __init__(self, blk, geom, job, encode):
        self.blk, self.geom, self.job, self.encode = blk, geom, job, encode

    def __iter__(self):
        return iter((self.blk, self.geom, self.job))


def prepare(R, seed):
    """Seeded requests: frames, RF banks, IAF parameters and decoders resident on the GPU.
    Each request is also encoded once here (untimed; the timed runs encode again)."""
    import torch
    from vtem_amd import synthetic as syn
    from vtem_amd.basis import Geometry
    from vtem_amd.batch import BlockJob
    from vtem_amd.batch_encode import BatchEncoder, EncodeJob, IAFDevice, max_spikes_rate
    from vtem_amd.block import Decoder, Encoder, dev, spatial_coeffs
    rng = np.random.default_rng(seed)
    items = []
    for i in range(R):
        side = int(rng.choice(syn.SERVE_SIDES))
        dur = float(rng.choice(syn.SERVE_DURATIONS))
        method = "rnn" if rng.random() < RNN_SHARE else "tikhonov"
        blk = syn.serve_block(side, dur, seed=seed * 100003 + i)
        geom = Geometry(side, side, blk.Lx, blk.Lx, blk.Lt, dur)
        enc = Encoder(geom, blk.bank)
        ej = EncodeJob(enc, dev(blk.video.frames.reshape(geom.frames, -1)),
                       IAFDevice.from_params(blk.iaf), max_spikes_rate(geom, blk.iaf))
        dec = Decoder(geom, spatial_coeffs(enc.RF, geom), blk.iaf.kappa, blk.iaf.delta,
                      blk.iaf.bias)
        job = BlockJob(dec, None, tag=i, method=method, rnn_steps=RNN_STEPS)
        items.append(Item(blk, geom, job, ej))
    be = BatchEncoder()
    flags = torch.zeros(R, dtype=torch.int32, device="cuda")
    sts = be.encode([it.encode for it in items], flags, carry=False)
    torch.cuda.synchronize()
    if int(flags.sum().item()):
        raise RuntimeError("spike buffer overflow in the untimed encode")
    for it, st in zip(items, sts):
        it.job = dataclasses.replace(it.job, spikes=st)
    return items


def make_requests(items, arrivals, form="kpcg"):
    """Encode+decode requests (the decode job's spikes come from the timed encode)."""
    from vtem_amd.serve import Request, footprint_bytes
    return [Request(i, float(a), dataclasses.replace(it.job, spikes=None), it.geom.frames,
                    footprint_bytes(it.job, form, encode=it.encode), encode=it.encode,
                    iaf=it.blk.iaf)
            for i, (it, a) in enumerate(zip(items, arrivals))]


def snr_db(u, urec):
    e = u - urec
    return 10 * math.log10(float(np.sum(u * u)) / float(np.sum(e * e)))


def dominant_kernel(items, max_batch, reps=200):
    """The served path's dominant kernel at its real shapes: the forward skinny product of
    the structured operator (Z_j = T_j (S X)_j, skinny_kernel + its split-K reduction)
    over the RNN requests of one batch, the product each of the RNN's 501 iterations runs
    once, with its operands in the MFMA tile layout as the iterations run it.  Timed with HIP events
    on the stream the kernels are launched on.  Algorithmic bytes per launch:
    8 (N Ds + Ds Dt + N Dt + N Dt^2) per block (S and X read, Z written, T_j read)."""
    import torch
    from vtem_amd._lib import stream_ptr
    from vtem_amd.batch import BatchDecoder
    jobs = [job for _, _, job in items if job.method == "rnn"][:max_batch]
    jobs = jobs or [job for _, _, job in items][:max_batch]
    bd = BatchDecoder(precision="kpcg")
    by_dt = {}
    for jb in jobs:
        by_dt.setdefault(jb.decoder.geom.Dt, []).append(jb)
    gs = [bd._prologue(js, Dt) for Dt, js in sorted(by_dt.items())]
    blocks, _, N, Ds, Dt = bd._joint_blocks(gs)
    D = Ds * Dt
    dbj = np.concatenate([[0], np.cumsum(D)])
    nzj = np.concatenate([[0], np.cumsum(N * Dt)])
    X = torch.randn(int(dbj[-1]), dtype=torch.float64, device="cuda")
    Q = torch.empty_like(X)
    tiles = bd.operand_tiles(N, Ds, Dt)
    tiles[0].normal_()
    fwd, _ = bd._operator(blocks, X, Q, tiles[2], dbj, nzj, tiled=tiles)
    s = stream_ptr()
    for _ in range(5):
        fwd.run(s)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(reps):
        fwd.run(s)
    e1.record()
    torch.cuda.synchronize()
    us = e0.elapsed_time(e1) * 1000.0 / reps
    nbytes = float(8 * np.sum(N * Ds + Ds * Dt + N * Dt + N * Dt * Dt))
    return {"kernel": "skinny_kernel (+ skinny_reduce_kernel), forward operator product",
            "blocks": len(blocks), "avg_us": us, "bytes": nbytes, "achieved_GBs": nbytes / us / 1e3}


def precision_report(seed=0):
    """Precision form (128x128 px, 2 s, Lx=Ly=12, Lt=8, D = 10625, 8192 neurons of the
    config-1 laws): Gram relative Frobenius error of each low-precision Gram against the
    FP64 Gram, SNR of each solve form against the video and against the plain FP64 path,
    and the Gram's MFMA throughput (HIP events around the device work: the FP64 Gram kernel;
    the T split and the Gram kernel for the low-precision forms).  gram_tflops counts the Kronecker form's
    algorithmic flops N D (D+1) (lower triangle).  The kernels compute only the temporal
    pairs pt <= pt' (T_j is symmetric) and mirror them, so they run about half of those;
    mfma_executed_* counts the MFMA work actually issued (tile padding included) against
    the raw MFMA peak (split-BF16 issues 3 products)."""
    import torch
    from vtem_amd import synthetic as syn
    from vtem_amd.basis import Geometry
    from vtem_amd.autotune import gram_tuner, tune_key
    from vtem_amd.batch import BatchDecoder, BlockJob
    from vtem_amd._lib import lib, ptr, stream_ptr
    from vtem_amd.block import (Decoder, Encoder, dev, gram_descs, gram_lp_ws_bytes, spatial_coeffs,
                                split_T_size)
    geom = Geometry(128, 128, 12, 12, 8, 2.0)
    vid = syn.make_video(128, 128, 2.0, 100.0, 12, 12, 8, seed=seed)
    bank = syn.bank_dyadic(128.0, 128.0)
    iaf = syn.iaf_params(bank.num_neurons, 2.5, 1.0, 0.04, 0.004, np.random.default_rng(seed + 1))
    enc = Encoder(geom, bank)
    st = enc.encode(dev(vid.frames.reshape(geom.frames, -1)), iaf)
    dec = Decoder(geom, spatial_coeffs(enc.RF, geom), iaf.kappa, iaf.delta, iaf.bias)
    u = vid.frames.reshape(geom.frames, -1)
    tk1, tk2, nid, off = dec.intervals(st)
    q, Tm, tv = dec.moments(tk1, tk2, nid, off)
    n = int(off[dec.N].item())
    D, N = geom.D, dec.N
    flops = float(N) * D * (D + 1)
    # MFMA work issued: workgroups x rows x the upper-triangle column tiles x padded neurons
    nct = -(-geom.Dt * (geom.Dt + 1) // 2 // 16)
    nct = 1 if nct <= 1 else 3 if nct <= 3 else 6 if nct <= 6 else 10
    nb16, nb8 = -(-geom.Ds // 16), -(-geom.Ds // 8)
    exec_lp = nb16 * (nb16 + 1) / 2 * 256 * nct * 16 * 2.0 * (-(-N // 32) * 32)
    exec_64 = nb8 * (nb8 + 1) / 2 * 64 * nct * 16 * 2.0 * (-(-N // 16) * 16)

    def timed(fn, reps=5):
        """Median of reps timed calls (HIP events), after one warm call."""
        out = fn()                             # warm (and, for a new shape, the tile tuning)
        torch.cuda.synchronize()
        ts = []
        for _ in range(reps):
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record()
            out = fn()
            e1.record()
            torch.cuda.synchronize()
            ts.append(e0.elapsed_time(e1))
        return out, sorted(ts)[reps // 2]

    A64, t64 = timed(lambda: dec.gram(Tm))
    L64 = torch.tril(A64)
    ref = float(torch.linalg.norm(L64))
    forms = {"fp64": {"gram_ms": round(t64, 3), "gram_tflops": round(flops / t64 / 1e9, 2),
                      "gram_frac_of_fp64_peak": round(flops / t64 / 1e9 / 78.6, 4),
                      "mfma_executed_tflops": round(exec_64 / t64 / 1e9, 2),
                      "mfma_executed_frac": round(exec_64 / t64 / 1e9 / 78.6, 4)}}
    del L64
    def lp_gram(mode, name):
        """The low-precision Gram as device work only: buffers and the descriptor are made
        once, the timed call is the T split (with the FP16 scales) and the Gram kernel."""
        n16 = split_T_size(N, geom.Dt)
        hi = torch.empty(n16, dtype=torch.int16, device="cuda")
        lo = torch.empty(n16 if mode == 0 else 1, dtype=torch.int16, device="cuda")
        scales = torch.ones(5, dtype=torch.float32, device="cuda")
        A = torch.empty((D, D), dtype=torch.float32, device="cuda")
        desc = gram_descs([(dec.S, hi, lo, A, N, geom.Ds, geom.Dt, scales)], "cuda")
        s = stream_ptr()
        wss = {t: torch.empty(max(1, gram_lp_ws_bytes(geom.Ds, geom.Dt, mode, t) // 4), dtype=torch.float32,
                              device="cuda") for t in (2, 4, 8)}

        def run():
            lib.vt_gram_split_T(ptr(Tm), N, geom.Dt, ptr(hi), ptr(lo), mode, ptr(dec.S), dec.S.numel(),
                                ptr(scales), s)
            gram_tuner().launch(lambda tile: lib.vt_gram_lp(ptr(desc), 1, geom.Ds, geom.Dt, mode, tile,
                                                            ptr(wss[tile]), 4 * wss[tile].numel(), s),
                                name, geom.Dt, geom.Ds, 1, N)
            run.keep = (hi, lo, scales, desc, wss)
            return A
        return run

    for mode, name, peak in ((0, "bf16x3", 2500.0 / 3), (1, "fp16", 2500.0)):
        A, t = timed(lp_gram(mode, name))
        err = float(torch.linalg.norm(torch.tril(A.double() - A64))) / ref
        forms[name] = {"gram_rel_frobenius_err": float(f"{err:.3g}"), "gram_ms": round(t, 3),
                       "gram_tflops": round(flops / t / 1e9, 2),
                       "gram_frac_of_mfma_peak": round(flops / t / 1e9 / peak, 4),
                       "peak_tflops": round(peak, 1),
                       "mfma_executed_tflops": round(exec_lp * (3 if mode == 0 else 1) / t / 1e9, 2),
                       "mfma_executed_frac": round(exec_lp * (3 if mode == 0 else 1) / t / 1e9 / 2500.0, 4)}
        # the autotuner's tiling for this shape (row tiles per wave) and, when it was tuned
        # in this process, the time of every tiling (ms, Gram kernel alone)
        key = tune_key(name, geom.Dt, geom.Ds, 1, N)
        forms[name]["tuned_tile"] = gram_tuner().choices.get(key)
        if key in gram_tuner().timings:
            forms[name]["tile_ms"] = {str(k): round(v, 3) for k, v in gram_tuner().timings[key].items()}
        del A
    del A64
    torch.cuda.empty_cache()
    plain = dec.decode(st, precision="fp64")
    fr64 = plain.frames.cpu().numpy()
    s64 = snr_db(u, fr64)
    forms["fp64"]["snr_db"] = round(s64, 4)
    for name in ("mixed", "bf16x3", "fp16"):
        r = dec.decode(st, precision=name, refine_iters=6)
        h = np.sqrt(r.resid.cpu().numpy()) if r.resid is not None else None
        f = forms.setdefault(name, {})
        f["snr_db"] = round(snr_db(u, r.frames.cpu().numpy()), 4)
        f["d_snr_vs_fp64_db"] = float(f"{abs(f['snr_db'] - s64):.3g}")
        f["final_rel_residual"] = None if h is None else float(f"{h[-1] / h[0]:.3g}")
        f["fp64_fallback"] = bool(r.fallback)
    torch.cuda.synchronize()
    res, tk = timed(lambda: BatchDecoder(precision="kpcg").decode([BlockJob(dec, st)]))
    r = res[0]
    h = r.resid.cpu().numpy()
    forms["kpcg"] = {"snr_db": round(snr_db(u, r.frames.cpu().numpy()), 4),
                     "d_snr_vs_fp64_db": float(f"{abs(snr_db(u, r.frames.cpu().numpy()) - s64):.3g}"),
                     "iterations": int(r.state.cpu().numpy()[3]),
                     "final_rel_residual": float(f"{math.sqrt(h[1] / h[0]):.3g}"),
                     "decode_ms": round(tk, 3)}
    return {"workload": "128x128 px, 2 s, Lx=Ly=12, Lt=8, D=10625, 8192 neurons (dyadic Gabor + DoG)",
            "intervals": n, "forms": forms}


def stitched_c3(seed=0, precision="kpcg", max_batch=1024):
    """Batching form: a 512 x 512 px, 4 s video (trigonometric polynomial with the frame's
    and the duration's periods, orders 48 / 24: bandwidths Wx = 2 pi 48/512 rad/px and
    Wt = 2 pi 24/4 rad/s) in 64 px tiles with 16 px overlap and 1 s windows with 0.25 s
    overlap: 11 x 11 tiles x 5 windows = 605 blocks of Lx = Ly = 8, Lt = 8, each decoded in
    the Dirichlet space whose periods follow from the orders and the bandwidth,
    2 pi M / W = 85.3 px and 1.33 s (the reference's rule), each tile with 2048 neurons of
    the config-1 laws, IAF b = 2.5, delta ~ N(0.04, 0.004^2), and a contrast gain
    log-uniform in [0.25, 4].  Timed (video, RF bank, IAF parameters and decoders
    resident): the batched encode of all tiles, the batched decode of all blocks and the
    raised-cosine blending.  SNR over the full video and its interior (16 px border and the
    first / last 0.25 s excluded)."""
    import torch
    from vtem_amd import synthetic as syn
    from vtem_amd.stitch import TiledCodec
    vid = syn.make_video(512, 512, 4.0, 100.0, 48, 48, 24, seed=seed)
    rng = np.random.default_rng(seed + 3)
    codec = TiledCodec(512, 512, 4.0, 24, syn.bank_dyadic(64.0, 64.0),
                       Wx=2 * math.pi * 48 / 512, Wt=2 * math.pi * 24 / 4.0)
    n = len(codec.plan.tiles)
    N = codec.encoder.RF.shape[0]
    iaf = [syn.iaf_params(N, 2.5, 1.0, 0.04, 0.004, rng) for _ in range(n)]
    contrast = np.exp(rng.uniform(math.log(0.25), math.log(4.0), n))
    decs = codec.decoders(iaf, contrast)
    devs = codec.iaf_devices(iaf, contrast)

    def run(frames):
        V = torch.tensor(frames, dtype=torch.float64, device="cuda")
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        spikes = codec.encode(V, iaf, contrast, devices=devs)
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        out = codec.decode(spikes, decs, precision=precision, max_batch=max_batch)
        torch.cuda.synchronize()
        t2 = time.perf_counter()
        u = out.cpu().numpy()
        nsp = int(sum(int(s_.counts.sum().item()) for s_ in spikes))
        del spikes, out, V
        torch.cuda.empty_cache()
        return u, t1 - t0, t2 - t1, nsp, dict(codec.last_stats)

    run(vid.frames[:, :, :])                              # warm: code objects, allocator
    u, te, td, nsp, dst = run(vid.frames)
    x = vid.frames
    inner = (slice(25, -25), slice(16, -16), slice(16, -16))
    res = {"workload": "512x512 px, 4 s (video orders 48/24); 64 px tiles / 16 px overlap x 1 s windows / "
                       "0.25 s overlap; Lx=Ly=8, Lt=8 with periods 2 pi M/W = 85.3 px, 1.33 s; 2048 neurons "
                       "per tile",
           "blocks": len(codec.plan.blocks), "tiles": n, "windows": len(codec.plan.f0),
           "precision": precision, "preconditioner": dst.get("preconditioner"), "spikes": nsp,
           "kpcg_retries": dst["kpcg_retries"], "fp64_fallbacks": dst["fallbacks"],
           "encode_ms": round(te * 1e3, 2), "decode_blend_ms": round(td * 1e3, 2),
           "frames_per_s_512x512": round(400 / (te + td), 2),
           "block_frames_per_s": round(len(codec.plan.blocks) * codec.plan.window / (te + td), 1),
           "snr_db": round(snr_db(x, u), 3), "snr_interior_db": round(snr_db(x[inner], u[inner]), 3)}
    return res


KERNEL_SOURCES = ("vtem_amd/csrc/vt_batch.hip", "vtem_amd/csrc/vt_skinny.hpp", "vtem_amd/csrc/vt_common.hpp")


def kernel_source_sha16() -> str:
    """Hash of the dominant kernel's sources: the committed HBM-traffic figure is used only
    while these files are unchanged since it was measured."""
    import hashlib
    h = hashlib.sha256()
    for f in KERNEL_SOURCES:
        with open(os.path.join(ROOT, f), "rb") as fh:
            h.update(fh.read())
    return h.hexdigest()[:16]


def cholesky_timing(batch=16, D=4913, reps=3, seed=0):
    """The batched FP64 Cholesky (vt_potrf_f64) of `batch` SPD matrices G G^T / D + I of
    order D (the plain FP64 path's factorisation at D = 17^3): seconds per call and FP64
    TFLOP/s at D^3 / 3 flops per matrix (best of `reps` calls on fresh copies)."""
    import torch
    from vtem_amd._lib import lib, ptr, stream_ptr
    from vtem_amd.block import mat_descs
    g = torch.Generator(device="cuda").manual_seed(seed)
    base = []
    for _ in range(batch):
        G = torch.randn(D, D, dtype=torch.float64, device="cuda", generator=g)
        base.append(G @ G.T / D + torch.eye(D, dtype=torch.float64, device="cuda"))
    nblk = (D + 63) // 64
    inv = torch.zeros(batch * nblk * 4096, dtype=torch.float64, device="cuda")
    info = torch.zeros(batch, dtype=torch.int32, device="cuda")
    best = float("inf")
    for _ in range(reps):
        mats = [x.clone() for x in base]
        desc = mat_descs([(m.data_ptr(), D, D) for m in mats], "cuda")
        info.zero_()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        lib.vt_potrf_f64(ptr(desc), batch, D, ptr(inv), nblk * 4096, ptr(info), stream_ptr())
        e1.record()
        torch.cuda.synchronize()
        best = min(best, e0.elapsed_time(e1) * 1e-3)
        if int(info.max()) != 0:
            raise RuntimeError("Cholesky of an SPD test matrix failed")
    return {"batch": batch, "D": D, "ms": round(best * 1e3, 3),
            "tflops": round(batch * D ** 3 / 3 / best / 1e12, 2), "frac_fp64_peak": round(batch * D ** 3 / 3 / best / 78.6e12, 4)}


def copy_calibration(nbytes=512 << 20, reps=5):
    """A plain device copy of known size (the scale reference for HBM byte counters)."""
    import torch
    a = torch.empty(nbytes // 8, dtype=torch.float64, device="cuda").fill_(1.0)
    b = torch.empty_like(a)
    for _ in range(reps):
        b.copy_(a)
    torch.cuda.synchronize()
    return nbytes


def blas_threads() -> int:
    """Threads of the CPU baseline: its BLAS pool (the serial IAF loop uses one)."""
    try:
        from threadpoolctl import threadpool_info
        return max([int(p.get("num_threads", 1)) for p in threadpool_info()] or [1])
    except Exception:
        return len(os.sched_getaffinity(0))


def cpu_baseline(items, budget_s):
    """Float64 oracle encode + decode of requests until ~budget_s seconds of work: RF bank,
    frame filtering, trigonometric upsampling to dt, the serial IAF loop, then the decode
    of the resulting spikes."""
    from oracle import vtem_oracle as O
    frames, t_total, used = 0, 0.0, 0
    for it in items:
        blk, geom, job = it
        if t_total >= budget_s:
            break
        t0 = time.perf_counter()
        RF = O.rf_bank(blk.bank, geom.Px, geom.Py)
        v = O.filter_video(blk.video.frames.reshape(geom.frames, -1), RF, geom.dx, geom.dy)
        vu = O.upsample(v, geom.Lt, geom.duration, geom.fps, geom.iaf_times())
        lists, *_ = O.iaf_encode(vu, geom.dt, blk.iaf.kappa, blk.iaf.bias, blk.iaf.delta,
                                 blk.iaf.v0, max_spike=4 * it.encode.max_spike)
        x, y = O._grid(geom.Px, geom.Py, geom.Sx, geom.Sy)
        B = O.spatial_basis(x, y, O.spatial_modes(geom.Lx, geom.Ly), geom.Ox, geom.Oy)
        S = O.spatial_coeffs(RF, B)
        tk1, tk2, nid, _ = O.select_intervals(lists, (0.0, geom.duration))
        N = blk.bank.num_neurons
        q = O.compute_q(tk1, tk2, nid, blk.iaf.kappa, blk.iaf.delta, blk.iaf.bias, np.ones(N))
        Tm, tv = O.neuron_T(O.temporal_factors(tk1, tk2, geom.Lt, geom.W), q, nid, N)
        A = O.gram_kron(S, Tm)
        lam = O.tikhonov_lambda(A, tk1.size)
        r = O.rhs_kron(S, tv)
        if job.method == "rnn":
            X = O.rnn_dspace(A, r, O.lanczos_eta(A, lam), job.rnn_steps, lam)
        else:
            X = O.solve_tikhonov(A, r, lam)
        O.synthesize(X, B, O.temporal_basis(geom.frame_times(), geom.Lt, geom.W))
        t_total += time.perf_counter() - t0
        frames += geom.frames
        used += 1
    return frames / t_total, used, t_total


def main():
    args = parse()
    if args.gpus != 1:
        print(json.dumps({"error": "this build measures one GPU; runs on several GPUs are not "
                                   "part of it (blocks are independent and would be assigned "
                                   "round-robin by cost to GPUs)", "n_gpus": args.gpus}))
        sys.exit(2)
    import torch
    from vtem_amd._lib import require_gpu
    from vtem_amd.serve import Server, poisson_arrivals
    require_gpu()
    if args.precision_only:
        print(json.dumps({"precision_c2": precision_report(args.seed)}), flush=True)
        return
    if args.c3_only:
        print(json.dumps({"stitched_c3": stitched_c3(args.seed)}), flush=True)
        return
    if args.cholesky_only:
        print(json.dumps({"cholesky_f64": cholesky_timing(seed=args.seed)}), flush=True)
        return
    items = prepare(args.requests, args.seed)
    if args.roofline_only:
        dk = dominant_kernel(items, args.max_batch)
        copy_calibration()
        print(json.dumps({"dominant_kernel": dk}), flush=True)
        return
    R = len(items)
    frames = sum(g.frames for _, g, _ in items)
    D = np.array([g.D for _, g, _ in items])
    zeros = np.zeros(R)                       # a closed batch: every request at t = 0

    def run_form(form):
        """Warm-up waves as closed batches; one more closed batch gives the capacity
        (requests/s); the timed steps are Poisson arrivals at args.overload x that capacity
        (a fresh seeded arrival sequence per step), served by continuous batching."""
        server = Server(max_batch=args.max_batch, streams=args.streams, precision=form,
                        rnn_lanes=args.rnn_lanes, graph_chunk=args.graph_chunk,
                        adaptive=form != "kpcg", max_delay=args.max_delay_ms * 1e-3,
                        rnn_max_delay=args.rnn_max_delay_ms * 1e-3,
                        rnn_lane_delays=[float(x) * 1e-3 for x in args.rnn_lane_delays_ms.split(",")]
                        if args.rnn_lane_delays_ms else None, rnn_persistent=args.rnn_persistent,
                        rnn_prep=not args.no_rnn_prep)
        for _ in range(args.warmup):           # allocator, code objects, precision policy
            server.serve(make_requests(items, zeros, form))
        torch.cuda.synchronize()
        closed = server.serve(make_requests(items, zeros, form))
        cap = R / closed["wall_s"]
        rate = args.overload * cap
        stats, reqs = [], None
        for k in range(args.steps):
            arr = poisson_arrivals(R, rate, np.random.default_rng(args.seed + 1000 + k))
            reqs = make_requests(items, arr, form)
            torch.cuda.synchronize()
            stats.append(server.serve(reqs))
        wall = sum(st["wall_s"] for st in stats)
        return server, stats, reqs, wall, {"frames_per_s": round(frames / closed["wall_s"], 2),
                                           "capacity_rps": round(cap, 2), "offered_rps": round(rate, 2)}

    # plain (per-block FP64) on the same wave
    plain_reqs = make_requests(items, zeros)
    Server().serve(plain_reqs, mode="plain")              # warm
    plain_reqs = make_requests(items, zeros)
    plain = Server().serve(plain_reqs, mode="plain")
    server, stats, chosen_reqs, wall, closed = run_form(args.form)
    value = frames * args.steps / wall

    def snr_diffs(reqs):
        d = []
        for (blk, geom, job), rc, rp in zip(items, reqs, plain_reqs):
            u = blk.video.frames.reshape(geom.frames, -1)
            d.append(abs(snr_db(u, rc.result.frames.cpu().numpy()) -
                         snr_db(u, rp.result.frames.cpu().numpy())))
        return d

    d_snr = snr_diffs(chosen_reqs)
    snrs = [snr_db(blk.video.frames.reshape(geom.frames, -1), rp.result.frames.cpu().numpy())
            for (blk, geom, _), rp in zip(items, plain_reqs)]
    compare = None
    if args.compare:
        _, cst, creqs, cwall, _ = run_form(args.compare)
        cd = snr_diffs(creqs)
        compare = {"form": args.compare, "frames_per_s": round(frames * args.steps / cwall, 2),
                   "p50_ms": round(cst[-1]["p50_ms"], 3), "p99_ms": round(cst[-1]["p99_ms"], 3),
                   "fp64_fallbacks": cst[-1]["fallbacks"],
                   "speedup_vs_plain": round(frames * args.steps / cwall / plain["frames_per_s"], 3),
                   "snr_max_abs_diff_vs_plain_db": float(f"{max(cd):.3g}")}
    # decode only (spikes resident -> frames resident), chosen vs plain, closed batches
    def decode_requests(form):
        from vtem_amd.serve import Request, footprint_bytes
        return [Request(i, 0.0, it.job, it.geom.frames, footprint_bytes(it.job, form))
                for i, it in enumerate(items)]
    dserver = Server(max_batch=args.max_batch, streams=args.streams, precision=args.form,
                     rnn_lanes=args.rnn_lanes, graph_chunk=args.graph_chunk, adaptive=args.form != "kpcg",
                     rnn_persistent=args.rnn_persistent, rnn_prep=not args.no_rnn_prep)
    dserver.serve(decode_requests(args.form))                     # warm
    dwall = sum(dserver.serve(decode_requests(args.form))["wall_s"] for _ in range(args.steps))
    Server().serve(decode_requests("fp64"), mode="plain")         # warm
    dplain = Server().serve(decode_requests("fp64"), mode="plain")
    decode_only = {"chosen_frames_per_s": round(frames * args.steps / dwall, 2),
                   "plain_frames_per_s": round(dplain["frames_per_s"], 2),
                   "speedup_vs_plain": round(frames * args.steps / dwall / dplain["frames_per_s"], 3),
                   "note": "spikes resident to frames resident, closed batch of the same requests"}
    # roofline of the dominant stage (HIP events on the launch stream, one instrumented wave)
    # one stream, so that a stage's events bracket only its own kernels
    iserver = Server(max_batch=args.max_batch, streams=1, precision=args.form,
                     graph_chunk=args.graph_chunk, adaptive=args.form != "kpcg",
                     rnn_persistent=args.rnn_persistent, rnn_prep=not args.no_rnn_prep)
    iserver.policy = server.policy
    timers = iserver.serve(make_requests(items, zeros), instrument=True)["timers"]
    torch.cuda.synchronize()
    stage = {k: (sum(a.elapsed_time(b) for a, b, _ in v), sum(w() if callable(w) else w for _, _, w in v), len(v))
             for k, v in timers.items()}
    dom = max(stage, key=lambda k: stage[k][0])
    t_ms, work, nlaunch = stage[dom]
    peak, unit, bound = PEAKS[dom]
    achieved = work / (t_ms * 1e-3) / (1e9 if unit == "GB/s" else 1e12)
    dk = dominant_kernel(items, args.max_batch)
    traffic = None
    tpath = os.path.join(ROOT, "profiles", "dominant_kernel_traffic.json")
    if os.path.exists(tpath):
        with open(tpath) as f:
            tr = json.load(f)
        if tr.get("blocks") == dk["blocks"] and tr.get("requests") == R and tr.get("seed") == args.seed \
                and tr.get("source_sha16") == kernel_source_sha16():
            traffic = tr.get("traffic_bytes")
    line = {
        "metric": "decoded frames/s at fixed SNR (chosen vs plain form) + p50/p99 block latency",
        "value": round(value, 2),
        "unit": "frames/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(wall * 1000.0 / args.steps, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if args.form == "kpcg" else "bf16x3+f32+f64",
        "data": "synthetic (seeded bandlimited videos resident in HBM; encode and decode timed)",
        "config": {"workload": "serving form: blocks {32,48,64,96}^2 px x {0.5,1,2} s, "
                               f"N=round(2D/9), Poisson arrivals at {args.overload:g}x the closed-batch "
                               "capacity, continuous batching",
                   "max_delay_ms": args.max_delay_ms, "rnn_max_delay_ms": args.rnn_max_delay_ms,
                   "requests_per_step": R, "frames_per_step": frames, "D_min": int(D.min()),
                   "D_max": int(D.max()), "max_batch": args.max_batch, "streams": args.streams,
                              "rnn_lanes": args.rnn_lanes, "graph_chunk": args.graph_chunk,
                   "rnn_lane_delays_ms": args.rnn_lane_delays_ms or None},
        "chosen": {"form": args.form, "frames_per_s": round(value, 2),
                   "p50_ms": round(stats[-1]["p50_ms"], 3), "p99_ms": round(stats[-1]["p99_ms"], 3),
                   "batches": stats[-1]["batches"], "fp64_fallbacks": stats[-1]["fallbacks"],
                   "rnn_preparations": stats[-1].get("rnn_preps", 0),
                   "host_launch_ms": round(stats[-1]["host_launch_s"] * 1e3, 2),
                   "host_finalize_ms": round(stats[-1]["host_finalize_s"] * 1e3, 2),
                   "wave_ms": round(stats[-1]["wall_s"] * 1e3, 2),
                   "closed_batch": closed},
        "decode_only": decode_only,
        "plain": {"frames_per_s": round(plain["frames_per_s"], 2),
                  "p50_ms": round(plain["p50_ms"], 3), "p99_ms": round(plain["p99_ms"], 3)},
        "speedup_vs_plain": round(value / plain["frames_per_s"], 3),
        "compare": compare,
        "snr": {"mean_db": round(float(np.mean(snrs)), 3), "min_db": round(float(np.min(snrs)), 3),
                "max_abs_diff_vs_plain_db": float(f"{max(d_snr):.3g}")},
        "stage_ms": {k: round(v[0], 3) for k, v in stage.items()},
        "stage_roofline": {k: {"achieved": round(v[1] / (v[0] * 1e-3) / (1e9 if PEAKS[k][1] == "GB/s" else 1e12), 3),
                               "unit": PEAKS[k][1], "peak": PEAKS[k][0],
                               "frac": round(v[1] / (v[0] * 1e-3) / (1e9 if PEAKS[k][1] == "GB/s" else 1e12)
                                             / PEAKS[k][0], 5)}
                           for k, v in stage.items() if k in PEAKS and v[1] > 0},
        "dominant_stage": {"stage": dom, "bound": bound, "achieved": round(achieved, 3), "peak": peak,
                           "unit": unit, "frac": round(achieved / peak, 5),
                           "work_note": WORK_NOTES.get(dom, "")},
        "roofline": {"kernel": dk["kernel"], "bound": "hbm", "achieved": round(dk["achieved_GBs"], 2),
                     "peak": 8000.0, "unit": "GB/s", "frac": round(dk["achieved_GBs"] / 8000.0, 5),
                     "traffic": traffic, "avg_us": round(dk["avg_us"], 3),
                     "algorithmic_bytes": dk["bytes"], "blocks": dk["blocks"]},
    }
    if not args.no_sweep:
        # p50/p99 over LAT_REQUESTS requests per rate point: the prepared blocks, cycled
        cap_rps = closed["capacity_rps"]
        sweep = []
        cyc = [items[i % R] for i in range(LAT_REQUESTS)]
        for frac in (0.25, 0.5, 0.75, 1.0):
            rate = frac * cap_rps
            arr = poisson_arrivals(LAT_REQUESTS, rate, np.random.default_rng(args.seed + 17))
            st = server.serve(make_requests(cyc, arr))
            sweep.append({"offered_rps": round(rate, 2), "requests": LAT_REQUESTS,
                          "frames_per_s": round(st["frames_per_s"], 2),
                          "p50_ms": round(st["p50_ms"], 3), "p99_ms": round(st["p99_ms"], 3)})
        line["latency_sweep"] = sweep
    if not args.no_precision:
        line["precision_c2"] = precision_report(args.seed)
    if not args.no_c3:
        line["stitched_c3"] = stitched_c3(args.seed)
    if not args.no_cpu:
        fps, used, t = cpu_baseline(items, args.cpu_budget)
        from threadpoolctl import threadpool_limits
        with threadpool_limits(limits=1):
            fps1, used1, t1 = cpu_baseline(items, args.cpu_budget / 2)
        line["cpu_baseline"] = {"value": round(fps, 3), "unit": "frames/s", "cores": blas_threads(),
                                "host_cpus": os.cpu_count(), "kind": "port",
                                "sample": f"float64 oracle encode (RF bank, frame filtering, "
                                          f"upsampling, serial IAF) + decode of the first {used} "
                                          f"requests ({t:.1f} s); cores = the BLAS pool's threads "
                                          f"(the serial IAF loop uses one)",
                                "one_thread": {"value": round(fps1, 3), "cores": 1,
                                               "sample": f"the first {used1} requests ({t1:.1f} s)"}}
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
