"""Serving runtime: continuous batching of encode+decode requests on one GPU.

Requests (video blocks of different sizes and durations: frames resident in HBM with the
block's receptive-field bank, IAF parameters and prepared decoder) arrive over time.  A
request either carries its video to encode (EncodeJob) or already-encoded spike trains.
The server:

* admits arrived requests into a ready queue, RNN-decoded requests into a queue of their
  own;
* whenever one of its streams is free, forms a batch from the head of the queue, bounded
  by a request count and by an HBM budget (each request's in-flight footprint: spike
  buffer and encoder workspaces, intervals, moments, the solver's state and the output
  frames);
* launches the whole batch on that stream without a host synchronisation -- the batched
  encode (filter products and time-parallel IAF of every block) and then the batched
  decode -- so batches formed from requests that arrive while others are on the GPU run
  concurrently on the device; the RNN queue feeds lanes of its own (rnn_lanes), whose
  batches split the waiting RNN requests between the free RNN lanes (up to
  rnn_max_batch each), so RNN requests share a few sequences of iteration launches; with
  staggered lanes (rnn_lane_delays) the first lane takes every waiting RNN request at
  once and a later lane holds its queue for a set time, gathering the RNN requests that
  arrive while the first lane's chain runs into one more chain that starts before it
  ends;
* retires a batch when its completion event fires: re-runs any block whose spike buffer
  overflowed (with an exact buffer) or whose low-precision solve missed tolerance,
  releases its memory and stamps each request's completion time.

Latency is arrival -> result resident in HBM.  The plain mode is the per-block FP64
path: one request at a time, in arrival order, no batching: the reference's encoder
(frame filtering, serial IAF loop) and the FP64 Gram + Cholesky decode.  Both modes take
crossing times from the cancellation-free root (root mode 1, see BatchEncoder).

The reference has no serving layer; its decode drivers run one spike file at a time
(vtem/vtdm.py:18-309).
"""

from __future__ import annotations

import collections
import dataclasses
import math
import time
from typing import Any, Dict, List, Optional, Sequence

import numpy as np
import torch

from .batch import BatchDecoder, BlockJob
from .batch_encode import BatchEncoder, EncodeJob, encode_footprint
from .block import DecodeResult, SpikeTrains


@dataclasses.dataclass
class Request:
    rid: int
    arrival: float                 # seconds after the start of the run
    job: BlockJob                  # decode part (spikes None when `encode` is given)
    frames: int                    # frames produced by the decode
    footprint: int                 # HBM bytes while in flight
    encode: Optional[EncodeJob] = None
    iaf: Any = None                # host IAF parameters (plain mode's encoder)
    start: Optional[float] = None
    done: Optional[float] = None
    result: Optional[DecodeResult] = None
    batch: Optional[int] = None
    spikes: Optional[SpikeTrains] = None


def footprint_bytes(job: BlockJob, form: str = "kpcg", encode: EncodeJob = None) -> int:
    """In-flight HBM of one decode (the decoder's S and the spike trains are resident
    before admission and not counted): intervals, moments T_j/t_j, output frames and the
    solver's own state -- for the Gram forms the FP32 D x D factor, its block inverses and
    the split moments; for "kpcg" the spatial preconditioner (two Ds x Ds) and about a
    dozen D-vectors; for the RNN a handful of D-vectors."""
    d = job.decoder
    g = d.geom
    D, Ds = g.D, g.Ds
    rows = encode.max_spike if encode is not None else job.spikes.spikes.shape[0]
    cap = rows * d.N
    common = 20 * cap + 8 * d.N * g.Dt * (g.Dt + 3) + 8 * g.frames * g.P + 8 * g.frames * g.Ds
    if encode is not None:
        common += encode_footprint(encode)
    if job.method == "rnn":
        return int(common + 8 * 6 * D)
    if form == "kpcg":
        return int(common + 16 * Ds * Ds + 8 * 64 * Ds + 8 * 12 * D)
    nblk = (D + 63) // 64
    return int(common + 4 * D * D + 4 * nblk * 4096 + 4 * d.N * g.Dt * g.Dt + 64 * D)


class PrecisionPolicy:
    """Per request class (spatial order Ds, temporal order Dt): start with the
    low-precision form and route the class to batched FP64 once its guard failures
    dominate (>= `min_fail` failures and a failure rate above `rate`).  Blocks of a
    class share bank density and duration, hence their conditioning."""

    def __init__(self, low: str = "bf16x3", min_fail: int = 2, rate: float = 0.5):
        self.low = low
        self.min_fail = min_fail
        self.rate = rate
        self.stats: Dict[tuple, List[int]] = {}      # class -> [attempts, failures]

    @staticmethod
    def key(job: BlockJob) -> tuple:
        g = job.decoder.geom
        return (g.Ds, g.Dt)

    def choose(self, job: BlockJob) -> str:
        a, f = self.stats.get(self.key(job), (0, 0))
        if f >= self.min_fail and f > self.rate * a:
            return "fp64"
        return self.low

    def record(self, job: BlockJob, precision: str, failed: bool):
        if precision == "fp64" or job.method != "tikhonov":
            return
        st = self.stats.setdefault(self.key(job), [0, 0])
        st[0] += 1
        st[1] += int(failed)


class Server:
    def __init__(self, hbm_budget: int = 64 << 30, max_batch: int = 32, streams: int = 2,
                 precision: str = "bf16x3", verify: bool = True, adaptive: bool = True,
                 rnn_lane: bool = True, rnn_max_batch: int = 64, rnn_lanes: int = 1,
                 graph_chunk: int = 25, max_delay: float = 0.0, rnn_max_delay: float = 0.0,
                 rnn_lane_delays: Optional[Sequence[float]] = None, rnn_persistent: bool = False,
                 rnn_prep: bool = True):
        self.hbm_budget = int(hbm_budget)
        self.max_batch = int(max_batch)
        self.n_streams = int(streams)
        # RNN requests (hundreds of dependent operator applications each) get a lane of
        # their own: one stream whose batch takes every RNN request waiting when it frees
        # up, so their iterations share launches instead of each batch running its own loop
        self.rnn_lane = bool(rnn_lane)
        self.rnn_max_batch = int(rnn_max_batch)
        self.rnn_lanes = max(1, int(rnn_lanes))
        self.precision = precision
        self.verify = verify
        # RNN and power loops replayed as HIP graphs of this many iterations (0: no graph)
        self.graph_chunk = int(graph_chunk)
        self.policy = PrecisionPolicy(precision) if adaptive else None
        # dynamic batching: a free lane holds its queue until the batch is full, its
        # oldest request has waited this long (seconds), or no further arrivals are due
        self.max_delay = float(max_delay)
        self.rnn_max_delay = float(rnn_max_delay)
        # staggered RNN lanes: lane k holds its queue until the oldest request has waited
        # rnn_lane_delays[k] (seconds) and takes every waiting RNN request (no even split),
        # so a later lane gathers the requests that arrive while the first lane's chain runs
        self.rnn_lane_delays = None if rnn_lane_delays is None else [float(x) for x in rnn_lane_delays]
        if self.rnn_lane_delays is not None:
            self.rnn_lanes = len(self.rnn_lane_delays)
        # RNN iterations in one persistent launch (default: two launches per iteration,
        # replayed as HIP graphs)
        self.rnn_persistent = bool(rnn_persistent)
        # RNN requests that arrive while the RNN lane runs a chain are prepared at once
        # (encode, prologue, step sizes), so the next chain starts with its iterations: on a
        # stream of their own with one RNN lane (without held queues), on the held lane's
        # stream with staggered lanes
        self.rnn_prep = bool(rnn_prep)
        self._streams: List[torch.cuda.Stream] = []

    def _lane_streams(self, n: int) -> List[torch.cuda.Stream]:
        """The server's n streams, made once and reused by every serve call, on distinct
        hardware queues where the device has enough (independent_streams)."""
        if len(self._streams) < n:
            self._streams = independent_streams(n)
        return self._streams[:n]

    # ----------------------------------------------------------------------------------
    def serve(self, requests: List[Request], mode: str = "served",
              clock=time.perf_counter, instrument: bool = False) -> Dict[str, Any]:
        """Serve the requests; with `instrument`, stats["timers"] holds HIP events of the
        batched stages (name -> [(start, end, algorithmic work)]; the work is a number, or a
        callable giving it once the stream has run)."""
        if mode == "plain":
            return self._serve_plain(requests, clock)
        for r in requests:
            if r.footprint > self.hbm_budget:
                raise ValueError(f"request {r.rid} needs {r.footprint} B, above the HBM budget")
        bd = BatchDecoder(precision=self.precision, verify=self.verify, graph_chunk=self.graph_chunk,
                          rnn_persistent=self.rnn_persistent)
        bd.instrument(instrument)
        be = BatchEncoder()
        overflows = 0
        # lanes: (stream, queue, max batch); lane queues hold arrived, not yet launched requests
        ready: collections.deque = collections.deque()
        ready_rnn: collections.deque = collections.deque()
        # The RNN lane comes first: when several lanes are free, its batch (the longest chain
        # of dependent launches) is enqueued before the host spends its time on the others.
        nr = self.rnn_lanes if self.rnn_lane else 0
        prep_on = (nr == 1 and self.rnn_prep and self.rnn_lane_delays is None and self.rnn_max_delay == 0.0)
        all_streams = self._lane_streams(nr + self.n_streams + int(prep_on))
        lanes = [(all_streams[k], ready_rnn, self.rnn_max_batch) for k in range(nr)]
        lanes += [(all_streams[nr + k], ready, self.max_batch) for k in range(self.n_streams)]
        streams = [ln[0] for ln in lanes]
        busy = [None] * len(lanes)              # (event, handle, reqs, bytes, [(reqs, overflow flags)])
        # streams that prepare waiting RNN requests for a lane (lane -> stream): one RNN lane
        # has a stream of its own for it; staggered lanes prepare on a held lane's stream
        prep_streams: Dict[int, torch.cuda.Stream] = {}
        if prep_on:
            prep_streams[0] = all_streams[-1]
        elif self.rnn_prep and self.rnn_lane_delays is not None and nr > 1:
            prep_streams = {k: streams[k] for k in range(1, nr) if self.rnn_lane_delays[k] > 0.0}
        held: Dict[int, List[tuple]] = {k: [] for k in prep_streams}   # (reqs, RnnPrep, bytes, flag sets)
        preps_made = 0
        pending = collections.deque(sorted(requests, key=lambda r: r.arrival))
        in_use = 0
        nbatch = 0
        fallbacks = 0
        host_launch = host_finalize = 0.0      # host seconds spent enqueueing / retiring
        torch.cuda.synchronize()
        t0 = clock()
        while pending or ready or ready_rnn or any(held.values()) or any(b is not None for b in busy):
            now = clock() - t0
            while pending and pending[0].arrival <= now:
                r = pending.popleft()
                (ready_rnn if self.rnn_lane and is_rnn(r) else ready).append(r)
            # retire finished batches
            for k, b in enumerate(busy):
                if b is not None and b[0].query():
                    ev, h, reqs, nbytes, flag_sets = b
                    tf = clock()
                    with torch.cuda.stream(streams[k]):
                        res = bd.finalize(h)
                        for enc, flags in flag_sets:
                            over = [r for r, f in zip(enc, flags.cpu().numpy()) if f]
                            for r in over:
                                i = reqs.index(r)
                                res[i] = self._redo_overflow(r, be, bd)
                            overflows += len(over)
                    fallbacks += bd.last_stats.get("fallbacks", 0)
                    if self.policy is not None:
                        failed = set(bd.last_stats.get("fallback_index", []))
                        for i, (r, pr) in enumerate(zip(reqs, h.precisions)):
                            self.policy.record(r.job, pr, i in failed)
                    tdone = clock() - t0
                    host_finalize += tdone + t0 - tf
                    for r, rr in zip(reqs, res):
                        r.result, r.done = rr, tdone
                    in_use -= nbytes
                    busy[k] = None
            # prepare the RNN requests waiting behind a running chain (one preparation in
            # flight per stream, so arrivals during it are prepared together)
            for k, ps in prep_streams.items():
                if not ready_rnn or busy[0] is None or (k > 0 and busy[k] is not None):
                    continue
                if held[k] and not held[k][-1][1].done.query():
                    continue
                batch, nbytes = form_batch(ready_rnn, self.rnn_max_batch, self.hbm_budget - in_use)
                if batch:
                    tstart = clock() - t0
                    with torch.cuda.stream(ps):
                        flags = self._encode(batch, be, bd)
                        prep = bd.prepare_rnn([r.job for r in batch])
                    host_launch += clock() - t0 - tstart
                    for r in batch:
                        r.start = tstart
                    in_use += nbytes
                    preps_made += 1
                    held[k].append((batch, prep, nbytes, self._flag_set(batch, flags)))
            # launch on a free lane
            launched = False
            wake = math.inf                     # earliest batching deadline of a held queue
            for k, (_, queue, cap) in enumerate(lanes):
                # an RNN lane with prepared requests: the first lane takes every preparation
                # at once, a held lane its own once its oldest request has waited its delay;
                # they go into one chain with any waiting requests (prepared on the lane's
                # own stream first)
                take = [] if k >= nr or busy[k] is not None else \
                    [j for j in held if held[j]] if k == 0 else ([k] if held.get(k) else [])
                if take:
                    oldest = min(r.arrival for j in take for p in held[j] for r in p[0])
                    delay = 0.0 if k == 0 or self.rnn_lane_delays is None else self.rnn_lane_delays[k]
                    count = sum(len(p[0]) for j in take for p in held[j]) + len(queue)
                    if not should_launch(count, cap, clock() - t0 - oldest, delay, bool(pending)):
                        wake = min(wake, oldest + delay)
                        continue
                    tstart = clock() - t0
                    batch, nbytes = form_batch(queue, self.rnn_max_batch, self.hbm_budget - in_use)
                    ready_preps = [p for j in take for p in held[j]]
                    reqs = [r for p in ready_preps for r in p[0]] + batch
                    flag_sets = [fs for p in ready_preps for fs in p[3]]
                    with torch.cuda.stream(streams[k]):
                        preps = [p[1] for p in ready_preps]
                        if batch:
                            flags = self._encode(batch, be, bd)
                            preps.append(bd.prepare_rnn([r.job for r in batch]))
                            flag_sets += self._flag_set(batch, flags)
                        h = bd.launch_rnn(preps)
                        ev = torch.cuda.Event()
                        ev.record()
                    host_launch += clock() - t0 - tstart
                    for r in batch:
                        r.start = tstart
                    for r in reqs:
                        r.batch = nbatch
                    nbatch += 1
                    in_use += nbytes
                    busy[k] = (ev, h, reqs, nbytes + sum(p[2] for p in ready_preps), flag_sets)
                    for j in take:
                        held[j] = []
                    launched = True
                    continue
                if busy[k] is not None or not queue:
                    continue
                if k < nr and self.rnn_lane_delays is None:
                    cap = split_cap(len(queue), cap, sum(1 for i in range(k, nr) if busy[i] is None))
                if k < nr:
                    delay = self.rnn_max_delay if self.rnn_lane_delays is None else self.rnn_lane_delays[k]
                else:
                    delay = self.max_delay
                now = clock() - t0
                if not should_launch(len(queue), cap, now - queue[0].arrival, delay, bool(pending)):
                    wake = min(wake, queue[0].arrival + delay)
                    continue
                batch, nbytes = form_batch(queue, cap, self.hbm_budget - in_use)
                if not batch:
                    continue
                tstart = clock() - t0
                precs = [self.policy.choose(r.job) if self.policy else self.precision
                         for r in batch]
                with torch.cuda.stream(streams[k]):
                    flags = self._encode(batch, be, bd)
                    h = bd.launch([r.job for r in batch], precs)
                    ev = torch.cuda.Event()
                    ev.record()
                host_launch += clock() - t0 - tstart
                for r in batch:
                    r.start, r.batch = tstart, nbatch
                nbatch += 1
                in_use += nbytes
                busy[k] = (ev, h, batch, nbytes, self._flag_set(batch, flags))
                launched = True
            if not launched:
                # nothing to launch: wait for the next arrival, a batching deadline or a
                # completion
                now = clock() - t0
                nxt = min(pending[0].arrival if pending else math.inf, wake) - now
                if any(b is not None for b in busy):
                    time.sleep(min(max(nxt, 0.0), 2e-4))
                elif nxt > 0 and nxt < math.inf:
                    time.sleep(min(nxt, 5e-3))
        torch.cuda.synchronize()
        wall = clock() - t0
        extra = {"batches": nbatch, "fallbacks": fallbacks, "overflows": overflows, "rnn_preps": preps_made,
                 "host_launch_s": host_launch, "host_finalize_s": host_finalize}
        if instrument:
            extra["timers"] = bd.timers
        return self._stats(requests, wall, extra)

    @staticmethod
    def _flag_set(batch: List[Request], flags) -> list:
        """[(the batch's requests that were encoded, their overflow flags)], or []."""
        return [] if flags is None else [([r for r in batch if r.encode is not None], flags)]

    @staticmethod
    def _encode(batch: List[Request], be: BatchEncoder, bd: BatchDecoder):
        """Batched encode of the requests that carry video; their decode jobs get the
        resident spike trains.  Returns the per-request overflow flags (device) or None.
        Instrumented as stages "encode_coeffs" (algorithmic HBM bytes: frames, RF banks and
        the two coefficient products) and "iaf" (algorithmic FP64 flops of the input's
        evaluation)."""
        enc = [r for r in batch if r.encode is not None]
        if not enc:
            return None
        flags = torch.zeros(len(enc), dtype=torch.int32, device=be.device)
        sts = be.encode([r.encode for r in enc], flags, carry=False,
                        timers=bd if bd.timers is not None else None)
        for r, st in zip(enc, sts):
            r.spikes = st
            r.job = dataclasses.replace(r.job, spikes=st)
        return flags

    @staticmethod
    def _redo_overflow(r: Request, be: BatchEncoder, bd: BatchDecoder) -> DecodeResult:
        """A block fired more often than its buffer held: encode it again with the buffer
        the counts call for and decode it alone (FP64-guarded solve of the batch's form)."""
        need = int(r.spikes.counts.max().item())
        while True:
            job = dataclasses.replace(r.encode, max_spike=2 * max(need, r.encode.max_spike))
            flags = torch.zeros(1, dtype=torch.int32, device=be.device)
            st = be.encode([job], flags, carry=False)[0]
            if not int(flags.item()):
                break
            need = int(st.counts.max().item())
        r.spikes = st
        r.job = dataclasses.replace(r.job, spikes=st)
        return bd.decode([r.job])[0]

    def _serve_plain(self, requests, clock):
        order = sorted(requests, key=lambda r: r.arrival)
        torch.cuda.synchronize()
        t0 = clock()
        for r in order:
            now = clock() - t0
            if r.arrival > now:
                time.sleep(r.arrival - now)
            r.start = clock() - t0
            jb = r.job
            spikes = jb.spikes
            if r.encode is not None:
                # the reference's encoder per block: every frame filtered, then the
                # temporal coefficients, then the serial one-lane-per-neuron IAF loop
                e = r.encode
                spikes = e.encoder.encode(e.frames, r.iaf, parallel=False, root_mode=1)
                r.spikes = spikes
            r.result = jb.decoder.decode(spikes, jb.t0, precision="fp64", method=jb.method,
                                         rnn_steps=jb.rnn_steps)
            torch.cuda.current_stream().synchronize()
            r.done = clock() - t0
        wall = clock() - t0
        return self._stats(requests, wall, {"batches": len(order), "fallbacks": 0})

    @staticmethod
    def _stats(requests, wall, extra):
        lat = np.array([r.done - r.arrival for r in requests]) * 1000.0
        frames = sum(r.frames for r in requests)
        first = min(r.arrival for r in requests)
        span = max(r.done for r in requests) - first
        out = {"requests": len(requests), "frames": frames, "wall_s": wall,
               "frames_per_s": frames / span if span > 0 else float("nan"),
               "p50_ms": float(np.percentile(lat, 50)), "p99_ms": float(np.percentile(lat, 99)),
               "mean_ms": float(lat.mean())}
        out.update(extra)
        return out


def streams_overlap(a: torch.cuda.Stream, b: torch.cuda.Stream, spin_ms: float = 2.0) -> bool:
    """True if work enqueued on b runs while a spin kernel occupies a (and the reverse).
    Two streams that HIP placed on one hardware queue run in submission order, so the
    second one's work waits for the spin."""
    spin = getattr(torch.cuda, "_sleep", None)   # torch's spin kernel (a test utility)
    if spin is None:
        return True                              # cannot probe: take the streams as they are
    x = torch.zeros(1, device=a.device)
    torch.cuda.synchronize()
    cycles = int(spin_ms * 1.6e6)              # the spin counts at about 1.6 GHz
    for s, t in ((a, b), (b, a)):
        with torch.cuda.stream(s):
            spin(cycles)
            e_spin = torch.cuda.Event()
            e_spin.record()
        with torch.cuda.stream(t):
            x.add_(1.0)
            e_work = torch.cuda.Event()
            e_work.record()
        t0 = time.perf_counter()
        while not e_work.query() and time.perf_counter() - t0 < 0.1:
            pass
        ran_beside = e_work.query() and not e_spin.query()
        torch.cuda.synchronize()
        if not ran_beside:
            return False
    return True


def independent_streams(n: int, tries: int = 16) -> List[torch.cuda.Stream]:
    """n streams on distinct hardware queues where the device has enough.  HIP gives a
    stream a hardware queue when the stream first runs work, and once every queue is taken
    further streams share them: two lanes on one queue run one after the other.  Streams
    are given work in order and tested against each kept stream (streams_overlap); one that
    waits for a kept stream is replaced by a new one, up to `tries` times, after which it
    is kept as it is."""
    kept: List[torch.cuda.Stream] = []
    spare = tries
    while len(kept) < n:
        s = torch.cuda.Stream()
        with torch.cuda.stream(s):
            torch.zeros(1, device=s.device).add_(1.0)   # the stream takes its queue now
        torch.cuda.synchronize()
        if spare > 0 and not all(streams_overlap(k, s) for k in kept):
            spare -= 1
            continue
        kept.append(s)
    return kept


def is_rnn(r: Request) -> bool:
    """Requests decoded by the gradient-flow RNN go to the server's RNN lane."""
    return r.job.method == "rnn"


def split_cap(waiting: int, cap: int, free_lanes: int) -> int:
    """Batch cap of one of `free_lanes` free lanes that share a queue of `waiting`
    requests: an even split, at most `cap`."""
    return max(1, min(cap, -(-waiting // max(1, free_lanes))))


def should_launch(queued: int, cap: int, oldest_wait: float, max_delay: float,
                  more_due: bool) -> bool:
    """Dynamic batching rule of a free lane: launch when the queue fills a batch, when its
    oldest request has waited max_delay seconds, or when no further arrivals are due."""
    return queued > 0 and (queued >= cap or oldest_wait >= max_delay or not more_due)


def form_batch(queue: collections.deque, cap: int, free_bytes: int):
    """Take requests from the head of `queue` (arrival order) while the batch holds fewer
    than `cap` and their footprints fit in `free_bytes`.  Returns (batch, bytes)."""
    batch, nbytes = [], 0
    while queue and len(batch) < cap and nbytes + queue[0].footprint <= free_bytes:
        r = queue.popleft()
        batch.append(r)
        nbytes += r.footprint
    return batch, nbytes


def poisson_arrivals(n: int, rate: float, rng: np.random.Generator) -> np.ndarray:
    """Arrival times of a Poisson process with `rate` requests/s (first at 0)."""
    if rate <= 0 or math.isinf(rate):
        return np.zeros(n)
    gaps = rng.exponential(1.0 / rate, n)
    gaps[0] = 0.0
    return np.cumsum(gaps)
