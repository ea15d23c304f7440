"""Batched decode of many independent blocks (tiles / windows / served requests).

The reference decodes stitched tiles and windows one after another
(vtem/vtdm.py:236-302 spatial tiles, :630-692 temporal windows), each with its own
VTDM_prep / compute_G / solve / reconstruct.  Here a group of blocks that share the
temporal order Dt runs every stage as one launch (or two) with per-block descriptors:

  intervals (global neuron numbering) -> q -> T_j, t_j -> r_b = S_b^T t_b -> lam_b
  -> low-precision Gram A_b (split-BF16 x3 or FP16, FP32 accumulation)
  -> A_b + lam_b I -> batched FP32 Cholesky -> batched triangular solves
  -> FP64 iterative refinement through each block's exact Kronecker operator
  -> frames_b = Theta_b X_b^T B_s,b^T.

Blocks keep their own sizes (neurons, spatial order, spike counts, duration); matrices
of different D share one launch sequence (a block that has finished its columns skips
the step).  Each block's result is the same as decoding it alone: no reduction order
depends on which other blocks are in the batch.
"""

from __future__ import annotations

import ctypes
import dataclasses
import math
from typing import Any, Dict, List, Optional

import numpy as np
import torch

from ._lib import VtemError, lib, ptr, stream_ptr
from .autotune import gram_tuner
from .block import (F64, I32, LANCZOS_STEPS, MAT_DTYPE, Decoder, DecodeResult, SpikeTrains, mat_descs,
                    split_T_size, tile_doubles)

SPIKE_DTYPE = np.dtype([("spikes", np.uint64), ("counts", np.uint64), ("ld", np.int64),
                        ("max_rows", np.int32), ("neuron_base", np.int32), ("t0", np.float64),
                        ("Tw", np.float64)])
TJ_DTYPE = np.dtype([("neuron_base", np.int32), ("num_neurons", np.int32), ("t_base", np.int64),
                     ("Lt", np.int32), ("pad0", np.int32), ("W", np.float64), ("T_base", np.int32),
                     ("pad1", np.int32)])
GEMM_DTYPE = np.dtype([("A", np.uint64), ("B", np.uint64), ("C", np.uint64), ("sam", np.int64),
                       ("sak", np.int64), ("sbk", np.int64), ("sbn", np.int64), ("ldc", np.int64),
                       ("M", np.int32), ("N", np.int32), ("K", np.int32), ("pad", np.int32),
                       ("alpha", np.float64), ("beta", np.float64)])
TN_DTYPE = np.dtype([("A", np.uint64), ("B", np.uint64), ("C", np.uint64), ("ws", np.uint64),
                     ("lda", np.int64), ("ldb", np.int64), ("ldc", np.int64), ("M", np.int32),
                     ("K", np.int32)])
DIAG_DTYPE = np.dtype([("A", np.uint64), ("lda", np.int64), ("D", np.int32), ("pad", np.int32)])
GRAM_LP_DTYPE = np.dtype([("S", np.uint64), ("Th", np.uint64), ("Tl", np.uint64), ("A", np.uint64),
                          ("lds", np.int64), ("lda", np.int64), ("N", np.int32), ("Ds", np.int32),
                          ("Dt", np.int32), ("pad", np.int32), ("scales", np.uint64)])
GRAM64_DTYPE = np.dtype([("S", np.uint64), ("Tm", np.uint64), ("A", np.uint64), ("lds", np.int64),
                         ("lda", np.int64), ("N", np.int32), ("Ds", np.int32), ("Dt", np.int32),
                         ("pad", np.int32)])
KPM_DTYPE = np.dtype([("S", np.uint64), ("c", np.uint64), ("M", np.uint64), ("lds", np.int64),
                      ("ldm", np.int64), ("N", np.int32), ("Ds", np.int32)])
SK_DTYPE = np.dtype([("A", np.uint64), ("B", np.uint64), ("C", np.uint64), ("T", np.uint64),
                     ("ws", np.uint64), ("active", np.uint64), ("sam", np.int64), ("sak", np.int64),
                     ("ldb", np.int64), ("ldc", np.int64), ("tstride", np.int64), ("M", np.int32),
                     ("K", np.int32), ("Nc", np.int32), ("pad", np.int32)])

assert SPIKE_DTYPE.itemsize == 48 and TJ_DTYPE.itemsize == 40 and GEMM_DTYPE.itemsize == 96
assert TN_DTYPE.itemsize == 64 and DIAG_DTYPE.itemsize == 24
assert GRAM_LP_DTYPE.itemsize == 72 and GRAM64_DTYPE.itemsize == 56 and KPM_DTYPE.itemsize == 48
assert SK_DTYPE.itemsize == 104
# SK_DTYPE "pad" bits: operands stored in the MFMA tile layout (vt_tile_f64)
SK_A_TILED, SK_B_TILED, SK_C_TILED = 1, 2, 4


def _to_dev(arr: np.ndarray, device) -> torch.Tensor:
    """Descriptor bytes to the device: staged in pinned host memory and copied without
    blocking the host (the caching host allocator keeps the staging block until the copy
    has run)."""
    host = torch.from_numpy(np.ascontiguousarray(arr).view(np.uint8).copy()).pin_memory()
    return host.to(device, non_blocking=True)


def _h2d(arr: np.ndarray, device) -> torch.Tensor:
    """A small index/offset array to the device (pinned staging, non-blocking)."""
    return torch.from_numpy(np.ascontiguousarray(arr)).pin_memory().to(device, non_blocking=True)


def skinny_workspace(M: int, K: int, Nc: int) -> int:
    """Doubles of split-K workspace one skinny problem needs (queried from the library)."""
    out = ctypes.c_longlong(0)
    lib.vt_gemm_skinny_workspace_size(int(M), int(K), int(Nc), ctypes.addressof(out))
    return int(out.value)


_SPLIT_RULE = None


def skinny_splits_array(K) -> np.ndarray:
    """K splits of skinny problems, by the rule the library reports (vt_skinny_split_rule:
    one split per chunks_per_split chunks of 16 along K, between 1 and max_splits)."""
    global _SPLIT_RULE
    if _SPLIT_RULE is None:
        c, m = ctypes.c_int(0), ctypes.c_int(0)
        lib.vt_skinny_split_rule(ctypes.addressof(c), ctypes.addressof(m))
        _SPLIT_RULE = (int(c.value), int(m.value))
    cps, mx = _SPLIT_RULE
    return np.clip(((np.asarray(K, np.int64) + 15) // 16) // cps, 1, mx)


def skinny_workspace_array(M, K, Nc) -> np.ndarray:
    """skinny_workspace for arrays of problems: ks M Nc doubles when a problem's K is split
    in ks > 1 parts."""
    M, K, Nc = (np.asarray(v, np.int64) for v in (M, K, Nc))
    ks = skinny_splits_array(K)
    return np.where(ks > 1, ks * M * Nc, 0)


@dataclasses.dataclass
class BlockJob:
    """One block to decode: a prepared Decoder (geometry, S, IAF parameters) and its
    resident spike trains; window starts at t0 and lasts the decoder's duration.
    method "tikhonov" (pseudoinverse, regularised) or "rnn" (rnn_steps + 1 gradient-flow
    iterations with step 1/lambda_max).  x0: a starting iterate for a "kpcg" solve (the
    D-space coefficients of an earlier solve of the same block): the solve then runs on
    the residual r - (A + lam I) x0 and adds its correction."""
    decoder: Decoder
    spikes: SpikeTrains
    t0: float = 0.0
    tag: Any = None
    method: str = "tikhonov"
    rnn_steps: int = 500
    x0: Optional[torch.Tensor] = None


@dataclasses.dataclass
class _Prologue:
    """Per-group products of the prologue (intervals -> q -> T_j, t_j -> r -> lam) in the
    group's global neuron / D-space numbering."""
    jobs: list
    decs: list
    Dt: int
    B: int
    N: np.ndarray
    Ds: np.ndarray
    D: np.ndarray
    nbase: np.ndarray
    dbase: np.ndarray
    Ntot: int
    Dtot: int
    Tm: torch.Tensor
    r: torch.Tensor
    lam: torch.Tensor
    n_b: torch.Tensor
    tn_ws: torch.Tensor
    ws_base: np.ndarray
    seg: torch.Tensor
    elem_job: torch.Tensor
    nbase_d: torch.Tensor
    keep: tuple


def _fill_gemm(arr, A, B, C, sam, sak, sbk, sbn, ldc, M, N, K, alpha):
    """Fill a GEMM_DTYPE descriptor array column by column (scalars or per-row arrays)."""
    for name, v in (("A", A), ("B", B), ("C", C)):
        arr[name] = np.asarray(v, np.int64).astype(np.uint64)
    for name, v in (("sam", sam), ("sak", sak), ("sbk", sbk), ("sbn", sbn), ("ldc", ldc), ("M", M), ("N", N),
                    ("K", K)):
        arr[name] = v
    arr["alpha"] = alpha
    arr["beta"] = 0.0


class _RnnBarrier:
    """Barrier workspace of a persistent RNN launch (word 0: the abort word)."""

    def __init__(self, bar):
        self.bar = bar


_RNN_CAP = None


def _rnn_capacity():
    """([workgroups that fit for <= 17 columns, for wider], largest group) of the
    persistent RNN launch on the current device (queried once)."""
    global _RNN_CAP
    if _RNN_CAP is None:
        w = (ctypes.c_int * 2)()
        mg = ctypes.c_int(0)
        lib.vt_kron_rnn_persistent_capacity(w, ctypes.byref(mg))
        _RNN_CAP = ([int(w[0]), int(w[1])], int(mg.value))
    return _RNN_CAP


@dataclasses.dataclass
class _Skinny:
    """Launch record of one skinny grouped product."""
    descs: torch.Tensor
    batch: int
    max_M: int
    max_K: int
    max_Nc: int
    ws: torch.Tensor

    def run(self, s):
        lib.vt_gemm_skinny(ptr(self.descs), self.batch, self.max_M, self.max_K, self.max_Nc, s)

    def args(self):
        return ptr(self.descs), self.max_M, self.max_K

    def args5(self):
        return ptr(self.descs), self.batch, self.max_M, self.max_K, self.max_Nc


@dataclasses.dataclass
class BatchHandle:
    jobs: List[BlockJob]
    results: List[Optional[DecodeResult]]
    keep: list
    checks: list
    precisions: list = dataclasses.field(default_factory=list)
    # (job indices, barrier workspace) of every persistent RNN launch: a nonzero abort word
    # means its iterates are invalid and the jobs are solved again by the launch form
    rnn_checks: list = dataclasses.field(default_factory=list)


@dataclasses.dataclass
class RnnPrep:
    """RNN jobs made ready for their iterations (BatchDecoder.prepare_rnn): per step count,
    the prologue groups [(job indices, prologue)] and the blocks' step sizes eta (joint
    order), with the buffers to keep; `done` is recorded after the last preparing launch."""
    jobs: List["BlockJob"]
    steps: Dict[int, tuple]
    done: Optional[torch.cuda.Event] = None


PRECISIONS = ("fp64", "mixed", "bf16x3", "fp16", "kpcg")


class BatchDecoder:
    """Decodes lists of BlockJobs, grouped by (temporal order Dt, precision).

    precision (solver form) per block: "fp64" (FP64 Gram + FP64 Cholesky), "mixed" (FP64
    Gram, FP32 Cholesky, FP64 refinement), "bf16x3" / "fp16" (low-precision Gram, FP32
    Cholesky, FP64 refinement), "kpcg" (no Gram: FP64 conjugate gradients through the
    exact Kronecker operator, preconditioned by (S^T C S + lam I) kron Tbar).  Blocks of
    the non-FP64 forms that fail the guard (relative residual of the FP64 system above
    `tol`, or a factorisation that lost positive definiteness) are redone together as
    one batched FP64 group; kpcg blocks first get one more CG solve of retry_iters
    iterations (0: none) that starts from their first iterate and solves for its
    correction (in finite precision the recursively updated CG residual drifts from the
    true one on ill-conditioned blocks; the correction solve restarts from the true
    residual).  The CG, RNN and power-iteration loops are replayed as HIP graphs of
    graph_chunk iterations (0: every kernel launched by the host; same results)."""

    def __init__(self, device="cuda", precision: str = "bf16x3", lam_rel: float = 1e-6,
                 refine_iters: int = 4, tol: float = 1e-10, verify: bool = True,
                 cg_iters: int = 32, cg_stop: float = 1e-12, retry_iters: Optional[int] = None,
                 graph_chunk: int = 25, preconditioner: str = "auto", prune: int = 0,
                 rnn_persistent: bool = False):
        if precision not in PRECISIONS:
            raise ValueError(f"batched precision must be one of {PRECISIONS}")
        self.device = device
        self.precision = precision
        self.lam_rel = lam_rel
        self.refine_iters = refine_iters
        self.tol = tol
        self.verify = verify
        self.cg_iters = cg_iters
        self.cg_stop = cg_stop
        # kpcg blocks that fail the guard are solved again with this many iterations
        # before the FP64 redo (CG that ran out of iterations on a harder system)
        self.retry_iters = 4 * cg_iters if retry_iters is None else retry_iters
        # CG, RNN and power loops: replay a HIP graph of this many iterations (0: launch each)
        self.graph_chunk = graph_chunk
        # KPCG preconditioner: "kron" (S^T C S + lam I) kron Tbar, "kron_sum"
        # (S^T C S) kron Tbar + lam I (exact in lam), "auto" (kron_sum for blocks decoded in a
        # period longer than their window)
        if preconditioner not in ("auto", "kron", "kron_sum"):
            raise ValueError("preconditioner must be auto, kron or kron_sum")
        self.preconditioner = preconditioner
        # prune > 0: a kpcg block whose residual decay predicts it cannot reach the stop
        # within cg_iters is stopped from iteration `prune` on (checked every 10), so it
        # reaches the FP64 redo without spending the rest of the iterations
        if prune and prune < 10:
            raise ValueError("prune must be 0 or at least 10")
        self.prune = int(prune)
        # RNN iterations in one persistent launch (vt_kron_rnn_persistent) where the batch
        # allows it (same iterates; measured no faster than the graph-replayed launch form on
        # the serving batches, so off by default); otherwise, and to redo an aborted
        # persistent launch, the launch form (vt_kron_rnn)
        self.rnn_persistent = bool(rnn_persistent)
        self.rnn_ring_bytes = 2 << 30          # ring slots for every iteration up to this size
        self.rnn_limit_s = 2.0                 # longest barrier wait of the persistent launch
        self._pre_used = None                  # the last KPCG solve's preconditioner
        self.last_stats: Dict[str, Any] = {}
        self.timers: Optional[Dict[str, list]] = None     # stage -> [(ev0, ev1, work)]

    def instrument(self, on: bool = True):
        """Record HIP events around the main stages (on the launch stream)."""
        self.timers = {} if on else None

    def _tic(self):
        if self.timers is None:
            return None
        e = torch.cuda.Event(enable_timing=True)
        e.record()
        return e

    def _toc(self, name, e0, work):
        """Close a stage: work is its algorithmic work, or a callable giving it once the
        stream has run (for work that depends on device state, such as CG iterations)."""
        if e0 is None:
            return
        e1 = torch.cuda.Event(enable_timing=True)
        e1.record()
        self.timers.setdefault(name, []).append((e0, e1, work))

    def decode(self, jobs: List[BlockJob], precisions=None) -> List[DecodeResult]:
        """Launch and finalize (synchronous on the current stream's results)."""
        return self.finalize(self.launch(jobs, precisions))

    def launch(self, jobs: List[BlockJob], precisions=None) -> "BatchHandle":
        """Enqueue every stage of every group on the current stream; no host sync.  The
        blocks of all "kpcg" groups share one joint CG solve."""
        if precisions is None:
            precisions = [self.precision] * len(jobs)
        h = BatchHandle(jobs, [None] * len(jobs), [], [])
        h.precisions = list(precisions)
        groups: Dict[tuple, List[int]] = {}
        for i, (jb, pr) in enumerate(zip(jobs, precisions)):
            if pr not in PRECISIONS:
                raise ValueError(f"unknown precision {pr}")
            if jb.method not in ("tikhonov", "rnn"):
                raise ValueError(f"unknown method {jb.method}")
            key = (jb.decoder.geom.Dt, "rnn", jb.rnn_steps) if jb.method == "rnn" else \
                (jb.decoder.geom.Dt, pr, 0)
            groups.setdefault(key, []).append(i)
        main = torch.cuda.current_stream()
        rnn_keys = [k for k in sorted(groups) if k[1] == "rnn"]
        other_keys = [k for k in sorted(groups) if k[1] != "rnn"]
        side = None
        if rnn_keys and other_keys:
            # The RNN groups (hundreds of small dependent launches) run on a side stream,
            # concurrently with the other groups' solves; the main stream waits for them
            # at the end, so the batch's completion covers both.
            side = self._side_stream(main)
            fork = torch.cuda.Event()
            fork.record(main)
            side.wait_event(fork)
            with torch.cuda.stream(side):
                self._launch_rnn(h, jobs, groups, rnn_keys)
        elif rnn_keys:
            self._launch_rnn(h, jobs, groups, rnn_keys)
        kp = []
        for key in other_keys:
            Dt, pr, _ = key
            idx = groups[key]
            g = self._prologue([jobs[i] for i in idx], Dt)
            if pr == "kpcg":
                kp.append((idx, g))
                continue
            X, nrm, info, skeep = self._solve_direct(g, pr)
            res, keep = self._synthesis(g, X, nrm, info, None, 0)
            for i, r in zip(idx, res):
                h.results[i] = r
            h.keep.append((g.keep, skeep, keep))
            h.checks.append((idx, (nrm, info, skeep[4]), pr))
        if kp:
            gs = [g for _, g in kp]
            X, nrm, info, st, boff, skeep = self._solve_kpcg(gs)
            all_idx = []
            for (idx, g), (b0, e0) in zip(kp, boff):
                res, keep = self._synthesis(g, X[e0:e0 + g.Dtot], nrm[b0:b0 + g.B], info[b0:b0 + g.B],
                                            st[4 * b0:4 * (b0 + g.B)], 0)
                for i, r in zip(idx, res):
                    h.results[i] = r
                h.keep.append((g.keep, keep))
                all_idx += idx
            h.keep.append(skeep)
            h.checks.append((all_idx, (nrm, info, None), "kpcg"))
        if side is not None:
            join = torch.cuda.Event()
            join.record(side)
            main.wait_event(join)
            for r in h.results:
                if r is not None:       # frames made on the side stream, used on main
                    r.frames.record_stream(main)
        return h

    def _side_stream(self, main) -> torch.cuda.Stream:
        ss = getattr(self, "_sides", None)
        if ss is None:
            ss = self._sides = {}
        key = main.cuda_stream
        if key not in ss:
            ss[key] = torch.cuda.Stream(device=main.device)     # default priority
        return ss[key]

    def _launch_rnn(self, h: "BatchHandle", jobs, groups, keys):
        """Prologue, step sizes, joint RNN iterations and synthesis of the RNN groups
        (current stream)."""
        self._iterate_prepared(h, [self._prepare_rnn_groups(jobs, groups, keys)])

    def prepare_rnn(self, jobs: List[BlockJob]) -> RnnPrep:
        """The first stage of RNN jobs, enqueued on the current stream with no host sync:
        intervals, moments and right-hand sides (the prologue) and the step sizes (50
        power steps).  launch_rnn runs their iterations later, on any stream, alone or
        joined with other prepared jobs; the results are those of launch() (every block's
        arithmetic is its own)."""
        groups: Dict[tuple, List[int]] = {}
        for i, jb in enumerate(jobs):
            if jb.method != "rnn":
                raise ValueError("prepare_rnn takes RNN jobs only")
            groups.setdefault((jb.decoder.geom.Dt, "rnn", jb.rnn_steps), []).append(i)
        prep = self._prepare_rnn_groups(jobs, groups, sorted(groups))
        prep.done = torch.cuda.Event()
        prep.done.record()
        return prep

    def launch_rnn(self, preps: List[RnnPrep]) -> "BatchHandle":
        """The iterations and synthesis of prepared RNN jobs on the current stream (after
        each preparation's event): one joint launch sequence per step count.  The handle's
        jobs are the preparations' jobs in order; finalize() retires it."""
        main = torch.cuda.current_stream()
        jobs = [jb for p in preps for jb in p.jobs]
        h = BatchHandle(jobs, [None] * len(jobs), [], [])
        h.precisions = [self.precision] * len(jobs)
        for p in preps:
            if p.done is not None:
                main.wait_event(p.done)
        self._iterate_prepared(h, preps)
        h.keep.append(preps)
        return h

    def _prepare_rnn_groups(self, jobs, groups, keys) -> RnnPrep:
        rn: Dict[int, list] = {}
        for key in keys:
            Dt, _, steps = key
            idx = groups[key]
            rn.setdefault(steps, []).append((idx, self._prologue([jobs[i] for i in idx], Dt)))
        out = {}
        for steps, parts in sorted(rn.items()):
            eta, keep = self._rnn_power([g for _, g in parts])
            out[steps] = (parts, eta, keep)
        return RnnPrep(list(jobs), out)

    def _iterate_prepared(self, h: "BatchHandle", preps: List[RnnPrep]):
        """Joint iterations per step count over the preparations' blocks (job indices of
        preparation k offset by the jobs of the ones before it), then each group's
        synthesis into h.results."""
        merged: Dict[int, tuple] = {}
        off = 0
        for p in preps:
            for steps, (parts, eta, keep) in p.steps.items():
                m = merged.setdefault(steps, ([], [], []))
                m[0].extend([([i + off for i in idx], g) for idx, g in parts])
                m[1].append(eta)
                m[2].append(keep)
            off += len(p.jobs)
        for steps, (parts, etas, pkeep) in sorted(merged.items()):
            gs = [g for _, g in parts]
            eta = torch.cat(etas) if len(etas) > 1 else etas[0]
            X, st, boff, skeep = self._rnn_iterate(gs, eta, steps)
            if skeep and isinstance(skeep[-1], _RnnBarrier):
                h.rnn_checks.append(([i for idx, _ in parts for i in idx], skeep[-1].bar))
            for (idx, g), (b0, e0) in zip(parts, boff):
                zero = torch.zeros((g.B, 2), dtype=F64, device=self.device)
                info = torch.zeros(g.B, dtype=I32, device=self.device)
                res, keep = self._synthesis(g, X[e0:e0 + g.Dtot], zero, info, st[4 * b0:4 * (b0 + g.B)], 0)
                for i, r in zip(idx, res):
                    h.results[i] = r
                h.keep.append((g.keep, keep, zero, info))
            h.keep.append((skeep, pkeep))

    def _guard(self, checks):
        """(job index, precision) of every non-FP64 result that failed the guard (factor
        lost positive definiteness, FP64 relative residual above tol, or a triangular
        solve whose bounded wait expired), and the number of FP64 factorisations that
        failed.  An expired wait in an FP64 group raises: its results are invalid."""
        redo, failed_fp64 = [], 0
        for idx, (nrm, info, sync), pr in checks:
            hn = nrm.cpu().numpy()
            inf = info.cpu().numpy()
            expired = sync is not None and int(sync[1].item()) != 0
            if expired and pr == "fp64":
                raise VtemError("a triangular solve's bounded wait expired in the FP64 path: "
                                "its solutions are invalid")
            for k, i in enumerate(idx):
                if pr == "fp64":
                    failed_fp64 += int(inf[k] != 0)
                    continue
                ok = not expired and inf[k] == 0 and np.isfinite(hn[k]).all() and \
                    math.sqrt(hn[k, -1]) <= self.tol * math.sqrt(hn[k, 0])
                if not ok:
                    redo.append((i, pr))
        return redo, failed_fp64

    def finalize(self, h: "BatchHandle") -> List[DecodeResult]:
        """Guard: kpcg blocks that did not reach `tol` get a correction solve of retry_iters
        CG iterations from their iterate; blocks still failing, and low-precision blocks whose factor failed or
        whose refinement did not reach `tol`, are redone as one batched FP64 group
        (flagged)."""
        redo, failed_fp64, retried = [], 0, []
        aborted = [i for idx, bar in h.rnn_checks if int(bar[0].item()) != 0 for i in idx]
        if aborted:
            # a persistent RNN launch whose grid was not co-resident: the launch form
            # computes the same iterates
            base = self.rnn_persistent
            self.rnn_persistent = False
            try:
                h3 = self.launch([h.jobs[i] for i in aborted], [h.precisions[i] for i in aborted])
            finally:
                self.rnn_persistent = base
            for i, rr in zip(aborted, h3.results):
                h.results[i] = rr
            h.keep.append(h3.keep)
        if not self.verify:
            for _, (_, _, sync), _ in h.checks:
                if sync is not None and int(sync[1].item()) != 0:
                    raise VtemError("a triangular solve's bounded wait expired: its solutions are invalid")
        if self.verify:
            redo, failed_fp64 = self._guard(h.checks)
            kp = [i for i, pr in redo if pr == "kpcg"]
            if kp and self.retry_iters > 0:
                base = self.cg_iters
                self.cg_iters = self.retry_iters
                try:
                    h1 = self.launch([dataclasses.replace(h.jobs[i], x0=h.results[i].X) for i in kp],
                                     ["kpcg"] * len(kp))
                finally:
                    self.cg_iters = base
                bad, _ = self._guard(h1.checks)
                bad = {j for j, _ in bad}
                for j, (i, rr) in enumerate(zip(kp, h1.results)):
                    if j not in bad:
                        h.results[i] = rr
                h.keep.append(h1.keep)
                retried = kp
                redo = [(i, pr) for i, pr in redo if pr != "kpcg"] + [(kp[j], "kpcg") for j in sorted(bad)]
        redo = [i for i, _ in redo]
        h2 = None
        if redo:
            h2 = self.launch([h.jobs[i] for i in redo], ["fp64"] * len(redo))
            for i, rr in zip(redo, h2.results):
                rr.fallback = True
                h.results[i] = rr
            h.keep.append(h2.keep)
        self.last_stats = {"groups": len(h.checks), "fallbacks": len(redo), "rnn_aborted": len(aborted),
                           "fp64_not_pd": failed_fp64, "fallback_index": redo,
                           "kpcg_retries": len(retried), "preconditioner": self._pre_used}
        torch.cuda.current_stream().synchronize()
        if h2 is not None:
            for _, (_, _, sync), _ in h2.checks:
                if sync is not None and int(sync[1].item()) != 0:
                    raise VtemError("a triangular solve's bounded wait expired in the FP64 redo: "
                                    "its solutions are invalid")
        h.keep = []
        return h.results

    # ----------------------------------------------------------------------------------
    def _prologue(self, jobs: List[BlockJob], Dt: int) -> _Prologue:
        """Intervals, q, per-neuron temporal moments T_j and t_j, r = S^T t and lam for a
        group of blocks of temporal order Dt (global neuron numbering)."""
        dev = self.device
        s = stream_ptr()
        B = len(jobs)
        decs = [j.decoder for j in jobs]
        N = np.array([d.N for d in decs], np.int64)
        Ds = np.array([d.geom.Ds for d in decs], np.int64)
        D = Ds * Dt
        nbase = np.concatenate([[0], np.cumsum(N)])
        dbase = np.concatenate([[0], np.cumsum(D)])
        Ntot, Dtot = int(nbase[-1]), int(dbase[-1])
        kappa = torch.cat([d.kappa for d in decs])
        delta = torch.cat([d.delta for d in decs])
        bias = torch.cat([d.bias for d in decs])
        norm = torch.cat([d.norm for d in decs])
        s2 = torch.cat([d.s2 for d in decs])
        job_of = _h2d(np.repeat(np.arange(B, dtype=np.int32), N), dev)
        e_pro = self._tic()
        # ---- intervals
        sd = np.zeros(B, SPIKE_DTYPE)
        cap = 0
        for b, jb in enumerate(jobs):
            st = jb.spikes
            sd[b] = (st.spikes.data_ptr(), st.counts.data_ptr(), st.spikes.stride(0),
                     st.spikes.shape[0], nbase[b], jb.t0, jb.decoder.geom.duration)
            cap += st.spikes.shape[0] * int(N[b])
        sd_d = _to_dev(sd, dev)
        ws = torch.empty(2 * Ntot, dtype=I32, device=dev)
        offsets = torch.empty(Ntot + 1, dtype=I32, device=dev)
        tk1 = torch.empty(cap, dtype=F64, device=dev)
        tk2 = torch.empty(cap, dtype=F64, device=dev)
        nid = torch.empty(cap, dtype=I32, device=dev)
        lib.vt_intervals_batched(ptr(sd_d), ptr(job_of), Ntot, ptr(ws), ptr(offsets), ptr(tk1),
                                 ptr(tk2), ptr(nid), cap, s)
        base_d = _h2d(nbase.astype(np.int32), dev)
        n_b = torch.empty(B, dtype=I32, device=dev)
        lib.vt_seg_counts(ptr(offsets), ptr(base_d), B, ptr(n_b), s)
        # ---- q, T_j, t_j
        q = torch.empty(cap, dtype=F64, device=dev)
        lib.vt_q(ptr(q), ptr(tk1), ptr(tk2), ptr(nid), ptr(kappa), ptr(delta), ptr(bias), ptr(norm),
                 ptr(offsets[Ntot:]), cap, s)
        Tm = torch.empty((Ntot, Dt, Dt), dtype=F64, device=dev)
        tv = torch.empty((Ntot, Dt), dtype=F64, device=dev)
        if Dt > 17:
            raise ValueError(f"temporal order Lt = {(Dt - 1) // 2} above 8: the T_j kernel holds "
                             "at most 17 x 17 moments per neuron")
        tj = np.zeros(B, TJ_DTYPE)
        for b, d in enumerate(decs):
            tj[b] = (nbase[b], N[b], 0, d.geom.Lt, 0, d.geom.W, nbase[b], 0)
        tj_d = _to_dev(tj, dev)
        lib.vt_neuron_T_batched(ptr(Tm), ptr(tv), ptr(tk1), ptr(tk2), ptr(q), ptr(norm),
                                ptr(offsets), ptr(tj_d), B, int(N.max()), s)
        # ---- rhs, lambda
        r = torch.empty(Dtot, dtype=F64, device=dev)
        tn_ws = torch.empty(int(16 * Dt * Ds.sum()), dtype=F64, device=dev)
        ws_base = np.concatenate([[0], np.cumsum(16 * Dt * Ds)])
        tn_r = np.zeros(B, TN_DTYPE)
        for b, d in enumerate(decs):
            tn_r[b] = (d.S.data_ptr(), tv.data_ptr() + 8 * int(nbase[b]) * Dt,
                       r.data_ptr() + 8 * int(dbase[b]), tn_ws.data_ptr() + 8 * int(ws_base[b]),
                       d.S.stride(0), Dt, Dt, Ds[b], N[b])
        tn_r_d = _to_dev(tn_r, dev)
        lib.vt_gemm_tn_batched(ptr(tn_r_d), B, int(Ds.max()), Dt, s)
        lam = torch.empty(2 * B, dtype=F64, device=dev)
        s2off = _h2d(nbase[:-1].astype(np.int64), dev)
        N_d = _h2d(N.astype(np.int32), dev)
        lib.vt_kron_trace_batched(ptr(s2), ptr(Tm), ptr(N_d), ptr(s2off), ptr(s2off), Dt, ptr(n_b), 1,
                                  self.lam_rel, ptr(lam), B, s)
        self._toc("prologue", e_pro, 0.0)
        seg = _h2d(dbase.astype(np.int64), dev)
        elem_job = _h2d(np.repeat(np.arange(B, dtype=np.int32), D), dev)
        keep = (N_d, sd_d, ws, offsets, tk1, tk2, nid, base_d, q, tv, tn_r_d, s2off, tj_d, kappa, delta,
                bias, norm, s2, job_of)
        return _Prologue(jobs, decs, Dt, B, N, Ds, D, nbase, dbase, Ntot, Dtot, Tm, r, lam, n_b, tn_ws,
                         ws_base, seg, elem_job, base_d, keep)

    def _synthesis(self, g: _Prologue, X, nrm, info, st, _unused):
        """frames_b = Theta_b X_b^T B_s,b^T for every block; DecodeResults of the group."""
        dev = self.device
        s = stream_ptr()
        B, decs, Ds, Dt, dbase = g.B, g.decs, g.Ds, g.Dt, g.dbase
        T_out = np.array([d.theta_out.shape[0] for d in decs], np.int64)
        P = np.array([d.geom.P for d in decs], np.int64)
        Y2 = torch.empty(int((T_out * Ds).sum()), dtype=F64, device=dev)
        y2base = np.concatenate([[0], np.cumsum(T_out * Ds)])
        frames = [torch.empty((int(T_out[b]), int(P[b])), dtype=F64, device=dev) for b in range(B)]
        gs1 = np.zeros(B, GEMM_DTYPE)
        gs2 = np.zeros(B, GEMM_DTYPE)
        for b, d in enumerate(decs):
            y2p = Y2.data_ptr() + 8 * int(y2base[b])
            gs1[b] = (d.theta_out.data_ptr(), X.data_ptr() + 8 * int(dbase[b]), y2p, Dt, 1, 1, Dt, Ds[b],
                      T_out[b], Ds[b], Dt, 0, 1.0, 0.0)
            gs2[b] = (y2p, d.Bs.data_ptr(), frames[b].data_ptr(), Ds[b], 1, 1, Ds[b], P[b],
                      T_out[b], P[b], Ds[b], 0, 1.0, 0.0)
        gs1_d, gs2_d = _to_dev(gs1, dev), _to_dev(gs2, dev)
        e_syn = self._tic()                     # the two launches only (host work before)
        lib.vt_gemm_grouped(ptr(gs1_d), B, int(T_out.max()), int(Ds.max()), s)
        lib.vt_gemm_grouped(ptr(gs2_d), B, int(T_out.max()), int(P.max()), s)
        # algorithmic FP64 flops of the two products (Theta X^T, then times B_s^T)
        self._toc("synthesis", e_syn, float(np.sum(2.0 * T_out * Ds * (Dt + P))))
        results = []
        for b in range(B):
            Xb = X[int(dbase[b]):int(dbase[b + 1])]
            st_b = None if st is None else st[4 * b:4 * b + 4]
            results.append(DecodeResult(frames[b], Xb, g.lam[2 * b:2 * b + 2], g.n_b[b:b + 1],
                                        info[b:b + 1], nrm[b], False, st_b))
        return results, (Y2, gs1_d, gs2_d, X)

    def _solve_direct(self, g: "_Prologue", prec: str):
        """Gram (FP64 or low precision) + Cholesky (FP64 or FP32) + FP64 refinement."""
        dev = self.device
        s = stream_ptr()
        B, decs, N, Ds, D, Dt = g.B, g.decs, g.N, g.Ds, g.D, g.Dt
        nbase, dbase, Ntot, Dtot, Tm, r, lam = g.nbase, g.dbase, g.Ntot, g.Dtot, g.Tm, g.r, g.lam
        tn_ws, ws_base, seg, elem_job = g.tn_ws, g.ws_base, g.seg, g.elem_job
        # ---- Gram -------------------------------------------------------------------------
        maxD = int(D.max())
        nblk = (maxD + 63) // 64
        gflops = float(np.sum(N * D * (D + 1)))
        low = prec in ("bf16x3", "fp16")
        f64_factor = prec == "fp64"
        fdt = F64 if f64_factor else torch.float32
        Apool = torch.empty(int((D * D).sum()), dtype=fdt, device=dev)
        abase = np.concatenate([[0], np.cumsum(D * D)])
        As = [Apool[int(abase[b]):int(abase[b + 1])].view(int(D[b]), int(D[b])) for b in range(B)]
        extra_keep = []
        if low:
            mode = 0 if prec == "bf16x3" else 1
            # per block: tiled 16-bit T (and, for FP16, the block's own power-of-two scales,
            # so a block's result does not depend on the batch)
            tsz = np.array([split_T_size(int(N[b]), Dt) for b in range(B)], np.int64)
            tbase = np.concatenate([[0], np.cumsum(tsz)])
            Th = torch.empty(int(tbase[-1]), dtype=torch.int16, device=dev)
            Tl = torch.empty(int(tbase[-1]) if mode == 0 else 1, dtype=torch.int16, device=dev)
            scales = torch.ones(5 * B, dtype=torch.float32, device=dev)
            E2 = Dt * Dt
            for b, d in enumerate(decs):
                lib.vt_gram_split_T(ptr(Tm) + 8 * int(nbase[b]) * E2, int(N[b]), Dt, ptr(Th) + 2 * int(tbase[b]),
                                    ptr(Tl) + (2 * int(tbase[b]) if mode == 0 else 0), mode, ptr(d.S),
                                    d.S.numel(), ptr(scales) + 20 * b, s)
            gd = np.zeros(B, GRAM_LP_DTYPE)
            for b, d in enumerate(decs):
                off = 2 * int(tbase[b])
                gd[b] = (d.S.data_ptr(), Th.data_ptr() + off, Tl.data_ptr() + (off if mode == 0 else 0),
                         As[b].data_ptr(), d.S.stride(0), D[b], N[b], Ds[b], Dt, 0,
                         scales.data_ptr() + 20 * b)
            gd_d = _to_dev(gd, dev)
            e0 = self._tic()
            gram_tuner().launch(lambda tile: lib.vt_gram_lp(ptr(gd_d), B, int(Ds.max()), Dt, mode, tile, 0, 0, s),
                                prec, Dt, int(Ds.max()), B, int(N.max()))
            self._toc("gram_lp", e0, gflops)
            extra_keep += [Th, Tl, scales, gd_d]
        else:
            # FP64 Gram (into the factor buffer for fp64, into a staging buffer for mixed)
            G64 = Apool if f64_factor else torch.empty(int((D * D).sum()), dtype=F64, device=dev)
            Gs = [G64[int(abase[b]):int(abase[b + 1])].view(int(D[b]), int(D[b])) for b in range(B)]
            g64 = np.zeros(B, GRAM64_DTYPE)
            for b, d in enumerate(decs):
                g64[b] = (d.S.data_ptr(), Tm.data_ptr() + 8 * int(nbase[b]) * Dt * Dt,
                          Gs[b].data_ptr(), d.S.stride(0), D[b], N[b], Ds[b], Dt, 0)
            g64_d = _to_dev(g64, dev)
            e0 = self._tic()
            lib.vt_gram_f64_batched(ptr(g64_d), B, int(Ds.max()), Dt, s)
            self._toc("gram_f64", e0, gflops)
            extra_keep += [G64, g64_d]
            if prec == "mixed":
                for b in range(B):
                    lib.vt_lower_to_f32(ptr(As[b]), int(D[b]), ptr(Gs[b]), int(D[b]), int(D[b]),
                                        ptr(lam[2 * b:2 * b + 1]), s)
        if prec != "mixed":
            dd = np.zeros(B, DIAG_DTYPE)
            for b in range(B):
                dd[b] = (As[b].data_ptr(), D[b], D[b], 0)
            dd_d = _to_dev(dd, dev)
            lib.vt_diag_add_batched(ptr(dd_d), B, 0 if f64_factor else 1, ptr(lam), maxD, s)
            extra_keep.append(dd_d)
        # ---- Cholesky + solves -------------------------------------------------------------
        inv = torch.empty(B * nblk * 4096, dtype=fdt, device=dev)
        info = torch.zeros(B, dtype=I32, device=dev)
        mats = mat_descs([(As[b].data_ptr(), int(D[b]), int(D[b])) for b in range(B)], dev)
        potrf = lib.vt_potrf_f64 if f64_factor else lib.vt_potrf_f32
        potrs = lib.vt_potrs_f64 if f64_factor else lib.vt_potrs_f32
        e0 = self._tic()
        potrf(ptr(mats), B, maxD, ptr(inv), nblk * 4096, ptr(info), s)
        self._toc("potrf_f64" if f64_factor else "potrf_f32", e0,
                  float(np.sum(D.astype(np.float64) ** 3 / 3)))
        sync = torch.zeros(2 + 2 * B * nblk, dtype=I32, device=dev)
        # vectors in a padded (B x maxD) layout for the solves, concatenated elsewhere
        Xp = torch.zeros((B, maxD), dtype=F64, device=dev)
        pad_idx = _h2d(np.concatenate([b * maxD + np.arange(D[b]) for b in range(B)]), dev)

        def to_padded(vec, out):
            out.view(-1).index_copy_(0, pad_idx, vec)

        def from_padded(padded):
            return padded.view(-1).index_select(0, pad_idx)

        e_ref = self._tic()
        to_padded(r, Xp)
        potrs(ptr(mats), B, maxD, ptr(inv), nblk * 4096, ptr(Xp), maxD, ptr(sync), s)
        X = from_padded(Xp)
        iters = 0 if f64_factor else self.refine_iters
        # ---- refinement through the exact operators ----------------------------------------
        Y = torch.empty((Ntot, Dt), dtype=F64, device=dev)
        Z = torch.empty_like(Y)
        AX = torch.empty(Dtot, dtype=F64, device=dev)
        res = torch.empty(Dtot, dtype=F64, device=dev)
        nrm = torch.zeros((B, iters + 2), dtype=F64, device=dev)
        lib.vt_seg_sumsq(ptr(r), ptr(seg), B, ptr(nrm), iters + 2, s)
        g1 = np.zeros(B, GEMM_DTYPE)
        tn_a = np.zeros(B, TN_DTYPE)
        for b, d in enumerate(decs):
            g1[b] = (d.S.data_ptr(), X.data_ptr() + 8 * int(dbase[b]), Y.data_ptr() + 8 * int(nbase[b]) * Dt,
                     d.S.stride(0), 1, Dt, 1, Dt, N[b], Dt, Ds[b], 0, 1.0, 0.0)
            tn_a[b] = (d.S.data_ptr(), Z.data_ptr() + 8 * int(nbase[b]) * Dt,
                       AX.data_ptr() + 8 * int(dbase[b]), tn_ws.data_ptr() + 8 * int(ws_base[b]),
                       d.S.stride(0), Dt, Dt, Ds[b], N[b])
        g1_d, tn_a_d = _to_dev(g1, dev), _to_dev(tn_a, dev)
        resp = torch.zeros((B, maxD), dtype=F64, device=dev)
        for it in range(iters + 1):
            lib.vt_gemm_grouped(ptr(g1_d), B, int(N.max()), Dt, s)
            lib.vt_tj_apply(ptr(Z), ptr(Tm), ptr(Y), Ntot, Dt, s)
            lib.vt_gemm_tn_batched(ptr(tn_a_d), B, int(Ds.max()), Dt, s)
            lib.vt_residual_batched(ptr(res), ptr(r), ptr(AX), ptr(X), ptr(lam), ptr(elem_job), Dtot, s)
            nv = nrm[:, it + 1:]
            lib.vt_seg_sumsq(ptr(res), ptr(seg), B, ptr(nv), iters + 2, s)
            if it == iters:
                break
            to_padded(res, resp)
            potrs(ptr(mats), B, maxD, ptr(inv), nblk * 4096, ptr(resp), maxD, ptr(sync), s)
            lib.vt_add_inplace(ptr(X), ptr(from_padded(resp)), Dtot, s)
        self._toc("solve_refine", e_ref, float(np.sum(D.astype(np.float64) ** 2)) * 2 * (iters + 1))
        return X, nrm, info, (extra_keep, Apool, inv, mats, sync, Xp, pad_idx, Y, Z, AX, res, g1_d,
                              tn_a_d, resp)

    def _skinny(self, rows):
        """Skinny grouped product: rows of (A, B, C, T, active, sam, sak, ldb, ldc, tstride,
        M, K, Nc[, layout]) with pointers as ints (T = 0: no row factor; active = 0: always
        run; layout: the SK_*_TILED bits, default 0).  Returns the launch record
        (descriptors, max_M, max_K, max_Nc, split-K workspace)."""
        rows = [tuple(r) + (0,) * (14 - len(r)) for r in rows]
        R = np.array(rows, dtype=np.int64).reshape(len(rows), 14)
        wsz = skinny_workspace_array(R[:, 10], R[:, 11], R[:, 12])
        ws = torch.empty(max(1, int(wsz.sum())), dtype=F64, device=self.device)
        woff = np.concatenate([[0], np.cumsum(wsz)[:-1]])
        d = np.zeros(len(rows), SK_DTYPE)
        for j, name in enumerate(("A", "B", "C", "T")):
            d[name] = R[:, j].astype(np.uint64)
        d["ws"] = np.where(wsz > 0, ws.data_ptr() + 8 * woff, 0).astype(np.uint64)
        d["active"] = R[:, 4].astype(np.uint64)
        for j, name in zip(range(5, 14), ("sam", "sak", "ldb", "ldc", "tstride", "M", "K", "Nc", "pad")):
            d[name] = R[:, j]
        return _Skinny(_to_dev(d, self.device), len(rows), int(R[:, 10].max()), int(R[:, 11].max()),
                       int(R[:, 12].max()), ws)

    def _operator(self, blocks, src, dst, Z, dbj, nzj, active=None, block_order=False, tiled=None):
        """dst_b = A_b src_b for every block (blocks: (S, Tm_ptr, N, Ds, Dt, decoder)) as two
        skinny launches: Z_j = T_j (S src)_j (T fused in the epilogue), dst = S^T Z -- 4 N D
        flops, S read twice, no Gram.  Both products read the decoder's resident S and S^T
        in the MFMA tile layout.  With `active` (per-block flag pointers) converged blocks
        skip both products.  block_order: the forward descriptors in block order too (as the
        persistent RNN launch indexes them).  tiled = (xt, xt offsets, zt, zt offsets): src
        is read from its tiled copy xt (kept by the RNN update) and Z is written and read
        tiled in zt (zeroed by the caller), each block's at its offset (doubles)."""
        act = active or [0] * len(blocks)
        # forward descriptors largest-first (their order changes no result; the backward
        # descriptors stay in block order, which the fused RNN update indexes)
        order = list(range(len(blocks))) if block_order else \
            sorted(range(len(blocks)), key=lambda b: -blocks[b][2] * blocks[b][3])
        rows = []
        for b in order:
            S, Tp, N, Ds, Dt, dec = blocks[b][:6]
            if tiled is None:
                B, ldb, C, ldc, lay = src.data_ptr() + 8 * int(dbj[b]), Dt, Z.data_ptr() + 8 * int(nzj[b]), Dt, 0
            else:
                xt, xoff, zt, zoff = tiled
                B, ldb = xt.data_ptr() + 8 * int(xoff[b]), -(-Ds // 16)
                C, ldc = zt.data_ptr() + 8 * int(zoff[b]), -(-N // 16)
                lay = SK_B_TILED | SK_C_TILED
            rows.append((dec.S_tiled.data_ptr(), B, C, Tp, act[b], -(-Ds // 16), 1, ldb, ldc, Dt * Dt, N, Ds, Dt,
                         SK_A_TILED | lay))
        fwd = self._skinny(rows)
        rows = []
        for b, blk in enumerate(blocks):
            S, Tp, N, Ds, Dt, dec = blk[:6]
            if tiled is None:
                B, ldb, lay = Z.data_ptr() + 8 * int(nzj[b]), Dt, 0
            else:
                B, ldb, lay = tiled[2].data_ptr() + 8 * int(tiled[3][b]), -(-N // 16), SK_B_TILED
            rows.append((dec.St_tiled.data_ptr(), B, dst.data_ptr() + 8 * int(dbj[b]), 0, act[b],
                         -(-N // 16), 1, ldb, Dt, 0, Ds, N, Dt, SK_A_TILED | lay))
        bwd = self._skinny(rows)
        return fwd, bwd

    def _mgram(self, Ss, N, Ds, c_ptrs, s):
        """K_b = S_b^T diag(c_b) S_b (Ds x Ds, FP64) for every block into one pool.  Returns
        (pool, offsets, keep)."""
        dev = self.device
        B = len(Ss)
        mbase = np.concatenate([[0], np.cumsum(Ds * Ds)])
        Kpool = torch.empty(int(mbase[-1]), dtype=F64, device=dev)
        kd = np.zeros(B, KPM_DTYPE)
        for b, S in enumerate(Ss):
            kd[b] = (S.data_ptr(), c_ptrs[b], Kpool.data_ptr() + 8 * int(mbase[b]), S.stride(0), Ds[b], N[b], Ds[b])
        kd_d = _to_dev(kd, dev)
        e0 = self._tic()
        lib.vt_kpcg_mgram(ptr(kd_d), B, int(Ds.max()), s)
        self._toc("kpcg_mgram", e0, float(np.sum(2.0 * N * Ds * Ds)))
        return Kpool, mbase, [kd_d]

    def _spd_inverse(self, pool, mbase, dims, s):
        """Every matrix of the pool (matrix i: dims[i] x dims[i] at pool + mbase[i],
        symmetric positive definite, lower triangle read) replaced by its full inverse, FP64:
        Cholesky, then L^-1 by blocked forward substitution (64-row block steps, two grouped
        GEMMs each, diagonal blocks from the factor's block inverses), then L^-T L^-1.
        Returns (info, keep)."""
        dev = self.device
        B = len(dims)
        dims = np.asarray(dims, np.int64)
        maxDs = int(dims.max())
        nblk = (maxDs + 63) // 64
        base = pool.data_ptr() + 8 * np.asarray(mbase[:-1], np.int64)
        inv = torch.empty(B * nblk * 4096, dtype=F64, device=dev)
        info = torch.zeros(B, dtype=I32, device=dev)
        md = np.zeros(B, MAT_DTYPE)
        md["A"], md["lda"], md["D"] = base.astype(np.uint64), dims, dims
        mats = _to_dev(md, dev)
        e0 = self._tic()
        lib.vt_potrf_f64(ptr(mats), B, maxDs, ptr(inv), nblk * 4096, ptr(info), s)
        self._toc("kpcg_potrf", e0, float(np.sum(dims.astype(np.float64) ** 3 / 3)))
        e0 = self._tic()
        # L^-1 (lower triangular, zero above the diagonal)
        Lpool = torch.zeros(int(mbase[-1]), dtype=F64, device=dev)
        xoff = _h2d(np.asarray(mbase[:-1], np.int64), dev)
        dim = _h2d(dims.astype(np.int32), dev)
        lib.vt_kpcg_place_diag(ptr(Lpool), ptr(xoff), ptr(dim), B, maxDs, ptr(inv), nblk * 4096, s)
        Tb = torch.empty(B * 64 * maxDs, dtype=F64, device=dev)
        keep = [inv, mats, Lpool, xoff, dim, Tb]
        lbase = Lpool.data_ptr() + 8 * np.asarray(mbase[:-1], np.int64)
        tp = Tb.data_ptr() + 8 * np.arange(B, dtype=np.int64) * 64 * maxDs
        for i in range(1, nblk):
            rows = np.clip(dims - 64 * i, 0, 64)
            # T = L[i, :i] Linv[:i, :i];  Linv[i, :i] = -Dinv_i T
            g1 = np.zeros(B, GEMM_DTYPE)
            g2 = np.zeros(B, GEMM_DTYPE)
            _fill_gemm(g1, base + 8 * 64 * i * dims, lbase, tp, dims, 1, dims, 1, maxDs, rows, 64 * i, 64 * i,
                       1.0)
            # K = rows: T's rows beyond the block's own were never written
            _fill_gemm(g2, inv.data_ptr() + 8 * (np.arange(B, dtype=np.int64) * nblk + i) * 4096, tp,
                       lbase + 8 * 64 * i * dims, 64, 1, maxDs, 1, dims, rows, 64 * i, rows, -1.0)
            g1_d, g2_d = _to_dev(g1, dev), _to_dev(g2, dev)
            lib.vt_gemm_grouped(ptr(g1_d), B, 64, 64 * i, s)
            lib.vt_gemm_grouped(ptr(g2_d), B, 64, 64 * i, s)
            keep += [g1_d, g2_d]
        # inverse = Linv^T Linv (into the pool: the factor is no longer needed)
        g3 = np.zeros(B, GEMM_DTYPE)
        _fill_gemm(g3, lbase, lbase, base, 1, dims, dims, 1, dims, dims, dims, dims, 1.0)
        g3_d = _to_dev(g3, dev)
        lib.vt_gemm_grouped(ptr(g3_d), B, maxDs, maxDs, s)
        self._toc("kpcg_inverse", e0, float(np.sum(dims.astype(np.float64) ** 3)))
        keep.append(g3_d)
        return info, keep

    def _spatial_inverse(self, Ss, N, Ds, c_ptrs, lam, s):
        """Mi_b = (S_b^T diag(c_b) S_b + lam_b I)^-1 for every block, FP64."""
        B = len(Ss)
        Mpool, mbase, keep = self._mgram(Ss, N, Ds, c_ptrs, s)
        dd = np.zeros(B, DIAG_DTYPE)
        for b in range(B):
            dd[b] = (Mpool.data_ptr() + 8 * int(mbase[b]), Ds[b], Ds[b], 0)
        dd_d = _to_dev(dd, self.device)
        lib.vt_diag_add_batched(ptr(dd_d), B, 0, ptr(lam), int(Ds.max()), s)
        info, ikeep = self._spd_inverse(Mpool, mbase, Ds, s)
        return Mpool, mbase, info, keep + [dd_d] + ikeep

    def _solve_kpcg(self, gs: List[_Prologue]):
        """Kronecker-preconditioned CG on (A + lam I) X = r for every block of the groups,
        as one joint solve (one launch sequence whatever the blocks' temporal orders).

        Setup: c_j = tr(T_j)/Dt, Tbar (trace Dt) and its inverse Ti per group; Mi =
        (S^T diag(c) S + lam I)^-1 in FP64 per block.  Each iteration applies A through its
        structure (two skinny products) and the preconditioner as (Mi R) Ti (one skinny
        product with a shared right factor).  A block stops once its recursive residual
        falls below cg_stop relative (its products are skipped from then on); the guard
        then checks the true FP64 residual.  Returns the joint X, residual norms, factor
        info, CG state, per-group (block, element) offsets and the buffers to keep."""
        dev = self.device
        s = stream_ptr()
        e_setup = self._tic()
        kron_sum = self._kron_sum(gs)
        self._pre_used = "kron_sum" if kron_sum else "kron"
        blocks, c_ptrs, TiT_ptrs, keep = [], [], [], []
        eig = []                                 # Kronecker sum: per group (theta, U, UT)
        boff, b0, e0 = [], 0, 0
        for g in gs:
            E = g.Dt * g.Dt
            c = torch.empty(g.Ntot, dtype=F64, device=dev)
            tws = torch.empty(16 * g.B * E, dtype=F64, device=dev)
            if kron_sum:
                th = torch.empty(g.B * g.Dt, dtype=F64, device=dev)
                U = torch.empty(g.B * E, dtype=F64, device=dev)
                UT = torch.empty(g.B * E, dtype=F64, device=dev)
                lib.vt_kpcg_tbar_eig(ptr(g.Tm), ptr(g.nbase_d), g.Dt, g.B, ptr(tws), ptr(c), ptr(th), ptr(U),
                                     ptr(UT), s)
                eig.append((th, U, UT))
                keep += [c, tws, th, U, UT]
            else:
                TiT = torch.empty(g.B * E, dtype=F64, device=dev)
                lib.vt_kpcg_tbar(ptr(g.Tm), ptr(g.nbase_d), g.Dt, g.B, ptr(tws), ptr(c), ptr(TiT), s)
                keep += [c, TiT, tws]
            for b, d in enumerate(g.decs):
                blocks.append((d.S, g.Tm.data_ptr() + 8 * int(g.nbase[b]) * E, int(g.N[b]), int(g.Ds[b]), g.Dt,
                               d))
                c_ptrs.append(c.data_ptr() + 8 * int(g.nbase[b]))
                if not kron_sum:
                    TiT_ptrs.append(TiT.data_ptr() + 8 * b * E)
            boff.append((b0, e0))
            b0 += g.B
            e0 += g.Dtot
        Bt = len(blocks)
        N = np.array([b[2] for b in blocks], np.int64)
        Ds = np.array([b[3] for b in blocks], np.int64)
        Dt = np.array([b[4] for b in blocks], np.int64)
        D = Ds * Dt
        dbj = np.concatenate([[0], np.cumsum(D)])
        nzj = np.concatenate([[0], np.cumsum(N * Dt)])
        Dtot = int(dbj[-1])
        r = torch.cat([g.r for g in gs]) if len(gs) > 1 else gs[0].r
        lam = torch.cat([g.lam for g in gs]) if len(gs) > 1 else gs[0].lam
        seg = _h2d(dbj.astype(np.int64), dev)
        elem_job = _h2d(np.repeat(np.arange(Bt, dtype=np.int32), D), dev)
        R = r.clone()
        Zp = torch.empty(Dtot, dtype=F64, device=dev)
        st = torch.zeros(4 * Bt, dtype=F64, device=dev)
        st.view(Bt, 4)[:, 2] = 1.0          # active for the first preconditioner application
        act = [st.data_ptr() + 8 * (4 * b + 2) for b in range(Bt)]
        if not kron_sum:
            Mi, mbase, info, mkeep = self._spatial_inverse([b[0] for b in blocks], N, Ds, c_ptrs, lam, s)
            pres = [self._skinny([(Mi.data_ptr() + 8 * int(mbase[b]), R.data_ptr() + 8 * int(dbj[b]),
                                   Zp.data_ptr() + 8 * int(dbj[b]), TiT_ptrs[b], act[b],
                                   Ds[b], 1, Dt[b], Dt[b], 0, Ds[b], Ds[b], Dt[b]) for b in range(Bt)])]
            setup_flops = np.sum(2.0 * N * Ds * Ds + 2.0 * Ds.astype(np.float64) ** 3)
        else:
            Mi, info, mkeep, pres = self._kron_sum_pre(gs, eig, blocks, N, Ds, Dt, dbj, c_ptrs, lam, R, Zp, act, s)
            setup_flops = np.sum(2.0 * N * Ds * Ds + 2.0 * Dt * Ds.astype(np.float64) ** 3)
        self._toc("kpcg_setup", e_setup, float(setup_flops))
        X = torch.empty(Dtot, dtype=F64, device=dev)
        Pv = torch.empty(Dtot, dtype=F64, device=dev)
        Q = torch.empty(Dtot, dtype=F64, device=dev)
        Z = torch.empty(int(nzj[-1]), dtype=F64, device=dev)
        # the CG direction and Z also in the MFMA tile layout (p's copy kept by the CG steps)
        tiles = self.operand_tiles(N, Ds, Dt)
        ptab = _h2d(np.stack([tiles[1][:-1], Dt]).T.reshape(-1).astype(np.int64), dev)
        fp, bp = self._operator(blocks, Pv, Q, tiles[2], dbj, nzj, act, tiled=tiles)
        fx, bx = self._operator(blocks, X, Q, Z, dbj, nzj)
        iters = int(self.cg_iters)
        hs = iters + 1
        hist = torch.zeros((Bt, hs), dtype=F64, device=dev)
        maxNc = int(Dt.max())
        x0s = [jb.x0 for g in gs for jb in g.jobs]
        X0 = None
        if any(x is not None for x in x0s):
            # warm start: solve (A + lam I) E = r - (A + lam I) x0, then X = x0 + E
            X0 = torch.zeros(Dtot, dtype=F64, device=dev)
            for b, x in enumerate(x0s):
                if x is not None:
                    if x.numel() != int(D[b]):
                        raise ValueError("x0 does not match its block's coefficient count")
                    X0[int(dbj[b]):int(dbj[b]) + int(D[b])].copy_(x)
            X.copy_(X0)
            fx.run(s)
            bx.run(s)
            lib.vt_residual_batched(ptr(R), ptr(r), ptr(Q), ptr(X), ptr(lam), ptr(elem_job), Dtot, s)
        e_cg = self._tic()                      # the solve's launches only (host work before)
        for pr in pres:
            pr.run(s)
        lib.vt_kpcg_init(ptr(R), ptr(Zp), ptr(Pv), ptr(X), ptr(seg), Bt, ptr(st), ptr(hist), hs, ptr(tiles[0]),
                         ptr(ptab), s)
        pargs = [pr.args5() for pr in pres] + [(0, 0, 0, 0, 0)] * (3 - len(pres))
        kcnt = torch.empty(Bt * (-(-int(Ds.max()) // 16) + 1), dtype=torch.int32, device=dev)
        lib.vt_kpcg_iterate(*fp.args(), *bp.args(), *pargs[0], *pargs[1], *pargs[2], Bt, maxNc, ptr(X), ptr(R),
                            ptr(Pv), ptr(Q), ptr(Zp), ptr(lam), ptr(seg), ptr(st), ptr(hist), hs, iters,
                            float(self.cg_stop) ** 2, self.prune, self.graph_chunk, ptr(kcnt), ptr(tiles[0]),
                            ptr(ptab), s)
        if X0 is not None:
            lib.vt_add_inplace(ptr(X), ptr(X0), Dtot, s)
        # true FP64 residual of the final iterate (the guard's input)
        nrm = torch.zeros((Bt, 2), dtype=F64, device=dev)
        lib.vt_seg_sumsq(ptr(r), ptr(seg), Bt, ptr(nrm), 2, s)
        fx.run(s)
        bx.run(s)
        res = torch.empty(Dtot, dtype=F64, device=dev)
        lib.vt_residual_batched(ptr(res), ptr(r), ptr(Q), ptr(X), ptr(lam), ptr(elem_job), Dtot, s)
        lib.vt_seg_sumsq(ptr(res), ptr(seg), Bt, ptr(nrm[:, 1:]), 2, s)
        # algorithmic HBM bytes per iteration and block: S twice, T_j once, Z written and
        # read, Mi once, about a dozen passes over the D-space vectors; a block that has
        # converged skips its products, so each block counts its own iterations (st[4b+3])
        per = 8.0 * (2.0 * N * Ds + N * Dt * Dt + 2.0 * N * Dt + (Dt if kron_sum else 1) * Ds * Ds + 12.0 * D)
        self._toc("kpcg_iterate", e_cg,
                  lambda: float(np.sum(per * st.view(Bt, 4)[:, 3].cpu().numpy())))
        keep += [r, lam, seg, elem_job, Mi, mkeep, R, Zp, pres, X, Pv, Q, Z, fp, bp, fx, bx, st, hist, res, X0, kcnt,
                 tiles, ptab]
        return X, nrm, info, st, boff, keep

    def _kron_sum(self, gs) -> bool:
        """The Kronecker-sum preconditioner for this solve?  "auto": when a block is
        decoded in a temporal period longer than its window (its Tbar is then badly
        conditioned and the Kronecker product misjudges lam)."""
        if self.preconditioner in ("kron", "kron_sum"):
            return self.preconditioner == "kron_sum"
        return any(d.geom.T_period > d.geom.duration * (1 + 1e-12) for g in gs for d in g.decs)

    def _kron_sum_pre(self, gs, eig, blocks, N, Ds, Dt, dbj, c_ptrs, lam, R, Zp, act, s):
        """P^-1 R = sum_k (theta_k K + lam I)^-1 (R u_k) u_k^T per block, Tbar = U diag(theta)
        U^T, K = S^T diag(c) S: the Dt shifted matrices of every block inverted in one batch,
        and three skinny launches per application (Rt = R U; column k of Zt = Mi_k column k
        of Rt; Z = Zt U^T).  Returns (the inverses' pool, factor info, keep, launches)."""
        dev = self.device
        Bt = len(blocks)
        Kpool, kb, kkeep = self._mgram([b[0] for b in blocks], N, Ds, c_ptrs, s)
        nmat = Ds * Ds * Dt                      # per block: Dt matrices of Ds^2
        obase = np.concatenate([[0], np.cumsum(nmat)])
        Opool = torch.empty(int(obase[-1]), dtype=F64, device=dev)
        keep = [Kpool, kkeep]
        b0 = 0
        for g, (th, U, UT) in zip(gs, eig):
            idx = slice(b0, b0 + g.B)
            kb_d = _h2d(np.asarray(kb[:-1][idx], np.int64), dev)
            ob_d = _h2d(np.asarray(obase[:-1][idx], np.int64), dev)
            ds_d = _h2d(Ds[idx].astype(np.int32), dev)
            lib.vt_kpcg_shift_pool(ptr(Kpool), ptr(kb_d), ptr(ob_d), ptr(ds_d), ptr(th),
                                   ptr(lam) + 16 * b0, g.Dt, g.B, int(Ds[idx].max()), ptr(Opool), s)
            keep += [kb_d, ob_d, ds_d]
            b0 += g.B
        mats = np.concatenate([obase[b] + np.arange(Dt[b], dtype=np.int64) * Ds[b] * Ds[b] for b in range(Bt)])
        dims = np.repeat(Ds, Dt)
        info_all, ikeep = self._spd_inverse(Opool, np.concatenate([mats, [obase[-1]]]), dims, s)
        keep += ikeep
        Rt = torch.empty_like(R)
        Zt = torch.empty_like(R)
        keep += [Rt, Zt]
        Uptr, UTptr = [], []
        for g, (th, U, UT) in zip(gs, eig):
            E = g.Dt * g.Dt
            Uptr += [U.data_ptr() + 8 * b * E for b in range(g.B)]
            UTptr += [UT.data_ptr() + 8 * b * E for b in range(g.B)]
        rp, rtp, ztp, zp = (R.data_ptr(), Rt.data_ptr(), Zt.data_ptr(), Zp.data_ptr())
        pa = self._skinny([(rp + 8 * int(dbj[b]), Uptr[b], rtp + 8 * int(dbj[b]), 0, act[b],
                            Dt[b], 1, Dt[b], Dt[b], 0, Ds[b], Dt[b], Dt[b]) for b in range(Bt)])
        pb = self._skinny([(Opool.data_ptr() + 8 * int(obase[b] + k * Ds[b] * Ds[b]), rtp + 8 * (int(dbj[b]) + k),
                            ztp + 8 * (int(dbj[b]) + k), 0, act[b], Ds[b], 1, Dt[b], Dt[b], 0, Ds[b], Ds[b], 1)
                           for b in range(Bt) for k in range(int(Dt[b]))])
        pc = self._skinny([(ztp + 8 * int(dbj[b]), UTptr[b], zp + 8 * int(dbj[b]), 0, act[b],
                            Dt[b], 1, Dt[b], Dt[b], 0, Ds[b], Dt[b], Dt[b]) for b in range(Bt)])
        # a block's factor info: the first failing of its Dt matrices
        owner = _h2d(np.repeat(np.arange(Bt, dtype=np.int64), Dt), dev)
        info = torch.zeros(Bt, dtype=I32, device=dev).scatter_reduce_(0, owner, info_all.view(-1), "amax")
        keep.append(owner)
        return Opool, info, keep, [pa, pb, pc]

    def _joint_blocks(self, gs: List[_Prologue]):
        """Blocks of several prologue groups in one joint numbering: (S, T_j pointer, N, Ds,
        Dt, S^T) per block."""
        blocks, boff, b0, e0 = [], [], 0, 0
        for g in gs:
            E = g.Dt * g.Dt
            for b, d in enumerate(g.decs):
                blocks.append((d.S, g.Tm.data_ptr() + 8 * int(g.nbase[b]) * E, int(g.N[b]), int(g.Ds[b]), g.Dt,
                               d))
            boff.append((b0, e0))
            b0 += g.B
            e0 += g.Dtot
        N = np.array([b[2] for b in blocks], np.int64)
        Ds = np.array([b[3] for b in blocks], np.int64)
        Dt = np.array([b[4] for b in blocks], np.int64)
        return blocks, boff, N, Ds, Dt

    def _rnn_power(self, gs: List[_Prologue]):
        """Step sizes eta_b = 1/theta_b of the groups' blocks (joint order), theta_b the
        largest Ritz value of A_b + lam_b I after LANCZOS_STEPS Lanczos steps through the
        structured operator (vt_kron_lanczos).  Returns (eta, keep)."""
        dev = self.device
        s = stream_ptr()
        e0 = self._tic()
        blocks, boff, N, Ds, Dt = self._joint_blocks(gs)
        Bt = len(blocks)
        D = Ds * Dt
        dbj = np.concatenate([[0], np.cumsum(D)])
        nzj = np.concatenate([[0], np.cumsum(N * Dt)])
        Dtot = int(dbj[-1])
        lam = torch.cat([g.lam for g in gs]) if len(gs) > 1 else gs[0].lam
        seg = _h2d(dbj.astype(np.int64), dev)
        V = torch.empty(Dtot, dtype=F64, device=dev)
        Vp = torch.empty(Dtot, dtype=F64, device=dev)
        Q = torch.empty(Dtot, dtype=F64, device=dev)
        # the Lanczos vector (kept tiled by the steps) and Z in the MFMA tile layout
        tiles = self.operand_tiles(N, Ds, Dt)
        Z = tiles[2]
        vtab = _h2d(np.stack([tiles[1][:-1], Dt]).T.reshape(-1).astype(np.int64), dev)
        fv, bv = self._operator(blocks, V, Q, Z, dbj, nzj, tiled=tiles)
        eta = torch.empty(Bt, dtype=F64, device=dev)
        ntile = -(-int(Ds.max()) // 16)
        cnt = torch.empty(Bt * ntile + 2 * Bt, dtype=torch.int32, device=dev)
        part = torch.empty(Bt * ntile, dtype=F64, device=dev)
        tri = torch.empty(2 * LANCZOS_STEPS * Bt, dtype=F64, device=dev)
        lib.vt_kron_lanczos(*fv.args(), *bv.args(), Bt, int(Dt.max()), ptr(V), ptr(Vp), ptr(Q), ptr(lam),
                            ptr(seg), LANCZOS_STEPS, ptr(eta), self.graph_chunk, ptr(cnt), ptr(part), ptr(tri),
                            ptr(tiles[0]), ptr(vtab), s)
        nbytes = 8.0 * np.sum(2.0 * N * Ds + N * Dt * Dt + 2.0 * N * Dt + 6.0 * D)
        self._toc("rnn_lanczos", e0, float(nbytes) * LANCZOS_STEPS)
        return eta, [lam, seg, V, Vp, Q, Z, fv, bv, eta, cnt, part, tri, tiles, vtab]

    def _rnn_iterate(self, gs: List[_Prologue], eta, steps: int):
        """Gradient-flow decoder for every block of the groups, as one joint launch
        sequence: steps + 1 iterations X <- X + eta_b (r - (A_b + lam_b I) X) from 0 through
        the structured operator, eta (joint order) from _rnn_power.  Returns the joint X,
        the per-block state [eta, 0, 0, iterations], the per-group offsets and the buffers
        to keep."""
        dev = self.device
        s = stream_ptr()
        e0 = self._tic()
        blocks, boff, N, Ds, Dt = self._joint_blocks(gs)
        Bt = len(blocks)
        D = Ds * Dt
        dbj = np.concatenate([[0], np.cumsum(D)])
        nzj = np.concatenate([[0], np.cumsum(N * Dt)])
        Dtot = int(dbj[-1])
        r = torch.cat([g.r for g in gs]) if len(gs) > 1 else gs[0].r
        lam = torch.cat([g.lam for g in gs]) if len(gs) > 1 else gs[0].lam
        seg = _h2d(dbj.astype(np.int64), dev)
        X = torch.empty(Dtot, dtype=F64, device=dev)
        Q = torch.empty(Dtot, dtype=F64, device=dev)
        st = torch.zeros(4 * Bt, dtype=F64, device=dev)
        maxNc = int(Dt.max())
        plan = self._rnn_plan(N, Ds, Dt) if self.rnn_persistent else None
        extra = []
        if plan is None:
            tiles = self.operand_tiles(N, Ds, Dt)
            Xt, xoff, Z, _ = tiles
            xoff_d = _h2d(xoff[:-1].astype(np.int64), dev)
            fx, bx = self._operator(blocks, X, Q, Z, dbj, nzj, tiled=tiles)
            cnt = torch.empty(Bt * -(-int(Ds.max()) // 16), dtype=torch.int32, device=dev)
            lib.vt_kron_rnn(*fx.args(), *bx.args(), Bt, maxNc, ptr(X), ptr(r), ptr(Q), ptr(lam), ptr(eta),
                            ptr(seg), Dtot, steps, self.graph_chunk, ptr(cnt), ptr(Xt), ptr(xoff_d),
                            int(xoff[-1]), s)
            extra = [cnt, Xt, xoff_d]
        else:
            Z = torch.empty(int(nzj[-1]), dtype=F64, device=dev)
            fx, bx = self._operator(blocks, X, Q, Z, dbj, nzj)
            gstart, grid = plan
            fpo, bpo = self._operator(blocks, X, Q, Z, dbj, nzj, block_order=True)
            gs_d = _h2d(gstart.astype(np.int32), dev)
            slot_x = -(-Dtot // 512) * 512
            slot_z = -(-int(nzj[-1]) // 512) * 512
            ring = steps + 1 if 8 * (steps + 1) * (slot_x + slot_z) <= self.rnn_ring_bytes else 32
            ring_x = torch.empty(ring * slot_x, dtype=F64, device=dev)
            ring_z = torch.empty(ring * slot_z, dtype=F64, device=dev)
            words = ctypes.c_longlong(0)
            lib.vt_kron_rnn_persistent_bar_words(Bt, ctypes.byref(words))
            bar = torch.empty(int(words.value), dtype=torch.int32, device=dev)
            lib.vt_kron_rnn_persistent(ptr(fpo.descs), ptr(bpo.descs), Bt, maxNc, ptr(gs_d), grid, ptr(X), Dtot,
                                       ptr(r), ptr(Z), ptr(lam), ptr(eta), ptr(seg), steps, ptr(ring_x), slot_x,
                                       ptr(ring_z), slot_z, ring, ptr(bar), float(self.rnn_limit_s), s)
            extra = [fpo, bpo, gs_d, ring_x, ring_z, _RnnBarrier(bar)]
        sv = st.view(Bt, 4)
        sv[:, 0] = eta
        sv[:, 3] = float(steps + 1)
        nbytes = 8.0 * np.sum(2.0 * N * Ds + N * Dt * Dt + 2.0 * N * Dt + 4.0 * D)
        self._toc("rnn_iterate", e0, float(nbytes) * (steps + 1))
        return X, st, boff, [r, lam, seg, X, Q, Z, fx, bx, st, eta] + extra

    def operand_tiles(self, N, Ds, Dt):
        """Buffers for the operator's B operands in the MFMA tile layout (the RNN's iterate
        or the CG direction, and Z): per block the tiles of X_b^T (Dt x Ds) and of Z_b^T
        (Dt x N).  Returns (Xt, X offsets, Zt, Z offsets) with the offsets in doubles (one
        more entry: the total); both zeroed, as their padding is never written."""
        xsz = np.array([tile_doubles(int(t), int(d)) for t, d in zip(Dt, Ds)], np.int64)
        zsz = np.array([tile_doubles(int(t), int(n)) for t, n in zip(Dt, N)], np.int64)
        xoff = np.concatenate([[0], np.cumsum(xsz)])
        zoff = np.concatenate([[0], np.cumsum(zsz)])
        Xt = torch.zeros(int(xoff[-1]), dtype=F64, device=self.device)
        Zt = torch.zeros(int(zoff[-1]), dtype=F64, device=self.device)
        return Xt, xoff, Zt, zoff

    def _rnn_plan(self, N, Ds, Dt):
        """Workgroup groups of the persistent RNN launch, or None where the launch form
        runs (forward products that split K, or more blocks than fit).  Each block's group
        gets one workgroup per tile of its larger phase (forward row tiles, or backward row
        tiles times K splits), all scaled down together when the grid would exceed what
        fits on the device at once.  Returns (gstart, grid)."""
        if np.any(Ds > 1008) or np.any(Dt > 32):
            return None
        cap, mg = _rnn_capacity()
        cap = cap[1 if int(Dt.max()) > 17 else 0]
        ks = skinny_splits_array(N)                           # the skinny product's split rule
        need = np.maximum(-(-N // 16), -(-Ds // 16) * ks).astype(np.int64)
        g = np.minimum(need, mg)
        if g.sum() > cap:
            g = np.maximum(1, np.floor(g * (cap / float(g.sum())))).astype(np.int64)
        if g.sum() > cap:
            return None
        gstart = np.concatenate([[0], np.cumsum(g)]).astype(np.int64)
        return gstart, int(gstart[-1])
