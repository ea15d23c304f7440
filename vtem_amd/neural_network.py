"""Recurrent (gradient-flow) decoder with the reference's API (vtem/neural_network.py).

rnn3(G, q, dt, alpha, steps, XOUTPUTSTEPS) runs steps+1 iterations of
c <- c + dt*alpha (q - G c) from c = 0 (vtem/neural_network.py:60-92, update kernel
:102-128) in vt_rnn_spike (a matrix-vector product and a fused update per iteration,
all iterations enqueued by one call).  With XOUTPUTSTEPS, XOUTPUTSTEPS snapshots of the
iterate are returned as well, as (c, output) like the reference.
"""

from __future__ import annotations

import numpy as np
import torch

from ._lib import lib, ptr, stream_ptr

F64 = torch.float64


def _dev(a, dtype=F64):
    if isinstance(a, torch.Tensor):
        return a.to(device="cuda", dtype=dtype).contiguous()
    return torch.as_tensor(np.ascontiguousarray(a), dtype=dtype, device="cuda")


def rnn3(G, q, dt=1e-6, alpha=5000, steps=4000, XOUTPUTSTEPS=None):
    Gd = _dev(G)
    qd = _dev(q).reshape(-1)
    n = qd.numel()
    if Gd.shape != (n, n):
        raise ValueError("G must be n x n with n = len(q)")
    c = torch.zeros(n, dtype=F64, device="cuda")
    w1 = torch.empty(n, dtype=F64, device="cuda")
    w2 = torch.empty(n, dtype=F64, device="cuda")
    eta = float(dt) * float(alpha)
    if not XOUTPUTSTEPS:
        lib.vt_rnn_spike(ptr(c), ptr(Gd), Gd.stride(0), ptr(qd), n, eta, int(steps), ptr(w1), ptr(w2),
                         stream_ptr())
        return c
    # snapshots: row k holds c after iteration (k+1)*x_steps, x_steps = steps // XOUTPUTSTEPS
    # (the reference also writes the iteration-0 iterate to the last row, which the final
    # snapshot then overwrites when steps is a multiple of XOUTPUTSTEPS)
    nout = int(XOUTPUTSTEPS)
    x_steps = max(1, int(steps) // nout)
    output = torch.empty((nout, n), dtype=F64, device="cuda")
    done = 0                                    # iterations run so far (indices 0..done-1)
    for k in range(nout):
        target = (k + 1) * x_steps              # run through iteration index `target`
        if target > int(steps):
            break
        cnt = target + 1 - done
        lib.vt_rnn_spike(ptr(c), ptr(Gd), Gd.stride(0), ptr(qd), n, eta, cnt - 1, ptr(w1), ptr(w2),
                         stream_ptr())
        done += cnt
        output[k].copy_(c)
    if done < int(steps) + 1:
        lib.vt_rnn_spike(ptr(c), ptr(Gd), Gd.stride(0), ptr(qd), n, eta, int(steps) - done, ptr(w1),
                         ptr(w2), stream_ptr())
    return c, output
