"""Linear algebra with the reference's API (vtem/utils/linalg.py): the truncated-eigen
pseudoinverse solve of the spike-space decoder.

solve_eq_sym(G, q, rcond) returns c = V diag(1/s_i if s_i > rcond s_0 else 0) V^T q for
symmetric G = V diag(w) V^T, s = |w| (the singular values of a symmetric matrix), as
vtem/utils/linalg.py:591-631 computes through CULA's SVD.  Here the eigendecomposition
is rocSOLVER's divide and conquer (vt_eig_sym), the filter is vt_eig_filter and the two
products with the eigenvector matrix are vt_gemm_f64.  G is overwritten (the reference
destroys G too, :598-599).
"""

from __future__ import annotations

import numpy as np
import torch

from .._lib import VtemError, lib, ptr, stream_ptr

F64 = torch.float64


def _dev(a, dtype=F64):
    if isinstance(a, torch.Tensor):
        return a.to(device="cuda", dtype=dtype).contiguous()
    return torch.as_tensor(np.ascontiguousarray(a), dtype=dtype, device="cuda")


def eig_sym(G):
    """(w ascending, V with eigenvectors in its rows); G is overwritten by V."""
    A = _dev(G) if not (isinstance(G, torch.Tensor) and G.is_cuda and G.dtype == F64 and G.is_contiguous()) else G
    n = A.shape[0]
    w = torch.empty(n, dtype=F64, device="cuda")
    E = torch.empty(n, dtype=F64, device="cuda")
    info = torch.zeros(1, dtype=torch.int32, device="cuda")
    lib.vt_eig_sym(ptr(A), A.stride(0), n, ptr(w), ptr(E), ptr(info), stream_ptr())
    if int(info.item()) != 0:
        raise VtemError(f"eigendecomposition did not converge (info {int(info.item())})")
    return w, A


def solve_eq_sym(G, q, rcond=1e-8):
    w, V = eig_sym(G)
    n = V.shape[0]
    qd = _dev(q).reshape(-1)
    y = torch.empty(n, dtype=F64, device="cuda")
    s = stream_ptr()
    # y = V q  (rows of V are eigenvectors)
    lib.vt_gemm_f64(n, 1, n, 1.0, ptr(V), n, 1, 0, ptr(qd), 1, 1, 0, 0.0, ptr(y), 1, 0, 1, s)
    lib.vt_eig_filter(ptr(y), ptr(w), n, float(rcond), s)
    c = torch.empty(n, dtype=F64, device="cuda")
    # c = V^T y
    lib.vt_gemm_f64(n, 1, n, 1.0, ptr(V), 1, n, 0, ptr(y), 1, 1, 0, 0.0, ptr(c), 1, 0, 1, s)
    return c
