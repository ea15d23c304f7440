"""Linear algebra with the reference's API (vtem/utils/linalg.py): the truncated-SVD
pseudoinverse solve of the spike-space decoder and the symmetric eigendecomposition.

solve_eq_sym(G, q, rcond) returns c = U diag(1/s_i if s_i > rcond s_0 else 0) U^T q, U the
left singular vectors and s the singular values of G, as vtem/utils/linalg.py:591-631
computes through CULA's SVD.  Here the SVD comes from a one-sided Jacobi iteration on the
rows of G (vt_svd_jacobi_sweep, sweeps until one changes nothing), and c from the
converged rows (vt_svd_pinv_apply).  G is overwritten (the reference destroys G too,
:598-599).  eig_sym accumulates the rotations as well and returns the eigenvalues in
ascending order with the eigenvectors (vt_svd_eig_finish).  Each sweep ends with a host
synchronisation (the convergence test), as the reference's solve returns host-visible
results.
"""

from __future__ import annotations

import numpy as np
import torch

from .._lib import VtemError, lib, ptr, stream_ptr

F64 = torch.float64
MAX_SWEEPS = 40
EPS = 2.220446049250313e-16


def _dev(a, dtype=F64):
    if isinstance(a, torch.Tensor):
        return a.to(device="cuda", dtype=dtype).contiguous()
    return torch.as_tensor(np.ascontiguousarray(a), dtype=dtype, device="cuda")


def _work_copy(G):
    """G itself when it is a contiguous FP64 device matrix (overwritten), else a copy."""
    if isinstance(G, torch.Tensor) and G.is_cuda and G.dtype == F64 and G.is_contiguous():
        return G
    return _dev(G)


def frobenius(A: torch.Tensor) -> float:
    """|A|_F of a contiguous FP64 device matrix (a host synchronisation)."""
    out = torch.zeros(1, dtype=F64, device=A.device)
    lib.vt_sumsq(ptr(A), A.numel(), ptr(out), stream_ptr())
    return float(out.item()) ** 0.5


def jacobi_sweeps(A: torch.Tensor, V: torch.Tensor = None, fro: float = None) -> int:
    """Sweeps of one-sided Jacobi on the rows of A (and V) until a sweep rotates nothing;
    returns the number of sweeps.  A pair is rotated when |a_p.a_q| > n eps |a_p| |a_q|;
    rows of norm at most n eps |A|_F (|A|_F is invariant under the rotations) are
    numerically zero: their directions are the rounding noise of the rotations.  That is
    far below any singular value the pseudoinverse keeps (rcond = 1e-8 of the largest)."""
    n = A.shape[0]
    if A.shape != (n, n):
        raise ValueError("G must be square matrix")
    tol = max(n, 16) * EPS
    small = max(n, 16) * EPS * (frobenius(A) if fro is None else fro)
    rot = torch.zeros(1, dtype=torch.int32, device=A.device)
    s = stream_ptr()
    for sweep in range(1, MAX_SWEEPS + 1):
        lib.vt_zero_int(ptr(rot), s)
        lib.vt_svd_jacobi_sweep(ptr(A), A.stride(0), ptr(V) if V is not None else 0,
                                V.stride(0) if V is not None else 0, n, tol, small * small, ptr(rot), s)
        if int(rot.item()) == 0:
            return sweep
    raise VtemError(f"Jacobi SVD did not converge in {MAX_SWEEPS} sweeps "
                    f"({int(rot.item())} rotations in the last)")


def eig_sym(G):
    """(w ascending, E with the eigenvectors in its rows) of the symmetric G (vtem/utils/
    linalg.py:634-718); G is overwritten."""
    A = _work_copy(G)
    n = A.shape[0]
    # G + shift I with shift > |G|_F >= |G|_2 is positive definite: its singular vectors
    # are eigenvectors, even where G has eigenvalues of equal size and opposite sign
    shift = 1.01 * frobenius(A) + 1e-300
    lib.vt_diag_add(ptr(A), 0, A.stride(0), n, shift, 0, stream_ptr())
    V = torch.eye(n, dtype=F64, device=A.device)
    jacobi_sweeps(A, V)
    w = torch.empty(n, dtype=F64, device=A.device)
    E = torch.empty((n, n), dtype=F64, device=A.device)
    ws = torch.empty(n, dtype=F64, device=A.device)
    lib.vt_svd_eig_finish(ptr(A), A.stride(0), ptr(V), V.stride(0), n, shift, ptr(w), ptr(E), E.stride(0),
                          ptr(ws), stream_ptr())
    return w, E


def solve_eq_sym(G, q, rcond=1e-8):
    """Truncated-SVD solve of G c = q for symmetric G (vtem/utils/linalg.py:591-631)."""
    A = _work_copy(G)
    n = A.shape[0]
    qd = _dev(q).reshape(-1)
    if qd.numel() != n:
        raise ValueError("number of columns of G must be the same of size of q")
    jacobi_sweeps(A)
    c = torch.empty(n, dtype=F64, device=A.device)
    ws = torch.empty(2 * n, dtype=F64, device=A.device)
    lib.vt_svd_pinv_apply(ptr(A), A.stride(0), n, ptr(qd), float(rcond), ptr(c), ptr(ws), stream_ptr())
    return c
