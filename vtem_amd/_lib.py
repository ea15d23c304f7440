"""ctypes binding of libvtem_hip.so (the C-ABI declared in include/vtem_hip.h).

Pointers are passed as integers (device addresses from torch tensors' data_ptr()),
sizes as C ints / int64, scalars as double.  Every entry returns an int status:
0 = OK, > 0 = hipError_t, < 0 = argument error; `check` maps it to an exception.
There is no fallback: if the library is missing the import of any compute module fails.
"""

from __future__ import annotations

import ctypes
import os

import torch

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libvtem_hip.so")

P = ctypes.c_void_p
I = ctypes.c_int
L = ctypes.c_longlong
D = ctypes.c_double
F = ctypes.c_float

# name -> argument types (return type is always c_int)
SIGNATURES = {
    # encode
    "vt_rf_bank": [P, I, L, I, P, P, P, P, P, P, D, D, D, I, I, D, D, D, D, P],
    "vt_rf_gabor": [P, I, L, I, P, P, P, P, P, D, I, I, D, D, D, D, P],
    "vt_rf_dog": [P, I, L, I, P, P, P, D, D, I, I, D, D, D, D, P],
    "vt_iaf_encode": [P, I, L, I, I, P, L, I, P, P, P, P, P, P, D, I, P],
    "vt_iaf_encode_trig": [P, L, P, I, I, I, P, L, I, P, P, P, P, P, P, D, I, P],
    "vt_iaf_encode_trig_par": [P, L, P, I, I, I, P, L, I, P, P, P, P, P, P, D, I, P, L, P],
    "vt_iaf_encode_par": [P, I, L, I, I, P, L, I, P, P, P, P, P, P, D, I, P, L, P],
    "vt_iaf_encode_trig_batched": [P, I, I, I, I, I, P],
    "vt_iaf_par_workspace_size": [I, I, P],
    "vt_iaf_encode_par_noalloc": [P, I, L, I, I, P, L, I, P, P, P, P, P, P, D, I, P],
    "vt_noalloc_scratch_bytes": [P],
    "vt_iaf_par_desc_size": [P],
    "vt_iaf_encode_rt": [P, L, I, I, P, L, I, P, P, P, P, P, P, D, P, P, I, P],
    "vt_intervals": [P, L, I, P, I, D, D, P, P, P, P, P, I, P],
    "vt_exclusive_scan": [P, I, P, P],
    # dense
    "vt_gemm_f64": [I, I, I, D, P, L, L, L, P, L, L, L, D, P, L, L, I, P],
    "vt_gemm_f32": [I, I, I, F, P, L, L, L, P, L, L, L, F, P, L, L, I, P],
    "vt_mfma_probe": [I, P, P, P, P],
    "vt_gemm_tn_small": [I, I, I, D, P, L, P, L, D, P, L, P, P],
    # decode
    "vt_q": [P, P, P, P, P, P, P, P, P, I, P],
    "vt_temporal_factors": [P, P, P, I, D, P, I, P],
    "vt_neuron_T": [P, P, P, P, P, P, P, I, I, D, P],
    "vt_neuron_T_batched": [P, P, P, P, P, P, P, P, I, I, P],
    "vt_tj_apply": [P, P, P, I, I, P],
    "vt_axpbypcz": [P, D, P, D, P, D, P, L, P],
    "vt_trace_lambda": [P, I, L, I, P, D, P, P],
    "vt_G_spike": [P, L, P, P, D, I, P, L, P, I, P],
    "vt_diag_add": [P, I, L, I, D, P, P],
    "vt_reconstruct_spike": [P, L, P, L, P, P, P, P, I, P, P, I, D, I, I, P],
    "vt_gemv": [P, P, L, P, I, P],
    "vt_residual": [P, P, P, P, P, L, P],
    "vt_sumsq": [P, L, P, P],
    "vt_rnn_spike": [P, P, L, P, I, D, I, P, P, P],
    # gram / solve
    "vt_gram_f64": [P, L, P, I, I, I, P, L, P],
    "vt_gram_f64_batched": [P, I, I, I, P],
    "vt_gram_split_T": [P, I, I, P, P, I, P, L, P, P],
    "vt_gram_split_T_size": [I, I, P],
    "vt_gram_lp": [P, I, I, I, I, I, P, L, P],
    "vt_gram_lp_ws_size": [I, I, I, I, I, P],
    "vt_kron_trace": [P, P, I, I, P, D, P, P],
    "vt_kron_trace_batched": [P, P, P, P, P, I, P, I, D, P, I, P],
    "vt_row_sumsq": [P, L, I, I, P, P],
    "vt_transpose_f64": [P, L, I, I, P, L, P],
    "vt_tile_f64": [P, L, I, I, I, P, P],
    "vt_symmetrize": [P, I, L, I, P],
    "vt_potrf_f64": [P, I, I, P, L, P, P],
    "vt_potrf_f32": [P, I, I, P, L, P, P],
    "vt_potrs_f64": [P, I, I, P, L, P, L, P, P],
    "vt_potrs_f32": [P, I, I, P, L, P, L, P, P],
    "vt_lower_to_f32": [P, L, P, L, I, P, P],
    # batched
    "vt_intervals_batched": [P, P, I, P, P, P, P, P, I, P],
    "vt_gemm_grouped": [P, I, I, I, P],
    "vt_gemm_tn_batched": [P, I, I, I, P],
    "vt_residual_batched": [P, P, P, P, P, P, L, P],
    "vt_seg_sumsq": [P, P, I, P, I, P],
    "vt_diag_add_batched": [P, I, I, P, I, P],
    "vt_add_inplace": [P, P, L, P],
    "vt_seg_counts": [P, P, I, P, P],
    # Kronecker-preconditioned CG
    "vt_kpcg_tbar": [P, P, I, I, P, P, P, P],
    "vt_kpcg_mgram": [P, I, I, P],
    "vt_kpcg_place_diag": [P, P, P, I, I, P, L, P],
    "vt_gemm_skinny": [P, I, I, I, I, P],
    "vt_gemm_skinny_workspace_size": [I, I, I, P],
    "vt_skinny_split_rule": [P, P],
    "vt_kpcg_init": [P, P, P, P, P, I, P, P, I, P, P, P],
    "vt_kpcg_step_a": [P, P, P, P, P, P, I, P, P, I, I, D, I, I, P],
    "vt_kpcg_step_b": [P, P, P, P, I, P, P, P, P],
    "vt_kpcg_iterate": [P, I, I, P, I, I, P, I, I, I, I, P, I, I, I, I, P, I, I, I, I, I, I, P, P, P, P, P, P,
                        P, P, P, I, I, D, I, I, P, P, P, P],
    "vt_kpcg_tbar_eig": [P, P, I, I, P, P, P, P, P, P],
    "vt_kpcg_shift_pool": [P, P, P, P, P, P, I, I, I, P, P],
    "vt_kron_lanczos": [P, I, I, P, I, I, I, I, P, P, P, P, P, I, P, I, P, P, P, P, P, P],
    # spatial Dirichlet side, pseudoinverse, stitching
    "vt_dirichlet_table": [P, L, I, I, D, D, D, D, I, I, D, D, P],
    "vt_ds_assemble": [P, P, L, I, I, I, I, D, P],
    "vt_scale_outer": [P, L, I, P, P],
    "vt_svd_jacobi_sweep": [P, L, P, L, I, D, D, P, P],
    "vt_zero_int": [P, P],
    "vt_svd_pinv_apply": [P, L, I, P, D, P, P, P],
    "vt_svd_eig_finish": [P, L, P, L, I, D, P, P, L, P, P],
    "vt_stitch_accumulate": [P, L, L, P, I, L, P],
    "vt_kron_rnn": [P, I, I, P, I, I, I, I, P, P, P, P, P, P, L, I, I, P, P, P, L, P],
    "vt_kron_rnn_persistent": [P, P, I, I, P, I, P, L, P, P, P, P, P, I, P, L, P, L, I, P, D, P],
    "vt_kron_rnn_persistent_capacity": [P, P],
    "vt_kron_rnn_persistent_bar_words": [I, P],
}


class VtemError(RuntimeError):
    pass


_ERRS = {-1: "invalid argument", -2: "unsupported", -3: "spike buffer overflow",
         -4: "matrix not positive definite", -5: "size beyond a documented limit"}


def check(status: int, name: str = "") -> None:
    if status == 0:
        return
    if status < 0:
        msg = _ERRS.get(status, "error")
        if status in (-1, -5):
            raise ValueError(f"{name}: {msg} (status {status})")
        raise VtemError(f"{name}: {msg} (status {status})")
    raise VtemError(f"{name}: HIP error {status}")


class _Lib:
    def __init__(self):
        self._cdll = None

    def load(self):
        if self._cdll is None:
            if not os.path.exists(LIB_PATH):
                raise ImportError(f"{LIB_PATH} is missing: build it with "
                                  "`python -m vtem_amd.build` (there is no CPU fallback)")
            lib = ctypes.CDLL(LIB_PATH)
            for name, args in SIGNATURES.items():
                fn = getattr(lib, name)
                fn.argtypes = args
                fn.restype = ctypes.c_int
            self._cdll = lib
        return self._cdll

    def __getattr__(self, name):
        if not name.startswith("vt_"):
            raise AttributeError(name)
        fn = getattr(self.load(), name)

        def call(*args):
            check(fn(*args), name)

        call.__name__ = name
        return call


lib = _Lib()


def ptr(t) -> int:
    """Device (or host) address of a tensor, or 0 for None."""
    if t is None:
        return 0
    return int(t.data_ptr())


def stream_ptr(stream=None) -> int:
    s = stream if stream is not None else torch.cuda.current_stream()
    return int(s.cuda_stream)


def require_gpu():
    if not torch.cuda.is_available():
        raise RuntimeError("vtem_amd needs a ROCm GPU (torch.cuda.is_available() is False)")
    lib.load()
