"""Spike-space Dirichlet decoder with the reference's API (vtem/dirichlet.py).

dirichlet(decode_neuron, tk1, tk2, neuron_ind, kappa, delta, bias, norm, Wt, Mt) holds
the intervals of the selected neurons; compute_q, compute_G(Dsw, lamb) and
reconstruct(dirich, time_frame, dt) run vt_q, vt_G_spike (+ vt_diag_add) and
vt_reconstruct_spike.  As in the reference, reconstruct reads the solution c from
self.d_q (the caller overwrites it with the solver's output, e.g. solve_eq_sym or
rnn3).  The Dsw / Ds / dirich arguments, file names in the reference, are file names
(array files, utils.simpleio) or the arrays themselves (the outputs of vrf.compute_Dsw /
compute_Ds / compute_dirich_space).
"""

from __future__ import annotations

import numpy as np
import torch

from ._lib import lib, ptr, stream_ptr
from .utils import simpleio

F64 = torch.float64


def _dev(a, dtype=F64, device="cuda"):
    if isinstance(a, (str, bytes)) or hasattr(a, "__fspath__"):
        a = simpleio.read_file(a)
    if isinstance(a, torch.Tensor):
        return a.to(device=device, dtype=dtype).contiguous()
    return torch.as_tensor(np.ascontiguousarray(a), dtype=dtype, device=device)


class dirichlet:
    def __init__(self, decode_neuron, tk1, tk2, neuron_ind, kappa, delta, bias, norm, Wt, Mt,
                 dtype=np.float64, device="cuda"):
        if np.dtype(dtype) != np.float64:
            raise ValueError("the spike-space decoder computes in float64")
        self.size = int(np.asarray(tk1).size if not isinstance(tk1, torch.Tensor) else tk1.numel())
        if self.size == 0:
            raise ValueError("no spikes to decode")
        self.decode_neuron = int(decode_neuron)
        self.Wt, self.Mt = float(Wt), int(Mt)
        self.device = device
        self.d_tk1 = _dev(tk1, F64, device)
        self.d_tk2 = _dev(tk2, F64, device)
        self.d_neuron_ind = _dev(neuron_ind, torch.int32, device)
        self.d_kappa = _dev(kappa, F64, device)
        self.d_delta = _dev(delta, F64, device)
        self.d_bias = _dev(bias, F64, device)
        self.d_norm = _dev(norm, F64, device)
        self._n = torch.tensor([self.size], dtype=torch.int32, device=device)

    def compute_q(self):
        """q_k = (kappa_j delta_j - b_j (tk2 - tk1)) norm_j (vtem/dirichlet.py:55-60, :204-222)."""
        self.d_q = torch.empty(self.size, dtype=F64, device=self.device)
        lib.vt_q(ptr(self.d_q), ptr(self.d_tk1), ptr(self.d_tk2), ptr(self.d_neuron_ind), ptr(self.d_kappa),
                 ptr(self.d_delta), ptr(self.d_bias), ptr(self.d_norm), ptr(self._n), self.size,
                 stream_ptr())
        return self.d_q

    def compute_G(self, Dsw, lamb=0.0):
        """G (n x n) = G_t(k, l) Dsw[j(k), j(l)] (+ lamb I) (vtem/dirichlet.py:62-85)."""
        W = _dev(Dsw, F64, self.device)
        self.d_G = torch.empty((self.size, self.size), dtype=F64, device=self.device)
        lib.vt_G_spike(ptr(self.d_G), self.size, ptr(self.d_tk1), ptr(self.d_tk2), self.Wt, self.Mt, ptr(W),
                       W.stride(0), ptr(self.d_neuron_ind), self.size, stream_ptr())
        if lamb != 0:
            lib.vt_diag_add(ptr(self.d_G), 0, self.size, self.size, float(lamb), 0, stream_ptr())
        return self.d_G

    def compute_Gb(self, Ds, lamb=0.0):
        """G from the RFs' Dirichlet coefficients Ds (N x modes, complex): Dsw =
        Re(Ds Ds^H) * norm norm^T with this decoder's norm, then compute_G
        (vtem/dirichlet.py:87-127)."""
        from .vrf import weighted_gram
        D = _dev(Ds, torch.complex128, self.device)
        return self.compute_G(weighted_gram(D.reshape(D.shape[0], -1), self.d_norm, self.device), lamb)

    def reconstruct(self, dirich, time_frame, dt):
        """u(t, x, y) = sum_k c_k norm_j psi_j(x, y) Phi_k(t) at t = arange(t0, t1, dt),
        c = self.d_q (vtem/dirichlet.py:165-189) -> (T, Py, Px) float64."""
        psi = _dev(dirich, F64, self.device)
        N, Py, Px = psi.shape
        t = np.arange(time_frame[0], time_frame[1], dt)
        d_t = _dev(t, F64, self.device)
        P = Py * Px
        u = torch.empty((t.size, Py, Px), dtype=F64, device=self.device)
        lib.vt_reconstruct_spike(ptr(u), P, ptr(psi), P, ptr(self.d_tk1), ptr(self.d_tk2), ptr(self.d_q),
                                 ptr(d_t), t.size, ptr(self.d_neuron_ind), ptr(self.d_norm), self.Mt,
                                 self.Wt / self.Mt, self.size, P, stream_ptr())
        return u

    def freeG(self):
        self.d_G = None
