"""Integrate-and-fire population encoders with the reference's API (vtem/ensemble_encode.py).

IAF_encode (deterministic thresholds) and IAF_encode_rt (per-spike random thresholds)
keep the reference's constructor, load_parameters, encode(neural_inputs, startbias,
avg_rate) -> (list of N spike arrays, int32 counts) and reset_timer.  IAF_encode runs
vt_iaf_encode, one lane per neuron, bit-identical to the reference's loop
(parallel=True: the time-parallel vt_iaf_encode_par, same counts, times to rounding, for
long inputs); IAF_encode_rt runs vt_iaf_encode_rt.  v0
and the time counter stay on the GPU between calls, as the reference's state does.  encode_device returns the
resident spike buffer instead of host lists (the decode path's input).

Defaults draw v0 (and the rt thresholds) from numpy's global generator, as the reference
does (vtem/ensemble_encode.py:64-65, :214-219); pass rng= for reproducible draws.
"""

from __future__ import annotations

import math
from typing import List, Tuple

import numpy as np
import torch

from ._lib import lib, ptr, stream_ptr
from .block import SpikeTrains, iaf_par_workspace

F64 = torch.float64


def _dev(a, dtype=F64, device="cuda"):
    if isinstance(a, torch.Tensor):
        return a.to(device=device, dtype=dtype).contiguous()
    return torch.as_tensor(np.ascontiguousarray(a), dtype=dtype, device=device)


class pop_encode:
    def __init__(self, num_neurons, dt, dtype=np.float64, device="cuda", rng=None, root_mode=0,
                 parallel=False):
        if np.dtype(dtype) != np.float64:
            raise ValueError("the encoder stores spike times in float64")
        self.num_neurons = int(num_neurons)
        self.dtype = np.dtype(dtype)
        self.dt = float(dt)
        self.device = device
        self.rng = rng
        self.root_mode = int(root_mode)
        # parallel=True: the time-parallel encoder (same counts, times to rounding);
        # default: the serial loop, bit-identical to the reference's
        self.parallel = bool(parallel)

    def _uniform(self, n):
        return (self.rng.uniform(0, 1, n) if self.rng is not None else np.random.uniform(0, 1, n))

    def _normal(self, shape):
        return (self.rng.normal(size=shape) if self.rng is not None else np.random.normal(size=shape))

    def set_time_count(self):
        self.h_time_count = np.zeros(self.num_neurons)

    def reset_timer(self):
        """Spike times of the next call start from 0 again (vtem/ensemble_encode.py:170-172)."""
        self.d_time_count.zero_()

    def _inputs(self, neural_inputs):
        x = _dev(neural_inputs, F64, self.device)
        if x.dim() != 2 or x.shape[1] != self.num_neurons:
            raise ValueError("input size should match number of neurons")
        return x

    @staticmethod
    def rearrange(spikes: torch.Tensor, counts: np.ndarray) -> List[np.ndarray]:
        """Host list of per-neuron spike arrays (vtem/ensemble_encode.py:335-340)."""
        h = spikes.cpu().numpy()
        return [h[:int(c), j].copy() for j, c in enumerate(counts)]


class IAF_encode(pop_encode):
    """Population encoding with IAF neurons (vtem/ensemble_encode.py:21-172)."""

    def set_parameters(self):
        n = self.num_neurons
        self.h_kappa = np.ones(n)
        self.h_delta = np.full(n, 0.03)
        self.h_bias = np.full(n, 0.8)
        self.h_sigma = np.zeros(n)

    def set_initial_value(self):
        self.h_v0 = self._uniform(self.num_neurons) * self.h_delta

    def load_parameters(self, h_kappa=None, h_bias=None, h_delta=None, h_sigma=None, h_time_count=None,
                        h_v0=None):
        if h_kappa is None:
            self.set_parameters()
        else:
            self.h_kappa = np.asarray(h_kappa, np.float64)
            self.h_delta = np.asarray(h_delta, np.float64)
            self.h_bias = np.asarray(h_bias, np.float64)
            self.h_sigma = np.zeros(self.h_kappa.shape)
        if h_time_count is None:
            self.set_time_count()
        else:
            self.h_time_count = np.asarray(h_time_count, np.float64).reshape(-1)
        if h_v0 is None:
            self.set_initial_value()
        else:
            self.h_v0 = np.asarray(h_v0, np.float64)
        self.d_kappa = _dev(self.h_kappa, F64, self.device)
        self.d_delta = _dev(self.h_delta, F64, self.device)
        self.d_bias = _dev(self.h_bias, F64, self.device)
        self.d_v0 = _dev(self.h_v0, F64, self.device)
        self.d_time_count = _dev(self.h_time_count, F64, self.device)

    def encode_device(self, neural_inputs, startbias=0, avg_rate=0.1) -> SpikeTrains:
        """Encode (T, N) inputs sampled every dt; spikes stay on the GPU.  Neuron j of the
        input uses the parameters of neuron startbias + j."""
        x = self._inputs(neural_inputs)
        T, N = x.shape
        max_spike = int(math.ceil(T / avg_rate))
        spikes = torch.empty((max_spike, N), dtype=F64, device=self.device)
        counts = torch.empty(N, dtype=torch.int32, device=self.device)
        o = int(startbias)
        args = (ptr(x), 0, x.stride(0), N, T, ptr(spikes), N, max_spike, ptr(self.d_v0) + 8 * o,
                ptr(self.d_kappa) + 8 * o, ptr(self.d_bias) + 8 * o, ptr(self.d_delta) + 8 * o,
                ptr(self.d_time_count) + 8 * o, ptr(counts), self.dt, self.root_mode)
        if self.parallel:
            ws = iaf_par_workspace(N, T, self.device)
            lib.vt_iaf_encode_par(*args, ptr(ws), ws.numel(), stream_ptr())
        else:
            lib.vt_iaf_encode(*args, stream_ptr())
        return SpikeTrains(spikes, counts)

    def encode(self, neural_inputs, startbias=0, avg_rate=0.1) -> Tuple[List[np.ndarray], np.ndarray]:
        st = self.encode_device(neural_inputs, startbias, avg_rate)
        counts = st.counts.cpu().numpy()
        if counts.max(initial=0) >= st.spikes.shape[0]:
            raise ValueError("number of spikes exceeded the limit of buffer")
        return self.rearrange(st.spikes, counts), counts


class IAF_encode_rt(pop_encode):
    """IAF neurons with a random threshold drawn per spike, N(delta, sigma^2)
    (vtem/ensemble_encode.py:175-333).  The N(0,1) draws are made on the host and passed in
    the spike buffer, as the reference does."""

    def set_parameters(self):
        n = self.num_neurons
        self.h_kappa = np.ones(n)
        self.h_delta = np.full(n, 0.03)
        self.h_bias = np.full(n, 0.8)
        self.h_sigma = np.full(n, 0.0003)
        self.h_delta_value = self._normal(n) * self.h_sigma + self.h_delta

    def set_initial_value(self):
        self.h_v0 = self._uniform(self.num_neurons) * self.h_delta_value

    def load_parameters(self, h_kappa=None, h_bias=None, h_delta=None, h_sigma=None, h_delta_value=None,
                        h_time_count=None, h_v0=None):
        if h_kappa is None:
            self.set_parameters()
        else:
            self.h_kappa = np.asarray(h_kappa, np.float64)
            self.h_delta = np.asarray(h_delta, np.float64)
            self.h_bias = np.asarray(h_bias, np.float64)
            self.h_sigma = np.asarray(h_sigma, np.float64)
            self.h_delta_value = np.asarray(h_delta_value, np.float64)
        if h_time_count is None:
            self.set_time_count()
        else:
            self.h_time_count = np.asarray(h_time_count, np.float64).reshape(-1)
        if h_v0 is None:
            self.set_initial_value()
        else:
            self.h_v0 = np.asarray(h_v0, np.float64)
        self.d_delta_value = _dev(self.h_delta_value, F64, self.device)
        self.d_kappa = _dev(self.h_kappa, F64, self.device)
        self.d_delta = _dev(self.h_delta, F64, self.device)
        self.d_bias = _dev(self.h_bias, F64, self.device)
        self.d_sigma = _dev(self.h_sigma, F64, self.device)
        self.d_v0 = _dev(self.h_v0, F64, self.device)
        self.d_time_count = _dev(self.h_time_count, F64, self.device)

    def encode_device(self, neural_inputs, startbias=0, avg_rate=0.1, randnum=None) -> SpikeTrains:
        x = self._inputs(neural_inputs)
        T, N = x.shape
        max_spike = int(math.ceil(T / avg_rate))
        draws = self._normal((max_spike, N)) if randnum is None else randnum
        spikes = _dev(draws, F64, self.device)
        if spikes.shape != (max_spike, N):
            raise ValueError("random draws must have shape (ceil(T/avg_rate), N)")
        counts = torch.empty(N, dtype=torch.int32, device=self.device)
        o = int(startbias)
        lib.vt_iaf_encode_rt(ptr(x), x.stride(0), N, T, ptr(spikes), N, max_spike, ptr(self.d_v0) + 8 * o,
                             ptr(self.d_kappa) + 8 * o, ptr(self.d_bias) + 8 * o, ptr(self.d_delta) + 8 * o,
                             ptr(self.d_time_count) + 8 * o, ptr(counts), self.dt,
                             ptr(self.d_delta_value) + 8 * o, ptr(self.d_sigma) + 8 * o, self.root_mode,
                             stream_ptr())
        return SpikeTrains(spikes, counts)

    def encode(self, neural_inputs, startbias=0, avg_rate=0.1, randnum=None):
        st = self.encode_device(neural_inputs, startbias, avg_rate, randnum)
        counts = st.counts.cpu().numpy()
        if counts.max(initial=0) >= st.spikes.shape[0]:
            raise ValueError("number of spikes exceeded the limit of buffer")
        return self.rearrange(st.spikes, counts), counts
