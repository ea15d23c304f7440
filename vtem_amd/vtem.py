"""Video time encoding drivers with the reference's API (vtem/vtem.py).

VTEM_IAF filters a video with a receptive-field bank and encodes the filter outputs with
IAF neurons in chunks of at most 101 frames that share one frame with the next chunk
(vtem/vtem.py:165-200): the neurons' membrane state carries over, the time counter is
reset per chunk and the chunk's spike times are shifted by its start, so the result
equals one uninterrupted encode of the whole input sampled at dt = 1/fps.

The video comes as an array (npinput=True), an array file (h5input=True, utils.simpleio)
or an uncompressed AVI file (utils.videoio); the spikes are returned and, when
output_filename is given, written as a spike file in the reference's node layout
(utils.spike_store; the container is numpy's .npz since HDF5 is not part of this build).
load_encoded reads such a file back into the same EncodedVideo.
"""

from __future__ import annotations

import dataclasses
import os
import types
from typing import Any, List, Optional

import numpy as np

from . import ensemble_encode as ee
from . import vrf as vrf_mod
from .utils import simpleio, spike_store, videoio


@dataclasses.dataclass
class EncodedVideo:
    """Spike trains of a VTEM run with the attributes the reference stores in its spike
    file (vtem/utils/spike_store.py:18-156)."""
    spikes: List[np.ndarray]          # per neuron, absolute times in seconds
    spike_count: np.ndarray
    rf: Any
    iaf: Any
    Wt: float
    Wx: float
    Wy: float
    Px: int
    Py: int
    dx: float
    dy: float
    fps: Optional[float]
    refresh_interval: float
    num_frames: Optional[int]
    total_intervals: int = 0

    @property
    def duration(self) -> float:
        """Encoded seconds: num_frames / fps, or (from a spike file, which stores neither)
        the number of stored time bins times their length."""
        if self.num_frames is not None and self.fps:
            return self.num_frames / self.fps
        return self.total_intervals * self.refresh_interval


def _rf_type(rf) -> str:
    if isinstance(rf, vrf_mod.vrf_gabor):
        return "gabor"
    if isinstance(rf, vrf_mod.vrf_cs):
        return "cs"
    raise TypeError("Receptive field type not recognized")


def load_encoded(spikefile, device="cuda") -> EncodedVideo:
    """The EncodedVideo of a spike file written by VTEM_IAF: the receptive fields rebuilt
    from the stored bank, the IAF parameters, and each neuron's spikes in absolute time
    (bin i's stored times + i * refresh_interval, as vtem/utils/spike_store.py:319 reads
    them).  Wx and Wy are read by node name (the reference stores them exchanged)."""
    rd = spike_store.vtem_storage_read(spikefile)
    N = rd.read_num_neurons()
    Wt, Wx, Wy, Px, Py, dx, dy = rd.read_video_attributes()
    kind = rd.read_vrf_type()
    if kind == "gabor":
        alpha, l, x0, y0, ab, KAPPA = rd.read_vrf()
        rf = vrf_mod.vrf_gabor((Py, Px), dx=dx, dy=dy, device=device)
        rf.load_parameters(N, alpha, l, x0, y0, ab, KAPPA=KAPPA)
    else:
        alpha, x0, y0, sc, ss = rd.read_vrf()
        rf = vrf_mod.vrf_cs((Py, Px), dx=dx, dy=dy, device=device)
        rf.load_parameters(N, alpha, x0, y0, sigma_center=sc, sigma_surround=ss)
    kappa, delta, bias, sigma = rd.read_iaf()
    iaf = types.SimpleNamespace(h_kappa=kappa, h_delta=delta, h_bias=bias, h_sigma=sigma, num_neurons=N)
    total = rd.read_total_intervals()
    rd.choose = np.arange(N)
    rd.decode_neurons = N
    spikes = rd.bin_spikes(range(total))
    counts = np.array([s.size for s in spikes], np.int32)
    return EncodedVideo(spikes, counts, rf, iaf, Wt, Wx, Wy, Px, Py, dx, dy, None, rd.read_refresh_interval(),
                        None, total)


def _read_input(videofile, h5input, npinput):
    if npinput:
        return videofile
    if h5input:
        return simpleio.read_file(videofile)
    return videoio.read_video(videofile)


def VTEM_IAF(videofile, output_filename, Wt, rf, Wx=2 * np.pi * 4, Wy=2 * np.pi * 4, start_frame=0,
             end_frame=None, dx=1.0 / 16, dy=1.0 / 16, fps=100, domain=None, h5input=True, npinput=False,
             iaf=None, chunk_frames: int = 101, avg_rate: float = 0.1) -> EncodedVideo:
    """Encode a video (array (frames, Py, Px) with npinput=True) with IAF neurons behind the
    receptive fields `rf` (a loaded vrf_gabor / vrf_cs).  `iaf` may be a loaded
    IAF_encode; otherwise one with the reference's default parameters is made."""
    video = _read_input(videofile, h5input, npinput)
    total, Py, Px = video.shape
    end_frame = total if end_frame is None else min(int(end_frame), total)
    dt = 1.0 / fps
    if iaf is None:
        iaf = ee.IAF_encode(rf.num_neurons, dt)
        iaf.load_parameters()
    chunk = max(2, min(int(chunk_frames), end_frame - start_frame + 1))
    refresh = (chunk - 1) * dt
    store = None
    if output_filename:
        store = spike_store.vtem_storage(output_filename, rf.num_neurons, _rf_type(rf))
        store.write_video_attributes(Wt, Wx, Wy, rf.Px, rf.Py, float(dx), float(dy))
        if store.filter_type == "gabor":
            store.write_gabor_parameters(rf)
        else:
            store.write_cs_parameters(rf)
        store.write_neuron_parameters(iaf)
        store.write_refresh_interval(refresh)
    per_neuron: List[List[np.ndarray]] = [[] for _ in range(rf.num_neurons)]
    counts = np.zeros(rf.num_neurons, np.int64)
    for k, i in enumerate(range(start_frame, end_frame, chunk - 1)):
        block = video[i:i + chunk]
        if block.shape[0] < 2:
            break
        v = rf.filter(block)
        sp, cnt = iaf.encode(v, avg_rate=avg_rate)
        # absolute time as a spike file reader forms it: bin time + k * refresh interval
        for j in range(rf.num_neurons):
            if cnt[j]:
                per_neuron[j].append(sp[j] + k * refresh)
        counts += cnt
        if store is not None:
            store.write_spike_in_timebin(np.asarray(cnt, np.int32),
                                         np.concatenate([np.asarray(sp[j])[:cnt[j]] for j in range(rf.num_neurons)]))
        iaf.reset_timer()
    spikes = [np.concatenate(p) if p else np.zeros(0) for p in per_neuron]
    out = EncodedVideo(spikes, counts.astype(np.int32), rf, iaf, float(Wt), float(Wx), float(Wy),
                       int(rf.Px), int(rf.Py), float(dx), float(dy), float(fps), refresh,
                       end_frame - start_frame, 0 if store is None else store.refresh_interval)
    if store is not None:
        store.close()
    return out


def VTEM_Gabor_IAF(videofile, output_filename, Wt, Wx=2 * np.pi * 4, Wy=2 * np.pi * 4, start_frame=0,
                   end_frame=None, dx=1.0 / 16, dy=1.0 / 16, fps=100, domain=None, h5input=True,
                   npinput=False, **kw) -> EncodedVideo:
    """VTEM_IAF with the default Gabor bank on the video's grid (vtem/vtem.py:209-317)."""
    Py, Px = videofile.shape[1:]
    rf = vrf_mod.vrf_gabor((Py, Px), dx=dx, dy=dy, domain=domain)
    rf.load_parameters()
    return VTEM_IAF(videofile, output_filename, Wt, rf, Wx, Wy, start_frame, end_frame, dx, dy, fps,
                    domain, h5input, npinput, **kw)


def VTEM_CS_IAF(videofile, output_filename, Wt, Wx=2 * np.pi * 4, Wy=2 * np.pi * 4, start_frame=0,
                end_frame=None, dx=1.0 / 16, dy=1.0 / 16, fps=100, domain=None, h5input=True,
                npinput=False, **kw) -> EncodedVideo:
    """VTEM_IAF with the default centre-surround bank (vtem/vtem.py:319-427)."""
    Py, Px = videofile.shape[1:]
    rf = vrf_mod.vrf_cs((Py, Px), dx=dx, dy=dy, domain=domain)
    rf.load_parameters()
    return VTEM_IAF(videofile, output_filename, Wt, rf, Wx, Wy, start_frame, end_frame, dx, dy, fps,
                    domain, h5input, npinput, **kw)
