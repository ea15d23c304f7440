"""Video time encoding drivers with the reference's API (vtem/vtem.py).

VTEM_IAF filters a video with a receptive-field bank and encodes the filter outputs with
IAF neurons in chunks of at most 101 frames that share one frame with the next chunk
(vtem/vtem.py:165-200): the neurons' membrane state carries over, the time counter is
reset per chunk and the chunk's spike times are shifted by its start, so the result
equals one uninterrupted encode of the whole input sampled at dt = 1/fps.

The reference reads the video from an HDF5 / AVI file and writes spikes to an HDF5
file; neither HDF5 nor a video codec library is part of this build, so the video is
passed as an array (npinput=True) and the spikes are returned (and written to an .npz
archive when output_filename is given).
"""

from __future__ import annotations

import dataclasses
from typing import Any, List

import numpy as np

from . import ensemble_encode as ee
from . import vrf as vrf_mod


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
    fps: float
    refresh_interval: float
    num_frames: int


def VTEM_IAF(videofile, output_filename, Wt, rf, Wx=2 * np.pi * 4, Wy=2 * np.pi * 4, start_frame=0,
             end_frame=None, dx=1.0 / 16, dy=1.0 / 16, fps=100, domain=None, h5input=True, npinput=False,
             iaf=None, chunk_frames: int = 101, avg_rate: float = 0.1) -> EncodedVideo:
    """Encode a video (array (frames, Py, Px) with npinput=True) with IAF neurons behind the
    receptive fields `rf` (a loaded vrf_gabor / vrf_cs).  `iaf` may be a loaded
    IAF_encode; otherwise one with the reference's default parameters is made."""
    if not npinput:
        raise NotImplementedError("video files need HDF5 / a video codec; pass the array with "
                                  "npinput=True")
    video = videofile
    total, Py, Px = video.shape
    end_frame = total if end_frame is None else min(int(end_frame), total)
    dt = 1.0 / fps
    if iaf is None:
        iaf = ee.IAF_encode(rf.num_neurons, dt)
        iaf.load_parameters()
    chunk = max(2, min(int(chunk_frames), end_frame - start_frame + 1))
    refresh = (chunk - 1) * dt
    per_neuron: List[List[np.ndarray]] = [[] for _ in range(rf.num_neurons)]
    counts = np.zeros(rf.num_neurons, np.int64)
    for i in range(start_frame, end_frame, chunk - 1):
        block = video[i:i + chunk]
        if block.shape[0] < 2:
            break
        v = rf.filter(block)
        sp, cnt = iaf.encode(v, avg_rate=avg_rate)
        offset = (i - start_frame) * dt
        for j in range(rf.num_neurons):
            if cnt[j]:
                per_neuron[j].append(sp[j] + offset)
        counts += cnt
        iaf.reset_timer()
    spikes = [np.concatenate(p) if p else np.zeros(0) for p in per_neuron]
    out = EncodedVideo(spikes, counts.astype(np.int32), rf, iaf, float(Wt), float(Wx), float(Wy),
                       int(rf.Px), int(rf.Py), float(dx), float(dy), float(fps), refresh,
                       end_frame - start_frame)
    if output_filename:
        cum = np.concatenate([[0], np.cumsum(out.spike_count)])
        np.savez(output_filename, spikes=np.concatenate(spikes) if spikes else np.zeros(0),
                 cum_count=cum, Wt=Wt, Wx=Wx, Wy=Wy, Px=rf.Px, Py=rf.Py, dx=dx, dy=dy, fps=fps)
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
