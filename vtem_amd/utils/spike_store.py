"""Spike files with the reference's API (vtem/utils/spike_store.py:18-351).

A spike file holds what a VTEM run needs for decoding, in the reference's node layout:

    video/Wt, video/Wx, video/Wy                 bandwidths (scalars)
    vrf/Px, vrf/Py, vrf/dx, vrf/dy               the RF grid
    vrf/alpha, vrf/x0, vrf/y0                    per neuron
    vrf/l, vrf/ab, vrf/KAPPA                     Gabor banks
    vrf/sigma_center, vrf/sigma_surround         centre-surround banks
    vrf/num_neurons
    iaf/kappa, iaf/delta, iaf/bias, iaf/sigma    per neuron
    iaf/refresh_interval                         seconds per stored time bin
    spikes/spike_count<i>, spikes/spikes<i>      bin i: counts per neuron, and the spike
                                                 times of neuron 0, then 1, ... relative to
                                                 the start of the bin (i * refresh_interval)
    spikes/total_intervals

The reference writes it with PyTables; here the nodes are the members of a zip of .npy
files (numpy's .npz container, under the caller's file name; see utils.simpleio).  As in
the reference, write_video_attributes stores Wy under video/Wx and Wx under video/Wy
(spike_store.py:85-89), and read_video_attributes reads them back by node name, so a
write / read round trip exchanges Wx and Wy.
"""

from __future__ import annotations

from typing import List

import numpy as np

from . import simpleio


class vtem_storage:
    """Writer (vtem/utils/spike_store.py:6-156).  Nodes are collected in memory and the
    file is written by close()."""

    def __init__(self, filename, num_neurons, filter_type, title: str = "VTEM_gabor_iaf"):
        if filter_type not in ("gabor", "cs"):
            raise ValueError("filter_type must be 'gabor' or 'cs'")
        self.filename = filename
        self.num_neurons = int(num_neurons)
        self.filter_type = filter_type
        self.title = title
        self.nodes = {}
        self.ext = {k: [] for k in ("vrf/alpha", "vrf/x0", "vrf/y0", "iaf/kappa", "iaf/delta", "iaf/bias",
                                    "iaf/sigma")}
        if filter_type == "gabor":
            self.ext["vrf/l"] = []
            self.ext["vrf/ab"] = []
        self.refresh_interval = 0          # number of stored time bins
        self.isopen = True

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def _scalar(self, path, v, dtype=np.float64):
        self.nodes[path] = np.asarray(v, dtype)

    def write_video_attributes(self, Wt, Wx, Wy, Px, Py, dx, dy):
        self._scalar("video/Wt", Wt)
        self._scalar("video/Wx", Wy)          # the reference's node assignment (:87-88)
        self._scalar("video/Wy", Wx)
        self._scalar("vrf/Px", Px, np.int64)
        self._scalar("vrf/Py", Py, np.int64)
        self._scalar("vrf/dx", dx)
        self._scalar("vrf/dy", dy)

    def write_gabor_parameters(self, filter):
        if "vrf/KAPPA" not in self.nodes:
            self._scalar("vrf/KAPPA", filter.KAPPA)
        self.ext["vrf/alpha"].append(np.asarray(filter.h_alpha, np.float64))
        self.ext["vrf/l"].append(np.asarray(filter.h_l, np.float64))
        self.ext["vrf/x0"].append(np.asarray(filter.h_x0, np.float64))
        self.ext["vrf/y0"].append(np.asarray(filter.h_y0, np.float64))
        self.ext["vrf/ab"].append(np.asarray(filter.h_ab, np.int32))

    def write_cs_parameters(self, filter):
        if "vrf/sigma_center" not in self.nodes:
            self._scalar("vrf/sigma_center", filter.sigma_center)
            self._scalar("vrf/sigma_surround", filter.sigma_surround)
        self.ext["vrf/alpha"].append(np.asarray(filter.h_alpha, np.float64))
        self.ext["vrf/x0"].append(np.asarray(filter.h_x0, np.float64))
        self.ext["vrf/y0"].append(np.asarray(filter.h_y0, np.float64))

    def write_neuron_parameters(self, iaf):
        self.ext["iaf/kappa"].append(np.asarray(iaf.h_kappa, np.float64))
        self.ext["iaf/delta"].append(np.asarray(iaf.h_delta, np.float64))
        self.ext["iaf/bias"].append(np.asarray(iaf.h_bias, np.float64))
        self.ext["iaf/sigma"].append(np.asarray(getattr(iaf, "h_sigma", np.zeros(len(iaf.h_kappa))), np.float64))

    def write_refresh_interval(self, time):
        self._scalar("iaf/refresh_interval", time)

    def write_spike_in_timebin(self, spike_count, spikes):
        """Bin number self.refresh_interval: counts per neuron and the neurons' spike times
        (relative to the bin's start) one neuron after another."""
        i = self.refresh_interval
        self.nodes[f"spikes/spike_count{i}"] = np.asarray(spike_count, np.int32)
        self.nodes[f"spikes/spikes{i}"] = np.asarray(spikes, np.float64)
        self.refresh_interval += 1

    def close(self):
        """Write num_neurons (the number of RF parameter rows) and the number of bins, then
        the file."""
        if not self.isopen:
            return
        for k, parts in self.ext.items():
            dt = np.int32 if k == "vrf/ab" else np.float64
            self.nodes[k] = np.concatenate(parts).astype(dt) if parts else np.zeros(0, dt)
        self._scalar("vrf/num_neurons", self.nodes["vrf/alpha"].size, np.int64)
        self._scalar("spikes/total_intervals", self.refresh_interval, np.int64)
        simpleio.write_nodes(self.filename, self.nodes)
        self.isopen = False


class vtem_storage_read:
    """Reader (vtem/utils/spike_store.py:159-351)."""

    def __init__(self, filename):
        self.filename = filename
        self.nodes = simpleio.read_nodes(filename)
        self.filter_type = None
        self.num_neurons = None

    def close(self):
        pass

    def _node(self, path):
        try:
            return self.nodes[path]
        except KeyError:
            raise KeyError(f"{self.filename} has no node {path}") from None

    def read_num_neurons(self) -> int:
        self.num_neurons = int(self._node("vrf/num_neurons"))
        return self.num_neurons

    def read_video_attributes(self):
        """(Wt, Wx, Wy, Px, Py, dx, dy) by node name."""
        n = self._node
        return (float(n("video/Wt")), float(n("video/Wx")), float(n("video/Wy")), int(n("vrf/Px")),
                int(n("vrf/Py")), float(n("vrf/dx")), float(n("vrf/dy")))

    def read_vrf_type(self):
        self.filter_type = "gabor" if "vrf/ab" in self.nodes else "cs"
        return self.filter_type

    def read_vrf(self):
        """Gabor: (alpha, l, x0, y0, ab, KAPPA); centre-surround: (alpha, x0, y0,
        sigma_center, sigma_surround)."""
        if self.filter_type is None:
            self.read_vrf_type()
        n = self._node
        if self.filter_type == "gabor":
            return (n("vrf/alpha"), n("vrf/l"), n("vrf/x0"), n("vrf/y0"), n("vrf/ab"), float(n("vrf/KAPPA")))
        return (n("vrf/alpha"), n("vrf/x0"), n("vrf/y0"), float(n("vrf/sigma_center")),
                float(n("vrf/sigma_surround")))

    def read_refresh_interval(self) -> float:
        return float(self._node("iaf/refresh_interval"))

    def read_total_intervals(self) -> int:
        return int(self._node("spikes/total_intervals"))

    def read_iaf(self):
        n = self._node
        return n("iaf/kappa"), n("iaf/delta"), n("iaf/bias"), n("iaf/sigma")

    def select_neurons(self, domain):
        """Neurons whose RF centre is strictly inside domain [x_start, x_end, y_start, y_end]
        widened by 0.01, centres shifted to the domain's centre (:243-277)."""
        from ..vtdm import select_neurons
        if self.num_neurons is None:
            self.read_num_neurons()
        if self.filter_type is None:
            self.read_vrf_type()
        v = self.read_vrf()
        if self.filter_type == "gabor":
            h_alpha, h_l, h_x0, h_y0, h_ab, self.KAPPA = v
        else:
            h_alpha, h_x0, h_y0, self.sigma_center, self.sigma_surround = v

        class _C:
            pass
        c = _C()
        c.h_x0, c.h_y0 = h_x0, h_y0
        self.choose, self.x0, self.y0 = select_neurons(c, domain)
        self.alpha = h_alpha[self.choose]
        if self.filter_type == "gabor":
            self.l = h_l[self.choose]
            self.ab = h_ab[self.choose]
        self.decode_neurons = int(self.choose.size)

    def bin_spikes(self, bins) -> List[np.ndarray]:
        """Per selected neuron, its spikes of the given time bins in absolute time: bin i's
        stored times + i * refresh_interval (the reference's expression, :319)."""
        refresh = self.read_refresh_interval()
        parts: List[list] = [[] for _ in range(self.decode_neurons)]
        for i in bins:
            cnt = self._node(f"spikes/spike_count{i}")
            cum = np.concatenate([[0], np.cumsum(cnt, dtype=np.int64)])
            sp = self._node(f"spikes/spikes{i}") + i * refresh
            for j, nj in enumerate(self.choose):
                parts[j].append(sp[cum[nj]:cum[nj + 1]])
        return [np.concatenate(p) if p else np.zeros(0) for p in parts]

    def read_select_spikes(self, time_frame):
        """Intervals of the selected neurons in time_frame = [start, end] from the spikes of
        bins floor(start/refresh) - 1 .. floor(end/refresh) + 1 (:279-351): for each neuron the
        spikes strictly inside the window plus the last one before and the first one after,
        times relative to start.  Returns (tk1, tk2, neuron_ind, kappa, delta, bias,
        sigma of all neurons); the selection runs in the HIP library (vt_intervals)."""
        from ..vtdm import _SpikeSource
        h_kappa, h_delta, h_bias, h_sigma = self.read_iaf()
        self.kappa = h_kappa[self.choose]
        self.delta = h_delta[self.choose]
        self.bias = h_bias[self.choose]
        refresh = self.read_refresh_interval()
        total = self.read_total_intervals()
        s0 = int(np.floor(time_frame[0] / refresh))
        s1 = int(np.floor(time_frame[1] / refresh))
        bins = range(max(0, s0 - 1), min(total, s1 + 2))
        src = _SpikeSource(self.bin_spikes(bins))
        tk1, tk2, nid, n = src.select(np.arange(self.decode_neurons), float(time_frame[0]),
                                      float(time_frame[1]))
        return (tk1.cpu().numpy(), tk2.cpu().numpy(), nid.cpu().numpy(), self.kappa, self.delta,
                self.bias, h_sigma)
