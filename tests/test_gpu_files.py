"""The file route of the reference's API on the GPU (vtem/vtem.py:165-200,
vtem/vtdm.py:18-309, :699-1071, vtem/utils/spike_store.py:279-351).

* VTEM_IAF with an output file writes the spike file; load_encoded reads it back into the
  same spike trains, bit for bit (absolute time = bin time + i * refresh interval).
* read_select_spikes of a window (bins floor(t0/T)-1 .. floor(t1/T)+1 read, intervals
  selected on the GPU) equals the oracle's restatement of the selection on the same
  spikes, bit for bit.
* decode_video from the file name (Dsw / dirich written to files, the video to an array
  file) equals decode_video from the in-memory result, bit for bit, and the written
  video equals the returned one.
* precompute=False (VTDM_prepb / VTDMb: Ds written, the weighted Gram formed at decode
  time) gives the same video bit for bit; with spatial stitching and write_blocks every
  tile's video is written as an AVI file.
"""

import math
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def torch_mod():
    import torch
    from vtem_amd._lib import require_gpu
    require_gpu()
    return torch


def _setup(torch_mod, tmp_path, chunk=30):
    from vtem_amd import synthetic as syn
    from vtem_amd import vrf
    from vtem_amd.ensemble_encode import IAF_encode
    from vtem_amd.vtem import VTEM_IAF
    vid = syn.make_video(48, 48, 1.0, 100.0, 4, 4, 2, seed=1)
    bank = syn.gabor_grid_bank(48.0, 48.0, 4.0, 6, n_orient=4)
    g = vrf.vrf_gabor((48, 48), dx=1.0, dy=1.0)
    g.load_parameters(bank.num_neurons, bank.h_alpha, bank.h_l, bank.h_x0, bank.h_y0, bank.h_ab,
                      KAPPA=bank.KAPPA)
    N = bank.num_neurons
    rng = np.random.default_rng(7)
    kappa, bias, delta = np.ones(N), np.full(N, 2.0), np.full(N, 0.04)
    iaf = IAF_encode(N, 0.01)
    iaf.load_parameters(kappa, bias, delta, h_v0=rng.uniform(0, 1, N) * delta)
    Mx = 4
    W = 2 * math.pi * Mx / 32
    Wt = 2 * math.pi * 10 / 3
    f = str(tmp_path / "spikes.h5")
    enc = VTEM_IAF(vid.frames, f, Wt, g, W, W, dx=1.0, dy=1.0, npinput=True, iaf=iaf, chunk_frames=chunk)
    return vid, bank, enc, f, Mx, Wt


def test_spike_file_round_trip_and_selection(torch_mod, oracle, tmp_path):
    from vtem_amd.utils import spike_store as ss
    from vtem_amd.vtem import load_encoded
    print("\n[case] spike file round trip, 48x48 x 1 s, 30-frame bins", flush=True)
    vid, bank, enc, f, Mx, Wt = _setup(torch_mod, tmp_path)
    back = load_encoded(f)
    assert back.total_intervals == enc.total_intervals > 3
    np.testing.assert_array_equal(back.spike_count, enc.spike_count)
    for a, b in zip(back.spikes, enc.spikes):
        np.testing.assert_array_equal(a, b)
    assert (back.Wx, back.Wy, back.Wt) == (enc.Wy, enc.Wx, enc.Wt)
    rd = ss.vtem_storage_read(f)
    rd.read_num_neurons()
    dom = [-16.0, 20.0, -24.0, 8.0]
    rd.select_neurons(dom)
    refresh = rd.read_refresh_interval()
    for t0, t1 in [(0.0, 0.4), (0.33, 0.71), (0.58, 1.0)]:
        tk1, tk2, nid, kappa, delta, bias, sigma = rd.read_select_spikes([t0, t1])
        bins = range(max(0, int(np.floor(t0 / refresh)) - 1),
                     min(rd.read_total_intervals(), int(np.floor(t1 / refresh)) + 2))
        o1, o2, on, _ = oracle.select_intervals(rd.bin_spikes(bins), (t0, t1))
        np.testing.assert_array_equal(tk1, o1)
        np.testing.assert_array_equal(tk2, o2)
        np.testing.assert_array_equal(nid, on)
        np.testing.assert_array_equal(kappa, np.ones(rd.decode_neurons))
        print(f"  window ({t0}, {t1}): {tk1.size} intervals of {rd.decode_neurons} neurons", flush=True)


def test_decode_from_files(torch_mod, oracle, tmp_path):
    from vtem_amd import vtdm
    from vtem_amd.utils import simpleio, videoio
    print("\n[case] decode_video from files, precompute True / False, write_blocks", flush=True)
    vid, bank, enc, f, Mx, Wt = _setup(torch_mod, tmp_path)
    kw = dict(Mt=1, stitching=True, stitch_interval=0.3)
    u_mem = vtdm.decode_video(enc, None, None, 0.0, 0.6, 0.01, Mx, output_format=2, **kw)
    out = str(tmp_path / "rec")
    dsw, dirich = str(tmp_path / "dsw.h5"), str(tmp_path / "dirich.h5")
    # the file route reads Wx / Wy by node name (exchanged); both are equal here
    u_file = vtdm.decode_video(f, dsw, dirich, 0.0, 0.6, 0.01, Mx, output_format=1, output=out, **kw)
    np.testing.assert_array_equal(u_file, u_mem)
    np.testing.assert_array_equal(simpleio.read_file(out + ".h5"), u_file)
    assert simpleio.read_file(dsw).shape[0] == simpleio.read_file(dirich).shape[0]
    ds = str(tmp_path / "ds.h5")
    u_b = vtdm.decode_video(f, ds, dirich, 0.0, 0.6, 0.01, Mx, precompute=False, output_format=2, **kw)
    assert np.iscomplexobj(simpleio.read_file(ds))
    np.testing.assert_array_equal(u_b, u_mem)
    s = oracle.snr_db(vid.frames[:60], u_mem)
    print(f"  SNR {s:.2f} dB", flush=True)
    blocks = str(tmp_path / "blk")
    u_t = vtdm.decode_video(f, ds, dirich, 0.0, 0.3, 0.01, Mx, Mt=1, precompute=False, spatial_stitching=True,
                            spatial_interval=[32, 32], write_blocks=True, output_format=0, output=blocks)
    tiles = sorted(p for p in os.listdir(tmp_path) if p.startswith("blk") and p[3:-4].isdigit())
    assert len(tiles) == 4
    frames, _ = videoio.read_avi(os.path.join(tmp_path, tiles[0]))
    assert frames.shape == (30, 32, 32, 3)
    whole, _ = videoio.read_avi(blocks + ".avi")
    assert whole.shape == (30, 48, 48, 3) and u_t.shape == (30, 48, 48)
