"""Array, spike and video files (host side, no GPU).

* Array files: write_memory_to_file / read_file round trips of real and complex arrays,
  'a' mode appending along the leading dimension, write_array / read_array, and no
  extension added to the caller's file name (vtem/utils/simpleio.py:10-162).
* Spike files: a writer / reader round trip of the node layout of
  vtem/utils/spike_store.py:18-241 -- bank, IAF parameters, refresh interval, per-bin
  counts and bin-relative times read back as absolute times (bin time + i * refresh, :319),
  Wx / Wy exchanged by node name as in the reference (:85-89), and the strict-bound neuron
  selection with its centre shift (:243-277) against the oracle's restatement.
* Video files: uncompressed AVI round trips (odd widths for the 4-byte row padding, the
  frame rate in the header), the reference's grey conversion sum(BGR w)/255 with
  w = [0.114, 0.587, 0.299] (videoio.py:55-60), and write_video's normalisation (a negative
  minimum shifted to 0, a maximum above 1 scaled to 1, 255 v truncated to uint8,
  videoio.py:98-112).  The container is not HDF5 / DIVX (neither library is present): its
  parity with the reference's files is unpinned; the layout and arithmetic are.
"""

import os
import types

import numpy as np
import pytest


def test_array_files_round_trip(tmp_path):
    from vtem_amd.utils import simpleio as io
    rng = np.random.default_rng(0)
    a = rng.standard_normal((5, 3, 4))
    f = str(tmp_path / "dsw.h5")
    io.write_memory_to_file(a, f)
    assert os.path.exists(f) and not os.path.exists(f + ".npz")
    np.testing.assert_array_equal(io.read_file(f), a)
    b = rng.standard_normal((2, 3, 4))
    io.write_memory_to_file(b, f, mode="a")
    np.testing.assert_array_equal(io.read_file(f), np.concatenate([a, b]))
    with pytest.raises(ValueError):
        io.write_memory_to_file(rng.standard_normal((2, 4, 3)), f, mode="a")
    c = rng.standard_normal((4, 6)) + 1j * rng.standard_normal((4, 6))
    g = str(tmp_path / "ds.h5")
    io.write_memory_to_file(c, g)
    got = io.read_file(g)
    assert got.dtype == np.complex128
    np.testing.assert_array_equal(got, c)
    assert sorted(io.read_nodes(g)) == ["imag", "real"]
    h = str(tmp_path / "arr.h5")
    io.write_array(c.astype(np.complex64), h)
    io.write_array(c.astype(np.complex64), h, mode="a")
    np.testing.assert_array_equal(io.read_array(h), np.concatenate([c, c]).astype(np.complex64))
    with pytest.raises(TypeError):
        io.write_memory_to_file(np.zeros(3, np.int16), f)


def _bank(N, rng, kind):
    b = types.SimpleNamespace(h_alpha=rng.uniform(0.5, 2, N), h_x0=np.round(rng.uniform(-8, 8, N), 2),
                              h_y0=np.round(rng.uniform(-8, 8, N), 2))
    if kind == "gabor":
        b.h_l = rng.uniform(0, np.pi, N)
        b.h_ab = rng.integers(0, 2, N).astype(np.int32)
        b.KAPPA = 2.5
    else:
        b.sigma_center, b.sigma_surround = 0.5, 0.8
    return b


@pytest.mark.parametrize("kind", ["gabor", "cs"])
def test_spike_file_round_trip(tmp_path, oracle, kind):
    from vtem_amd.utils import spike_store as ss
    rng = np.random.default_rng(1)
    N, bins, refresh = 40, 4, 0.3
    bank = _bank(N, rng, kind)
    # domain bounds placed exactly on two centres +- 0.01: those are not selected
    bank.h_x0[0], bank.h_x0[1] = -4.01, 4.01
    iaf = types.SimpleNamespace(h_kappa=rng.uniform(0.5, 1.5, N), h_delta=rng.uniform(0.02, 0.05, N),
                                h_bias=rng.uniform(1, 2, N), h_sigma=np.zeros(N))
    f = str(tmp_path / "spikes.h5")
    st = ss.vtem_storage(f, N, kind)
    st.write_video_attributes(2 * np.pi * 4, 1.25, 2.5, 64, 48, 0.5, 0.25)
    (st.write_gabor_parameters if kind == "gabor" else st.write_cs_parameters)(bank)
    st.write_neuron_parameters(iaf)
    st.write_refresh_interval(refresh)
    rel = []
    for i in range(bins):
        cnt = rng.integers(0, 6, N).astype(np.int32)
        sp = [np.sort(rng.uniform(0, refresh, c)) for c in cnt]
        rel.append(sp)
        st.write_spike_in_timebin(cnt, np.concatenate(sp))
    st.close()
    rd = ss.vtem_storage_read(f)
    assert rd.read_num_neurons() == N and rd.read_total_intervals() == bins
    Wt, Wx, Wy, Px, Py, dx, dy = rd.read_video_attributes()
    assert (Wt, Wx, Wy, Px, Py, dx, dy) == (2 * np.pi * 4, 2.5, 1.25, 64, 48, 0.5, 0.25)
    assert rd.read_vrf_type() == kind and rd.read_refresh_interval() == refresh
    v = rd.read_vrf()
    np.testing.assert_array_equal(v[0], bank.h_alpha)
    for a, b in zip(rd.read_iaf(), (iaf.h_kappa, iaf.h_delta, iaf.h_bias, iaf.h_sigma)):
        np.testing.assert_array_equal(a, b)
    dom = [-4.0, 4.0, -3.0, 5.0]
    rd.select_neurons(dom)
    ch, x0, y0 = oracle.select_neurons(bank.h_x0, bank.h_y0, dom)
    np.testing.assert_array_equal(rd.choose, ch)
    assert 0 not in ch and 1 not in ch
    np.testing.assert_array_equal(rd.x0, x0)
    np.testing.assert_array_equal(rd.y0, y0)
    got = rd.bin_spikes(range(1, 3))
    for j, nj in enumerate(ch):
        want = np.concatenate([rel[i][nj] + i * refresh for i in (1, 2)])
        np.testing.assert_array_equal(got[j], want)


@pytest.mark.parametrize("w", [7, 8])
def test_avi_round_trip_and_grey(tmp_path, w):
    from vtem_amd.utils import videoio as vio
    rng = np.random.default_rng(2)
    bgr = rng.integers(0, 256, (5, 6, w, 3)).astype(np.uint8)
    f = str(tmp_path / "v.avi")
    vio.write_avi(bgr, f, fps=25)
    back, fps = vio.read_avi(f)
    np.testing.assert_array_equal(back, bgr)
    assert fps == 25
    grey = vio.read_video(f)
    want = (bgr[..., 0] * 0.114 + bgr[..., 1] * 0.587 + bgr[..., 2] * 0.299) / 255
    np.testing.assert_allclose(grey, want, rtol=0, atol=1e-15)
    np.testing.assert_array_equal(vio.read_video(f, color=True), bgr.astype(np.float64))
    vio.read_video(f, write_to_file=True)
    from vtem_amd.utils import simpleio
    np.testing.assert_array_equal(simpleio.read_file(str(tmp_path / "v.h5")), grey)


def test_write_video_normalisation(tmp_path):
    from vtem_amd.utils import videoio as vio
    rng = np.random.default_rng(3)
    u = rng.standard_normal((4, 5, 9)) * 3
    f = str(tmp_path / "rec.avi")
    vio.write_video(u, f, fps=10)
    back, fps = vio.read_avi(f)
    v = u - u.min()
    v = v / v.max()
    q = (255 * v).astype(np.uint8)
    for c in range(3):
        np.testing.assert_array_equal(back[..., c], q)
    assert fps == 10
    with pytest.raises(NotImplementedError):
        vio.write_video(u, f, fourcc=0x58564944)


def test_avi_rejects_compressed(tmp_path):
    import struct
    from vtem_amd.utils import videoio as vio
    f = str(tmp_path / "c.avi")
    vio.write_avi(np.zeros((1, 2, 2, 3), np.uint8), f)
    data = bytearray(open(f, "rb").read())
    p = data.find(b"strf") + 8
    struct.pack_into("<I", data, p + 16, 0x58564944)      # biCompression 'DIVX'
    open(f, "wb").write(bytes(data))
    with pytest.raises(NotImplementedError):
        vio.read_avi(f)
