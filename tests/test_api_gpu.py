"""The reference-shaped API (vrf, ensemble_encode, vtem, dirichlet, neural_network,
utils.linalg, vtdm) on the GPU against the float64 oracle on the same inputs."""

import math

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def torch_mod():
    import torch
    from vtem_amd._lib import require_gpu
    require_gpu()
    return torch


def _banner(msg):
    print(f"\n[case] {msg}", flush=True)


def _rel(a, b):
    return float(np.abs(np.asarray(a) - np.asarray(b)).max() / max(1e-300, np.abs(np.asarray(b)).max()))


def _c0():
    from vtem_amd import synthetic as syn
    vid = syn.make_video(32, 32, 1.0, 100.0, 5, 5, 4, seed=0)
    bank = syn.bank_config0(32.0, 32.0)
    return vid, bank


def _gabor(bank, sel=None, shape=(32, 32)):
    from vtem_amd.vrf import vrf_gabor
    sel = np.arange(bank.num_neurons) if sel is None else np.asarray(sel)
    g = vrf_gabor(shape, dx=1.0, dy=1.0)
    g.load_parameters(len(sel), bank.h_alpha[sel], bank.h_l[sel], bank.h_x0[sel], bank.h_y0[sel],
                      bank.h_ab[sel], KAPPA=bank.KAPPA)
    return g


def test_vrf_gabor_api(torch_mod, oracle):
    """RFs, filtering, Ds (ellipse and rect), Dsw and psi of vrf_gabor."""
    _banner("vrf_gabor api")
    vid, bank = _c0()
    g = _gabor(bank)
    rf = g.generate_visual_receptive_fields().cpu().numpy()
    rf_o = oracle.rf_gabor(bank.h_alpha, bank.h_l, bank.h_x0, bank.h_y0, bank.h_ab, bank.KAPPA, 32, 32,
                           32.0, 32.0)
    assert _rel(rf, rf_o) < 1e-13
    v = g.filter(vid.frames).cpu().numpy()
    assert _rel(v, oracle.filter_video(vid.frames, rf_o)) < 1e-12
    W = 2 * math.pi * 5 / 32
    for mask in ("ellipse", "rect"):
        Ds = g.compute_Ds(5, 5, mask=mask)
        Ds_o = oracle.compute_Ds_fft(rf_o, 1.0, 1.0, 5, 5, mask)
        assert _rel(Ds.cpu().numpy(), Ds_o) < 1e-12, mask
    Dsw = g.compute_Dsw(Ds, 5, 5, np.ones(256)).cpu().numpy()
    assert _rel(Dsw, oracle.compute_Dsw(Ds_o, np.ones(256))) < 1e-12
    psi = g.compute_dirich_space(Ds, 5, 5, 32, 32, 32.0, 32.0, W, W).cpu().numpy()
    psi_o = oracle.compute_psi(Ds_o, 32, 32, 32.0, 32.0, W / 5, W / 5, mask="rect")
    assert _rel(psi, psi_o) < 1e-11
    # the reference's FFT form of psi (vrf.py:437-499) gives the same functions
    psi_f = g.compute_dirich_space_fft(Ds, 5, 5, 32, 32, 32.0, 32.0, W, W).cpu().numpy()
    assert _rel(psi_f, psi_o) < 1e-11


def test_vrf_cs_api(torch_mod, oracle):
    _banner("vrf_cs api")
    from vtem_amd import synthetic as syn
    from vtem_amd.vrf import vrf_cs
    bank = syn.dog_grid_bank(32.0, 32.0, 4)
    c = vrf_cs((32, 32), dx=1.0, dy=1.0)
    c.load_parameters(bank.num_neurons, bank.h_alpha, bank.h_x0, bank.h_y0, sigma_center=bank.sigma_center,
                      sigma_surround=bank.sigma_surround)
    rf = c.generate_visual_receptive_fields().cpu().numpy()
    rf_o = oracle.rf_dog(bank.h_alpha, bank.h_x0, bank.h_y0, bank.sigma_center, bank.sigma_surround, 32, 32,
                         32.0, 32.0)
    assert _rel(rf, rf_o) < 1e-13
    Ds = c.compute_Ds(4, 4).cpu().numpy()
    assert _rel(Ds, oracle.compute_Ds_fft(rf_o, 1.0, 1.0, 4, 4, "ellipse")) < 1e-12


def test_iaf_encode_api(torch_mod, oracle):
    """IAF_encode.encode equals the oracle bit for bit; two calls with reset_timer in
    between continue the membrane state (times relative to each call)."""
    _banner("IAF_encode api")
    from vtem_amd.ensemble_encode import IAF_encode
    vid, bank = _c0()
    v = oracle.filter_video(vid.frames, oracle.rf_bank(bank, 32, 32))
    N = v.shape[1]
    rng = np.random.default_rng(3)
    kappa, bias, delta = np.ones(N), np.full(N, 2.0), np.full(N, 0.04)
    v0 = rng.uniform(0, 1, N) * delta
    iaf = IAF_encode(N, 0.01)
    iaf.load_parameters(kappa, bias, delta, h_v0=v0)
    lists, cnt = iaf.encode(v)
    lo, co, *_ = oracle.iaf_encode(v, 0.01, kappa, bias, delta, v0)
    np.testing.assert_array_equal(cnt, co)
    for a, b in zip(lists, lo):
        np.testing.assert_array_equal(a, b)
    # chunked: samples [0, 60) then [59, 100) with the timer reset in between
    iaf2 = IAF_encode(N, 0.01)
    iaf2.load_parameters(kappa, bias, delta, h_v0=v0)
    l1, c1 = iaf2.encode(v[:60])
    iaf2.reset_timer()
    l2, c2 = iaf2.encode(v[59:])
    np.testing.assert_array_equal(c1 + c2, cnt)
    for j in range(N):
        joined = np.concatenate([l1[j], l2[j] + 0.59])
        np.testing.assert_allclose(joined, lists[j], rtol=0, atol=1e-12)
    with pytest.raises(ValueError):
        iaf.encode(v, avg_rate=50.0)
    # time-parallel form: same counts, times within 1e-9 s
    iaf3 = IAF_encode(N, 0.01, parallel=True)
    iaf3.load_parameters(kappa, bias, delta, h_v0=v0)
    l3, c3 = iaf3.encode(v)
    np.testing.assert_array_equal(c3, co)
    assert max(np.abs(a - b).max(initial=0) for a, b in zip(l3, lo)) < 1e-9


def test_vtem_iaf_chunks(torch_mod, oracle):
    """VTEM_IAF in 30-frame chunks equals one encode of the whole filtered video."""
    _banner("VTEM_IAF chunks")
    from vtem_amd.ensemble_encode import IAF_encode
    from vtem_amd.vtem import VTEM_IAF
    vid, bank = _c0()
    g = _gabor(bank)
    N = bank.num_neurons
    rng = np.random.default_rng(4)
    kappa, bias, delta = np.ones(N), np.full(N, 2.0), np.full(N, 0.04)
    v0 = rng.uniform(0, 1, N) * delta
    iaf = IAF_encode(N, 0.01)
    iaf.load_parameters(kappa, bias, delta, h_v0=v0)
    enc = VTEM_IAF(vid.frames, None, 2 * math.pi * 4, g, dx=1.0, dy=1.0, npinput=True, iaf=iaf,
                   chunk_frames=30)
    lo, co, *_ = oracle.iaf_encode(oracle.filter_video(vid.frames, oracle.rf_bank(bank, 32, 32)), 0.01,
                                   kappa, bias, delta, v0)
    np.testing.assert_array_equal(enc.spike_count, co)
    for a, b in zip(enc.spikes, lo):
        np.testing.assert_allclose(a, b, rtol=0, atol=1e-12)


@pytest.mark.parametrize("kind", ["gabor", "cs"])
def test_vtem_default_bank_drivers(torch_mod, kind):
    """VTEM_Gabor_IAF / VTEM_CS_IAF (vtem/vtem.py:209-427) build the reference's default
    bank on the video's grid and encode with it: the same spikes as VTEM_IAF called with
    that bank and the same IAF state."""
    _banner(f"VTEM default-bank driver {kind}")
    from vtem_amd import vrf as vrf_mod
    from vtem_amd.ensemble_encode import IAF_encode
    from vtem_amd.vtem import VTEM_CS_IAF, VTEM_Gabor_IAF, VTEM_IAF
    vid, _ = _c0()
    frames = vid.frames[:40]
    drv = VTEM_Gabor_IAF if kind == "gabor" else VTEM_CS_IAF
    cls = vrf_mod.vrf_gabor if kind == "gabor" else vrf_mod.vrf_cs
    rf = cls((32, 32), dx=1.0 / 16, dy=1.0 / 16)
    rf.load_parameters()
    N = rf.num_neurons
    rng = np.random.default_rng(8)
    v0 = rng.uniform(0, 1, N) * 0.03

    def iaf():
        e = IAF_encode(N, 0.01)
        e.load_parameters(np.ones(N), np.full(N, 0.8), np.full(N, 0.03), h_v0=v0)
        return e

    a = drv(frames, None, 2 * math.pi * 4, npinput=True, h5input=False, iaf=iaf())
    b = VTEM_IAF(frames, None, 2 * math.pi * 4, rf, dx=1.0 / 16, dy=1.0 / 16, npinput=True,
                 h5input=False, iaf=iaf())
    assert N > 0 and a.spike_count.sum() > 0
    np.testing.assert_array_equal(a.spike_count, b.spike_count)
    for x, y in zip(a.spikes, b.spikes):
        np.testing.assert_array_equal(x, y)


def _spike_space_inputs(oracle, sub):
    vid, bank = _c0()
    RF = oracle.rf_bank(bank, 32, 32)
    v = oracle.filter_video(vid.frames, RF)
    N = v.shape[1]
    kappa, bias, delta = np.ones(N), np.full(N, 2.0), np.full(N, 0.04)
    v0 = np.random.default_rng(5).uniform(0, 1, N) * delta
    lists, *_ = oracle.iaf_encode(v, 0.01, kappa, bias, delta, v0)
    return vid, bank, RF, lists, kappa, bias, delta


def test_dirichlet_pinv_rnn_api(torch_mod, oracle):
    """dirichlet (q, G, reconstruct), solve_eq_sym and rnn3 on 16 neurons vs the oracle."""
    _banner("dirichlet / solve_eq_sym / rnn3 api")
    torch = torch_mod
    from vtem_amd import dirichlet as dr
    from vtem_amd.neural_network import rnn3
    from vtem_amd.utils.linalg import solve_eq_sym
    sub = np.arange(0, 256, 16)
    vid, bank, RF, lists, kappa, bias, delta = _spike_space_inputs(oracle, sub)
    ls_ = [lists[j] for j in sub]
    tk1, tk2, nid, _ = oracle.select_intervals(ls_, (0.0, 1.0))
    norm = np.ones(sub.size)
    Ds = oracle.compute_Ds_fft(RF[sub], 1.0, 1.0, 5, 5, "ellipse")
    Dsw = oracle.compute_Dsw(Ds, norm)
    Wt, Mt = 2 * math.pi * 4, 4
    cp = dr.dirichlet(sub.size, tk1, tk2, nid, kappa[sub], delta[sub], bias[sub], norm, Wt, Mt)
    q = cp.compute_q().cpu().numpy()
    q_o = oracle.compute_q(tk1, tk2, nid, kappa[sub], delta[sub], bias[sub], norm)
    assert _rel(q, q_o) < 1e-14
    G = cp.compute_G(Dsw, 0.0)
    G_o = oracle.compute_G_spike(tk1, tk2, nid, Dsw, Wt, Mt)
    assert _rel(G.cpu().numpy(), G_o) < 1e-13
    c = solve_eq_sym(G.clone(), cp.d_q).cpu().numpy()
    c_o = oracle.solve_eq_sym(G_o, q_o)
    # compare in the range of G (the truncated pseudoinverse is unique there)
    assert np.linalg.norm(G_o @ (c - c_o)) <= 1e-8 * np.linalg.norm(G_o @ c_o)
    cr = rnn3(G, cp.d_q, dt=1e-6, alpha=5000, steps=150).cpu().numpy()
    assert _rel(cr, oracle.rnn3(G_o, q_o, 1e-6, 5000, 150)) < 1e-12
    c_last, snaps = rnn3(G, cp.d_q, dt=1e-6, alpha=5000, steps=150, XOUTPUTSTEPS=5)
    assert _rel(c_last.cpu().numpy(), cr) < 1e-14
    assert _rel(snaps[2].cpu().numpy(), oracle.rnn3(G_o, q_o, 1e-6, 5000, 90)) < 1e-12
    cp.d_q = torch.as_tensor(c_o, device="cuda")
    psi_o = oracle.compute_psi(Ds, 32, 32, 32.0, 32.0, 2 * math.pi / 32, 2 * math.pi / 32)
    t = np.arange(0, 1.0, 0.01)
    u = cp.reconstruct(psi_o, [0, 1.0], 0.01).cpu().numpy()
    u_o = oracle.reconstruct_spike(c_o, tk1, tk2, nid, norm, psi_o, t, Mt, Wt / Mt)
    assert _rel(u, u_o) < 1e-11


def _oracle_window_decoder(oracle, lists, sel, kappa, delta, bias, Dsw, psi, Wt, Mt, dt, norm=None):
    norm = np.ones(len(sel)) if norm is None else norm

    def decode(start, end):
        ls_ = [lists[j] for j in sel]
        tk1, tk2, nid, _ = oracle.select_intervals(ls_, (start, end))
        q = oracle.compute_q(tk1, tk2, nid, kappa[sel], delta[sel], bias[sel], norm)
        G = oracle.compute_G_spike(tk1, tk2, nid, Dsw, Wt, Mt)
        c = oracle.solve_eq_sym(G, q)
        t = np.arange(0, round((end - start) / dt) * dt, dt)
        return oracle.reconstruct_spike(c, tk1, tk2, nid, norm, psi, t, Mt, Wt / Mt)
    return decode


def test_vtdm_windows_vs_oracle(torch_mod, oracle):
    """VTDM with temporal stitching (0.4 s windows, Hann crossfade) on 64 neurons of the
    32x32 bank equals the oracle's window loop with the oracle's spike-space decode.
    (64 receptive fields cannot resolve the 81 spatial modes, so the reconstruction SNR
    itself is low -- in the oracle too; the test is about parity.)"""
    _banner("VTDM windows")
    from vtem_amd import vtdm
    from vtem_amd.ensemble_encode import IAF_encode
    from vtem_amd.vtem import VTEM_IAF
    vid, bank = _c0()
    sel_all = np.arange(0, 256, 4)
    g = _gabor(bank, sel_all)
    N = sel_all.size
    rng = np.random.default_rng(6)
    kappa, bias, delta = np.ones(N), np.full(N, 2.0), np.full(N, 0.04)
    iaf = IAF_encode(N, 0.01)
    iaf.load_parameters(kappa, bias, delta, h_v0=rng.uniform(0, 1, N) * delta)
    W = 2 * math.pi * 5 / 32
    Wt = 2 * math.pi * 4
    enc = VTEM_IAF(vid.frames, None, Wt, g, W, W, dx=1.0, dy=1.0, npinput=True, iaf=iaf)
    Dsw, dirich, sel = vtdm.VTDM_prep(enc, None, None, 5, 5)
    assert sel.size == N
    u = vtdm.VTDM(enc, Dsw, dirich, 0.0, 1.0, 0.01, 5, 5, Mt=4, stitching=True, stitch_interval=0.4,
                  output_format=2, neurons=sel)
    rf_o = oracle.rf_gabor(bank.h_alpha[sel_all], bank.h_l[sel_all], bank.h_x0[sel_all],
                           bank.h_y0[sel_all], bank.h_ab[sel_all], bank.KAPPA, 32, 32, 32.0, 32.0)
    Ds_o = oracle.compute_Ds_fft(rf_o, 1.0, 1.0, 5, 5, "ellipse")
    Dsw_o = oracle.compute_Dsw(Ds_o, np.ones(N))
    psi_o = oracle.compute_psi(Ds_o, 32, 32, 32.0, 32.0, W / 5, W / 5)
    assert _rel(Dsw.cpu().numpy(), Dsw_o) < 1e-12 and _rel(dirich.cpu().numpy(), psi_o) < 1e-11
    dec = _oracle_window_decoder(oracle, enc.spikes, np.arange(N), kappa, delta, bias, Dsw_o, psi_o, Wt, 4,
                                 0.01)
    u_o = oracle.vtdm_windows(dec, 0.0, 1.0, 0.01, True, 0.4)
    s, s_o = oracle.snr_db(vid.frames, u), oracle.snr_db(vid.frames, u_o)
    print(f"VTDM windows: SNR gpu {s:.4f} dB oracle {s_o:.4f} dB, rel diff {_rel(u, u_o):.2e}")
    assert abs(s - s_o) < 0.1
    assert np.linalg.norm(u - u_o) <= 1e-6 * np.linalg.norm(u_o)


def test_decode_video_tiles_vs_oracle(torch_mod, oracle):
    """decode_video with spatial stitching (32-px tiles on a 48x48 video, 20 % overlap,
    Hann windows) equals the oracle's tile loop with the oracle's decode per tile."""
    _banner("decode_video tiles")
    from vtem_amd import synthetic as syn
    from vtem_amd import vtdm
    from vtem_amd.ensemble_encode import IAF_encode
    from vtem_amd.vtem import VTEM_IAF
    vid = syn.make_video(48, 48, 0.3, 100.0, 4, 4, 2, seed=1)
    bank = syn.gabor_grid_bank(48.0, 48.0, 4.0, 6, n_orient=4)
    g = _gabor(bank, shape=(48, 48))
    N = bank.num_neurons
    rng = np.random.default_rng(7)
    kappa, bias, delta = np.ones(N), np.full(N, 2.0), np.full(N, 0.04)
    iaf = IAF_encode(N, 0.01)
    iaf.load_parameters(kappa, bias, delta, h_v0=rng.uniform(0, 1, N) * delta)
    Mx, Wt, Mt = 4, 2 * math.pi * 10 / 3, 1
    W = 2 * math.pi * Mx / 32
    enc = VTEM_IAF(vid.frames, None, Wt, g, W, W, dx=1.0, dy=1.0, npinput=True, iaf=iaf)
    u = vtdm.decode_video(enc, None, None, 0.0, 0.3, 0.01, Mx, Mt=Mt, spatial_stitching=True,
                          spatial_interval=[32, 32], output_format=2)

    def tile(dom):
        sel, x0, y0 = oracle.select_neurons(bank.h_x0, bank.h_y0, dom)
        rf_o = oracle.rf_gabor(bank.h_alpha[sel], bank.h_l[sel], x0, y0, bank.h_ab[sel], bank.KAPPA, 32, 32,
                               32.0, 32.0)
        Ds_o = oracle.compute_Ds_fft(rf_o, 1.0, 1.0, Mx, Mx, "ellipse")
        Dsw_o = oracle.compute_Dsw(Ds_o, np.ones(sel.size))
        px, py = int(round(dom[1] - dom[0])), int(round(dom[3] - dom[2]))
        psi_o = oracle.compute_psi(Ds_o, px, py, float(px), float(py), W / Mx, W / Mx)
        dec = _oracle_window_decoder(oracle, enc.spikes, sel, kappa, delta, bias, Dsw_o, psi_o, Wt, Mt, 0.01)
        return oracle.vtdm_windows(dec, 0.0, 0.3, 0.01)

    u_o = oracle.decode_tiles(tile, [-24.0, 24.0, -24.0, 24.0], 1.0, 1.0, [32, 32], 30)
    print(f"decode_video tiles: rel diff {_rel(u, u_o):.2e} SNR gpu {oracle.snr_db(vid.frames, u):.3f} "
          f"oracle {oracle.snr_db(vid.frames, u_o):.3f}")
    assert np.linalg.norm(u - u_o) <= 1e-6 * np.linalg.norm(u_o)


def test_iaf_rt_encode_and_weighted_decode(torch_mod, oracle):
    """Random thresholds end to end.  IAF_encode_rt (sigma != 0, per-neuron kappa) equals
    the oracle's restatement of vtem/ensemble_encode.py:555-674 on the same N(0,1) draws:
    counts identical, times within 1e-9 s (bit for bit in fact, same loop).  Then VTDM_prep
    + VTDM over a sub-domain decode those spikes with the weights 1/(kappa sigma) normalised
    over ALL neurons and baked into Dsw (vtem/vtdm.py:408-413, :444) and applied to q and
    the reconstruction (:634-638): equal to the oracle's window decode with the same
    weights."""
    _banner("IAF_encode_rt + weighted decode")
    from vtem_amd import vtdm
    from vtem_amd.ensemble_encode import IAF_encode_rt
    from vtem_amd.vtem import VTEM_IAF
    vid, bank = _c0()
    g = _gabor(bank)
    N = bank.num_neurons
    rng = np.random.default_rng(12)
    kappa = rng.uniform(0.8, 1.25, N)
    bias = np.full(N, 2.0) * kappa
    mu = np.full(N, 0.04)
    sigma = rng.uniform(0.002, 0.006, N)
    dv0 = mu + sigma * rng.standard_normal(N)
    v0 = rng.uniform(0, 1, N) * dv0
    # encoder API vs the oracle on one call of filtered frames
    v = oracle.filter_video(vid.frames, oracle.rf_bank(bank, 32, 32))
    iaf = IAF_encode_rt(N, 0.01, rng=np.random.default_rng(40))
    iaf.load_parameters(kappa, bias, mu, sigma, dv0, h_v0=v0)
    lists, cnt = iaf.encode(v)
    draws = np.random.default_rng(40).normal(size=(int(np.ceil(v.shape[0] / 0.1)), N))
    lo, co, v0o, tco, dvo = oracle.iaf_encode_rt(v, 0.01, kappa, bias, mu, sigma, dv0, v0, draws)
    np.testing.assert_array_equal(cnt, co)
    assert max(np.abs(a - b).max(initial=0) for a, b in zip(lists, lo)) <= 1e-9
    np.testing.assert_allclose(iaf.d_delta_value.cpu().numpy(), dvo, rtol=0, atol=0)
    # VTEM_IAF with the rt population, then a weighted decode of a sub-domain
    iaf2 = IAF_encode_rt(N, 0.01, rng=np.random.default_rng(41))
    iaf2.load_parameters(kappa, bias, mu, sigma, dv0, h_v0=v0)
    W = 2 * math.pi * 5 / 32
    Wt = 2 * math.pi * 4
    enc = VTEM_IAF(vid.frames, None, Wt, g, W, W, dx=1.0, dy=1.0, npinput=True, iaf=iaf2)
    dom = [-16.0, 8.0, -16.0, 16.0]
    Dsw, dirich, sel = vtdm.VTDM_prep(enc, None, None, 5, 5, domain=dom)
    ch, x0, y0 = oracle.select_neurons(bank.h_x0, bank.h_y0, dom)
    np.testing.assert_array_equal(sel, ch)
    assert 0 < sel.size < N
    u = vtdm.VTDM(enc, Dsw, dirich, 0.0, 1.0, 0.01, 5, 5, Mt=4, domain=dom, output_format=2, neurons=sel)
    norm = oracle.decode_norm(kappa, sigma)[sel]
    assert norm.max() < 1.0 or np.argmax(oracle.decode_norm(kappa, sigma)) in sel
    rf_o = oracle.rf_gabor(bank.h_alpha[sel], bank.h_l[sel], x0, y0, bank.h_ab[sel], bank.KAPPA,
                           32, 32, 32.0, 32.0)
    Ds_o = oracle.compute_Ds_fft(rf_o, 1.0, 1.0, 5, 5, "ellipse")
    Dsw_o = oracle.compute_Dsw(Ds_o, norm)
    assert _rel(Dsw.cpu().numpy(), Dsw_o) < 1e-12
    psi_o = oracle.compute_psi(Ds_o, 24, 32, 24.0, 32.0, W / 5, W / 5)
    assert _rel(dirich.cpu().numpy(), psi_o) < 1e-11
    dec = _oracle_window_decoder(oracle, enc.spikes, sel, kappa, mu, bias, Dsw_o, psi_o, Wt, 4, 0.01,
                                 norm=norm)
    u_o = oracle.vtdm_windows(dec, 0.0, 1.0, 0.01)
    s, s_o = oracle.snr_db(vid.frames[:, :, :24], u), oracle.snr_db(vid.frames[:, :, :24], u_o)
    print(f"rt decode: rel diff {_rel(u, u_o):.2e}, SNR gpu {s:.4f} oracle {s_o:.4f} dB", flush=True)
    # the truncated pseudoinverse keeps singular values down to 1e-8 of the largest, so two
    # correct SVDs (one-sided Jacobi here, numpy's eigh in the oracle) differ by up to 1e8
    # times their rounding in those directions: 1e-5 relative, and the 0.1 dB bar
    assert np.linalg.norm(u - u_o) <= 1e-5 * np.linalg.norm(u_o)
    assert abs(s - s_o) < 0.1
