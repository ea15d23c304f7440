"""Stitched encode/decode (the batching form) on the GPU.

Blocks are 64 px tiles (16 px overlap) x 1 s windows (0.25 s overlap) with 2048 neurons
per tile (config-1 laws) and per-tile contrast multipliers (input gains in [0.25, 4]).
Each block is decoded in the Dirichlet space of orders 8, 8, 8 whose periods follow from
the orders and the video's bandwidth, 2 pi M / W (vtem/vtdm.py:385-388, :964-967): 4/3 of
the tile and of the window (85.3 px, 1.33 s), so a block sits inside one period with
margin.

* 128 x 128 px, 4 s, a video inside every block's space (orders 8, 8, 8 with the decoder's
  periods; 3 temporal periods over the 4 s, so the encoder's trigonometric input of order
  24 over 4 s holds it exactly).  The blocks decoded together by the batched KPCG decoder and blended equal
  every block decoded by the plain per-block FP64 path and blended the same way (within
  0.1 dB), and the blend reconstructs the video (SNR floor).
* The full 512 x 512 px, 4 s plan (605 blocks) of a video with orders 48, 48, 24 (the
  bandwidth of the blocks' space, not periodic in any block): two tiles' encodes and
  three blocks' decodes against the oracle, and the interior SNR of the blended video."""

import math

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _snr(u, v):
    return 10 * math.log10(float(np.sum(u * u)) / float(np.sum((u - v) ** 2)))


def _blend(codec, blocks, like):
    """Blend {(iy, ix, it): frames} with the codec's weights, in the plan's passes."""
    from vtem_amd.stitch import stitch_accumulate
    import torch
    p = codec.plan
    out = torch.zeros_like(like)
    keep = []
    for grp in p.passes():
        items = [(blocks[b], codec.wt[b[2]], codec.wy[b[0]], codec.wx[b[1]], p.f0[b[2]], p.y0[b[0]],
                  p.x0[b[1]]) for b in grp]
        keep.append(stitch_accumulate(out, items))
    torch.cuda.synchronize()
    return out


def test_stitched_codec_matches_plain_blocks():
    import torch
    from vtem_amd._lib import require_gpu
    require_gpu()
    from vtem_amd import synthetic as syn
    from vtem_amd.stitch import TiledCodec
    print("\n[case] stitched codec 128^2 x 4 s, 45 blocks", flush=True)
    rng = np.random.default_rng(9)
    # periods 256/3 px and 4/3 s (the decoder's); over the 4 s encode the temporal
    # harmonics m 0.75 Hz are the order-24 harmonics of 0.25 Hz
    c = syn.video_coefficients(8, 8, 8, rng)
    frames = syn.evaluate_video(c, 128, 128, np.arange(400) / 100.0, Sx=256 / 3, Sy=256 / 3, T=4 / 3)
    frames /= np.abs(frames).max()
    codec = TiledCodec(128, 128, 4.0, 24, syn.bank_dyadic(64.0, 64.0),
                       Wx=2 * math.pi * 8 / (256 / 3), Wt=2 * math.pi * 8 / (4 / 3))
    assert abs(codec.dec_geom.Ox - 2 * math.pi * 3 / 256) < 1e-15 and abs(codec.dec_geom.W - 1.5 * math.pi) < 1e-14
    ntile = len(codec.plan.tiles)
    iaf = [syn.iaf_params(codec.encoder.RF.shape[0], 2.5, 1.0, 0.04, 0.004, rng) for _ in range(ntile)]
    contrast = np.exp(rng.uniform(math.log(0.25), math.log(4.0), ntile))
    V = torch.tensor(frames, dtype=torch.float64, device="cuda")
    spikes = codec.encode(V, iaf, contrast)
    # resident IAF parameters (built once) encode the same spikes
    again = codec.encode(V, iaf, contrast, devices=codec.iaf_devices(iaf, contrast))
    for a, b in zip(spikes, again):
        ca = a.counts.cpu().numpy()
        np.testing.assert_array_equal(ca, b.counts.cpu().numpy())
        sa, sb = a.spikes.cpu().numpy(), b.spikes.cpu().numpy()
        for j in range(ca.size):
            np.testing.assert_array_equal(sa[:ca[j], j], sb[:ca[j], j])
    decs = codec.decoders(iaf, contrast)
    got = codec.decode(spikes, decs).cpu().numpy()
    print(f"decode stats {codec.last_stats}", flush=True)
    # blocks decoded inside a longer temporal period get the Kronecker-sum preconditioner
    assert codec.last_stats["preconditioner"] == "kron_sum"
    counts = [int(s.counts.sum().item()) for s in spikes]
    print(f"spikes per tile {min(counts)}..{max(counts)}", flush=True)
    # plain per-block FP64 decodes, blended the same way
    p = codec.plan
    T, W = p.tile, p.window
    blocks = {}
    for k, (iy, ix) in enumerate(p.tiles):
        for it, f0 in enumerate(p.f0):
            blocks[(iy, ix, it)] = decs[k].decode(spikes[k], f0 / 100.0, precision="fp64").frames.view(W, T, T)
    ref = _blend(codec, blocks, V).cpu().numpy()
    s_got, s_ref = _snr(frames, got), _snr(frames, ref)
    print(f"SNR stitched kpcg {s_got:.3f} dB, plain fp64 {s_ref:.3f} dB", flush=True)
    assert abs(s_got - s_ref) < 0.1
    assert s_ref > 60.0


def _oracle_tile_encode(oracle, codec, frames, iy, ix, pk):
    g = codec.enc_geom
    p = codec.plan
    y0, x0 = p.y0[iy], p.x0[ix]
    tv = frames[:, y0:y0 + p.tile, x0:x0 + p.tile].reshape(frames.shape[0], -1)
    RF = oracle.rf_bank(codec.bank, p.tile, p.tile)
    vu = oracle.upsample(oracle.filter_video(tv, RF), g.Lt, g.duration, g.fps, g.iaf_times())
    lists, cnt, *_ = oracle.iaf_encode(vu, g.dt, pk.kappa, pk.bias, pk.delta, pk.v0, root_mode=1,
                                       max_spike=4000)
    return RF, lists, cnt


def _oracle_block_decode(oracle, codec, RF, lists, pk, it):
    """The oracle's FP64 D-space Tikhonov decode of one (tile, window) block."""
    g = codec.dec_geom
    t0 = codec.plan.f0[it] / codec.fps
    tk1, tk2, nid, _ = oracle.select_intervals(lists, (t0, t0 + g.duration))
    N = RF.shape[0]
    q = oracle.compute_q(tk1, tk2, nid, pk.kappa, pk.delta, pk.bias, np.ones(N))
    x, y = oracle._grid(g.Px, g.Py, g.Sx, g.Sy)
    B = oracle.spatial_basis(x, y, oracle.spatial_modes(g.Lx, g.Ly), g.Ox, g.Oy)
    S = oracle.spatial_coeffs(RF, B)
    Tm, tv = oracle.neuron_T(oracle.temporal_factors(tk1, tk2, g.Lt, g.W), q, nid, N)
    A = oracle.gram_kron(S, Tm)
    X = oracle.solve_tikhonov(A, oracle.rhs_kron(S, tv), oracle.tikhonov_lambda(A, tk1.size))
    return oracle.synthesize(X, B, oracle.temporal_basis(g.frame_times(), g.Lt, g.W)).reshape(
        g.frames, g.Py, g.Px)


def test_config3_full_plan_vs_oracle(oracle):
    import torch
    from vtem_amd._lib import require_gpu
    require_gpu()
    from vtem_amd import synthetic as syn
    from vtem_amd.batch import BatchDecoder, BlockJob
    from vtem_amd.stitch import TiledCodec
    print("\n[case] stitched codec 512^2 x 4 s, 605 blocks", flush=True)
    vid = syn.make_video(512, 512, 4.0, 100.0, 48, 48, 24, seed=0)
    rng = np.random.default_rng(3)
    codec = TiledCodec(512, 512, 4.0, 24, syn.bank_dyadic(64.0, 64.0),
                       Wx=2 * math.pi * 48 / 512, Wt=2 * math.pi * 24 / 4.0)
    p = codec.plan
    assert len(p.blocks) == 605
    n = len(p.tiles)
    N = codec.encoder.RF.shape[0]
    iaf = [syn.iaf_params(N, 2.5, 1.0, 0.04, 0.004, rng) for _ in range(n)]
    contrast = np.exp(rng.uniform(math.log(0.25), math.log(4.0), n))
    pks = codec.tile_params(iaf, contrast)
    V = torch.tensor(vid.frames, dtype=torch.float64, device="cuda")
    spikes = codec.encode(V, iaf, contrast)
    decs = codec.decoders(iaf, contrast)
    got = codec.decode(spikes, decs).cpu().numpy()
    st = codec.last_stats
    print(f"blocks {st['blocks']}, kpcg retries {st['kpcg_retries']}, fp64 fallbacks {st['fallbacks']}",
          flush=True)
    inner = (slice(25, -25), slice(16, -16), slice(16, -16))
    s_all, s_in = _snr(vid.frames, got), _snr(vid.frames[inner], got[inner])
    print(f"stitched SNR {s_all:.2f} dB, interior {s_in:.2f} dB", flush=True)
    assert s_in > 30.0
    assert st["preconditioner"] == "kron_sum"
    # two tiles' encodes (a corner tile with the clamped last column, an interior tile)
    for (iy, ix) in [(0, len(p.x0) - 1), (5, 4)]:
        k = p.tiles.index((iy, ix))
        RF, lists, cnt = _oracle_tile_encode(oracle, codec, vid.frames, iy, ix, pks[k])
        gc, gs = spikes[k].counts.cpu().numpy(), spikes[k].spikes.cpu().numpy()
        np.testing.assert_array_equal(gc, cnt)
        err = max(float(np.abs(gs[:cnt[j], j] - lists[j]).max()) for j in range(N) if cnt[j])
        print(f"  tile {(iy, ix)} gain {contrast[k]:.3f}: {int(cnt.sum())} spikes, max |dt| {err:.2e} s",
              flush=True)
        assert err <= 1e-9
    # three blocks: FP64 per-block and batched kpcg against the oracle's FP64 decode
    T, W = p.tile, p.window
    for (iy, ix, it) in [(0, len(p.x0) - 1, 0), (5, 4, 2), (10, 0, 4)]:
        k = p.tiles.index((iy, ix))
        lists = spikes[k].to_lists()
        RF = oracle.rf_bank(codec.bank, T, T)
        u_o = _oracle_block_decode(oracle, codec, RF, lists, pks[k], it)
        t0 = p.f0[it] / codec.fps
        u64 = decs[k].decode(spikes[k], t0, precision="fp64").frames.view(W, T, T).cpu().numpy()
        ukp = BatchDecoder(precision="kpcg", cg_iters=160).decode(
            [BlockJob(decs[k], spikes[k], t0=t0)])[0].frames.view(W, T, T).cpu().numpy()
        truth = vid.frames[p.f0[it]:p.f0[it] + W, p.y0[iy]:p.y0[iy] + T, p.x0[ix]:p.x0[ix] + T]
        so, s64, skp = _snr(truth, u_o), _snr(truth, u64), _snr(truth, ukp)
        rel = float(np.linalg.norm(u64 - u_o) / np.linalg.norm(u_o))
        print(f"  block {(iy, ix, it)}: SNR oracle {so:.3f} fp64 {s64:.3f} kpcg {skp:.3f} dB, "
              f"fp64 vs oracle rel {rel:.2e}", flush=True)
        assert abs(s64 - so) < 0.1 and abs(skp - so) < 0.1
        assert rel < 1e-6
