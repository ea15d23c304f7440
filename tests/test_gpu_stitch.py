"""Stitched encode/decode (the batching form) on the GPU: a 128 x 128 px, 2 s video in
64 px tiles (16 px overlap) and 1 s windows (0.25 s overlap), 27 blocks of D = 4913 with
2048 neurons per tile, per-tile contrast multipliers (input gains in [0.25, 4]).  The blocks decoded together by the
batched KPCG decoder and blended give the same video, within 0.1 dB SNR, as every block
decoded by the plain per-block FP64 path and blended the same way.  The video has the
blocks' own periods (64 px, 1 s), so each block's basis holds it exactly and the blended
video must reconstruct it (SNR floor); a misplaced tile or window would break that."""

import math

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _snr(u, v):
    return 10 * math.log10(float(np.sum(u * u)) / float(np.sum((u - v) ** 2)))


def test_stitched_codec_matches_plain_blocks():
    import torch
    from vtem_amd._lib import require_gpu
    require_gpu()
    from vtem_amd import synthetic as syn
    from vtem_amd.stitch import TiledCodec, stitch_accumulate
    print("\n[case] stitched codec 128^2 x 2 s, 27 blocks", flush=True)
    # a video with the blocks' own periods (64 px, 1 s), so that every block's basis holds
    # it exactly and the blended result must reconstruct it well
    rng = np.random.default_rng(9)
    c = syn.video_coefficients(8, 8, 8, rng)
    frames = syn.evaluate_video(c, 128, 128, np.arange(200) / 100.0, Sx=64.0, Sy=64.0, T=1.0)
    frames /= np.abs(frames).max()
    codec = TiledCodec(128, 128, 2.0, 16, syn.bank_dyadic(64.0, 64.0))
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
    counts = [int(s.counts.sum().item()) for s in spikes]
    print(f"spikes per tile {min(counts)}..{max(counts)}", flush=True)
    # plain per-block FP64 decodes, blended the same way
    p = codec.plan
    T, W = p.tile, p.window
    ref = torch.zeros_like(V)
    blocks = []
    for k, (iy, ix) in enumerate(p.tiles):
        for it, f0 in enumerate(p.f0):
            r = decs[k].decode(spikes[k], f0 / 100.0, precision="fp64")
            blocks.append(((iy, ix, it), r.frames.view(W, T, T)))
    for par in range(8):
        items = [(fr, codec.wt[it], codec.wy[iy], codec.wx[ix], p.f0[it], p.y0[iy], p.x0[ix])
                 for (iy, ix, it), fr in blocks
                 if (iy & 1) == (par & 1) and (ix & 1) == ((par >> 1) & 1) and (it & 1) == ((par >> 2) & 1)]
        if items:
            stitch_accumulate(ref, items)
    torch.cuda.synchronize()
    ref = ref.cpu().numpy()
    s_got, s_ref = _snr(frames, got), _snr(frames, ref)
    print(f"SNR stitched kpcg {s_got:.3f} dB, plain fp64 {s_ref:.3f} dB", flush=True)
    assert abs(s_got - s_ref) < 0.1
    assert s_ref > 25.0
