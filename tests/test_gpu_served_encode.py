"""The served encoder against the oracle.

BatchEncoder is called exactly as the server calls it (Server._encode: one batched call
over the batch's blocks, carry=False, an overflow flag per block, the stored-input walk)
on one block of every serving class -- {32, 48, 64, 96}^2 px x {0.5, 1, 2} s, Lx = Ly =
round(side/8), Lt = round(4 dur), N = round(2D/9) neurons of the config-1 laws, IAF
b = 2.5, kappa = 1, delta ~ N(0.04, 0.004^2) -- and on the 64^2 px, 2 s, 2048-neuron block
of the plain-method configuration (Lx = Ly = Lt = 8).  The oracle filters every frame
with the receptive fields, upsamples the filter outputs by exact trigonometric
interpolation and runs the serial IAF loop in C (vtem/ensemble_encode.py:486-545), in the
same root mode (1, the cancellation-free form the served path uses).

Bar: per-neuron spike counts identical; spike times within 1e-9 s."""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _banner(msg):
    print(f"\n[case] {msg}", flush=True)


def _blocks():
    from vtem_amd import synthetic as syn
    out = []
    k = 0
    for side in syn.SERVE_SIDES:
        for dur in syn.SERVE_DURATIONS:
            blk = syn.serve_block(side, dur, seed=7000 + k)
            out.append((f"serve {side}^2 x {dur} s", blk.video, blk.bank, blk.iaf, blk.Lx, blk.Lt, dur))
            k += 1
    vid = syn.make_video(64, 64, 2.0, 100.0, 8, 8, 8, seed=31)
    bank = syn.bank_dyadic(64.0, 64.0)
    iaf = syn.iaf_params(bank.num_neurons, 2.5, 1.0, 0.04, 0.004, np.random.default_rng(32))
    out.append(("plain-method 64^2 x 2 s, 2048 neurons", vid, bank, iaf, 8, 8, 2.0))
    return out


def test_served_encoder_vs_oracle(oracle):
    import torch
    from vtem_amd._lib import require_gpu
    from vtem_amd.basis import Geometry
    from vtem_amd.batch_encode import BatchEncoder, EncodeJob, IAFDevice, max_spikes_rate
    from vtem_amd.block import Encoder, dev
    require_gpu()
    blocks = _blocks()
    jobs = []
    for name, vid, bank, iaf, L, Lt, dur in blocks:
        side = vid.Px
        geom = Geometry(side, side, L, L, Lt, dur)
        enc = Encoder(geom, bank)
        jobs.append(EncodeJob(enc, dev(vid.frames.reshape(geom.frames, -1)), IAFDevice.from_params(iaf),
                              max_spikes_rate(geom, iaf)))
    _banner(f"served encoder, {len(jobs)} blocks in one batched call")
    be = BatchEncoder()
    assert be.root_mode == 1
    flags = torch.zeros(len(jobs), dtype=torch.int32, device="cuda")
    sts = be.encode(jobs, flags, carry=False)
    torch.cuda.synchronize()
    assert int(flags.sum().item()) == 0
    worst = 0.0
    for (name, vid, bank, iaf, L, Lt, dur), job, st in zip(blocks, jobs, sts):
        g = job.encoder.geom
        RF = oracle.rf_bank(bank, g.Px, g.Py)
        v = oracle.filter_video(vid.frames.reshape(g.frames, -1), RF, g.dx, g.dy)
        vu = oracle.upsample(v, g.Lt, g.duration, g.fps, g.iaf_times())
        lists, cnt, *_ = oracle.iaf_encode(vu, g.dt, iaf.kappa, iaf.bias, iaf.delta, iaf.v0,
                                           root_mode=1, max_spike=2 * job.max_spike)
        gc = st.counts.cpu().numpy()
        gs = st.spikes.cpu().numpy()
        np.testing.assert_array_equal(gc, cnt, err_msg=name)
        err = max((float(np.abs(gs[:cnt[j], j] - lists[j]).max()) for j in range(cnt.size) if cnt[j]),
                  default=0.0)
        print(f"  {name}: N={cnt.size} spikes={int(cnt.sum())} max |dt| {err:.2e} s", flush=True)
        assert err <= 1e-9, (name, err)
        worst = max(worst, err)
    print(f"served encoder: all counts identical, worst spike-time difference {worst:.2e} s", flush=True)
