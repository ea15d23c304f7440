"""Time-parallel IAF encoder (vt_iaf_encode_par, vt_iaf_encode_trig_par and the batched
form) against the oracle's serial restatement of the reference's loop
(vtem/ensemble_encode.py:438-552) and against the serial HIP kernel.

Bar: spike counts identical per neuron; spike times within 1e-9 s; carried v0 within
1e-12 (volts, thresholds ~0.04) and time_count within 1e-9 s."""

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


def _c0(oracle):
    from vtem_amd import synthetic as syn
    vid = syn.make_video(32, 32, 1.0, 100.0, 5, 5, 4, seed=0)
    bank = syn.bank_config0()
    iafp = syn.iaf_params(bank.num_neurons, 2.0, 1.0, 0.04, 0.0, np.random.default_rng(1))
    RF = oracle.rf_bank(bank, 32, 32)
    v = oracle.filter_video(vid.frames, RF)
    vu = oracle.upsample(v, 4, 1.0, 100.0, np.arange(10001) * 1e-4)
    return vid, bank, iafp, vu


class _State:
    def __init__(self, torch, iafp, N):
        d = lambda a: torch.tensor(np.ascontiguousarray(a[:N]), dtype=torch.float64, device="cuda")
        self.v0, self.kap, self.bia, self.dlt = d(iafp.v0), d(iafp.kappa), d(iafp.bias), d(iafp.delta)
        self.tc = torch.zeros(N, dtype=torch.float64, device="cuda")


def _run_sampled(torch, x, ld, N, st, max_spike, root_mode, parallel):
    """One encoder call on rows of x (device, T x ld); returns (spikes, counts) on host."""
    from vtem_amd._lib import lib, ptr, stream_ptr
    from vtem_amd.block import iaf_par_workspace
    T = x.shape[0]
    spikes = torch.full((max_spike, N), -1.0, dtype=torch.float64, device="cuda")
    counts = torch.zeros(N, dtype=torch.int32, device="cuda")
    args = (ptr(x), 0, ld, N, T, ptr(spikes), N, max_spike, ptr(st.v0), ptr(st.kap), ptr(st.bia),
            ptr(st.dlt), ptr(st.tc), ptr(counts), 1e-4, root_mode)
    if parallel:
        ws = iaf_par_workspace(N, T)
        lib.vt_iaf_encode_par(*args, ptr(ws), ws.numel(), stream_ptr())
    else:
        lib.vt_iaf_encode(*args, stream_ptr())
    torch.cuda.synchronize()
    return spikes.cpu().numpy(), counts.cpu().numpy()


def _check(gs, gc, lists, cnt):
    np.testing.assert_array_equal(gc, cnt)
    err = 0.0
    for j in range(len(lists)):
        if cnt[j]:
            err = max(err, np.abs(gs[:cnt[j], j] - lists[j]).max())
    assert err < 1e-9, err
    return err


@pytest.mark.parametrize("root_mode", [0, 1])
def test_iaf_par_sampled_vs_oracle(torch_mod, oracle, root_mode):
    torch = torch_mod
    _banner(f"iaf_par sampled c0 root_mode={root_mode}")
    vid, bank, iafp, vu = _c0(oracle)
    N = bank.num_neurons
    lists, cnt, v0o, tco, _ = oracle.iaf_encode(vu, 1e-4, iafp.kappa, iafp.bias, iafp.delta,
                                                iafp.v0, root_mode=root_mode, max_spike=200)
    assert cnt.sum() > 1000
    st = _State(torch, iafp, N)
    x = torch.tensor(vu, dtype=torch.float64, device="cuda")
    gs, gc = _run_sampled(torch, x, N, N, st, 200, root_mode, True)
    err = _check(gs, gc, lists, cnt)
    print(f"max spike-time error {err:.3e} s", flush=True)
    assert np.abs(st.v0.cpu().numpy() - v0o).max() < 1e-12
    assert np.abs(st.tc.cpu().numpy() - tco).max() < 1e-9


def test_iaf_par_chained_ragged(torch_mod, oracle):
    """Calls of 1, 2, 65, 449 and the remaining samples carrying v0/time_count, on the
    first 200 of 256 neurons (row stride 256): every call agrees with the oracle's call on
    the same chunk and carried state."""
    torch = torch_mod
    _banner("iaf_par chained ragged")
    vid, bank, iafp, vu = _c0(oracle)
    N, ld = 200, vu.shape[1]
    st = _State(torch, iafp, N)
    v0 = iafp.v0[:N].copy()
    tc = np.zeros(N)
    x_all = torch.tensor(vu, dtype=torch.float64, device="cuda")
    bounds = [0, 1, 3, 68, 517, vu.shape[0]]
    for a, b in zip(bounds[:-1], bounds[1:]):
        lists, cnt, v0, tc, _ = oracle.iaf_encode(vu[a:b, :N], 1e-4, iafp.kappa[:N], iafp.bias[:N],
                                                  iafp.delta[:N], v0, time_count=tc, max_spike=200)
        gs, gc = _run_sampled(torch, x_all[a:b], ld, N, st, 200, 0, True)
        _check(gs, gc, lists, cnt)
        assert np.abs(st.v0.cpu().numpy() - v0).max() < 1e-12, (a, b)
        assert np.abs(st.tc.cpu().numpy() - tc).max() < 1e-9, (a, b)


def test_iaf_par_overflow_matches_serial(torch_mod, oracle):
    """A buffer of 5 rows: the serial loop stops at max_spike + 1 spikes with the last row
    holding spike max_spike; the parallel encoder leaves the same counts, rows and state."""
    torch = torch_mod
    _banner("iaf_par overflow")
    vid, bank, iafp, vu = _c0(oracle)
    N = bank.num_neurons
    x = torch.tensor(vu, dtype=torch.float64, device="cuda")
    st_s, st_p = _State(torch, iafp, N), _State(torch, iafp, N)
    ss, cs = _run_sampled(torch, x, N, N, st_s, 5, 0, False)
    sp, cp = _run_sampled(torch, x, N, N, st_p, 5, 0, True)
    assert (cs == 6).sum() > N // 2
    np.testing.assert_array_equal(cp, cs)
    assert np.abs(sp - ss).max() < 1e-9
    assert np.abs(st_p.v0.cpu().numpy() - st_s.v0.cpu().numpy()).max() < 1e-12


def test_iaf_par_float_input(torch_mod, oracle):
    """FP32 input (in_dtype 1): same as the serial kernel on the same FP32 samples."""
    torch = torch_mod
    _banner("iaf_par fp32 input")
    vid, bank, iafp, vu = _c0(oracle)
    N = bank.num_neurons
    xf = torch.tensor(vu, dtype=torch.float32, device="cuda")
    from vtem_amd._lib import lib, ptr, stream_ptr
    from vtem_amd.block import iaf_par_workspace
    res = []
    for parallel in (False, True):
        st = _State(torch, iafp, N)
        spikes = torch.zeros((200, N), dtype=torch.float64, device="cuda")
        counts = torch.zeros(N, dtype=torch.int32, device="cuda")
        args = (ptr(xf), 1, N, N, xf.shape[0], ptr(spikes), N, 200, ptr(st.v0), ptr(st.kap),
                ptr(st.bia), ptr(st.dlt), ptr(st.tc), ptr(counts), 1e-4, 0)
        if parallel:
            ws = iaf_par_workspace(N, xf.shape[0])
            lib.vt_iaf_encode_par(*args, ptr(ws), ws.numel(), stream_ptr())
        else:
            lib.vt_iaf_encode(*args, stream_ptr())
        torch.cuda.synchronize()
        res.append((spikes.cpu().numpy(), counts.cpu().numpy()))
    (ss, cs), (sp, cp) = res
    np.testing.assert_array_equal(cp, cs)
    for j in range(N):
        assert np.abs(sp[:cs[j], j] - ss[:cs[j], j]).max(initial=0) < 1e-9


def test_encoder_trig_par_vs_oracle(torch_mod, oracle):
    """Encoder (filter + trig coefficients + time-parallel IAF) vs the oracle on c0."""
    torch = torch_mod
    from vtem_amd.basis import Geometry
    from vtem_amd.block import Encoder, dev
    _banner("encoder trig par c0")
    vid, bank, iafp, vu = _c0(oracle)
    lists, cnt, *_ = oracle.iaf_encode(vu, 1e-4, iafp.kappa, iafp.bias, iafp.delta, iafp.v0)
    enc = Encoder(Geometry(32, 32, 5, 5, 4, 1.0), bank)
    st = enc.encode(dev(vid.frames.reshape(100, -1)), iafp)
    torch.cuda.synchronize()
    _check(st.spikes.cpu().numpy(), st.counts.cpu().numpy(), lists, cnt)


def _serve_jobs(torch, specs):
    from vtem_amd import synthetic as syn
    from vtem_amd.basis import Geometry
    from vtem_amd.block import Encoder, dev
    from vtem_amd.batch_encode import EncodeJob, IAFDevice, max_spikes_rate
    out = []
    for P, Ls, Lt, dur, bank, seed in specs:
        geom = Geometry(P, P, Ls, Ls, Lt, dur)
        vid = syn.make_video(P, P, dur, 100.0, Ls, Ls, Lt, seed=seed)
        iafp = syn.iaf_params(bank.num_neurons, 2.5, 1.0, 0.04, 0.004, np.random.default_rng(seed))
        enc = Encoder(geom, bank)
        V = dev(vid.frames.reshape(geom.frames, -1))
        ms = max_spikes_rate(geom, iafp)
        out.append((EncodeJob(enc, V, IAFDevice.from_params(iafp), ms), iafp))
    return out


def test_batched_encode(torch_mod):
    """Blocks of different geometry (Dt 9 and 17, 1 s and 2 s, 256 and 2048 neurons):
    (a) one batched launch gives bit-identical spikes, counts and carried state to each
    block encoded alone by the same batched encoder; (b) counts equal, and times within
    1e-9 s of, the single-block Encoder (frame filtering, then coefficients; serial IAF);
    (c) the overflow flag is raised for exactly the block given too small a buffer.

    (b) uses the cancellation-free root (root_mode 1): the two paths form the coefficients
    in different association orders (1e-15 apart), and the reference's root formula
    (sqrt(y1^2 + 2 dv a) - y1) / a turns that into spike-time differences of ~1e-8 s on
    neurons whose input is nearly constant (|a| just above its 1e-12 switch), in float64
    NumPy as much as here."""
    torch = torch_mod
    from vtem_amd import synthetic as syn
    from vtem_amd.batch_encode import BatchEncoder, EncodeJob, IAFDevice
    _banner("batched encode")
    specs = [(32, 5, 4, 1.0, syn.bank_config0(), 0), (64, 8, 8, 2.0, syn.bank_dyadic(64., 64.), 1),
             (32, 5, 4, 1.0, syn.bank_config0(), 2)]
    jobs = _serve_jobs(torch, specs)
    be = BatchEncoder()
    flags = torch.zeros(len(jobs), dtype=torch.int32, device="cuda")
    out = be.encode([j for j, _ in jobs], flags)
    alone = []
    for job, iafp in jobs:
        j2 = EncodeJob(job.encoder, job.frames, IAFDevice.from_params(iafp), job.max_spike)
        alone.append((j2, be.encode([j2])[0]))
    stable = []
    for job, iafp in jobs:
        j3 = EncodeJob(job.encoder, job.frames, IAFDevice.from_params(iafp), job.max_spike)
        stable.append(BatchEncoder(root_mode=1).encode([j3])[0])
    serial = [job.encoder.encode(job.frames, iafp, root_mode=1, max_spike=job.max_spike,
                                 parallel=False) for job, iafp in jobs]
    torch.cuda.synchronize()
    assert flags.sum().item() == 0
    for (job, _), got, (j2, one), stb, ser in zip(jobs, out, alone, stable, serial):
        c = got.counts.cpu().numpy()
        assert c.max() <= job.max_spike
        np.testing.assert_array_equal(c, one.counts.cpu().numpy())
        g, o = got.spikes.cpu().numpy(), one.spikes.cpu().numpy()
        for j in range(c.size):
            np.testing.assert_array_equal(g[:c[j], j], o[:c[j], j])
        torch.testing.assert_close(job.iaf.v0, j2.iaf.v0, rtol=0, atol=0)
        torch.testing.assert_close(job.iaf.time_count, j2.iaf.time_count, rtol=0, atol=0)
        c1 = stb.counts.cpu().numpy()
        np.testing.assert_array_equal(c1, ser.counts.cpu().numpy())
        g1, s = stb.spikes.cpu().numpy(), ser.spikes.cpu().numpy()
        err = max(np.abs(g1[:c1[j], j] - s[:c1[j], j]).max(initial=0) for j in range(c1.size))
        assert err < 1e-9, err
    # (c) overflow: the middle block gets 3 rows
    small = [(EncodeJob(j.encoder, j.frames, IAFDevice.from_params(p), 3 if i == 1 else j.max_spike), p)
             for i, (j, p) in enumerate(jobs)]
    flags.zero_()
    be.encode([j for j, _ in small], flags)
    torch.cuda.synchronize()
    assert flags.cpu().numpy().tolist() == [0, 1, 0]


def test_batched_encode_stored_input(torch_mod):
    """The crossing walk reading the input stored by the first pass gives bit-identical
    counts, spike times and carried state to the walk that evaluates the trigonometric input
    again (Dt 5, 9, 17 and 65 in one launch; a non-power-of-two kappa, so the stored value
    is a rounded quotient); with the first pass on MFMA (the inputs' sums in the MFMA's
    order) the counts are identical and the times within 1e-10 s."""
    torch = torch_mod
    from vtem_amd import synthetic as syn
    from vtem_amd.batch_encode import BatchEncoder, EncodeJob, IAFDevice
    _banner("batched encode, stored input")
    specs = [(32, 5, 4, 1.0, syn.bank_config0(), 3), (64, 8, 8, 2.0, syn.bank_dyadic(64., 64.), 4),
             (32, 4, 2, 0.5, syn.bank_config0(), 5), (32, 5, 32, 1.0, syn.bank_config0(), 6)]
    jobs = [j for j, _ in _serve_jobs(torch, specs)]
    jobs[1].iaf.kappa.mul_(0.7)
    jobs[1].iaf.bias.mul_(0.7)
    be = BatchEncoder()
    a = be.encode(jobs, carry=False, store_input=True, mfma=False)
    sa = [(v.clone(), t.clone()) for v, t in be.last_states]
    b = be.encode(jobs, carry=False, store_input=False)
    sb = [(v.clone(), t.clone()) for v, t in be.last_states]
    m = be.encode(jobs, carry=False, store_input=True, mfma=True)
    sm = be.last_states
    torch.cuda.synchronize()
    worst = 0.0
    for x, y, z, (v1, t1), (v2, t2), (v3, t3) in zip(a, b, m, sa, sb, sm):
        c = x.counts.cpu().numpy()
        assert c.sum() > 0
        np.testing.assert_array_equal(c, y.counts.cpu().numpy())
        g, o = x.spikes.cpu().numpy(), y.spikes.cpu().numpy()
        for j in range(c.size):
            np.testing.assert_array_equal(g[:c[j], j], o[:c[j], j])
        torch.testing.assert_close(v1, v2, rtol=0, atol=0)
        torch.testing.assert_close(t1, t2, rtol=0, atol=0)
        # the first pass on MFMA: the same sums in another order
        np.testing.assert_array_equal(c, z.counts.cpu().numpy())
        gm = z.spikes.cpu().numpy()
        worst = max([worst] + [float(np.abs(gm[:c[j], j] - g[:c[j], j]).max(initial=0)) for j in range(c.size)])
        torch.testing.assert_close(v3, v1, rtol=0, atol=1e-12)
        torch.testing.assert_close(t3, t1, rtol=0, atol=0)
    print(f"MFMA first pass: counts identical, max |dt| {worst:.2e} s")
    assert worst < 1e-10
