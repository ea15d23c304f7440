"""Host logic of the serving runtime (no GPU): arrival process, footprints, precision
policy, configuration laws."""

import numpy as np

from vtem_amd import synthetic as syn


def test_poisson_arrivals_rate_and_order():
    from vtem_amd.serve import poisson_arrivals
    a = poisson_arrivals(20000, 50.0, np.random.default_rng(0))
    assert a[0] == 0 and np.all(np.diff(a) >= 0)
    assert abs((len(a) - 1) / a[-1] - 50.0) < 1.5
    assert np.all(poisson_arrivals(5, float("inf"), np.random.default_rng(0)) == 0)


class _G:
    def __init__(self, Ds, Dt):
        self.Ds, self.Dt = Ds, Dt


class _Dec:
    def __init__(self, Ds, Dt):
        self.geom = _G(Ds, Dt)


class _Job:
    def __init__(self, Ds, Dt, method="tikhonov"):
        self.decoder = _Dec(Ds, Dt)
        self.method = method


def test_precision_policy_routes_failing_class_to_fp64():
    from vtem_amd.serve import PrecisionPolicy
    p = PrecisionPolicy("bf16x3", min_fail=2, rate=0.5)
    bad, good = _Job(81, 5), _Job(289, 17)
    assert p.choose(bad) == "bf16x3"
    p.record(bad, "bf16x3", True)
    assert p.choose(bad) == "bf16x3"          # one failure is not enough
    p.record(bad, "bf16x3", True)
    assert p.choose(bad) == "fp64"
    for _ in range(5):
        p.record(good, "bf16x3", False)
    p.record(good, "bf16x3", True)
    p.record(good, "bf16x3", True)
    assert p.choose(good) == "bf16x3"         # 2 of 7 failures: keep low precision
    p.record(bad, "fp64", True)               # fp64 outcomes do not count
    assert p.stats[(81, 5)] == [2, 2]


def test_serve_block_law():
    for side in syn.SERVE_SIDES:
        for dur in syn.SERVE_DURATIONS:
            b = syn.serve_block(side, dur, 7)
            L, Lt = round(side / 8), round(4 * dur)
            assert (b.Lx, b.Lt) == (L, Lt)
            D = (2 * L + 1) ** 2 * (2 * Lt + 1)
            assert b.D == D and b.bank.num_neurons == round(2 * D / 9)
            n_dog = int((b.bank.kind == 1).sum())
            assert n_dog == round(b.bank.num_neurons / 8)
            assert b.video.frames.shape == (int(round(dur * 100)), side, side)


def test_dyadic_bank_counts():
    b1 = syn.bank_dyadic(64.0, 64.0)
    assert b1.num_neurons == 2048 and int((b1.kind == 0).sum()) == 1792
    b2 = syn.bank_dyadic(128.0, 128.0)
    assert b2.num_neurons == 8192 and int((b2.kind == 0).sum()) == 7168
    b0 = syn.bank_config0()
    assert b0.num_neurons == 256 and set(np.round(b0.h_l / (np.pi / 8)).astype(int)) == set(range(8))


def test_skinny_workspace_query():
    """The split-K workspace query (host-only entry): a reduction splits only beyond 1008
    terms (one split per 512 terms), in at most 8 parts, and the size scales with M and
    Nc."""
    from vtem_amd.batch import skinny_workspace
    assert skinny_workspace(100, 1008, 17) == 0
    assert skinny_workspace(100, 1009, 17) == 2 * 100 * 17
    assert skinny_workspace(625, 625, 17) == 0
    assert skinny_workspace(625, 2361, 17) == 4 * 625 * 17
    assert skinny_workspace(3, 10 ** 6, 5) == 8 * 3 * 5


def test_skinny_workspace_array_matches_the_library_query():
    """The vectorised split rule the descriptor builders use gives the library's own answer
    on every problem shape, across the split boundaries."""
    from vtem_amd.batch import skinny_workspace, skinny_workspace_array
    rng = np.random.default_rng(5)
    K = np.concatenate([np.arange(1, 40), [1007, 1008, 1009, 1024, 1025, 1535, 1536, 1537, 4095, 4096, 4097,
                                           8191, 8192, 8193, 10 ** 6], rng.integers(1, 20000, 60)])
    M = rng.integers(1, 700, K.size)
    Nc = rng.integers(1, 33, K.size)
    got = skinny_workspace_array(M, K, Nc)
    want = [skinny_workspace(m, k, n) for m, k, n in zip(M, K, Nc)]
    assert got.tolist() == want


def test_iaf_par_host_queries():
    """Host-only entries of the time-parallel encoder: the descriptor layout the batched
    launch reads matches the host dtype, and the workspace holds (segments + 2) x N pairs
    of doubles with one segment per 64 sample intervals."""
    import ctypes
    from vtem_amd._lib import lib
    from vtem_amd.batch_encode import IAFP_DTYPE
    from vtem_amd.block import iaf_par_workspace_bytes
    b = ctypes.c_int(0)
    lib.vt_iaf_par_desc_size(ctypes.addressof(b))
    assert b.value == IAFP_DTYPE.itemsize == 176
    assert iaf_par_workspace_bytes(100, 0) == 2 * 100 * 16
    assert iaf_par_workspace_bytes(100, 1) == 2 * 100 * 16
    assert iaf_par_workspace_bytes(100, 65) == 3 * 100 * 16
    assert iaf_par_workspace_bytes(100, 66) == 4 * 100 * 16
    assert iaf_par_workspace_bytes(2361, 20001) == (313 + 2) * 2361 * 16


def test_batch_formation_and_rnn_lane_routing():
    """Batches take requests in arrival order up to the request cap and the free HBM;
    RNN-decoded requests are routed to their own lanes, which split the waiting ones."""
    import collections
    from vtem_amd.serve import Request, form_batch, is_rnn, split_cap

    class _J:
        def __init__(self, method):
            self.method = method

    reqs = [Request(i, 0.001 * i, _J("rnn" if i % 4 == 3 else "tikhonov"), 100, 10 * (i + 1))
            for i in range(12)]
    assert [is_rnn(r) for r in reqs[:4]] == [False, False, False, True]
    q = collections.deque(reqs)
    b, nb = form_batch(q, 3, 10 ** 9)
    assert [r.rid for r in b] == [0, 1, 2] and nb == 60 and q[0].rid == 3
    b, nb = form_batch(q, 10, 95)                  # footprints 40 + 50 fit, 60 does not
    assert [r.rid for r in b] == [3, 4] and nb == 90
    b, nb = form_batch(q, 10, 10)                  # nothing fits
    assert b == [] and nb == 0 and len(q) == 7
    # several free RNN lanes split the waiting RNN requests evenly, within the cap
    assert split_cap(13, 64, 1) == 13
    assert split_cap(13, 64, 2) == 7 and split_cap(6, 64, 2) == 3
    assert split_cap(13, 4, 2) == 4 and split_cap(0, 64, 3) == 1
    # staggered lanes: one lane per hold time
    from vtem_amd.serve import Server
    s = Server(rnn_lanes=1, rnn_lane_delays=(0, 0.016))
    assert s.rnn_lanes == 2 and s.rnn_lane_delays == [0.0, 0.016]
    assert Server(rnn_lanes=3).rnn_lane_delays is None


def test_spike_buffer_rate_rows():
    """Spike rows from the fastest neuron's bias-driven rate b / (kappa delta) over the
    window, times the margin, plus 16."""
    from vtem_amd.basis import Geometry
    from vtem_amd.batch_encode import max_spikes_rate
    from vtem_amd.synthetic import IAFParams
    g = Geometry(32, 32, 4, 4, 4, 2.0)
    p = IAFParams(np.ones(3), np.full(3, 2.5), np.array([0.05, 0.025, 0.04]), np.zeros(3), np.zeros(3))
    assert max_spikes_rate(g, p, margin=2.0) == int(np.ceil(2.0 * (2.5 / 0.025) * 2.0)) + 16
    assert max_spikes_rate(g, p, margin=1.0) == 200 + 16


def test_guard_reads_the_triangular_solve_expiry_word():
    """The batched decoder's guard reads word 1 of each direct solve's flag workspace: an
    expired bounded wait sends a low-precision group's blocks to the FP64 redo, and
    raises for an FP64 group (its solutions are invalid).  Forced from the host with CPU
    tensors; no kernel runs."""
    import pytest
    import torch
    from vtem_amd._lib import VtemError
    from vtem_amd.batch import BatchDecoder
    bd = BatchDecoder(device="cpu", precision="bf16x3")
    nrm = torch.tensor([[1.0, 1e-30], [1.0, 1e-30]], dtype=torch.float64)
    info = torch.zeros(2, dtype=torch.int32)
    ok = torch.zeros(6, dtype=torch.int32)
    bad = ok.clone()
    bad[1] = 1
    assert bd._guard([([3, 4], (nrm, info, ok), "bf16x3")]) == ([], 0)
    assert bd._guard([([3, 4], (nrm, info, bad), "bf16x3")]) == ([(3, "bf16x3"), (4, "bf16x3")], 0)
    assert bd._guard([([3, 4], (nrm, info, None), "kpcg")]) == ([], 0)
    with pytest.raises(VtemError):
        bd._guard([([3, 4], (nrm, info, bad), "fp64")])


def test_dynamic_batching_rule():
    from vtem_amd.serve import should_launch
    assert not should_launch(0, 8, 1.0, 0.0, False)           # nothing queued
    assert should_launch(3, 8, 0.0, 0.0, True)                # no delay: launch at once
    assert not should_launch(3, 8, 0.002, 0.005, True)        # hold: not full, young, more due
    assert should_launch(8, 8, 0.0, 0.005, True)              # full batch
    assert should_launch(3, 8, 0.005, 0.005, True)            # oldest waited the delay
    assert should_launch(3, 8, 0.0, 0.005, False)             # no further arrivals
