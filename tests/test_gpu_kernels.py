"""GPU parity of each HIP entry point against the float64 oracle (same seeded inputs).

Integer/index outputs must match exactly; floating-point tolerances are written per test.
"""


import dataclasses
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


def test_mfma_layouts(torch_mod):
    """Accumulator/operand lane layout of every MFMA form the kernels use."""
    torch = torch_mod
    from vtem_amd._lib import lib, ptr, stream_ptr
    rng = np.random.default_rng(0)
    for kind, (K, dt) in enumerate([(4, torch.float64), (4, torch.float32), (32, torch.float32),
                                    (32, torch.float32)]):
        _banner(f"mfma probe kind={kind}")
        A = rng.integers(-4, 5, (16, K)).astype(np.float64)
        B = rng.integers(-4, 5, (K, 16)).astype(np.float64)
        a = torch.tensor(A, dtype=dt, device="cuda")
        b = torch.tensor(B, dtype=dt, device="cuda")
        out = torch.zeros((16, 16), dtype=dt, device="cuda")
        lib.vt_mfma_probe(kind, ptr(a), ptr(b), ptr(out), stream_ptr())
        torch.cuda.synchronize()
        np.testing.assert_array_equal(out.cpu().numpy().astype(np.float64), A @ B)


@pytest.mark.parametrize("M,N,K,ta,tb", [(64, 64, 16, False, False), (100, 37, 53, False, True),
                                         (17, 129, 200, True, False), (5, 3, 1, True, True)])
@pytest.mark.parametrize("dtype", ["f64", "f32"])
def test_gemm(torch_mod, M, N, K, ta, tb, dtype):
    torch = torch_mod
    from vtem_amd.block import gemm
    _banner(f"gemm {dtype} M={M} N={N} K={K} ta={ta} tb={tb}")
    rng = np.random.default_rng(M * N + K)
    A = rng.standard_normal((K, M) if ta else (M, K))
    B = rng.standard_normal((N, K) if tb else (K, N))
    C0 = rng.standard_normal((M, N))
    tdt = torch.float64 if dtype == "f64" else torch.float32
    a = torch.tensor(A, dtype=tdt, device="cuda")
    b = torch.tensor(B, dtype=tdt, device="cuda")
    c = torch.tensor(C0, dtype=tdt, device="cuda")
    sam, sak = (1, M) if ta else (K, 1)
    sbk, sbn = (1, K) if tb else (N, 1)
    gemm(M, N, K, 0.5, a, sam, sak, b, sbk, sbn, 2.0, c, N)
    ref = 0.5 * ((A.T if ta else A) @ (B.T if tb else B)) + 2.0 * C0
    tol = 1e-12 if dtype == "f64" else 2e-4
    err = np.abs(c.cpu().numpy() - ref).max() / max(1.0, np.abs(ref).max())
    assert err < tol, err


def test_gemm_skinny_route_invariance(torch_mod):
    """A problem's route through the skinny product is fixed by its own column count
    (<= 16: one MFMA tile; 17: one tile plus the vector-ALU column 16; more: two tiles),
    so its result is bit-identical whatever the launch's widest problem is; each result
    also matches float64 torch (row factor applied)."""
    torch = torch_mod
    from vtem_amd._lib import lib, ptr, stream_ptr
    from vtem_amd.batch import SK_DTYPE, _to_dev, skinny_workspace
    _banner("skinny route invariance (Nc 9 / 17 / 25 launches)")
    g = torch.Generator().manual_seed(11)
    probs = {}
    for name, (M, K, Nc) in {"p9": (300, 625, 9), "p17": (2361, 625, 17), "p17s": (100, 2361, 17),
                             "p25": (200, 289, 25)}.items():
        A = torch.randn(M, K, generator=g, dtype=torch.float64)
        Bm = torch.randn(K, Nc, generator=g, dtype=torch.float64)
        Tn = torch.randn(M, Nc, Nc, generator=g, dtype=torch.float64)
        ref = torch.einsum("mab,mb->ma", Tn, A @ Bm)
        probs[name] = (M, K, Nc, A.cuda(), Bm.cuda(), Tn.contiguous().cuda(), ref)

    def run(names):
        rows, outs, keep = [], [], []
        for nm in names:
            M, K, Nc, Ad, Bd, Td, _ = probs[nm]
            Cd = torch.full((M, Nc), float("nan"), dtype=torch.float64, device="cuda")
            wn = skinny_workspace(M, K, Nc)
            Wd = torch.empty(max(1, wn), dtype=torch.float64, device="cuda")
            rows.append((Ad.data_ptr(), Bd.data_ptr(), Cd.data_ptr(), Td.data_ptr(), Wd.data_ptr() if wn else 0,
                         0, K, 1, Nc, Nc, Nc * Nc, M, K, Nc, 0))
            outs.append(Cd)
            keep.append(Wd)
        d = np.zeros(len(rows), SK_DTYPE)
        for i, r in enumerate(rows):
            d[i] = r
        dd = _to_dev(d, "cuda")
        lib.vt_gemm_skinny(ptr(dd), len(rows), max(probs[n][0] for n in names), max(probs[n][1] for n in names),
                           max(probs[n][2] for n in names), stream_ptr())
        torch.cuda.synchronize()
        return [o.cpu() for o in outs]

    alone = {nm: run([nm])[0] for nm in probs}
    for nm, out in alone.items():
        ref = probs[nm][6]
        assert (out - ref).abs().max().item() < 1e-13 * max(1.0, ref.abs().max().item()), nm
    mixed = run(["p9", "p17", "p17s", "p25"])
    for nm, out in zip(["p9", "p17", "p17s", "p25"], mixed):
        assert torch.equal(out, alone[nm]), nm
    pair = run(["p9", "p17"])
    assert torch.equal(pair[0], alone["p9"]) and torch.equal(pair[1], alone["p17"])


def test_gemm_skinny_tiled_layouts(torch_mod):
    """The MFMA tile layout (vt_tile_f64) of A, B and C changes no value or sum order: the
    skinny product with A tiled, A and B tiled, and A, B and C tiled equals the row-major
    one bit for bit, on ragged shapes, Nc = 9 / 17 / 25 (one tile, the vector column, two
    tiles), a K split and the row factor.  vt_tile_f64 itself matches the layout built
    entry by entry from its index formula, zeros in the padding, in both orientations."""
    torch = torch_mod
    from vtem_amd._lib import lib, ptr, stream_ptr
    from vtem_amd.batch import SK_A_TILED, SK_B_TILED, SK_C_TILED, SK_DTYPE, _to_dev, skinny_workspace
    from vtem_amd.block import tile_doubles
    _banner("skinny product: tiled operand layouts == row-major")
    g = torch.Generator().manual_seed(12)

    def tile_index(m, k, kc):
        return ((m >> 4) * kc + (k >> 4)) * 256 + ((k & 3) >> 1) * 128 + (16 * ((k & 15) >> 2) + (m & 15)) * 2 + (k & 1)

    def tiled(X, trans=0):
        rows, cols = X.shape
        M, K = (cols, rows) if trans else (rows, cols)
        out = torch.full((tile_doubles(M, K),), float("nan"), dtype=torch.float64, device="cuda")
        lib.vt_tile_f64(ptr(X), X.stride(0), rows, cols, trans, ptr(out), stream_ptr())
        return out

    A = torch.randn(37, 53, generator=g, dtype=torch.float64).cuda()
    m = torch.arange(37).view(-1, 1).expand(37, 53)
    k = torch.arange(53).view(1, -1).expand(37, 53)
    idx = tile_index(m, k, -(-53 // 16))
    T = tiled(A).cpu()
    assert torch.equal(T[idx], A.cpu())
    pad = torch.ones(T.numel(), dtype=torch.bool)
    pad[idx.reshape(-1)] = False
    assert bool((T[pad] == 0).all())
    assert torch.equal(tiled(A, trans=1), tiled(A.t().contiguous()))

    def product(M, K, Nc, Ad, Bd, Td, lay):
        wn = skinny_workspace(M, K, Nc)
        Wd = torch.empty(max(1, wn), dtype=torch.float64, device="cuda")
        kc, mc = -(-K // 16), -(-M // 16)
        A_, sam = (tiled(Ad), kc) if lay & SK_A_TILED else (Ad, K)
        B_, ldb = (tiled(Bd, trans=1), kc) if lay & SK_B_TILED else (Bd, Nc)
        if lay & SK_C_TILED:
            C_, ldc = torch.zeros(tile_doubles(Nc, M), dtype=torch.float64, device="cuda"), mc
        else:
            C_, ldc = torch.full((M, Nc), float("nan"), dtype=torch.float64, device="cuda"), Nc
        d = np.zeros(1, SK_DTYPE)
        d[0] = (A_.data_ptr(), B_.data_ptr(), C_.data_ptr(), Td.data_ptr() if Td is not None else 0,
                Wd.data_ptr() if wn else 0, 0, sam, 1, ldb, ldc, Nc * Nc if Td is not None else 0, M, K, Nc, lay)
        dd = _to_dev(d, "cuda")
        lib.vt_gemm_skinny(ptr(dd), 1, M, K, Nc, stream_ptr())
        torch.cuda.synchronize()
        C = C_.cpu()
        if lay & SK_C_TILED:
            mm = torch.arange(M).view(-1, 1).expand(M, Nc)
            nn = torch.arange(Nc).view(1, -1).expand(M, Nc)
            C = C[tile_index(nn, mm, mc)]
        return C

    for M, K, Nc, with_t in ((300, 625, 9, True), (2361, 625, 17, True), (100, 2361, 17, False),
                             (200, 289, 25, True), (37, 53, 17, False)):
        Ad = torch.randn(M, K, generator=g, dtype=torch.float64).cuda()
        Bd = torch.randn(K, Nc, generator=g, dtype=torch.float64).cuda()
        Tn = torch.randn(M, Nc, Nc, generator=g, dtype=torch.float64)
        ref = torch.einsum("mab,mb->ma", Tn, Ad.cpu() @ Bd.cpu()) if with_t else Ad.cpu() @ Bd.cpu()
        Td = Tn.contiguous().cuda() if with_t else None
        row_major = product(M, K, Nc, Ad, Bd, Td, 0)
        assert (row_major - ref).abs().max().item() < 1e-13 * max(1.0, ref.abs().max().item()), (M, K, Nc)
        for lay in (SK_A_TILED, SK_A_TILED | SK_B_TILED, SK_A_TILED | SK_B_TILED | SK_C_TILED):
            assert torch.equal(product(M, K, Nc, Ad, Bd, Td, lay), row_major), (M, K, Nc, lay)


@pytest.mark.parametrize("tmode", ["none", "row", "shared"])
def test_gemm_skinny(torch_mod, tmode):
    """Skinny grouped product (both A orientations, ragged problems, Nc 5/17/32) with the
    optional row factor, against float64 torch."""
    torch = torch_mod
    from vtem_amd._lib import lib, ptr, stream_ptr
    from vtem_amd.batch import SK_DTYPE, _to_dev, skinny_workspace
    _banner(f"skinny {tmode}")
    g = torch.Generator().manual_seed(7)
    for Nc in (5, 17, 32):
        probs = [(300, 625, False), (64, 81, True), (1, 5, False), (2361, 625, True), (70, 1000, False),
                 (625, 2361, True), (17, 5000, False), (100, 1009, True)]
        rows, refs, keep = [], [], []
        for M, K, trans in probs:
            A = torch.randn(K, M, generator=g, dtype=torch.float64) if trans else \
                torch.randn(M, K, generator=g, dtype=torch.float64)
            Bm = torch.randn(K, Nc, generator=g, dtype=torch.float64)
            Tn = {"none": None, "row": torch.randn(M, Nc, Nc, generator=g, dtype=torch.float64),
                  "shared": torch.randn(1, Nc, Nc, generator=g, dtype=torch.float64)}[tmode]
            Am = A.t() if trans else A
            ref = Am @ Bm
            if Tn is not None:
                ref = torch.einsum("mab,mb->ma", Tn.expand(M, Nc, Nc), ref)
            Ad, Bd = A.cuda(), Bm.cuda()
            Cd = torch.full((M, Nc), float("nan"), dtype=torch.float64, device="cuda")
            Td = None if Tn is None else Tn.contiguous().cuda()
            sam, sak = (1, M) if trans else (K, 1)
            tstride = {"none": 0, "row": Nc * Nc, "shared": 0}[tmode]
            wn = skinny_workspace(M, K, Nc)
            Wd = torch.empty(max(1, wn), dtype=torch.float64, device="cuda")
            rows.append((Ad.data_ptr(), Bd.data_ptr(), Cd.data_ptr(), 0 if Td is None else Td.data_ptr(),
                         Wd.data_ptr() if wn else 0, 0, sam, sak, Nc, Nc, tstride, M, K, Nc, 0))
            keep.append(Wd)
            refs.append((ref, Cd))
            keep += [Ad, Bd, Cd, Td]
        d = np.zeros(len(rows), SK_DTYPE)
        for i, r in enumerate(rows):
            d[i] = r
        dd = _to_dev(d, "cuda")
        lib.vt_gemm_skinny(ptr(dd), len(rows), max(p[0] for p in probs), max(p[1] for p in probs), Nc,
                           stream_ptr())
        torch.cuda.synchronize()
        for (ref, Cd), (M, K, tr) in zip(refs, probs):
            err = (Cd.cpu() - ref).abs().max().item() / max(1.0, ref.abs().max().item())
            assert err < 1e-13, (Nc, M, K, tr, err)
    # one launch with per-problem column counts; a problem whose active flag is 0 is skipped
    flags = torch.tensor([1.0, 0.0, 1.0], dtype=torch.float64, device="cuda")
    rows, refs, keep = [], [], []
    for i, (M, K, Nc) in enumerate([(500, 1500, 5), (40, 300, 17), (130, 2000, 32)]):
        A = torch.randn(M, K, generator=g, dtype=torch.float64)
        Bm = torch.randn(K, Nc, generator=g, dtype=torch.float64)
        Ad, Bd = A.cuda(), Bm.cuda()
        Cd = torch.full((M, Nc), float("nan"), dtype=torch.float64, device="cuda")
        wn = skinny_workspace(M, K, Nc)
        Wd = torch.empty(max(1, wn), dtype=torch.float64, device="cuda")
        rows.append((Ad.data_ptr(), Bd.data_ptr(), Cd.data_ptr(), 0, Wd.data_ptr() if wn else 0,
                     flags.data_ptr() + 8 * i, K, 1, Nc, Nc, 0, M, K, Nc, 0))
        refs.append((A @ Bm, Cd))
        keep += [Ad, Bd, Cd, Wd]
    d = np.zeros(len(rows), SK_DTYPE)
    for i, r in enumerate(rows):
        d[i] = r
    dd = _to_dev(d, "cuda")
    lib.vt_gemm_skinny(ptr(dd), 3, 500, 2000, 32, stream_ptr())
    torch.cuda.synchronize()
    for i, (ref, Cd) in enumerate(refs):
        if i == 1:
            assert torch.isnan(Cd).all()
        else:
            assert (Cd.cpu() - ref).abs().max().item() < 1e-12 * max(1.0, ref.abs().max().item())


@pytest.mark.parametrize("M,Nc,K", [(289, 17, 2048), (5, 1, 3), (40, 32, 999)])
def test_gemm_tn_small(torch_mod, M, Nc, K):
    torch = torch_mod
    from vtem_amd._lib import lib, ptr, stream_ptr
    _banner(f"gemm_tn_small M={M} Nc={Nc} K={K}")
    rng = np.random.default_rng(M + K)
    A = rng.standard_normal((K, M + 3))
    B = rng.standard_normal((K, Nc))
    C0 = rng.standard_normal((M, Nc))
    a, b, c = (torch.tensor(v, device="cuda") for v in (A, B, C0))
    ws = torch.empty(16 * M * Nc, dtype=torch.float64, device="cuda")
    lib.vt_gemm_tn_small(M, Nc, K, 2.0, ptr(a), M + 3, ptr(b), Nc, -1.0, ptr(c), Nc, ptr(ws),
                         stream_ptr())
    ref = 2.0 * A[:, :M].T @ B - C0
    assert np.abs(c.cpu().numpy() - ref).max() / np.abs(ref).max() < 1e-13


def test_rf_bank(torch_mod, oracle):
    torch = torch_mod
    from vtem_amd import synthetic as syn
    from vtem_amd.block import rf_bank
    for name, bank, P in [("c0", syn.bank_config0(), 32), ("dyadic64", syn.bank_dyadic(64., 64.), 64)]:
        _banner(f"rf_bank {name}")
        ref = oracle.rf_bank(bank, P, P).reshape(bank.num_neurons, -1)
        got = rf_bank(bank, P, P).cpu().numpy()
        err = np.abs(got - ref).max() / np.abs(ref).max()
        assert err < 1e-13, err


def _c0_inputs(oracle):
    from vtem_amd import synthetic as syn
    vid = syn.make_video(32, 32, 1.0, 100.0, 5, 5, 4, seed=0)
    bank = syn.bank_config0()
    iafp = syn.iaf_params(bank.num_neurons, 2.0, 1.0, 0.04, 0.0, np.random.default_rng(1))
    RF = oracle.rf_bank(bank, 32, 32)
    v = oracle.filter_video(vid.frames, RF)
    t = np.arange(10001) * 1e-4
    vu = oracle.upsample(v, 4, 1.0, 100.0, t)
    return vid, bank, iafp, RF, v, vu


@pytest.mark.parametrize("root_mode", [0, 1])
def test_iaf_bitexact(torch_mod, oracle, root_mode):
    """Same dt-rate input -> identical spike counts and bit-identical spike times."""
    torch = torch_mod
    from vtem_amd._lib import lib, ptr, stream_ptr
    _banner(f"iaf bitexact root_mode={root_mode}")
    vid, bank, iafp, RF, v, vu = _c0_inputs(oracle)
    lists, cnt, v0o, tco, buf = oracle.iaf_encode(vu, 1e-4, iafp.kappa, iafp.bias, iafp.delta,
                                                  iafp.v0, root_mode=root_mode, max_spike=200)
    N = bank.num_neurons
    d = lambda a: torch.tensor(np.ascontiguousarray(a), dtype=torch.float64, device="cuda")
    inp = d(vu)
    spikes = torch.zeros((200, N), dtype=torch.float64, device="cuda")
    counts = torch.zeros(N, dtype=torch.int32, device="cuda")
    v0 = d(iafp.v0)
    tc = torch.zeros(N, dtype=torch.float64, device="cuda")
    kap, bia, dlt = d(iafp.kappa), d(iafp.bias), d(iafp.delta)
    lib.vt_iaf_encode(ptr(inp), 0, N, N, vu.shape[0], ptr(spikes), N, 200, ptr(v0), ptr(kap),
                      ptr(bia), ptr(dlt), ptr(tc), ptr(counts), 1e-4, root_mode, stream_ptr())
    torch.cuda.synchronize()
    gc = counts.cpu().numpy()
    np.testing.assert_array_equal(gc, cnt)
    gs = spikes.cpu().numpy()
    for j in range(N):
        np.testing.assert_array_equal(gs[:gc[j], j], lists[j])
    np.testing.assert_array_equal(v0.cpu().numpy(), v0o)
    np.testing.assert_array_equal(tc.cpu().numpy(), tco)


def test_encode_trig_c0(torch_mod, oracle):
    """GPU filter + trig upsampling + IAF vs oracle: counts identical, times within 1e-9 s."""
    torch = torch_mod
    from vtem_amd.basis import Geometry
    from vtem_amd.block import Encoder, dev
    _banner("encode c0 (filter + trig IAF)")
    vid, bank, iafp, RF, v, vu = _c0_inputs(oracle)
    lists, cnt, *_ = oracle.iaf_encode(vu, 1e-4, iafp.kappa, iafp.bias, iafp.delta, iafp.v0)
    geom = Geometry(32, 32, 5, 5, 4, 1.0)
    enc = Encoder(geom, bank)
    V = dev(vid.frames.reshape(100, -1))
    gv = enc.filter(V).cpu().numpy()
    assert np.abs(gv - v).max() < 1e-12 * max(1, np.abs(v).max())
    st = enc.encode(V, iafp)
    torch.cuda.synchronize()
    gl = st.to_lists()
    np.testing.assert_array_equal(st.counts.cpu().numpy(), cnt)
    err = max(np.abs(gl[j] - lists[j]).max() for j in range(len(lists)) if cnt[j])
    assert err < 1e-9, err


def _c0_decode_inputs(oracle):
    vid, bank, iafp, RF, v, vu = _c0_inputs(oracle)
    lists, cnt, *_ = oracle.iaf_encode(vu, 1e-4, iafp.kappa, iafp.bias, iafp.delta, iafp.v0)
    return vid, bank, iafp, RF, lists, cnt


def _spike_buffer(torch, lists):
    N = len(lists)
    m = max(len(s) for s in lists) + 2
    buf = np.zeros((m, N))
    for j, s in enumerate(lists):
        buf[:len(s), j] = s
    from vtem_amd.block import SpikeTrains
    return SpikeTrains(torch.tensor(buf, dtype=torch.float64, device="cuda"),
                       torch.tensor([len(s) for s in lists], dtype=torch.int32, device="cuda"))


def test_intervals_q_moments(torch_mod, oracle):
    torch = torch_mod
    from vtem_amd.basis import Geometry
    from vtem_amd.block import Decoder, spatial_coeffs, rf_bank
    _banner("intervals / q / T_j")
    vid, bank, iafp, RF, lists, cnt = _c0_decode_inputs(oracle)
    geom = Geometry(32, 32, 5, 5, 4, 1.0)
    S = spatial_coeffs(rf_bank(bank, 32, 32), geom)
    dec = Decoder(geom, S, iafp.kappa, iafp.delta, iafp.bias)
    st = _spike_buffer(torch, lists)
    for window in [(0.0, 1.0), (0.25, 0.5)]:
        tk1, tk2, nid, off = dec.intervals(st, window[0], window[1] - window[0])
        o1, o2, oi, oc = oracle.select_intervals(lists, window)
        n = int(off[-1].item())
        assert n == o1.size
        np.testing.assert_array_equal(nid[:n].cpu().numpy(), oi)
        np.testing.assert_array_equal(tk1[:n].cpu().numpy(), o1)
        np.testing.assert_array_equal(tk2[:n].cpu().numpy(), o2)
        np.testing.assert_array_equal(np.diff(off.cpu().numpy()), oc)
    tk1, tk2, nid, off = dec.intervals(st, 0.0, 1.0)
    q, Tm, tv = dec.moments(tk1, tk2, nid, off)
    o1, o2, oi, oc = oracle.select_intervals(lists, (0.0, 1.0))
    oq = oracle.compute_q(o1, o2, oi, iafp.kappa, iafp.delta, iafp.bias, np.ones(256))
    n = o1.size
    np.testing.assert_allclose(q[:n].cpu().numpy(), oq, rtol=0, atol=1e-15)
    al = oracle.temporal_factors(o1, o2, 4, 2 * np.pi)
    oT, ot = oracle.neuron_T(al, oq, oi, 256)
    eT = np.abs(Tm.cpu().numpy() - oT).max() / np.abs(oT).max()
    et = np.abs(tv.cpu().numpy() - ot).max() / np.abs(ot).max()
    assert eT < 1e-13 and et < 1e-13, (eT, et)


@pytest.mark.parametrize("cfg", ["c0", "c1small"])
def test_gram_f64(torch_mod, oracle, cfg):
    """Kronecker Gram on FP64 MFMA vs oracle: relative Frobenius error < 1e-13 (lower triangle)."""
    torch = torch_mod
    from vtem_amd.block import dev
    from vtem_amd._lib import lib, ptr, stream_ptr
    _banner(f"gram f64 {cfg}")
    rng = np.random.default_rng(3)
    N, Ds, Dt = (256, 121, 9) if cfg == "c0" else (300, 289, 17)
    S = rng.standard_normal((N, Ds))
    al = rng.standard_normal((4000, Dt)) * 0.01
    nid = np.sort(rng.integers(0, N, 4000)).astype(np.int32)
    Tm, _ = oracle.neuron_T(al, np.zeros(4000), nid, N)
    Aref = oracle.gram_kron(S, Tm)
    D = Ds * Dt
    A = torch.zeros((D, D), dtype=torch.float64, device="cuda")
    S_d, T_d = dev(S), dev(Tm)      # keep the buffers alive until the launch is enqueued
    lib.vt_gram_f64(ptr(S_d), Ds, ptr(T_d), N, Ds, Dt, ptr(A), D, stream_ptr())
    got = np.tril(A.cpu().numpy())
    ref = np.tril(Aref)
    err = np.linalg.norm(got - ref) / np.linalg.norm(ref)
    assert err < 1e-13, err


@pytest.mark.parametrize("D", [1, 63, 64, 65, 200, 256, 257, 1089, 2100])
@pytest.mark.parametrize("prec", ["f64", "f32"])
def test_potrf_potrs(torch_mod, D, prec):
    """Blocked Cholesky (64-column steps inside 256-column outer panels, one trailing
    update per outer panel) and the triangular solves, against LAPACK's factor and solve;
    sizes on both sides of the 64- and 256-column boundaries."""
    torch = torch_mod
    from vtem_amd.block import mat_descs
    from vtem_amd._lib import lib, ptr, stream_ptr
    _banner(f"potrf/potrs {prec} D={D}")
    rng = np.random.default_rng(D)
    G = rng.standard_normal((D, D + 5))
    A = G @ G.T / D + np.eye(D)
    b = rng.standard_normal(D)
    tdt = torch.float64 if prec == "f64" else torch.float32
    a = torch.tensor(A, dtype=tdt, device="cuda")
    nblk = (D + 63) // 64
    inv = torch.zeros(nblk * 4096, dtype=tdt, device="cuda")
    info = torch.zeros(1, dtype=torch.int32, device="cuda")
    mats = mat_descs([(a.data_ptr(), D, D)], "cuda")
    x = torch.tensor(b, dtype=torch.float64, device="cuda")
    f = lib.vt_potrf_f64 if prec == "f64" else lib.vt_potrf_f32
    g = lib.vt_potrs_f64 if prec == "f64" else lib.vt_potrs_f32
    sync = torch.zeros(2 + 2 * nblk, dtype=torch.int32, device="cuda")
    f(ptr(mats), 1, D, ptr(inv), inv.numel(), ptr(info), stream_ptr())
    g(ptr(mats), 1, D, ptr(inv), inv.numel(), ptr(x), D, ptr(sync), stream_ptr())
    # a second solve reuses the flag workspace (epoch advanced on device)
    x2 = torch.tensor(b, dtype=torch.float64, device="cuda")
    g(ptr(mats), 1, D, ptr(inv), inv.numel(), ptr(x2), D, ptr(sync), stream_ptr())
    torch.cuda.synchronize()
    assert int(info.item()) == 0
    assert int(sync[1].item()) == 0 and int(sync[0].item()) == 2
    np.testing.assert_array_equal(x.cpu().numpy(), x2.cpu().numpy())
    L = np.tril(a.cpu().numpy().astype(np.float64))
    Lref = np.linalg.cholesky(A)
    tol = 1e-12 if prec == "f64" else 1e-5
    assert np.abs(L - Lref).max() / np.abs(Lref).max() < tol
    xr = np.linalg.solve(A, b)
    assert np.abs(x.cpu().numpy() - xr).max() / np.abs(xr).max() < (1e-11 if prec == "f64" else 1e-4)


def test_potrf_batched_ragged(torch_mod):
    """One launch sequence factors matrices of different sizes (serving form), some ending
    inside an outer panel whose trailing update the others still receive."""
    torch = torch_mod
    from vtem_amd.block import mat_descs
    from vtem_amd._lib import lib, ptr, stream_ptr
    _banner("potrf batched ragged")
    sizes = [45, 130, 64, 300, 7, 513, 700, 256]
    rng = np.random.default_rng(5)
    mats, As, bs = [], [], []
    maxD = max(sizes)
    for D in sizes:
        G = rng.standard_normal((D, D))
        A = G @ G.T / D + np.eye(D)
        As.append(A)
        mats.append(torch.tensor(A, dtype=torch.float64, device="cuda"))
        bs.append(rng.standard_normal(D))
    nblk = (maxD + 63) // 64
    inv = torch.zeros(len(sizes) * nblk * 4096, dtype=torch.float64, device="cuda")
    info = torch.zeros(len(sizes), dtype=torch.int32, device="cuda")
    desc = mat_descs([(m.data_ptr(), D, D) for m, D in zip(mats, sizes)], "cuda")
    x = torch.zeros(len(sizes) * maxD, dtype=torch.float64, device="cuda")
    for i, b in enumerate(bs):
        x[i * maxD:i * maxD + len(b)] = torch.tensor(b, device="cuda")
    lib.vt_potrf_f64(ptr(desc), len(sizes), maxD, ptr(inv), nblk * 4096, ptr(info), stream_ptr())
    sync = torch.zeros(2 + 2 * len(sizes) * nblk, dtype=torch.int32, device="cuda")
    lib.vt_potrs_f64(ptr(desc), len(sizes), maxD, ptr(inv), nblk * 4096, ptr(x), maxD, ptr(sync),
                     stream_ptr())
    torch.cuda.synchronize()
    assert (info.cpu().numpy() == 0).all()
    assert int(sync[1].item()) == 0
    xs = x.cpu().numpy()
    for i, (A, b) in enumerate(zip(As, bs)):
        xr = np.linalg.solve(A, b)
        got = xs[i * maxD:i * maxD + len(b)]
        assert np.abs(got - xr).max() / np.abs(xr).max() < 1e-11


@pytest.mark.parametrize("mode,tol", [(0, 2e-5), (1, 3e-3)])
def test_gram_lowprec(torch_mod, oracle, mode, tol):
    """Split-BF16 (3 products) and FP16 Gram vs the FP64 oracle Gram: relative Frobenius
    error of the lower triangle below the stated tolerance."""
    torch = torch_mod
    from vtem_amd.basis import Geometry
    from vtem_amd.block import Decoder, dev
    _banner(f"gram low precision mode={mode}")
    rng = np.random.default_rng(11)
    N, L, Lt = 300, 8, 8
    geom = Geometry(64, 64, L, L, Lt, 2.0)
    S = rng.standard_normal((N, geom.Ds)) * 3.0
    al = rng.standard_normal((9000, geom.Dt)) * 0.01
    nid = np.sort(rng.integers(0, N, 9000)).astype(np.int32)
    Tm, _ = oracle.neuron_T(al, np.zeros(9000), nid, N)
    ref = np.tril(oracle.gram_kron(S, Tm))
    dec = Decoder(geom, dev(S), np.ones(N), np.ones(N), np.ones(N))
    Td = dev(Tm)
    A32 = dec.gram_lp(Td, mode)
    got = np.tril(A32.cpu().numpy()).astype(np.float64)      # the upper triangle is never written
    err = np.linalg.norm(got - ref) / np.linalg.norm(ref)
    print(f"gram mode {mode} rel Frobenius error {err:.3e}")
    assert err < tol, err


@pytest.mark.parametrize("mode", [0, 1])
def test_gram_tilings_bit_identical(torch_mod, oracle, mode, tmp_path):
    """The three tilings of the low-precision Gram (2, 4, 8 row tiles per wave) write the
    same bits, for temporal orders filling 1, 3, 6 and 10 column tiles and a spatial order
    that leaves partial 16-blocks; the autotuner's launch leaves the same Gram again, and
    one tiling agrees with the FP64 oracle Gram."""
    torch = torch_mod
    from vtem_amd._lib import lib, ptr, stream_ptr
    from vtem_amd.autotune import GramTuner, TILES
    from vtem_amd.block import dev, gram_descs, split_T_size
    _banner(f"gram tilings mode={mode}")
    rng = np.random.default_rng(23)
    N, Ds = 150, 75
    tuner = GramTuner(path=str(tmp_path / "tiles.json"), enabled=True, reps=1)
    for Dt in (5, 9, 13, 17):
        D = Ds * Dt
        S = rng.standard_normal((N, Ds))
        al = rng.standard_normal((2000, Dt)) * 0.1
        nid = np.sort(rng.integers(0, N, 2000)).astype(np.int32)
        Tm, _ = oracle.neuron_T(al, np.zeros(2000), nid, N)
        Sd, Td = dev(S), dev(Tm)
        n = split_T_size(N, Dt)
        hi = torch.empty(n, dtype=torch.int16, device="cuda")
        lo = torch.empty(n if mode == 0 else 1, dtype=torch.int16, device="cuda")
        scales = torch.ones(5, dtype=torch.float32, device="cuda")
        s = stream_ptr()
        lib.vt_gram_split_T(ptr(Td), N, Dt, ptr(hi), ptr(lo), mode, ptr(Sd), Sd.numel(), ptr(scales), s)
        outs = []
        for tile in TILES + (0,):
            A = torch.full((D, D), float("nan"), dtype=torch.float32, device="cuda")
            desc = gram_descs([(Sd, hi, lo, A, N, Ds, Dt, scales)], "cuda")
            lib.vt_gram_lp(ptr(desc), 1, Ds, Dt, mode, tile, 0, 0, s)
            outs.append((A, desc))
        A = torch.full((D, D), float("nan"), dtype=torch.float32, device="cuda")
        desc = gram_descs([(Sd, hi, lo, A, N, Ds, Dt, scales)], "cuda")
        chosen = tuner.launch(lambda t: lib.vt_gram_lp(ptr(desc), 1, Ds, Dt, mode, t, 0, 0, s),
                              "bf16x3" if mode == 0 else "fp16", Dt, Ds, 1, N)
        outs.append((A, desc))
        torch.cuda.synchronize()
        bits = [a.view(torch.int32) for a, _ in outs]
        for b in bits[1:]:
            assert torch.equal(bits[0], b), f"Dt={Dt}"
        assert chosen in TILES
        ref = np.tril(oracle.gram_kron(S, Tm))
        got = np.tril(np.nan_to_num(outs[0][0].cpu().numpy().astype(np.float64)))
        err = np.linalg.norm(got - ref) / np.linalg.norm(ref)
        print(f"Dt={Dt}: tilings identical, tuned tile {chosen} "
              f"({', '.join(f'{t}: {ms:.3f} ms' for t, ms in tuner.timings[next(iter(tuner.timings))].items())}), "
              f"rel Frobenius error {err:.2e}")
        tuner.timings.clear()
        assert err < (2e-5 if mode == 0 else 3e-3)


@pytest.mark.parametrize("mode", [0, 1])
def test_gram_tail_split(torch_mod, oracle, mode):
    """A single-problem Gram whose last round of pair blocks would leave most CUs idle
    splits those blocks over neuron ranges (FP32 partials summed in order by a second
    kernel).  For the first spatial orders whose plan splits on this device: every element
    of the lower triangle is written, and A equals the unsplit Gram (itself pinned against
    the FP64 oracle) within FP32 reassociation, for N whose ranges are all non-empty and
    for N with empty ranges."""
    torch = torch_mod
    from vtem_amd._lib import lib, ptr, stream_ptr
    from vtem_amd.block import dev, gram_descs, gram_lp_ws_bytes, split_T_size
    _banner(f"gram tail split mode={mode}")
    rng = np.random.default_rng(41)
    Dt = 9
    tile = 2 if mode == 0 else 4
    Ds = next((d for d in range(16 * 20 + 8, 16 * 64, 16) if gram_lp_ws_bytes(d, Dt, mode, tile) > 0), None)
    assert Ds is not None, "no spatial order up to 1024 splits"
    for N in (1000, 300):
        D = Ds * Dt
        S = rng.standard_normal((N, Ds))
        al = rng.standard_normal((4 * N, Dt)) * 0.1
        nid = np.sort(rng.integers(0, N, 4 * N)).astype(np.int32)
        Tm, _ = oracle.neuron_T(al, np.zeros(4 * N), nid, N)
        Sd, Td = dev(S), dev(Tm)
        n = split_T_size(N, Dt)
        hi = torch.empty(n, dtype=torch.int16, device="cuda")
        lo = torch.empty(n if mode == 0 else 1, dtype=torch.int16, device="cuda")
        scales = torch.ones(5, dtype=torch.float32, device="cuda")
        s = stream_ptr()
        lib.vt_gram_split_T(ptr(Td), N, Dt, ptr(hi), ptr(lo), mode, ptr(Sd), Sd.numel(), ptr(scales), s)
        wsb = gram_lp_ws_bytes(Ds, Dt, mode, tile)
        ws = torch.empty(wsb // 4, dtype=torch.float32, device="cuda")
        outs = []
        for w, wb in ((0, 0), (ptr(ws), wsb)):
            A = torch.full((D, D), float("nan"), dtype=torch.float32, device="cuda")
            desc = gram_descs([(Sd, hi, lo, A, N, Ds, Dt, scales)], "cuda")
            lib.vt_gram_lp(ptr(desc), 1, Ds, Dt, mode, tile, w, wb, s)
            outs.append((A, desc))
        torch.cuda.synchronize()
        a, b = (torch.tril(x.double()) for x, _ in outs)
        assert not torch.isnan(b).any() and not torch.isnan(a).any()
        err = float((a - b).abs().max() / a.abs().max())
        print(f"Ds={Ds} N={N}: workspace {wsb / 2**20:.1f} MiB, split vs unsplit max rel diff {err:.2e}")
        assert err < 2e-6


@pytest.mark.parametrize("precision", ["fp64", "mixed", "bf16x3", "fp16"])
def test_decode_c0_end_to_end(torch_mod, oracle, precision):
    """Config-0 decode on the GPU vs the oracle: SNR within 0.1 dB, coefficients close."""
    torch = torch_mod
    from vtem_amd.basis import Geometry
    from vtem_amd.block import Decoder, spatial_coeffs, rf_bank
    _banner(f"decode c0 end to end ({precision})")
    vid, bank, iafp, RF, lists, cnt = _c0_decode_inputs(oracle)
    geom = Geometry(32, 32, 5, 5, 4, 1.0)
    # oracle
    o1, o2, oi, oc = oracle.select_intervals(lists, (0.0, 1.0))
    oq = oracle.compute_q(o1, o2, oi, iafp.kappa, iafp.delta, iafp.bias, np.ones(256))
    x, y = oracle._grid(32, 32, 32.0, 32.0)
    Bs = oracle.spatial_basis(x, y, oracle.spatial_modes(5, 5), 2 * np.pi / 32, 2 * np.pi / 32)
    S = oracle.spatial_coeffs(RF, Bs)
    al = oracle.temporal_factors(o1, o2, 4, 2 * np.pi)
    oT, ot = oracle.neuron_T(al, oq, oi, 256)
    A = oracle.gram_kron(S, oT)
    r = oracle.rhs_kron(S, ot)
    lam = oracle.tikhonov_lambda(A, o1.size)
    X = oracle.solve_tikhonov(A, r, lam)
    urec = oracle.synthesize(X, Bs, oracle.temporal_basis(np.arange(100) / 100, 4, 2 * np.pi))
    snr_o = oracle.snr_db(vid.frames, urec)
    # gpu
    Sg = spatial_coeffs(rf_bank(bank, 32, 32), geom)
    assert np.abs(Sg.cpu().numpy() - S).max() / np.abs(S).max() < 1e-13
    dec = Decoder(geom, Sg, iafp.kappa, iafp.delta, iafp.bias)
    res = dec.decode(_spike_buffer(torch, lists), precision=precision, refine_iters=6)
    torch.cuda.synchronize()
    # config 0's bank leaves modes at lam/|A| ~ 5e-9: the low-precision Gram cannot stay
    # positive definite there, and the guarded path falls back to FP64
    print(f"fallback={res.fallback}")
    assert res.info is None or int(res.info.item()) == 0
    lam_g = res.lam.cpu().numpy()
    assert abs(lam_g[0] - lam) / lam < 1e-12
    fr = res.frames.cpu().numpy().reshape(100, 32, 32)
    snr_g = oracle.snr_db(vid.frames, fr)
    print(f"SNR oracle {snr_o:.4f} dB  gpu {snr_g:.4f} dB  ({precision})")
    assert abs(snr_g - snr_o) < 0.1
    assert oracle.snr_db(urec, fr) > 60.0


def test_spike_space_reference_path(torch_mod, oracle):
    """Reference spike-space G (closed form), RNN and reconstruction on the GPU vs oracle."""
    torch = torch_mod
    from vtem_amd.block import dev
    from vtem_amd._lib import lib, ptr, stream_ptr
    _banner("spike-space G / rnn / reconstruct")
    vid, bank, iafp, RF, lists, cnt = _c0_decode_inputs(oracle)
    sub = list(range(0, 256, 16))       # 16 neurons -> n ~ 800 intervals
    ls = [lists[j] for j in sub]
    o1, o2, oi, oc = oracle.select_intervals(ls, (0.0, 1.0))
    n = o1.size
    norm = np.ones(len(sub))
    RF4 = RF[sub]
    Ds_ = oracle.compute_Ds_fft(RF4, 1.0, 1.0, 5, 5, "ellipse")
    Dsw = oracle.compute_Dsw(Ds_, norm)
    G = oracle.compute_G_spike(o1, o2, oi, Dsw, 2 * np.pi * 4, 4)
    Gg = torch.zeros((n, n), dtype=torch.float64, device="cuda")
    d1, d2, dw, di = dev(o1), dev(o2), dev(Dsw), dev(oi, torch.int32)
    lib.vt_G_spike(ptr(Gg), n, ptr(d1), ptr(d2), 2 * np.pi * 4, 4, ptr(dw), len(sub), ptr(di), n,
                   stream_ptr())
    e = np.linalg.norm(Gg.cpu().numpy() - G) / np.linalg.norm(G)
    assert e < 1e-13, e
    q = oracle.compute_q(o1, o2, oi, iafp.kappa[sub], iafp.delta[sub], iafp.bias[sub], norm)
    c = oracle.rnn3(G, q, dt=1e-6, alpha=5000, steps=200)
    cg = torch.zeros(n, dtype=torch.float64, device="cuda")
    w1 = torch.empty(n, dtype=torch.float64, device="cuda")
    w2 = torch.empty(n, dtype=torch.float64, device="cuda")
    dq = dev(q)
    lib.vt_rnn_spike(ptr(cg), ptr(Gg), n, ptr(dq), n, 1e-6 * 5000, 200, ptr(w1), ptr(w2),
                     stream_ptr())
    e = np.abs(cg.cpu().numpy() - c).max() / np.abs(c).max()
    assert e < 1e-12, e
    psi = oracle.compute_psi(Ds_, 32, 32, 32.0, 32.0, 2 * np.pi / 32, 2 * np.pi / 32)
    t = np.arange(0, 1.0, 0.01)
    u = oracle.reconstruct_spike(c, o1, o2, oi, norm, psi, t, 4, 2 * np.pi)
    ug = torch.zeros((t.size, 1024), dtype=torch.float64, device="cuda")
    dpsi, dt_, dn = dev(psi.reshape(len(sub), -1)), dev(t), dev(norm)
    lib.vt_reconstruct_spike(ptr(ug), 1024, ptr(dpsi), 1024, ptr(d1), ptr(d2), ptr(cg), ptr(dt_),
                             t.size, ptr(di), ptr(dn), 4, 2 * np.pi, n, 1024, stream_ptr())
    e = np.abs(ug.cpu().numpy().reshape(u.shape) - u).max() / np.abs(u).max()
    assert e < 1e-11, e


def test_decode_c1_refinement_and_snr(torch_mod, oracle):
    """64x64 px, 2 s, D = 4913, 2048 neurons: every precision form reaches the FP64
    solution (refinement to 1e-10 relative residual) and its SNR is within 0.1 dB of the
    float64 oracle."""
    torch = torch_mod
    from vtem_amd import synthetic as syn
    from vtem_amd.basis import Geometry
    from vtem_amd.block import Decoder, Encoder, dev, spatial_coeffs
    _banner("decode c1 all precisions")
    geom = Geometry(64, 64, 8, 8, 8, 2.0)
    vid = syn.make_video(64, 64, 2.0, 100.0, 8, 8, 8, seed=0)
    bank = syn.bank_dyadic(64.0, 64.0)
    iafp = syn.iaf_params(bank.num_neurons, 2.5, 1.0, 0.04, 0.004, np.random.default_rng(1))
    enc = Encoder(geom, bank)
    st = enc.encode(dev(vid.frames.reshape(geom.frames, -1)), iafp)
    dec = Decoder(geom, spatial_coeffs(enc.RF, geom), iafp.kappa, iafp.delta, iafp.bias)
    # oracle on the same spikes
    lists = st.to_lists()
    RF = oracle.rf_bank(bank, 64, 64)
    B = geom.spatial().reshape(64, 64, -1)
    S = oracle.spatial_coeffs(RF, B)
    tk1, tk2, nid, _ = oracle.select_intervals(lists, (0.0, 2.0))
    q = oracle.compute_q(tk1, tk2, nid, iafp.kappa, iafp.delta, iafp.bias, np.ones(2048))
    Tm, tv = oracle.neuron_T(oracle.temporal_factors(tk1, tk2, 8, geom.W), q, nid, 2048)
    A = oracle.gram_kron(S, Tm)
    X = oracle.solve_tikhonov(A, oracle.rhs_kron(S, tv), oracle.tikhonov_lambda(A, tk1.size))
    u_o = oracle.synthesize(X, B, oracle.temporal_basis(geom.frame_times(), 8, geom.W))
    snr_o = oracle.snr_db(vid.frames, u_o)
    for prec in ["fp64", "mixed", "bf16x3", "fp16"]:
        res = dec.decode(st, precision=prec, refine_iters=6)
        torch.cuda.synchronize()
        fr = res.frames.cpu().numpy().reshape(vid.frames.shape)
        snr = oracle.snr_db(vid.frames, fr)
        line = f"{prec}: SNR {snr:.4f} dB (oracle {snr_o:.4f})"
        if res.resid is not None:
            h = np.sqrt(res.resid.cpu().numpy())
            rel = h[1:] / h[0]
            line += " residuals " + " ".join(f"{v:.1e}" for v in rel)
            assert rel[-1] < 1e-10, rel
        print(line)
        assert abs(snr - snr_o) < 0.1
    from vtem_amd.batch import BatchDecoder, BlockJob
    bd = BatchDecoder(precision="kpcg", verify=True)
    res = bd.decode([BlockJob(dec, st)])[0]
    h = res.resid.cpu().numpy()
    fr = res.frames.cpu().numpy().reshape(vid.frames.shape)
    snr = oracle.snr_db(vid.frames, fr)
    its = int(res.state.cpu().numpy()[3])
    print(f"kpcg: SNR {snr:.4f} dB (oracle {snr_o:.4f}) iterations {its} "
          f"residual {np.sqrt(h[1] / h[0]):.1e}")
    assert not res.fallback and np.sqrt(h[1] / h[0]) < 1e-10
    assert abs(snr - snr_o) < 0.1


def _encode_block(torch, blk):
    from vtem_amd.basis import Geometry
    from vtem_amd.block import Decoder, Encoder, dev, spatial_coeffs
    geom = Geometry(blk.side, blk.side, blk.Lx, blk.Lx, blk.Lt, blk.duration)
    enc = Encoder(geom, blk.bank)
    st = enc.encode(dev(blk.video.frames.reshape(geom.frames, -1)), blk.iaf)
    dec = Decoder(geom, spatial_coeffs(enc.RF, geom), blk.iaf.kappa, blk.iaf.delta, blk.iaf.bias)
    return geom, st, dec


def test_batch_equals_alone(torch_mod, oracle):
    """Ragged batch (sides 32/48, durations 0.5/1/2 s): every block's frames are
    bit-identical alone and in the mix (split-BF16 path), and SNR matches the oracle."""
    torch = torch_mod
    from vtem_amd import synthetic as syn
    from vtem_amd.batch import BatchDecoder, BlockJob
    _banner("batch == alone (ragged)")
    specs = [(32, 0.5), (48, 1.0), (32, 1.0), (48, 0.5), (32, 2.0), (48, 2.0)]
    blocks = [syn.serve_block(s, d, seed=100 + i) for i, (s, d) in enumerate(specs)]
    jobs, alone = [], []
    for blk in blocks:
        geom, st, dec = _encode_block(torch, blk)
        jobs.append(BlockJob(dec, st))
        alone.append(dec.decode(st, precision="bf16x3", verify=False).frames.cpu().numpy())
    bd = BatchDecoder(precision="bf16x3", verify=False)
    res = bd.decode(jobs)
    torch.cuda.synchronize()
    for i, (r, a) in enumerate(zip(res, alone)):
        np.testing.assert_array_equal(r.frames.cpu().numpy(), a, err_msg=f"block {i} {specs[i]}")
    # reversed order gives the same per-block results
    res2 = bd.decode(jobs[::-1])
    for r, a in zip(res2, alone[::-1]):
        np.testing.assert_array_equal(r.frames.cpu().numpy(), a)
    # with the guard on (as served), every block's SNR matches its FP64 decode; blocks
    # whose low-precision factor is not positive definite (short blocks with barely more
    # neurons than spatial modes) are redone in FP64 and flagged
    res = BatchDecoder(precision="bf16x3", verify=True).decode(jobs)
    for i, (blk, jb, r) in enumerate(zip(blocks, jobs, res)):
        f64 = jb.decoder.decode(jb.spikes, precision="fp64").frames.cpu().numpy()
        u = blk.video.frames.reshape(f64.shape)
        s_b, s_64 = oracle.snr_db(u, r.frames.cpu().numpy()), oracle.snr_db(u, f64)
        print(f"block {specs[i]}: D={blk.D} SNR batched {s_b:.3f} dB fp64 {s_64:.3f} dB "
              f"fallback={r.fallback}")
        assert abs(s_b - s_64) < 0.1


@pytest.mark.parametrize("pre", ["kron", "kron_sum"])
def test_kpcg_batch(torch_mod, oracle, pre):
    """Kronecker-preconditioned CG over a ragged batch, including short blocks whose FP32
    Cholesky fails: every block reaches 1e-10 relative residual of the FP64 system, its
    SNR is within 0.1 dB of the FP64 Cholesky decode, and its frames are bit-identical
    alone and in the batch -- with the Kronecker-product preconditioner and with the
    Kronecker-sum one (S^T C S) kron Tbar + lam I."""
    torch = torch_mod
    from vtem_amd import synthetic as syn
    from vtem_amd.batch import BatchDecoder, BlockJob
    _banner(f"kpcg batch, preconditioner {pre}")
    specs = [(32, 0.5), (48, 1.0), (32, 1.0), (48, 0.5), (32, 2.0), (48, 2.0), (64, 0.5)]
    blocks = [syn.serve_block(s, d, seed=300 + i) for i, (s, d) in enumerate(specs)]
    jobs = []
    for blk in blocks:
        geom, st, dec = _encode_block(torch, blk)
        jobs.append(BlockJob(dec, st))
    bd = BatchDecoder(precision="kpcg", verify=False, preconditioner=pre)
    res = bd.decode(jobs)
    assert bd.last_stats["preconditioner"] == pre
    its = [float(r.state.cpu().numpy()[3]) for r in res]
    for i, (blk, jb, r) in enumerate(zip(blocks, jobs, res)):
        h = r.resid.cpu().numpy()
        rel = float(np.sqrt(h[1] / h[0]))
        f64 = jb.decoder.decode(jb.spikes, precision="fp64").frames.cpu().numpy()
        u = blk.video.frames.reshape(f64.shape)
        s_k, s_64 = oracle.snr_db(u, r.frames.cpu().numpy()), oracle.snr_db(u, f64)
        print(f"block {specs[i]}: D={blk.D} iterations {int(its[i])} residual {rel:.1e} "
              f"SNR kpcg {s_k:.3f} dB fp64 {s_64:.3f} dB")
        assert rel < 1e-10
        assert abs(s_k - s_64) < 0.1
    for i, jb in enumerate(jobs):
        alone = bd.decode([jb])[0]
        np.testing.assert_array_equal(alone.frames.cpu().numpy(), res[i].frames.cpu().numpy(),
                                      err_msg=f"block {i} {specs[i]}")


def test_kpcg_more_blocks_than_row_tiles(torch_mod):
    """The fused CG iteration's counters (one per row tile of every block, then one per
    block) fill batch * (row tiles + 1) words: with more blocks (12) than row tiles (6 for
    Ds = 81) and the allocator's small blocks left holding all-ones words, every block
    still iterates and reaches the 1e-10 relative residual."""
    torch = torch_mod
    from vtem_amd import synthetic as syn
    from vtem_amd.batch import BatchDecoder, BlockJob
    _banner("kpcg: 12 blocks of 6 row tiles on reused memory")
    jobs = []
    for i in range(12):
        geom, st, dec = _encode_block(torch, syn.serve_block(32, 0.5, seed=800 + i))
        jobs.append(BlockJob(dec, st))
    junk = [torch.full((64 * 1024,), -1, dtype=torch.int32, device="cuda") for _ in range(256)]
    torch.cuda.synchronize()
    del junk
    res = BatchDecoder(precision="kpcg", verify=False, preconditioner="kron").decode(jobs)
    for i, r in enumerate(res):
        its = float(r.state.cpu().numpy()[3])
        h = r.resid.cpu().numpy()
        rel = float(np.sqrt(h[1] / h[0]))
        assert its > 0 and rel < 1e-10, (i, its, rel)


def test_kpcg_retry_guard(torch_mod, oracle):
    """With too few CG iterations every block misses the 1e-10 guard: the guard gives each
    a correction solve of retry_iters iterations from its iterate (no FP64 redo), which
    reaches the guard, equals -- bit for bit -- a solve warm-started by hand from the short
    solve's coefficients, and decodes within 0.1 dB of the FP64 path; with the retry
    switched off (retry_iters 0) the same blocks go to the batched FP64 path and equal its
    direct result."""
    torch = torch_mod
    from vtem_amd import synthetic as syn
    from vtem_amd.batch import BatchDecoder, BlockJob
    _banner("kpcg retry guard")
    specs = [(32, 1.0), (48, 0.5), (48, 1.0)]
    blocks = [syn.serve_block(s, d, seed=900 + i) for i, (s, d) in enumerate(specs)]
    jobs = []
    for blk in blocks:
        geom, st, dec = _encode_block(torch, blk)
        jobs.append(BlockJob(dec, st))
    bd = BatchDecoder(precision="kpcg", cg_iters=2, retry_iters=64)
    res = bd.decode(jobs)
    assert bd.last_stats["kpcg_retries"] == len(jobs) and bd.last_stats["fallbacks"] == 0
    short = BatchDecoder(precision="kpcg", cg_iters=2, verify=False).decode(jobs)
    warm = [dataclasses.replace(jb, x0=r.X.clone()) for jb, r in zip(jobs, short)]
    ref = BatchDecoder(precision="kpcg", cg_iters=64, verify=False).decode(warm)
    f64 = BatchDecoder(precision="fp64").decode(jobs)
    for i, (r, q, f) in enumerate(zip(res, ref, f64)):
        assert not r.fallback
        h = r.resid.cpu().numpy()
        assert math.sqrt(h[1] / h[0]) <= 1e-10
        np.testing.assert_array_equal(r.frames.cpu().numpy(), q.frames.cpu().numpy(),
                                      err_msg=f"block {i} {specs[i]}")
        u = blocks[i].video.frames.reshape(f.frames.shape)
        assert abs(oracle.snr_db(u, r.frames.cpu().numpy()) - oracle.snr_db(u, f.frames.cpu().numpy())) < 0.1
    bd0 = BatchDecoder(precision="kpcg", cg_iters=2, retry_iters=0)
    res0 = bd0.decode(jobs)
    assert bd0.last_stats["kpcg_retries"] == 0 and bd0.last_stats["fallbacks"] == len(jobs)
    for i, (r, q) in enumerate(zip(res0, f64)):
        assert r.fallback
        np.testing.assert_array_equal(r.frames.cpu().numpy(), q.frames.cpu().numpy(),
                                      err_msg=f"block {i} {specs[i]}")


def test_kpcg_prune(torch_mod, oracle):
    """Pruning: with a cap of 18 iterations and the decay test from iteration 10, blocks
    whose last 10 iterations predict no stop by iteration 18 end at iteration 10 and are
    redone in FP64 (equal to the direct FP64 decode, bit for bit); the others are those
    of the unpruned solve, bit for bit; a block decoded alone is pruned the same way."""
    torch = torch_mod
    from vtem_amd import synthetic as syn
    from vtem_amd.batch import BatchDecoder, BlockJob
    _banner("kpcg prune")
    specs = [(32, 0.5), (48, 1.0), (32, 1.0), (48, 2.0), (64, 0.5)]
    jobs = []
    for i, (sd, du) in enumerate(specs):
        blk = syn.serve_block(sd, du, seed=500 + i)
        geom, st, dec = _encode_block(torch, blk)
        jobs.append(BlockJob(dec, st))
    pruned = BatchDecoder(precision="kpcg", cg_iters=18, retry_iters=0, prune=10)
    res = pruned.decode(jobs)
    plain = BatchDecoder(precision="kpcg", cg_iters=18, retry_iters=0).decode(jobs)
    f64 = BatchDecoder(precision="fp64").decode(jobs)
    n_pruned = 0
    for i, (r, q, f) in enumerate(zip(res, plain, f64)):
        if r.fallback:
            n_pruned += 1
            np.testing.assert_array_equal(r.frames.cpu().numpy(), f.frames.cpu().numpy())
        else:
            np.testing.assert_array_equal(r.frames.cpu().numpy(), q.frames.cpu().numpy())
        alone = pruned.decode([jobs[i]])[0]
        assert alone.fallback == r.fallback
    its = [int(r.state.cpu().numpy()[3]) for r in BatchDecoder(precision="kpcg", cg_iters=18, retry_iters=0,
                                                              prune=10, verify=False).decode(jobs)]
    print(f"iterations {its}, pruned {n_pruned} of {len(jobs)}")
    assert n_pruned >= 1 and 10 in its


def test_rnn_batch(torch_mod, oracle):
    """Batched gradient-flow decoder (blocks of different temporal orders in one joint
    launch sequence) against the per-block path and the oracle's coefficient-space RNN
    with the same step: iterates agree to 1e-9 relative, SNR within 0.1 dB."""
    torch = torch_mod
    from vtem_amd import synthetic as syn
    from vtem_amd.batch import BatchDecoder, BlockJob
    _banner("rnn batch")
    specs = [(32, 0.5), (32, 2.0), (48, 1.0)]
    steps = 60
    blocks = [syn.serve_block(s, d, seed=700 + i) for i, (s, d) in enumerate(specs)]
    jobs = []
    for blk in blocks:
        geom, st, dec = _encode_block(torch, blk)
        jobs.append(BlockJob(dec, st, method="rnn", rnn_steps=steps))
    res = BatchDecoder(precision="kpcg").decode(jobs)
    for i, (blk, jb, r) in enumerate(zip(blocks, jobs, res)):
        eta = float(r.state.cpu().numpy()[0])
        alone = jb.decoder.decode(jb.spikes, method="rnn", rnn_steps=steps)
        Xa, Xb = alone.X.cpu().numpy(), r.X.cpu().numpy()
        assert np.abs(Xa - Xb).max() <= 1e-9 * np.abs(Xa).max()
        # oracle: same spikes, same eta
        d = jb.decoder
        g = d.geom
        lists = jb.spikes.to_lists()
        S = d.S.cpu().numpy()
        tk1, tk2, nid, _ = oracle.select_intervals(lists, (0.0, g.duration))
        q = oracle.compute_q(tk1, tk2, nid, blk.iaf.kappa, blk.iaf.delta, blk.iaf.bias, np.ones(d.N))
        Tm, tv = oracle.neuron_T(oracle.temporal_factors(tk1, tk2, g.Lt, g.W), q, nid, d.N)
        A = oracle.gram_kron(S, Tm)
        lam = oracle.tikhonov_lambda(A, tk1.size)
        lam_max = np.linalg.eigvalsh(A + lam * np.eye(A.shape[0]))[-1]
        Xo = oracle.rnn_dspace(A, oracle.rhs_kron(S, tv), eta, steps, lam)
        err = np.abs(Xb - Xo).max() / np.abs(Xo).max()
        u = blk.video.frames.reshape(g.frames, -1)
        s_b = oracle.snr_db(u, r.frames.cpu().numpy())
        s_a = oracle.snr_db(u, alone.frames.cpu().numpy())
        print(f"block {specs[i]}: D={g.D} eta*lam_max={eta * lam_max:.4f} |X-X_oracle|/|X| {err:.1e} "
              f"SNR batched {s_b:.3f} alone {s_a:.3f}")
        # the Rayleigh quotient approaches lambda_max from below: eta*lam_max >= 1, and the
        # iteration is stable for eta*lam_max < 2
        assert 0.999 < eta * lam_max < 1.2
        assert err < 1e-9
        assert abs(s_b - s_a) < 0.1


def test_rnn_graph_replay(torch_mod):
    """RNN and power loops replayed as HIP graphs of 8 iterations (61 and 50 iterations:
    replays plus a directly launched remainder) give bit-identical iterates to the
    launch-by-launch loop, on every one of three decodes with eager work between them."""
    torch = torch_mod
    from vtem_amd import synthetic as syn
    from vtem_amd.batch import BatchDecoder, BlockJob
    _banner("rnn graph replay")
    specs = [(32, 0.5), (32, 2.0), (48, 1.0)]
    jobs = []
    for i, (sd, du) in enumerate(specs):
        blk = syn.serve_block(sd, du, seed=720 + i)
        geom, st, dec = _encode_block(torch, blk)
        jobs.append(BlockJob(dec, st, method="rnn", rnn_steps=60))
    ref = BatchDecoder(precision="kpcg", graph_chunk=0, rnn_persistent=False).decode(jobs)
    bd = BatchDecoder(precision="kpcg", graph_chunk=8, rnn_persistent=False)
    for rep in range(3):
        res = bd.decode(jobs)
        torch.cuda.synchronize()
        for a, b in zip(ref, res):
            assert torch.equal(a.X, b.X), f"replay {rep}"
            assert torch.equal(a.state, b.state)
        # eager work on the same stream between the graph launches
        torch.ones(1 << 20, dtype=torch.float64, device="cuda").sum().item()


def test_rnn_persistent(torch_mod):
    """The persistent RNN launch (every iteration in one launch, per-block workgroup
    groups with their own barriers, X and Z through ring slots) gives the launch form's
    iterates bit for bit: temporal orders 5, 9 and 17 (the vector-column route), backward
    products that split K, a ring holding every iteration and a 32-slot ring with cache
    invalidation, groups scaled down to a few workgroups each (many tiles per
    workgroup), a block alone and in the batch, and a launch whose barrier waits are
    given no time (it aborts, and the jobs are solved again by the launch form)."""
    torch = torch_mod
    from vtem_amd import batch as vb
    from vtem_amd import synthetic as syn
    from vtem_amd.batch import BatchDecoder, BlockJob
    _banner("rnn persistent")
    specs = [(32, 0.5), (48, 1.0), (64, 2.0), (96, 1.0), (32, 2.0)]
    steps = 60
    jobs = []
    for i, (sd, du) in enumerate(specs):
        blk = syn.serve_block(sd, du, seed=790 + i)
        geom, st, dec = _encode_block(torch, blk)
        jobs.append(BlockJob(dec, st, method="rnn", rnn_steps=steps))
    ns = [int(jb.decoder.N) for jb in jobs]
    assert max(ns) > 1008                      # some backward products split K
    ref = BatchDecoder(precision="kpcg", graph_chunk=0, rnn_persistent=False).decode(jobs)

    def same(res, what):
        torch.cuda.synchronize()
        for i, (a, b) in enumerate(zip(ref, res)):
            assert torch.equal(a.X, b.X), (what, specs[i])
            assert torch.equal(a.state, b.state), (what, specs[i])
            assert torch.equal(a.frames, b.frames), (what, specs[i])
        print(f"{what}: bit-identical on {len(res)} blocks", flush=True)

    bd = BatchDecoder(precision="kpcg", rnn_persistent=True)
    print("case: full ring", flush=True)
    same(bd.decode(jobs), "full ring")
    assert not bd.last_stats.get("rnn_aborted")
    print("case: 32-slot ring", flush=True)
    bd.rnn_ring_bytes = 0
    same(bd.decode(jobs), "32-slot ring")
    bd.rnn_ring_bytes = 2 << 30
    print("case: block alone", flush=True)
    alone = bd.decode([jobs[2]])
    torch.cuda.synchronize()
    assert torch.equal(alone[0].X, ref[2].X)
    saved = vb._rnn_capacity()
    try:
        print("case: 8 workgroups in all", flush=True)
        vb._RNN_CAP = ([8, 8], saved[1])
        N = np.array(ns, np.int64)
        Ds = np.array([jb.decoder.geom.Ds for jb in jobs], np.int64)
        Dt = np.array([jb.decoder.geom.Dt for jb in jobs], np.int64)
        plan = bd._rnn_plan(N, Ds, Dt)
        assert plan is not None and plan[1] <= 8
        same(bd.decode(jobs), "8 workgroups")
    finally:
        vb._RNN_CAP = saved
    print("case: barrier waits given no time", flush=True)
    bd.rnn_limit_s = 1e-12
    res = bd.decode(jobs)
    assert bd.last_stats.get("rnn_aborted") == len(jobs)
    same(res, "aborted and redone")


def test_kpcg_graph_replay(torch_mod):
    """The CG loop replayed as HIP graphs (each block's iteration index from its own
    count) gives bit-identical iterates, residual histories and iteration counts to the
    launch-by-launch loop, on three decodes in a row, with blocks that stop at different
    iterations and with a retry solve (more iterations) for blocks that miss the guard."""
    torch = torch_mod
    from vtem_amd import synthetic as syn
    from vtem_amd.batch import BatchDecoder, BlockJob
    _banner("kpcg graph replay")
    specs = [(32, 0.5), (32, 2.0), (48, 1.0), (64, 2.0)]
    jobs = []
    for i, (sd, du) in enumerate(specs):
        blk = syn.serve_block(sd, du, seed=760 + i)
        geom, st, dec = _encode_block(torch, blk)
        jobs.append(BlockJob(dec, st))
    for iters, retry in ((32, None), (3, 40)):
        ref = BatchDecoder(precision="kpcg", graph_chunk=0, cg_iters=iters, retry_iters=retry).decode(jobs)
        bd = BatchDecoder(precision="kpcg", graph_chunk=8, cg_iters=iters, retry_iters=retry)
        for rep in range(3):
            res = bd.decode(jobs)
            torch.cuda.synchronize()
            for a, b in zip(ref, res):
                assert torch.equal(a.X, b.X), (iters, rep)
                assert torch.equal(a.state, b.state)
                assert torch.equal(a.resid, b.resid)
            torch.ones(1 << 20, dtype=torch.float64, device="cuda").sum().item()
        print(f"cg_iters={iters}: iterations {[int(r.state.cpu().numpy()[3]) for r in ref]}, "
              f"retries {bd.last_stats.get('kpcg_retries')}")


def test_server_streams_run_side_by_side(torch_mod):
    """The server's streams (independent_streams) sit on distinct hardware queues: work on
    each runs while a spin kernel occupies any other, both ways; a server reuses the same
    streams on every serve call."""
    torch = torch_mod
    from vtem_amd.serve import Server, independent_streams, streams_overlap
    _banner("server streams on distinct hardware queues")
    ss = independent_streams(3)
    for i in range(3):
        for j in range(i + 1, 3):
            assert streams_overlap(ss[i], ss[j]), (i, j)
    srv = Server()
    a = srv._lane_streams(3)
    assert srv._lane_streams(3) == a and srv._lane_streams(2) == a[:2]


def test_server_alone_equals_mix(torch_mod):
    """Requests served with Poisson arrivals on 2 streams (continuous batching) give the
    same frames as each request decoded alone with the same precision decision."""
    torch = torch_mod
    import numpy as np
    from vtem_amd import synthetic as syn
    from vtem_amd.batch import BatchDecoder, BlockJob
    from vtem_amd.serve import Request, Server, footprint_bytes, poisson_arrivals
    _banner("server: alone == in the mix")
    specs = [(32, 2.0), (48, 2.0), (32, 1.0), (48, 1.0), (32, 2.0), (48, 2.0), (64, 1.0), (32, 0.5)]
    jobs = []
    for i, (sd, du) in enumerate(specs):
        blk = syn.serve_block(sd, du, seed=500 + i)
        geom, st, dec = _encode_block(torch, blk)
        jobs.append(BlockJob(dec, st, tag=i))
    arr = poisson_arrivals(len(jobs), 200.0, np.random.default_rng(3))
    reqs = [Request(i, float(a), jb, jb.decoder.geom.frames, footprint_bytes(jb))
            for i, (jb, a) in enumerate(zip(jobs, arr))]
    server = Server(max_batch=3, streams=2, adaptive=False)
    st = server.serve(reqs)
    print(f"served: batches={st['batches']} fallbacks={st['fallbacks']} p50={st['p50_ms']:.2f} ms")
    bd = BatchDecoder(precision="bf16x3", verify=True)
    for r in reqs:
        alone = bd.decode([r.job])[0]
        assert alone.fallback == r.result.fallback
        np.testing.assert_array_equal(r.result.frames.cpu().numpy(), alone.frames.cpu().numpy())


@pytest.mark.parametrize("lane_delays", [None, (0.0, 0.004), "prep", "prep_staggered"])
def test_server_encode_decode_alone_equals_mix(torch_mod, lane_delays):
    """Encode+decode requests (video resident, encoded by the server's batched encoder)
    served with Poisson arrivals on 2 streams: every request's spikes and frames are
    bit-identical to the request encoded and decoded alone (KPCG form, Tikhonov and RNN
    mixed; one RNN lane, or two staggered lanes holding 0 and 4 ms).  One request gets a
    4-row spike buffer: the server detects the overflow and redoes it with an exact
    buffer, giving the alone result with a large buffer.  "prep": mostly RNN requests
    arriving faster than a chain runs, so the ones behind a running chain are prepared
    (encode, prologue, step sizes) on the preparation stream and join the next chain;
    the overflowing request is one of them.  "prep_staggered": the same requests with two
    staggered lanes, the waiting ones prepared on the held lane's stream."""
    torch = torch_mod
    import dataclasses
    import numpy as np
    from vtem_amd import synthetic as syn
    from vtem_amd.basis import Geometry
    from vtem_amd.batch import BatchDecoder, BlockJob
    from vtem_amd.batch_encode import BatchEncoder, EncodeJob, IAFDevice, max_spikes_rate
    from vtem_amd.block import Decoder, Encoder, dev, spatial_coeffs
    from vtem_amd.serve import Request, Server, footprint_bytes, poisson_arrivals
    _banner(f"server: encode+decode alone == in the mix, overflow redo, RNN lane holds {lane_delays}")
    specs = [(32, 2.0), (48, 1.0), (32, 1.0), (48, 2.0), (64, 1.0), (32, 0.5), (48, 0.5)]
    prep = lane_delays in ("prep", "prep_staggered")
    if prep:
        lane_delays = (0.0, 0.004) if lane_delays == "prep_staggered" else None
    reqs, parts = [], []
    arr = poisson_arrivals(len(specs), 2000.0 if prep else 200.0, np.random.default_rng(5))
    for i, ((sd, du), a) in enumerate(zip(specs, arr)):
        blk = syn.serve_block(sd, du, seed=700 + i)
        geom = Geometry(sd, sd, blk.Lx, blk.Lx, blk.Lt, du)
        enc = Encoder(geom, blk.bank)
        ms = 4 if i == (5 if prep else 3) else max_spikes_rate(geom, blk.iaf)
        ej = EncodeJob(enc, dev(blk.video.frames.reshape(geom.frames, -1)),
                       IAFDevice.from_params(blk.iaf), ms)
        dec = Decoder(geom, spatial_coeffs(enc.RF, geom), blk.iaf.kappa, blk.iaf.delta,
                      blk.iaf.bias)
        rnn = (i != 1) if prep else (i % 3 == 2)
        job = BlockJob(dec, None, tag=i, method="rnn" if rnn else "tikhonov", rnn_steps=100)
        parts.append((ej, job, max_spikes_rate(geom, blk.iaf)))
        reqs.append(Request(i, float(a), job, geom.frames, footprint_bytes(job, "kpcg", encode=ej),
                            encode=ej, iaf=blk.iaf))
    server = Server(max_batch=3, streams=2, precision="kpcg", adaptive=False, rnn_lane_delays=lane_delays)
    st = server.serve(reqs)
    print(f"served: batches={st['batches']} overflows={st['overflows']} rnn_preps={st['rnn_preps']} "
          f"p50={st['p50_ms']:.2f} ms", flush=True)
    assert st["overflows"] == 1
    if prep:
        assert st["rnn_preps"] >= 1
    be, bd = BatchEncoder(), BatchDecoder(precision="kpcg")
    for r, (ej, job, ms_ok) in zip(reqs, parts):
        ej1 = dataclasses.replace(ej, max_spike=max(ej.max_spike, ms_ok)) if ej.max_spike == 4 else ej
        s1 = be.encode([ej1], carry=False)[0]
        alone = bd.decode([dataclasses.replace(job, spikes=s1)])[0]
        c = s1.counts.cpu().numpy()
        np.testing.assert_array_equal(r.spikes.counts.cpu().numpy(), c)
        g, o = r.spikes.spikes.cpu().numpy(), s1.spikes.cpu().numpy()
        for j in range(c.size):                 # rows beyond a neuron's count are unwritten
            np.testing.assert_array_equal(g[:c[j], j], o[:c[j], j])
        np.testing.assert_array_equal(r.result.frames.cpu().numpy(), alone.frames.cpu().numpy())
