"""The hand-written one-sided Jacobi SVD behind utils.linalg.solve_eq_sym / eig_sym, and
the plain-method configuration's decode by the truncated pseudoinverse of the Gram.

* solve_eq_sym on an indefinite symmetric matrix with a truncated part equals the oracle's
  restatement of the reference's SVD formula U diag(1/s) U^T q (vtem/utils/linalg.py:
  622-630) to 1e-10 relative; on a rank-deficient Gram it equals numpy's pinv in the range.
* eig_sym: eigenvalues within 1e-12 of numpy's (relative to the largest), ascending, and
  eigenvectors that diagonalise the matrix.
* 32 x 32 px, 1 s, Lx = Ly = 5, Lt = 4 (D = 1089), 256 neurons: the D-space pseudoinverse
  decode (rcond 1e-8, no lambda; vtem/vtdm.py:451) against the oracle's truncated
  pseudoinverse of its own Gram: equal in the range of the Gram, SNR within 0.1 dB."""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _banner(msg):
    print(f"\n[case] {msg}", flush=True)


@pytest.fixture(scope="module")
def torch_mod():
    import torch
    from vtem_amd._lib import require_gpu
    require_gpu()
    return torch


def test_solve_eq_sym_matches_svd_formula(torch_mod, oracle):
    torch = torch_mod
    from vtem_amd.utils.linalg import solve_eq_sym
    _banner("solve_eq_sym: Jacobi SVD vs the SVD formula")
    rng = np.random.default_rng(3)
    for n, case in ((301, "indefinite"), (256, "gram")):
        if case == "indefinite":
            Q, _ = np.linalg.qr(rng.standard_normal((n, n)))
            w = np.concatenate([rng.uniform(0.1, 5.0, 150), -rng.uniform(0.1, 5.0, 120),
                                1e-13 * rng.standard_normal(n - 270)])
            G = (Q * w) @ Q.T
            G = (G + G.T) / 2
        else:
            Phi = rng.standard_normal((n, 180))
            G = Phi @ Phi.T                       # rank 180
        q = rng.standard_normal(n)
        c = solve_eq_sym(torch.tensor(G, device="cuda"), q).cpu().numpy()
        c_o = oracle.solve_eq_sym(G, q, rcond=1e-8)
        rel = np.linalg.norm(c - c_o) / np.linalg.norm(c_o)
        print(f"  {case} n={n}: rel diff {rel:.2e}", flush=True)
        assert rel <= 1e-10
        if case == "gram":
            p = np.linalg.pinv(G, rcond=1e-8, hermitian=True) @ q
            assert np.linalg.norm(G @ (c - p)) <= 1e-9 * np.linalg.norm(G @ p)


def test_eig_sym(torch_mod):
    torch = torch_mod
    from vtem_amd.utils.linalg import eig_sym
    _banner("eig_sym: Jacobi")
    rng = np.random.default_rng(4)
    n = 200
    B = rng.standard_normal((n, n))
    G = (B + B.T) / 2
    w, E = eig_sym(torch.tensor(G, device="cuda"))
    w, E = w.cpu().numpy(), E.cpu().numpy()
    wn = np.linalg.eigvalsh(G)
    assert np.all(np.diff(w) >= 0)
    assert np.abs(w - wn).max() <= 1e-12 * np.abs(wn).max()
    R = E @ G @ E.T
    off = R - np.diag(np.diag(R))
    assert np.abs(off).max() <= 1e-11 * np.abs(wn).max()
    np.testing.assert_allclose(E @ E.T, np.eye(n), atol=1e-12)


def test_config0_pinv_decode_vs_oracle(torch_mod, oracle):
    torch = torch_mod
    from vtem_amd import synthetic as syn
    from vtem_amd.basis import Geometry
    from vtem_amd.block import Decoder, Encoder, dev, spatial_coeffs
    _banner("config 0: pseudoinverse decode, D = 1089")
    geom = Geometry(32, 32, 5, 5, 4, 1.0)
    vid = syn.make_video(32, 32, 1.0, 100.0, 5, 5, 4, seed=0)
    bank = syn.bank_config0()
    iafp = syn.iaf_params(bank.num_neurons, 2.0, 1.0, 0.04, 0.0, np.random.default_rng(1))
    enc = Encoder(geom, bank)
    st = enc.encode(dev(vid.frames.reshape(100, -1)), iafp, root_mode=1)
    dec = Decoder(geom, spatial_coeffs(enc.RF, geom), iafp.kappa, iafp.delta, iafp.bias)
    res = dec.decode(st, method="pinv")
    torch.cuda.synchronize()
    X_g = res.X.cpu().numpy()
    fr = res.frames.cpu().numpy().reshape(vid.frames.shape)
    lists = st.to_lists()
    N = bank.num_neurons
    RF = oracle.rf_bank(bank, 32, 32)
    B = geom.spatial().reshape(32, 32, -1)
    S = oracle.spatial_coeffs(RF, B)
    tk1, tk2, nid, _ = oracle.select_intervals(lists, (0.0, 1.0))
    q = oracle.compute_q(tk1, tk2, nid, iafp.kappa, iafp.delta, iafp.bias, np.ones(N))
    Tm, tv = oracle.neuron_T(oracle.temporal_factors(tk1, tk2, 4, geom.W), q, nid, N)
    A = oracle.gram_kron(S, Tm)
    A = (A + A.T) / 2
    r = oracle.rhs_kron(S, tv)
    X_o = oracle.solve_eq_sym(A, r, rcond=1e-8)
    u_o = oracle.synthesize(X_o, B, oracle.temporal_basis(geom.frame_times(), 4, geom.W))
    s_o, s_g = oracle.snr_db(vid.frames, u_o), oracle.snr_db(vid.frames, fr)
    rng_err = np.linalg.norm(A @ (X_g - X_o)) / np.linalg.norm(A @ X_o)
    print(f"  SNR gpu {s_g:.4f} oracle {s_o:.4f} dB; range rel diff {rng_err:.2e}", flush=True)
    assert rng_err <= 1e-8
    assert abs(s_g - s_o) < 0.1
