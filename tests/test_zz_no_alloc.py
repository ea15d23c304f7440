"""vt_iaf_encode_par_noalloc (the time-parallel encoder with its workspace in a static
per-GPU array) equals vt_iaf_encode_par with a caller-allocated workspace, bit for bit:
two calls in a row, the second on a smaller input, so the second runs over the array the
first call left behind.  A size beyond the array is rejected with -5."""

import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _encode(lib, ptr, stream_ptr, torch, x, N, T, params, noalloc):
    from vtem_amd.block import iaf_par_workspace
    v0, kap, bia, dlt = (torch.tensor(a, dtype=torch.float64, device="cuda") for a in params)
    tc = torch.zeros(N, dtype=torch.float64, device="cuda")
    spikes = torch.full((400, N), -1.0, dtype=torch.float64, device="cuda")
    counts = torch.zeros(N, dtype=torch.int32, device="cuda")
    args = (ptr(x), 0, x.stride(0), N, T, ptr(spikes), N, 400, ptr(v0), ptr(kap), ptr(bia), ptr(dlt),
            ptr(tc), ptr(counts), 1e-4, 1)
    if noalloc:
        lib.vt_iaf_encode_par_noalloc(*args, stream_ptr())
    else:
        ws = iaf_par_workspace(N, T)
        lib.vt_iaf_encode_par(*args, ptr(ws), ws.numel(), stream_ptr())
    torch.cuda.synchronize()
    return spikes.cpu().numpy(), counts.cpu().numpy(), v0.cpu().numpy(), tc.cpu().numpy()


def test_noalloc_equals_workspace_entry():
    import torch
    from vtem_amd._lib import lib, ptr, require_gpu, stream_ptr
    require_gpu()
    print("\n[case] iaf par noalloc vs workspace entry", flush=True)
    rng = np.random.default_rng(5)
    from vtem_amd._lib import LIB_PATH, SIGNATURES
    raw = ctypes.CDLL(LIB_PATH)
    for name in ("vt_noalloc_scratch_bytes", "vt_iaf_encode_par_noalloc"):
        getattr(raw, name).argtypes = SIGNATURES[name]
        getattr(raw, name).restype = ctypes.c_int
    b = ctypes.c_longlong()
    assert raw.vt_noalloc_scratch_bytes(ctypes.byref(b)) == 0 and b.value == 64 << 20
    for N, T in ((3000, 20001), (700, 5001)):
        t = np.arange(T)[:, None] * 1e-4
        x = torch.tensor(0.5 * np.sin(2 * np.pi * (3 + rng.uniform(0, 5, N)) * t + rng.uniform(0, 6, N)),
                         dtype=torch.float64, device="cuda")
        delta = rng.normal(0.04, 0.004, N)
        params = (rng.uniform(0, 1, N) * delta, np.ones(N), np.full(N, 2.5), delta)
        a = _encode(lib, ptr, stream_ptr, torch, x, N, T, params, False)
        c = _encode(lib, ptr, stream_ptr, torch, x, N, T, params, True)
        assert a[1].sum() > 10 * N
        for u, v in zip(a, c):
            np.testing.assert_array_equal(u, v)
        print(f"  N={N} T={T}: {int(a[1].sum())} spikes, identical", flush=True)
    # beyond the array: rejected before any launch
    N, T = 100000, 20001
    dummy = torch.zeros(1, dtype=torch.float64, device="cuda")
    p = ptr(dummy)
    r = raw.vt_iaf_encode_par_noalloc(p, 0, N, N, T, p, N, 1, p, p, p, p, p, p, 1e-4, 0, 0)
    assert r == -5
