"""GPU: the allocation-free dispersion image (msfdm_dispersion_image_noalloc)
called through the C ABI on caller-owned device buffers equals the allocating
entry point (msfdm_dispersion_image) bit for bit, on two calls in a row, the
second with a smaller gather; a gather beyond the scratch size is refused."""
import ctypes as C

import numpy as np
import pytest
import torch

import msfdm
from msfdm._lib import EINVAL, MsfdmError

pytestmark = pytest.mark.gpu


def _allocating(g, x, fs, vel):
    nt, ntr = g.shape
    nf = nt // 2 + 1
    amp = np.zeros((nf, vel.size))
    freqs = np.zeros(nf)
    dp = C.POINTER(C.c_double)
    rc = msfdm.lib().msfdm_dispersion_image(g.ctypes.data_as(C.POINTER(C.c_float)), nt, ntr, x.ctypes.data_as(dp),
                                            fs, vel.ctypes.data_as(dp), vel.size, 0, freqs.ctypes.data_as(dp),
                                            amp.ctypes.data_as(dp))
    assert rc == 0, msfdm._lib.last_error()
    return amp


def _noalloc(g, x, fs, vel):
    nt, ntr = g.shape
    nf = nt // 2 + 1
    dev = torch.device("cuda:0")
    dg = torch.from_numpy(g).to(dev)
    dx = torch.from_numpy(x).to(dev)
    dv = torch.from_numpy(vel).to(dev)
    da = torch.full((nf, vel.size), np.nan, dtype=torch.float64, device=dev)   # poisoned output
    stream = torch.cuda.current_stream(dev)
    torch.cuda.synchronize()
    rc = msfdm.lib().msfdm_dispersion_image_noalloc(dg.data_ptr(), nt, ntr, dx.data_ptr(), fs, dv.data_ptr(),
                                                    vel.size, da.data_ptr(), C.c_void_p(stream.cuda_stream))
    assert rc == 0, msfdm._lib.last_error()
    stream.synchronize()
    return da.cpu().numpy()


def _case(seed, nt, ntr, nc):
    rng = np.random.default_rng(seed)
    g = rng.standard_normal((nt, ntr)).astype(np.float32)
    x = (np.arange(ntr) - ntr // 3) * 2.5
    vel = np.linspace(100.0, 3000.0, nc)
    return g, x.astype(np.float64), 1000.0, vel


def test_noalloc_equals_allocating_entry_point_twice():
    torch.cuda.init()
    for seed, (nt, ntr, nc) in ((1, (512, 301, 181)), (2, (200, 97, 64))):
        print(f"[noalloc] nt={nt} ntr={ntr} nc={nc}", flush=True)
        g, x, fs, vel = _case(seed, nt, ntr, nc)
        want = _allocating(g, x, fs, vel)
        got = _noalloc(g, x, fs, vel)
        assert np.isfinite(got).all()
        assert np.array_equal(got, want)


def test_noalloc_refuses_gathers_beyond_its_scratch():
    torch.cuda.init()
    limit = msfdm._lib.DISP_SCRATCH_BYTES
    nt = 4096
    ntr = limit // ((nt // 2 + 1) * 16) + 1          # one trace more than fits
    dev = torch.device("cuda:0")
    dummy = torch.zeros(16, dtype=torch.float64, device=dev)
    rc = msfdm.lib().msfdm_dispersion_image_noalloc(dummy.data_ptr(), nt, ntr, dummy.data_ptr(), 1000.0,
                                                    dummy.data_ptr(), 4, dummy.data_ptr(), None)
    assert rc == EINVAL
    assert "too large" in msfdm._lib.last_error()
