"""The allocation-free decode entry point (brkga_evaluate_device_noalloc)
against the allocating one (brkga_evaluate_device), through the C ABI."""
import ctypes
import sys

import numpy as np
import pytest

from oracle import oracle
from tests.helpers import oracle_threads, syn_instance, syn_keys

pytestmark = pytest.mark.gpu

_P = ctypes.c_void_p


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _bind(lib):
    f = lib.brkga_evaluate_device_noalloc
    f.argtypes = [_P, _P, ctypes.c_int, ctypes.c_int64, ctypes.c_int64, _P, _P, _P, ctypes.c_int]
    f.restype = ctypes.c_int
    return f


def _noalloc(dec, keys, out, log, stream):
    f = _bind(dec._lib)
    return f(dec._h, keys.data_ptr(), 0, keys.shape[0], keys.shape[1], out.data_ptr(),
             log.data_ptr() if log is not None else None, ctypes.c_void_p(stream.cuda_stream), 1)


@pytest.mark.parametrize("cfg,sizes", [("C2", (300, 37)), ("C3", (160, 9))])
def test_noalloc_matches_evaluate_device(cfg, sizes):
    """Two calls in a row on one handle, the second with a smaller population:
    makespans and placement logs bit-identical to brkga_evaluate_device's, and
    the smaller call's makespans equal to the oracle's."""
    import torch
    from brkga_fft2_amd.native import Decoder
    _say("no-alloc decode %s %s" % (cfg, sizes))
    prob, n, M, packed = syn_instance(cfg)
    dec = Decoder(packed, max_bins=n)
    stream = torch.cuda.current_stream()
    for N in sizes:
        keys_h = syn_keys(cfg, N, n, seed=100 + N)
        keys = torch.from_numpy(keys_h).cuda()
        ref = torch.empty(N, dtype=torch.float64, device="cuda")
        ref_log = torch.empty((N, n, 6), dtype=torch.int32, device="cuda")
        dec.evaluate_device(keys, ref, log=ref_log, stream=stream)
        got = torch.full((N,), -7.0, dtype=torch.float64, device="cuda")
        got_log = torch.full((N, n, 6), -9, dtype=torch.int32, device="cuda")
        assert _noalloc(dec, keys, got, got_log, stream) == 0, dec._lib.brkga_last_error(dec._h)
        torch.cuda.synchronize()
        assert torch.equal(got, ref)
        assert torch.equal(got_log, ref_log)
    want, _, _ = oracle.decode(packed, keys_h, n_threads=oracle_threads())
    np.testing.assert_array_equal(got.cpu().numpy(), want)
    dec.close()


def test_noalloc_refuses_oversized_populations():
    """More contexts than the scratch array holds: a clear error and no launch;
    the handle still decodes afterwards."""
    import torch
    from brkga_fft2_amd.native import Decoder
    _say("no-alloc size limit")
    prob, n, M, packed = syn_instance("C2")
    dec = Decoder(packed, max_bins=n)
    N = 65536 // M + 1
    keys = torch.zeros((N, 2 * n), dtype=torch.float32, device="cuda")
    out = torch.empty(N, dtype=torch.float64, device="cuda")
    stream = torch.cuda.current_stream()
    assert _noalloc(dec, keys, out, None, stream) == -1
    assert b"exceeds" in dec._lib.brkga_last_error(dec._h)
    small = torch.from_numpy(syn_keys("C2", 8, n, seed=3)).cuda()
    got = torch.empty(8, dtype=torch.float64, device="cuda")
    assert _noalloc(dec, small, got, None, stream) == 0
    torch.cuda.synchronize()
    want, _, _ = oracle.decode(packed, small.cpu().numpy())
    np.testing.assert_array_equal(got.cpu().numpy(), want)
    dec.close()
