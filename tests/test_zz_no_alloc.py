"""tfx_mttkrp_noalloc (the split16 MTTKRP with its partials in a static device array) against
tfx_mttkrp_mixed with TFX_PREC_SPLIT16 (the same kernels with the context's workspace), through
the C ABI, on calls in a row, each on a smaller problem than the one before, and shapes that need
more than the static array are refused.

The last two modes (tile GEMMs accumulated in registers, partials summed in a fixed order) must
agree bit for bit.  The leading modes add the tile rows' weighted sums with LDS atomics, whose
order varies from call to call, so two calls of either entry point differ there by FP32 rounding
of those sums: they must agree to 1e-6 of the largest entry (the split16 result itself is within
3e-5 of the FP64 product)."""

import ctypes

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def _flat(F):
    return np.concatenate([f.flatten('F') for f in F])


def _call_both(dev, dims, R, mask, seed):
    from TensorFox import _lib
    rs = np.random.RandomState(seed)
    T = rs.standard_normal(dims).astype(np.float32)
    F = [rs.standard_normal((d, R)) for d in dims]
    Td = torch.from_numpy(T).to(dev.dev).view(-1)
    Fd = dev.upload(_flat(F), torch.float64)
    n = R * sum(dims)
    ref = dev.zeros(n, torch.float64)
    got = torch.full((n,), 7.0, dtype=torch.float64, device=dev.dev)
    ref.copy_(got)
    dev.mttkrp(Td, dims, Fd, R, ref, mask=mask, prec=_lib.TFX_PREC_SPLIT16)
    rc = dev.lib.tfx_mttkrp_noalloc(dev.ctx, ctypes.c_void_p(Td.data_ptr()), len(dims),
                                    _lib.dims_arg(dims), ctypes.c_void_p(Fd.data_ptr()), R,
                                    ctypes.c_uint(mask), ctypes.c_void_p(got.data_ptr()),
                                    dev.stream())
    _lib.check(rc, dev.ctx)
    torch.cuda.synchronize()
    return ref.cpu().numpy(), got.cpu().numpy()


def test_noalloc_mttkrp_equals_workspace_mttkrp():
    from TensorFox.device import get_device
    dev = get_device(0)
    for dims, R, mask, seed in (((96, 80, 200), 50, 7, 1), ((40, 36, 64), 20, 7, 2),
                                ((24, 20, 16, 12), 70, 15, 3), ((30, 28), 9, 3, 4)):
        print('noalloc mttkrp', dims, R, mask, flush=True)
        ref, got = _call_both(dev, dims, R, mask, seed)
        L = len(dims)
        lead = R * sum(dims[:L - 2])          # flat entries of modes 0..L-3
        assert np.array_equal(ref[lead:], got[lead:]), (dims, np.abs(ref - got).max())
        scale = np.abs(ref).max()
        assert scale > 0
        if lead:
            assert np.abs(ref[:lead] - got[:lead]).max() <= 1e-6 * scale, (dims, np.abs(ref - got).max())


def test_noalloc_mttkrp_refuses_oversized_scratch():
    from TensorFox import _lib
    from TensorFox.device import get_device
    dev = get_device(0)
    dims, R = (4, 256, 65536), 16        # K-partials alone: 8 j-blocks x 65536 x 64 x 4 B = 128 MiB
    Td = torch.zeros(int(np.prod(dims)), dtype=torch.float32, device=dev.dev)
    Fd = torch.zeros(R * sum(dims), dtype=torch.float64, device=dev.dev)
    out = torch.zeros_like(Fd)
    rc = dev.lib.tfx_mttkrp_noalloc(dev.ctx, ctypes.c_void_p(Td.data_ptr()), len(dims),
                                    _lib.dims_arg(dims), ctypes.c_void_p(Fd.data_ptr()), R,
                                    ctypes.c_uint(7), ctypes.c_void_p(out.data_ptr()), dev.stream())
    assert rc == -4
    assert b'scratch' in dev.lib.tfx_last_error(dev.ctx)
    rc = dev.lib.tfx_mttkrp_noalloc(dev.ctx, ctypes.c_void_p(Td.data_ptr()), 3,
                                    _lib.dims_arg((4, 256, 65534)), ctypes.c_void_p(Fd.data_ptr()),
                                    R, ctypes.c_uint(7), ctypes.c_void_p(out.data_ptr()),
                                    dev.stream())
    assert rc == -4                      # K % 4 != 0: outside the tiled kernel
