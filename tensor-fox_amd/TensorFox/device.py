"""Thin typed wrappers over libtfxhip entry points.

PyTorch-ROCm tensors are used only as device containers and for the current HIP stream; every
arithmetic kernel is a libtfxhip kernel.  All wrappers are asynchronous on the current stream.
"""

import ctypes

import torch

from . import _lib

_DT = {torch.float32: _lib.TFX_F32, torch.float64: _lib.TFX_F64}


def _p(t):
    return ctypes.c_void_p(t.data_ptr()) if t is not None else ctypes.c_void_p(0)


def dtype_code(t):
    try:
        return _DT[t.dtype]
    except KeyError:
        raise TypeError('unsupported dtype %s (float32 or float64 expected)' % t.dtype)


class Device:
    """A GPU and its libtfxhip context."""

    def __init__(self, device=0):
        if not torch.cuda.is_available():
            raise RuntimeError('the TensorFox MI355X path needs a ROCm GPU (torch.cuda.is_available() '
                               'is False)')
        self.index = int(device)
        self.dev = torch.device('cuda', self.index)
        self.lib = _lib.load()
        with torch.cuda.device(self.dev):
            self.ctx = _lib.context(self.index)
        cu, ma, mi = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
        self._call('tfx_device_info', ctypes.byref(cu), ctypes.byref(ma), ctypes.byref(mi))
        self.cu_count = cu.value

    # -- plumbing ------------------------------------------------------------------------------
    def stream(self):
        return ctypes.c_void_p(torch.cuda.current_stream(self.dev).cuda_stream)

    def _call(self, name, *args):
        rc = getattr(self.lib, name)(self.ctx, *args)
        _lib.check(rc, self.ctx)

    def empty(self, shape, dtype):
        return torch.empty(shape, dtype=dtype, device=self.dev)

    def zeros(self, shape, dtype):
        return torch.zeros(shape, dtype=dtype, device=self.dev)

    def upload(self, a, dtype):
        return torch.as_tensor(a).to(device=self.dev, dtype=dtype).contiguous()

    # -- kernels -------------------------------------------------------------------------------
    def mttkrp(self, T, dims, F, R, out, mask=None, prec=_lib.TFX_PREC_NATIVE):
        L = len(dims)
        if mask is None:
            mask = (1 << L) - 1
        self._call('tfx_mttkrp_mixed', _p(T), dtype_code(T), L, _lib.dims_arg(dims), _p(F),
                   dtype_code(F), int(R), ctypes.c_uint(mask), _p(out), int(prec), self.stream())

    def residual_sq(self, T, dims, F, R, out):
        """out (float64 device scalar) = ||T - [[F]]||^2."""
        self._call('tfx_cpd_residual_mixed', _p(T), dtype_code(T), len(dims), _lib.dims_arg(dims),
                   _p(F), dtype_code(F), int(R), _p(out), self.stream())

    def tensor_stats(self, X, out):
        """out[0] = sum|X|, out[1] = sum X^2 (float64 device)."""
        self._call('tfx_tensor_stats', _p(X), dtype_code(X), ctypes.c_int64(X.numel()), _p(out),
                   self.stream())

    def gramians(self, dims, R, F, Gr, P1, P2, gamma):
        self._call('tfx_gramians', dtype_code(F), len(dims), _lib.dims_arg(dims), int(R), _p(F),
                   _p(Gr), _p(P1), _p(P2), _p(gamma), self.stream())

    def mttkrp_neg_grad(self, T, dims, F, R, P1, N, grad, prec=_lib.TFX_PREC_NATIVE):
        """N = the MTTKRP of every mode, grad = N - F_l P1_l (minus the gradient)."""
        self._call('tfx_mttkrp_neg_grad', _p(T), dtype_code(T), len(dims), _lib.dims_arg(dims), _p(F),
                   dtype_code(F), int(R), _p(P1), _p(N), _p(grad), int(prec), self.stream())

    def neg_grad(self, dims, R, F, P1, N, grad):
        self._call('tfx_neg_grad', dtype_code(F), len(dims), _lib.dims_arg(dims), int(R), _p(F),
                   _p(P1), _p(N), _p(grad), self.stream())

    def cg(self, dims, R, F, P1, P2, gamma, damp, grad, maxiter, tol, y, jtjg, scal):
        self._call('tfx_cg', dtype_code(F), len(dims), _lib.dims_arg(dims), int(R), _p(F), _p(P1),
                   _p(P2), _p(gamma), float(damp), _p(grad), int(maxiter), float(tol), _p(y),
                   _p(jtjg), _p(scal), self.stream())

    def dgn_update(self, dims, R, x, y, alpha, grad, damp, eq, symm, factors_norm, x_new, F, scal):
        self._call('tfx_dgn_update', dtype_code(x), len(dims), _lib.dims_arg(dims), int(R), _p(x),
                   _p(y), float(alpha), _p(grad), float(damp), int(bool(eq)), int(symm),
                   float(factors_norm), _p(x_new), _p(F), _p(scal), self.stream())

    def dgn_run(self, T, prec, dims, R, F0, opts, inner, damp_sched, mt_key, mt_pos, Tsize, work,
                F_best, hist, info):
        """tfx_dgn_run: inner (int8 NumPy), damp_sched (float64 NumPy or None), mt_key (uint32
        NumPy, advanced in place), mt_pos (ctypes.c_int), hist (float64 NumPy 4 x maxiter), info
        (ctypes int array of 3)."""
        ptr = lambda a: ctypes.c_void_p(a.ctypes.data) if a is not None else ctypes.c_void_p(0)
        self._call('tfx_dgn_run', _p(T), dtype_code(T), int(prec), len(dims), _lib.dims_arg(dims),
                   int(R), _p(F0), ctypes.byref(opts), ptr(inner), ptr(damp_sched), ptr(mt_key),
                   ctypes.byref(mt_pos), float(Tsize), _p(work), ctypes.c_int64(work.numel()),
                   _p(F_best), ptr(hist), info, self.stream())

    def dgn_direct(self, dims, R, F, P1, P2, gamma, damp, grad, y, scal):
        self._call('tfx_dgn_direct', len(dims), _lib.dims_arg(dims), int(R), _p(F), _p(P1), _p(P2),
                   _p(gamma), float(damp), _p(grad), _p(y), _p(scal), self.stream())

    def als_solve(self, dims, R, mode, P1, N, F, status):
        """F's block of `mode` <- N's block times pinv(P1[mode]); status: int32 device scalar."""
        self._call('tfx_als_solve', dtype_code(F), len(dims), _lib.dims_arg(dims), int(R), int(mode),
                   _p(P1), _p(N), _p(F), _p(status), self.stream())

    def dogleg_dots(self, y, g, jg, out):
        self._call('tfx_dogleg_dots', dtype_code(y), ctypes.c_int64(y.numel()), _p(y), _p(g), _p(jg),
                   _p(out), self.stream())

    def axpby(self, a, y, b, g, out):
        self._call('tfx_axpby', dtype_code(y), ctypes.c_int64(y.numel()), float(a), _p(y), float(b),
                   _p(g), _p(out), self.stream())

    def gram_mode(self, X, dims, mode, G, prec=_lib.TFX_PREC_NATIVE):
        self._call('tfx_gram_mode', _p(X), dtype_code(X), len(dims), _lib.dims_arg(dims), int(mode),
                   _p(G), int(prec), self.stream())

    def eig_top(self, G, n, Q0, p, k, iters, U, lam):
        self._call('tfx_eig_top', _p(G), int(n), _p(Q0), int(p), int(k), int(iters), _p(U), _p(lam),
                   self.stream())

    def ttm(self, X, dims, mode, U, k, Y, sig2=None, prec=_lib.TFX_PREC_NATIVE):
        self._call('tfx_ttm_mixed', _p(X), dtype_code(X), len(dims), _lib.dims_arg(dims), int(mode),
                   _p(U), int(k), _p(Y), dtype_code(Y), _p(sig2), int(prec), self.stream())

    def gemm_small(self, transA, m, n, kk, A, B, C):
        self._call('tfx_gemm_small', dtype_code(B), int(bool(transA)), int(m), int(n), int(kk), _p(A),
                   _p(B), _p(C), self.stream())

    def sparse_pack(self, coords, ordering, sorted_dims, idx, bad):
        nnz = coords.shape[0] if coords.dim() == 2 else 0
        order = (ctypes.c_int32 * len(ordering))(*[int(o) for o in ordering])
        self._call('tfx_sparse_pack', ctypes.c_int64(nnz), len(sorted_dims), _p(coords), order,
                   _lib.dims_arg(sorted_dims), _p(idx), _p(bad), self.stream())

    def sparse_gram(self, nnz, dims, idx, data, mode, G):
        self._call('tfx_sparse_gram', ctypes.c_int64(nnz), len(dims), _lib.dims_arg(dims), _p(idx),
                   _p(data), int(mode), _p(G), self.stream())

    def sparse_ttm0(self, nnz, dims, idx, data, U, k, Y):
        self._call('tfx_sparse_ttm0', ctypes.c_int64(nnz), len(dims), _lib.dims_arg(dims), _p(idx),
                   _p(data), _p(U), int(k), _p(Y), self.stream())

    def sparse_residual(self, nnz, dims, idx, data, F, R, out):
        self._call('tfx_sparse_residual', ctypes.c_int64(nnz), len(dims), _lib.dims_arg(dims),
                   _p(idx), _p(data), _p(F), int(R), _p(out), self.stream())

    def convert(self, src, dst):
        self._call('tfx_convert', ctypes.c_int64(src.numel()), _p(src), dtype_code(src), _p(dst),
                   dtype_code(dst), self.stream())


_devices = {}


def get_device(index=0):
    d = _devices.get(index)
    if d is None:
        d = Device(index)
        _devices[index] = d
    return d
