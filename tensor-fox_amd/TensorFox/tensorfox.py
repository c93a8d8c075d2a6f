"""Drop-in `cpd(T, R, options)` for dense tensors on one MI355X (TensorFox.py:69-365 -> tricpd
:424-655): sort dims -> MLSVD -> initialization -> dGN (or ALS, method='als') on the core ->
uncompression -> optional refinement on the full tensor -> final error -> original mode order.

Precision forms (option `precision`):
  'fp64'  -- the tensor is stored in FP64 and every product is FP64 (default for float64 input);
  'mixed' -- the tensor is stored in FP32 (half the HBM traffic of every pass over it); the
             passes over it (Gram, mode product, MTTKRP, residual) read FP32 and accumulate in
             FP64, with the split 16-bit MFMA form where it is built; the latency-bound stages
             (core dGN, CG, eigensolve, factors) run in FP64 (default for float32 input).
"""

import math
import sys
import time
from copy import deepcopy

import numpy as np
import torch

from . import _lib, sparse
from .auxiliar import Output, consistency, make_options, sort_dims, unsort_dims
from .als import als
from .compression import MAX_MODE_DIM, mlsvd
from .device import get_device
from .gauss_newton import Problem, dGN
from .initialization import smart, smart_random

_CLEAN_RNG = np.random.RandomState(98765)


def _equalize(F, R):
    """Conversion.py:361-387 on host factors: every rank-one term's columns rescaled to the
    geometric mean of their norms.  Column norms are sqrt(dot) of contiguous copies and the
    products run left to right, which is what np.linalg.norm and np.prod compute for these
    sizes, so the result is bitwise that of the per-column loop; the scaling is applied to all
    columns at once."""
    L = len(F)
    Ft = [np.ascontiguousarray(f.T) for f in F]
    scale = np.ones((L, R))
    for r in range(R):
        nr = [math.sqrt(Ft[l][r].dot(Ft[l][r])) for l in range(L)]
        p = nr[0]
        for l in range(1, L):
            p = p * nr[l]
        if p != 0.0:
            num = p ** (1 / L)
            for l in range(L):
                scale[l, r] = num / nr[l]
    for l in range(L):
        F[l] = F[l] * scale[l][None, :]
    return F


def _transform(F, symm, factors_norm):
    """Conversion.py:412-444 on host factors."""
    L = len(F)
    if symm:
        s = F[0]
        for l in range(1, L):
            s = s + F[l]
        F[0] = s / L
        for l in range(1, L):
            F[l] = F[0]
    if factors_norm > 0:
        for l in range(L):
            F[l] = factors_norm * (1 / np.linalg.norm(F[l])) * F[l]
    return F


def _clean_zeros(F, R):
    """Initialization.py:361-386; the reference draws these under Numba's own RNG, so a private
    generator is used to leave the global NumPy stream untouched."""
    for f in F:
        idx = np.argwhere(f == 0.0)
        for i, r in idx:
            f[i, r] = 1e-4 * _CLEAN_RNG.randn() * np.linalg.norm(f)
    return F


def _flat(F):
    return np.concatenate([f.flatten('F') for f in F])


def _unflat(x, dims, R):
    out, s = [], 0
    for d in dims:
        out.append(np.array(x[s:s + R * d].reshape(R, d).T))
        s += R * d
    return out


def _precision_of(T, options):
    if options.precision is not None:
        if options.precision not in ('mixed', 'fp64'):
            raise ValueError("options.precision must be 'mixed' or 'fp64'")
        return options.precision
    dt = T.dtype
    is32 = (dt == np.float32) if isinstance(T, np.ndarray) else (dt == torch.float32)
    return 'mixed' if is32 else 'fp64'


def _to_device(dev, T, ordering, tdt):
    """Sorted, contiguous copy of T on the device in the storage dtype."""
    L = T.ndim
    identity = all(int(ordering[l]) == l for l in range(L))
    if isinstance(T, torch.Tensor):
        X = T if T.is_cuda else T.to(dev.dev)
        if not identity:
            X = X.permute(*[int(o) for o in ordering])
        return X.to(tdt).contiguous().view(-1)
    A = np.asarray(T)
    if not identity:
        A = np.moveaxis(A, ordering, np.arange(L))
    A = np.ascontiguousarray(A, dtype=np.float32 if tdt == torch.float32 else np.float64)
    return torch.from_numpy(A).to(dev.dev).view(-1)


def starting_point(dev, S, S_dims, U, dims, R, ordering, options, wdt):
    """Flat starting factors on the core (Initialization.py:20-94): the user's factors projected
    by U_l^T, or 'random' / 'smart_random' / 'smart' followed by clean-up, equalization and
    transform.  Draws from the global NumPy generator as the reference does."""
    L = len(S_dims)
    init = options.initialization
    if type(init) == list:
        F0 = dev.empty(R * sum(S_dims), wdt)
        off = 0
        for l in range(L):
            B = dev.upload(np.asarray(init[ordering[l]], dtype=np.float64).flatten('F'), wdt)
            dev.gemm_small(True, S_dims[l], R, dims[l], U[l], B, F0[off:off + R * S_dims[l]])
            off += R * S_dims[l]
        return F0
    if init in ('random', 'smart_random', 'smart'):
        if init == 'random':
            F = [np.random.randn(S_dims[l], R) for l in range(L)]
        else:
            # index selection on the core (Initialization.py:97-358), read back once
            S_host = S.to(torch.float64).cpu().numpy().reshape(S_dims)
            F = (smart_random if init == 'smart_random' else smart)(S_host, S_dims, R)
        F = _clean_zeros(F, R)
        F = _equalize(F, R)
        F = _transform(F, options.symm, options.factors_norm)
        return dev.upload(_flat(F), wdt)
    sys.exit('Error with init parameter initialization.')


def _sync():
    torch.cuda.synchronize()


def tricpd(T, R, options):
    """Dense dGN path of TensorFox.py:424-655 on the GPU."""
    t_all = time.perf_counter()
    timings = {}
    dev = get_device(options.device)
    display = options.display
    wdt = torch.float64                                   # factors, CG, core
    prec = _lib.TFX_PREC_NATIVE
    is_sp = sparse.is_sparse(T)
    if is_sp:
        # stored entries in FP64 (TensorFox.py:476-498); the passes over them are FP64
        data, idxs, dims_in = sparse.coo_arrays(T)
        if len(data) == 0:
            raise ValueError('sparse tensor with no stored entries')
        sparse.report(data, dims_in, display)
        precision = 'fp64'
        L = len(dims_in)
        dims, ordering = sort_dims(tuple(dims_in))
        Xsp = sparse.SparseTensor(dev, data, idxs, dims_in, ordering)
        Tsize = Xsp.Tsize
        sparse.fix_tol(options)
    else:
        precision = _precision_of(T, options)
        tdt = torch.float32 if precision == 'mixed' else torch.float64   # the tensor in HBM
        # compression of an FP32 tensor: bf16x3 MFMA Gram, FP32 MFMA mode product
        mprec = _lib.TFX_PREC_SPLIT16 if precision == 'mixed' else _lib.TFX_PREC_NATIVE
        L = T.ndim
        dims, ordering = sort_dims(T.shape)
        Tdev = _to_device(dev, T, ordering, tdt)
        N = int(np.prod(dims))
        stats = dev.zeros(2, torch.float64)
        dev.tensor_stats(Tdev, stats)
        st = stats.cpu().numpy()
        Tsize = float(np.sqrt(st[1]))
        mean_abs = float(st[0]) / N
    t0 = time.perf_counter()
    timings['upload'] = t0 - t_all

    # compression
    if is_sp:
        if display != 0:
            print('-----------------------------------------------------------------------------------------------')
            print('Computing MLSVD')
        S, S_dims, U, sigmas = sparse.mlsvd(dev, Xsp, R, options)
    else:
        S, S_dims, U, sigmas = mlsvd(dev, Tdev, dims, R, options, wdt, mprec)
    if options.symm:
        R_min = min(S_dims)
        view = S.view(*S_dims)
        for l in range(L):
            view = view.narrow(l, 0, R_min)
        S = view.contiguous().view(-1)
        S_dims = [R_min] * L
        U = [u[:R_min].contiguous() for u in U]
    _sync()
    t1 = time.perf_counter()
    timings['mlsvd'] = t1 - t0
    if display > 0:
        if not is_sp:
            print('-----------------------------------------------------------------------------------------------')
            print('Computing MLSVD')
        if tuple(S_dims) == tuple(dims):
            print('    No compression detected')
            print('    Working with dimensions', tuple(dims))
        else:
            print('    Compression detected')
            print('    Compressing from', tuple(dims), 'to', tuple(S_dims))

    # initialization (Initialization.py:20-94)
    init = options.initialization
    F0 = starting_point(dev, S, S_dims, U, dims, R, ordering, options, wdt)
    Sstats = dev.zeros(2, torch.float64)
    dev.tensor_stats(S, Sstats)
    sst = Sstats.cpu().numpy()
    Ssize = float(np.sqrt(sst[1]))
    S_mean_abs = float(sst[0]) / int(np.prod(S_dims))
    t2 = time.perf_counter()
    timings['init'] = t2 - t1

    # main stage on the core: ALS for method 'als', dGN otherwise (TensorFox.py:563-568; the
    # tensor-train path calls this for its first core with method 'ttcpd')
    prob = Problem(dev, S, S_dims, R, wdt, prec)
    if options.method == 'als':
        Fm, ss_m, e_m, im_m, g_m, stop_m = als(prob, F0, options, Ssize)
    else:
        Fm, ss_m, e_m, im_m, g_m, stop_m = dGN(prob, F0, options, Ssize, S_mean_abs,
                                               aliased=type(init) != list)
    _sync()
    t3 = time.perf_counter()
    timings['main'] = t3 - t2

    # uncompression (TensorFox.py:573-574)
    Ffull = dev.empty(R * sum(dims), wdt)
    off_s, off_f = 0, 0
    for l in range(L):
        dev.gemm_small(False, dims[l], R, S_dims[l], U[l], Fm[off_s:off_s + R * S_dims[l]],
                       Ffull[off_f:off_f + R * dims[l]])
        off_s += R * S_dims[l]
        off_f += R * dims[l]
    _sync()
    t4 = time.perf_counter()
    timings['uncompress'] = t4 - t3

    # refinement (TensorFox.py:583-611); none for sparse input (TensorFox.py:577-581)
    if is_sp:
        options.refine = False
    if options.refine:
        if display > 0:
            print('Computing refinement of solution')
        # the refinement streams the full tensor: the mixed form runs its MTTKRP on the bf16
        # matrix cores (split FP32 operands), the FP64 form with FP64 FMA
        rprec = _lib.TFX_PREC_SPLIT16 if precision == 'mixed' else _lib.TFX_PREC_NATIVE
        probT = Problem(dev, Tdev, dims, R, wdt, rprec)
        if options.method == 'als':
            Ffull, ss_r, e_r, im_r, g_r, stop_r = als(probT, Ffull, options, Tsize)
        else:
            Ffull, ss_r, e_r, im_r, g_r, stop_r = dGN(probT, Ffull, options, Tsize, mean_abs)
    else:
        ss_r, e_r, im_r, g_r, stop_r = np.array([0]), np.array([0]), np.array([0]), np.array([0]), 8
    _sync()
    t5 = time.perf_counter()
    timings['refine'] = t5 - t4

    # final error in FP64 and host factors
    F64 = dev.empty(R * sum(dims), torch.float64)
    dev.convert(Ffull, F64)
    if is_sp:
        # error over the stored entries (Auxiliar.py:150-159, Critical.py:25-53)
        err2 = sparse.residual_sq(dev, Xsp, F64, R)
    else:
        err2 = dev.zeros(1, torch.float64)
        dev.residual_sq(Tdev, dims, F64, R, err2)
    final_err = float(np.sqrt(err2.cpu().numpy()[0])) / Tsize
    factors = _unflat(F64.cpu().numpy(), dims, R)
    factors = unsort_dims(factors, ordering)
    t6 = time.perf_counter()
    timings['final'] = t6 - t5
    timings['total'] = t6 - t_all
    if options.refine:
        num_steps = ss_m.size + ss_r.size
        rel_error = float(e_r[-1])
    else:
        num_steps = ss_m.size
        rel_error = final_err
    out = Output(num_steps, rel_error, [ss_m, ss_r], [e_m, e_r], [im_m, im_r], [g_m, g_r],
                 [stop_m, stop_r], options, timings=timings, precision=precision)
    out.final_rel_error = final_err
    out.compressed_dims = tuple(S_dims)
    out.dogleg_calls = prob.dogleg_calls + (probT.dogleg_calls if options.refine else 0)
    if display > 0:
        print('Final results')
        print('    Number of steps =', out.num_steps)
        print('    Relative error =', out.rel_error)
    return factors, out


def check_supported(dims):
    """Shapes that pass the reference's validation but not this path's limits raise
    NotImplementedError before anything is uploaded: orders above the library's TFX_MAXL (the
    reference accepts up to 12, Auxiliar.py:31-33) and modes longer than the device MLSVD's
    Gram size."""
    L = len(dims)
    if L > _lib.TFX_MAXL:
        raise NotImplementedError('tensors of order %d are not on the MI355X path (orders 2-%d)'
                                  % (L, _lib.TFX_MAXL))
    big = [int(d) for d in dims if int(d) > MAX_MODE_DIM]
    if big:
        raise NotImplementedError('mode sizes above %d (%s) are not on the MI355X path: the MLSVD '
                                  'forms the I x I Gram matrix of every mode' % (MAX_MODE_DIM, big))


def cpd(T, R, options=False):
    """Canonical polyadic decomposition of a dense tensor, same call and outputs as
    TensorFox.cpd (TensorFox.py:69-365).  T: NumPy array or torch tensor (CPU or GPU), float32
    or float64; R: int; options: any object whose attributes override the defaults."""
    dims_orig = tuple(int(d) for d in T[2]) if sparse.is_sparse(T) else tuple(T.shape)
    options = make_options(options)
    consistency(R, dims_orig, options)
    if sparse.is_sparse(T):
        sparse.check_supported(dims_orig, R, options.svd_n_components)
    else:
        check_supported(dims_orig)
    if options.method in ('dGN', 'als'):
        return tricpd(T, R, options)
    from .tt import ttcpd
    return ttcpd(T, R, options)
