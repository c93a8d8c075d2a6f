"""Tensor-train CPD, method 'ttcpd' (TensorFox.py:252-421, 658-817; Auxiliar.py:365-527;
Conversion.py:460-495), on the GPU.

cpd with method 'ttcpd' compresses T by the MLSVD (dense or sparse input), pads core modes
shorter than R with 1e-12 noise, and takes the CPD of the core from a tensor train
G_0 (I_0 x R), G_l (R x I_l x R), G_{L-1} (R x I_{L-1}):
- the train: for l = 0 .. L-2 the current remainder Y (r1 x I_l x rest, C order) has its first
  two modes merged; the top-R eigenvectors U of that merged mode's Gram matrix (tfx_gram_mode,
  tfx_eig_top) are the left singular vectors the reference takes from a randomized SVD with all
  components (n_iter=0, so an exact SVD up to rounding), with the same sign rule (largest entry
  of each column positive); G_l = U reshaped, Y <- U^T Y (tfx_ttm).  The merged row index
  a I_l + i of C order is a permutation of the reference's Fortran-order a + r1 i, which leaves
  the eigenvectors and their signs unchanged;
- the cores: tricpd of G_1 (the GPU pipeline, best of `trials`), then bicpd of G_2 .. G_{L-2}
  with the first factor fixed to pinv(Z_prev^T) (forward sweeps; odd epochs also sweep
  backwards with the third factor fixed to pinv(X_next^T)), each bicpd being the MLSVD of the
  core, a random start with the fixed mode's start U^T F_fixed, and ALS (the default
  bi_method) or dGN with that mode held (device ALS / dGN with a fixed mode);
- the factors: G_0 X_1, the middle factors, G_{L-1}^T Z_last, equalized; deflated to the core
  dims, U_l W_l, error against T.
The reference's draws from the global NumPy generator are replayed in order (the inflation
noise, the randomized SVD sketches of the train, every starting point and CG cap of the core
CPDs), so a seeded call follows the reference's random choices.

Small R x R and core-sized products run in tfx_gemm_small; pinv(B^T) of a square B is
B (B^T B)^+ with the ALS pseudo-inverse (tfx_als_solve).
"""

import sys

import numpy as np
import torch

from . import _lib, sparse
from .auxiliar import consistency, sort_dims, unsort_dims
from .compression import mlsvd
from .device import get_device
from .gauss_newton import Problem, dGN
from .als import als

F64 = torch.float64
_DIRECT_EIG = 512          # largest Gram solved by the direct (full) eigensolver
_TT_ITERS = 8              # subspace steps when the merged mode is longer
_TT_RNG = np.random.RandomState(31415)


class FinalOutputs:
    """make_final_outputs (Auxiliar.py:223-239)."""

    def __init__(self, num_steps, rel_error, accuracy, outputs, options):
        self.num_steps = num_steps
        self.rel_error = rel_error
        self.accuracy = accuracy
        self.cpd_output = outputs
        self.options = options


def _colmajor(a):
    """Host matrix -> device FP64 flat column-major."""
    a = np.asarray(a, dtype=np.float64)
    return torch.from_numpy(np.asfortranarray(a).ravel(order='F').copy())


def _gemm(dev, m, n, k, A_cm, B_cm):
    """C = A B with A (m x k) and B (k x n) device column-major flats; C column-major."""
    C = dev.empty(m * n, F64)
    dev.gemm_small(False, m, n, k, A_cm, B_cm, C)
    return C


def _pinv_t(dev, B_cm, R):
    """pinv(B^T) of a square R x R matrix (column-major flat) as B (B^T B)^+."""
    BtB = dev.empty(R * R, F64)
    dev.gemm_small(True, R, R, R, B_cm, B_cm, BtB)
    P1 = BtB.view(1, R, R).expand(3, R, R).contiguous()     # als_solve reads P1[mode]
    out = B_cm.clone()
    status = dev.zeros(1, torch.int32)
    dev.als_solve((R, R, R), R, 0, P1, B_cm, out, status)
    return out


def cpdtt(dev, S, dims, R):
    """The CPD tensor train of the core S (device C-order flat, dims).  Returns the cores as
    device tensors in the reference's shapes and the column-major top eigenvector blocks."""
    L = len(dims)
    rng = np.random.mtrand._rand
    G = []
    Y, r1 = S, 1
    for l in range(L - 1):
        n = r1 * dims[l]
        rest = int(np.prod(dims[l + 1:]))
        cur = [n] + list(dims[l + 1:])
        m = min(n, rest)
        rng.normal(size=(m, m + 10))     # the randomized SVD's sketch (Auxiliar.py:371)
        if R > n:
            raise NotImplementedError('tensor train: merged mode of size %d below R = %d' % (n, R))
        Gm = dev.empty((n, n), F64)
        dev.gram_mode(Y, cur, 0, Gm)
        U = dev.empty((R, n), F64)       # column-major n x R
        lam = dev.empty(R, F64)
        if n <= _DIRECT_EIG:
            dev.eig_top(Gm, n, None, n, R, 0, U, lam)
        else:
            p = min(R + 10, n)
            Q0 = dev.upload(_TT_RNG.normal(size=(n, p)).ravel(order='F'), F64)
            dev.eig_top(Gm, n, Q0, p, R, _TT_ITERS, U, lam)
        g = U.t().contiguous()           # n x R row-major: g[a I_l + i, b]
        G.append(g.view(dims[0], R) if l == 0 else g.view(r1, dims[l], R))
        Ynew = dev.empty(R * rest, F64)
        dev.ttm(Y, cur, 0, U, R, Ynew, None)
        Y, r1 = Ynew, R
    G.append(Y.view(R, dims[-1]))
    return G


def bicpd(dev, Tg, R, fixed_cm, mode, options):
    """Third-order CPD of the core Tg (device (R, I, R) tensor) with factor `mode` fixed to the
    column-major block fixed_cm.  Follows TensorFox.py:658-817.  Returns host X, Y, Z (the fixed
    one in the compressed space, as the reference returns it) and the output."""
    from .tensorfox import starting_point, _unflat
    from .auxiliar import Output
    bi_method = options.bi_method_parameters[0]
    dims = list(Tg.shape)
    consistency(R, tuple(dims), options)
    X = Tg.contiguous().view(-1)
    st = dev.zeros(2, F64)
    dev.tensor_stats(X, st)
    Tsize = float(np.sqrt(st.cpu().numpy()[1]))
    S, S_dims, U, _ = mlsvd(dev, X, dims, R, options, F64, _lib.TFX_PREC_NATIVE)
    if options.symm:
        R_min = min(S_dims)
        view = S.view(*S_dims)
        for l in range(3):
            view = view.narrow(l, 0, R_min)
        S, S_dims = view.contiguous().view(-1), [R_min] * 3
        U = [u[:R_min].contiguous() for u in U]
    F0 = starting_point(dev, S, S_dims, U, dims, R, [0, 1, 2], options, F64)
    off = R * sum(S_dims[:mode])
    dev.gemm_small(True, S_dims[mode], R, dims[mode], U[mode], fixed_cm,
                   F0[off:off + R * S_dims[mode]])
    block = F0[off:off + R * S_dims[mode]].clone()
    sst = dev.zeros(2, F64)
    dev.tensor_stats(S, sst)
    ss_ = sst.cpu().numpy()
    Ssize, S_mean_abs = float(np.sqrt(ss_[1])), float(ss_[0]) / int(np.prod(S_dims))
    prob = Problem(dev, S, S_dims, R, F64, _lib.TFX_PREC_NATIVE)
    if bi_method == 'als':
        Fm, ss, e, im, g, stop = als(prob, F0, options, Ssize, fixed=(mode, block))
    else:
        Fm, ss, e, im, g, stop = dGN(prob, F0, options, Ssize, S_mean_abs,
                                     aliased=type(options.initialization) != list, fixed=(mode, block))
    Ffull = dev.empty(R * sum(dims), F64)
    o_s, o_f = 0, 0
    for l in range(3):
        if l == mode:
            Ffull[o_f:o_f + R * dims[l]].copy_(fixed_cm)
        else:
            dev.gemm_small(False, dims[l], R, S_dims[l], U[l], Fm[o_s:o_s + R * S_dims[l]],
                           Ffull[o_f:o_f + R * dims[l]])
        o_s += R * S_dims[l]
        o_f += R * dims[l]
    err2 = dev.zeros(1, F64)
    dev.residual_sq(X, dims, Ffull, R, err2)
    rel_error = float(np.sqrt(err2.cpu().numpy()[0])) / Tsize
    full = _unflat(Ffull.cpu().numpy(), dims, R)
    full[mode] = _unflat(Fm.cpu().numpy(), S_dims, R)[mode]
    z = np.array([0])
    out = Output(ss.size, rel_error, [ss, z], [e, z], [im, z], [g, z], [stop, 5], options)
    if options.display > 0:
        print('Final results of bicpd')
        print('    Number of steps =', out.num_steps)
        print('    Relative error =', out.rel_error)
    return full[0], full[1], full[2], out


def cpd_cores(dev, G, max_trials, epochs, R, display, options):
    """CPDs of the train's cores (Auxiliar.py:424-527)."""
    from .tensorfox import tricpd
    L = len(G)
    if epochs % 2 == 0:
        epochs += 1
    cpd_list = list(range(L - 2))
    outputs = list(range(L - 2))
    best_error = np.inf
    for _ in range(max_trials):
        if display > 0:
            print()
            print('CPD 1')
        factors, output = tricpd(G[1], R, options)
        X, Y, Z = factors
        if output.rel_error < best_error:
            best_output, best_error = output, output.rel_error
            best_X, best_Y, best_Z = X, Y, Z
            if best_error < 1e-4:
                break
    outputs[0] = best_output
    cpd_list[0] = [best_X, best_Y, best_Z]
    if display < 0:
        print('CPD 1 error =', best_error)
    low, upp = 2, L - 2
    for epoch in range(epochs):
        if epoch % 2 == 0:
            for l in range(low, L - 1):
                best_error = np.inf
                fixed_cm = _pinv_t(dev, _colmajor(best_Z).to(dev.dev), R)
                fixed_X = fixed_cm.cpu().numpy().reshape(R, R).T
                for _ in range(max_trials):
                    if display > 0:
                        print()
                        print('CPD', l)
                    X, Y, Z, output = bicpd(dev, G[l], R, fixed_cm, 0, options)
                    if output.rel_error < best_error:
                        best_output, best_error = output, output.rel_error
                        best_X, best_Y, best_Z = fixed_X, Y, Z
                        if best_error < 1e-4:
                            break
                if epoch == epochs - 1:
                    outputs[l - 1] = best_output
                    cpd_list[l - 1] = [fixed_X, best_Y, best_Z]
                if display < 0:
                    print('CPD', l, 'error =', best_error)
        else:
            if L == 3:
                low, upp = 1, L - 1
            for l in reversed(range(1, upp)):
                best_error = np.inf
                fixed_cm = _pinv_t(dev, _colmajor(best_X).to(dev.dev), R)
                fixed_Z = fixed_cm.cpu().numpy().reshape(R, R).T
                for _ in range(max_trials):
                    if display > 0:
                        print()
                        print('CPD', l)
                    X, Y, Z, output = bicpd(dev, G[l], R, fixed_cm, 2, options)
                    if output.rel_error < best_error:
                        best_output, best_error = output, output.rel_error
                        best_X, best_Y, best_Z = X, Y, fixed_Z
                        if best_error < 1e-4:
                            break
                if epoch == epochs - 2:
                    outputs[l - 1] = best_output
                    cpd_list[l - 1] = [best_X, best_Y, fixed_Z]
                if display < 0:
                    print('CPD', l, 'error =', best_error)
    return cpd_list, outputs, best_Z


def highcpd(dev, S, dims, R, options):
    """TensorFox.py:368-421: flat equalized factors of the core (device, column-major blocks)."""
    L = len(dims)
    options.refine = False
    G = cpdtt(dev, S, dims, R)
    if options.display != 0:
        print('Total of', L - 2, 'third order CPDs to be computed:')
        print('===============================================================================================')
    cpd_list, outputs, best_Z = cpd_cores(dev, G, options.trials, options.epochs, R,
                                          options.display, options)
    blocks = [_gemm(dev, dims[0], R, R, G[0].t().contiguous().view(-1),
                    _colmajor(cpd_list[0][0]).to(dev.dev))]
    for l in range(L - 2):
        blocks.append(_colmajor(cpd_list[l][1]).to(dev.dev))
    # G_{L-1} is R x I_{L-1} in C order: as a column-major flat it is G_{L-1}^T
    blocks.append(_gemm(dev, dims[-1], R, R, G[-1].contiguous().view(-1),
                        _colmajor(best_Z).to(dev.dev)))
    x = torch.cat(blocks)
    zero = torch.zeros_like(x)
    F = torch.empty_like(x)
    x_new = torch.empty_like(x)
    sc = dev.zeros(4, F64)
    dev.dgn_update(dims, R, x, zero, 0.0, zero, 0.0, True, 0, 0.0, x_new, F, sc)   # equalize
    return F, outputs


def ttcpd(T, R, options):
    """cpd with method 'ttcpd' (TensorFox.py:256-365), dense or sparse T."""
    from .tensorfox import _precision_of, _to_device, _unflat
    dev = get_device(options.device)
    display = options.display
    is_sp = sparse.is_sparse(T)
    if is_sp:
        data, idxs, dims_in = sparse.coo_arrays(T)
        if len(data) == 0:
            raise ValueError('sparse tensor with no stored entries')
        sparse.report(data, dims_in, display)
        L = len(dims_in)
        dims, ordering = sort_dims(tuple(dims_in))
        Xsp = sparse.SparseTensor(dev, data, idxs, dims_in, ordering)
        Tsize = Xsp.Tsize
        sparse.fix_tol(options)
    else:
        precision = _precision_of(T, options)
        tdt = torch.float32 if precision == 'mixed' else F64
        mprec = _lib.TFX_PREC_SPLIT16 if precision == 'mixed' else _lib.TFX_PREC_NATIVE
        L = T.ndim
        dims, ordering = sort_dims(T.shape)
        Tdev = _to_device(dev, T, ordering, tdt)
        st = dev.zeros(2, F64)
        dev.tensor_stats(Tdev, st)
        Tsize = float(np.sqrt(st.cpu().numpy()[1]))
    if display != 0:
        print('-----------------------------------------------------------------------------------------------')
        print('Computing MLSVD')
    if is_sp:
        S, S_dims, U, _ = sparse.mlsvd(dev, Xsp, R, options)
    else:
        S, S_dims, U, _ = mlsvd(dev, Tdev, dims, R, options, F64, mprec)
    if display != 0:
        if tuple(S_dims) == tuple(dims):
            print('    No compression detected')
            print('    Working with dimensions', tuple(dims))
        else:
            print('    Compression detected')
            print('    Compressing from', tuple(dims), 'to', tuple(S_dims))
        print()
    S_orig = list(S_dims)
    inflate_status = R > min(S_dims)
    if inflate_status:
        new_dims = [max(int(d), R) for d in S_dims]
        Sinf = dev.upload(1e-12 * np.random.randn(*new_dims), F64).view(*new_dims)
        Sinf[tuple(slice(d) for d in S_dims)].copy_(S.view(*S_dims))
        S, S_dims = Sinf.contiguous().view(-1), new_dims
    options.trunc_dims = 0
    Fc, outputs = highcpd(dev, S, S_dims, R, options)
    # deflate (Conversion.py:483-495) and U_l W_l
    Ffull = dev.empty(R * sum(dims), F64)
    o_c, o_f = 0, 0
    for l in range(L):
        W = Fc[o_c:o_c + R * S_dims[l]].view(R, S_dims[l])[:, :S_orig[l]].contiguous().view(-1)
        dev.gemm_small(False, dims[l], R, S_orig[l], U[l], W, Ffull[o_f:o_f + R * dims[l]])
        o_c += R * S_dims[l]
        o_f += R * dims[l]
    err2 = dev.zeros(1, F64)
    if is_sp:
        err2 = sparse.residual_sq(dev, Xsp, Ffull, R)
    else:
        dev.residual_sq(Tdev, dims, Ffull, R, err2)
    rel_error = float(np.sqrt(err2.cpu().numpy()[0])) / Tsize
    factors = unsort_dims(_unflat(Ffull.cpu().numpy(), dims, R), ordering)
    num_steps = sum(o.num_steps for o in outputs)
    accuracy = max(0, 100 * (1 - rel_error))
    if display != 0:
        print()
        print('===============================================================================================')
        print('Final results')
        print('    Number of steps =', num_steps)
        print('    Relative error =', rel_error)
    return factors, FinalOutputs(num_steps, rel_error, accuracy, outputs, options)
