"""MLSVD compression on the GPU (Compression.py:48-301).

For each mode the reference computes a randomized SVD of the unfolding (Gaussian sketch with
R + 10 columns, two un-normalised power iterations, QR, small SVD).  Here the unfolding's Gram
matrix G = X_(l) X_(l)^T is formed on the device, its top eigenpairs are found in FP64 by block
subspace iteration started from the same Gaussian sketch (drawn from the global NumPy RNG in the
reference's order and shape) followed by Rayleigh-Ritz, and the core is the exact projection
U_l^T X_(l).  The singular values used for truncation are the FP64 norms of the projected rows,
so exactly low-rank inputs truncate as in the reference (Compression.py:261-301).
"""

import numpy as np
import torch

from . import _lib

SUBSPACE_ITERS = 2
# largest mode size of the device MLSVD: the mode Gram is formed as an I x I matrix and its
# eigensolver stages I x p blocks (tfx_gram_mode, tfx_eig_top)
MAX_MODE_DIM = 8192
_PRIVATE_RNG_SEED = 20240611


def _pick_tol(tol_mlsvd, L):
    if type(tol_mlsvd) == list:
        return tol_mlsvd[0] if L > 3 else tol_mlsvd[1]
    return tol_mlsvd


def clean_compression(sigma, tol_mlsvd, L):
    """Smallest i with sum(sigma[i:]^2)/sum(sigma^2) < tol_mlsvd/L (Compression.py:261-301).

    The tail sums come from one cumulative sum; an index whose ratio lies within 1e-9 (relative)
    of the threshold is decided with the reference's own np.sum expression, so the result is the
    reference's (the two sums differ by rounding only, about 1e-14 relative)."""
    eps = tol_mlsvd / L
    n = len(sigma)
    if n <= 1:
        return n
    sq = sigma ** 2
    total = np.sum(sq)
    with np.errstate(invalid='ignore', divide='ignore'):
        approx = np.cumsum(sq[::-1])[::-1] / total
    for i in range(1, n):
        a = approx[i]
        if a >= eps * (1 + 1e-9):
            continue
        if a < eps * (1 - 1e-9) or np.sum(sigma[i:] ** 2) / np.sum(sigma ** 2) < eps:
            return i
    return n


def _data_prec(X, prec):
    """The split precision applies to FP32 data only; FP64 intermediates run FP64 MFMA."""
    return prec if X.dtype == torch.float32 else _lib.TFX_PREC_NATIVE


def _top_eigvecs(dev, X, dims, l, n_comp, prec, draw):
    """U_l (I_l x n_comp, FP64 device, column-major) and eigenvalues of the mode-l Gram matrix.
    `draw` is the sketch drawn from the global RNG for this mode (shape as the reference's)."""
    I = dims[l]
    p = n_comp + 10
    G = dev.empty((I, I), torch.float64)
    dev.gram_mode(X, dims, l, G, prec=_data_prec(X, prec))
    U = dev.empty((n_comp, I), torch.float64)       # column-major I x n_comp
    lam = dev.empty(n_comp, torch.float64)
    if I <= p:
        dev.eig_top(G, I, None, p, n_comp, 0, U, lam)
    else:
        if draw.shape[0] == I:
            Q0 = draw
        else:
            Q0 = np.random.RandomState(_PRIVATE_RNG_SEED + l).normal(size=(I, p))
        Q0d = dev.upload(np.asfortranarray(Q0).ravel(order='F'), torch.float64)
        dev.eig_top(G, I, Q0d, p, n_comp, SUBSPACE_ITERS, U, lam)
    return U, lam


def mlsvd(dev, T, dims, R, options, wdt, prec):
    """Returns (S, S_dims, U list (FP64 device, column-major I_l x k_l), sigmas list (host))."""
    L = len(dims)
    tol_mlsvd = _pick_tol(options.tol_mlsvd, L)
    if tol_mlsvd == -1:
        U = []
        for l in range(L):
            U.append(dev.upload(np.identity(dims[l]).ravel(), torch.float64))
        return T, list(dims), U, [np.ones(d) for d in dims]
    method = options.mlsvd_method
    if method == 'gpu':
        method = 'classic'
    svd_n = options.svd_n_components
    U_list, sig_list = [], []
    if method == 'seq':
        rng = np.random.mtrand._rand

        def draw_shape(l, cur_dims):
            # the reference's sketch shape for mode l (Compression.py:208-230)
            n_comp0 = min(R, dims[l]) if svd_n is None else svd_n
            n_rest = int(np.prod(cur_dims)) // cur_dims[l]
            transpose = cur_dims[l] < n_rest
            return (cur_dims[l] if transpose else n_rest, n_comp0 + 10)

        S, S_dims = T, list(dims)
        pending = rng.normal(size=draw_shape(0, S_dims))
        for l in range(L):
            n_comp = min(R, dims[l]) if svd_n is None else svd_n
            n_rest = int(np.prod(S_dims)) // S_dims[l]
            draw = pending
            n_comp = min(n_comp, S_dims[l], n_rest)
            U, lam = _top_eigvecs(dev, S, S_dims, l, n_comp, prec, draw)
            new_dims = list(S_dims)
            new_dims[l] = n_comp
            Snew = dev.empty(int(np.prod(new_dims)), wdt)
            sig2 = dev.empty(n_comp, torch.float64)
            dev.ttm(S, S_dims, l, U, n_comp, Snew, sig2, prec=_data_prec(S, prec))
            # next mode's sketch, drawn while the device works, for the shape it has when this
            # mode keeps n_comp components; the RNG state is restored and the sketch drawn again
            # when the truncation below changes that shape (the draws stay the reference's)
            if l + 1 < L:
                spec_shape = draw_shape(l + 1, new_dims)
                spec_state = rng.get_state()
                pending = rng.normal(size=spec_shape)
            sigma = np.sqrt(np.maximum(sig2.cpu().numpy(), 0.0))
            dim = clean_compression(sigma, tol_mlsvd, L)
            if dim < n_comp:
                U = U[:dim].contiguous()
                new_dims[l] = dim
                Snew = dev.empty(int(np.prod(new_dims)), wdt)
                dev.ttm(S, S_dims, l, U, dim, Snew, None, prec=_data_prec(S, prec))
                sigma = sigma[:dim]
            U_list.append(U)
            sig_list.append(sigma)
            S, S_dims = Snew, new_dims
            if l + 1 < L and draw_shape(l + 1, S_dims) != spec_shape:
                rng.set_state(spec_state)
                pending = rng.normal(size=draw_shape(l + 1, S_dims))
    elif method == 'classic':
        # every mode's eigensolve is queued before any eigenvalue is read back (the sketches are
        # drawn in the reference's order; their shapes do not depend on the results)
        queued = []
        for l in range(L):
            n_comp = min(R, dims[l]) if svd_n is None else svd_n
            n_rest = int(np.prod(dims)) // dims[l]
            transpose = dims[l] < n_rest
            draw = np.random.mtrand._rand.normal(size=(dims[l] if transpose else n_rest, n_comp + 10))
            n_comp = min(n_comp, dims[l], n_rest)
            queued.append(_top_eigvecs(dev, T, dims, l, n_comp, prec, draw))
        for U, lam in queued:
            sigma = np.sqrt(np.maximum(lam.cpu().numpy(), 0.0))
            dim = clean_compression(sigma, tol_mlsvd, L)
            U_list.append(U[:dim].contiguous())
            sig_list.append(sigma[:dim])
        S, S_dims = T, list(dims)
        for l in range(L):
            k = U_list[l].shape[0]
            new_dims = list(S_dims)
            new_dims[l] = k
            Snew = dev.empty(int(np.prod(new_dims)), wdt)
            dev.ttm(S, S_dims, l, U_list[l], k, Snew, None, prec=_data_prec(S, prec))
            S, S_dims = Snew, new_dims
    else:
        raise ValueError("mlsvd_method must be 'seq', 'classic' or 'gpu' on the dense path")
    td = options.trunc_dims
    if type(td) == list:
        view = S.view(*S_dims)
        for l in range(L):
            k = min(td[l], S_dims[l])
            view = view.narrow(l, 0, k)
            U_list[l] = U_list[l][:min(td[l], U_list[l].shape[0])].contiguous()
        S = view.contiguous().view(-1)
        S_dims = list(view.shape)
    return S, S_dims, U_list, sig_list
