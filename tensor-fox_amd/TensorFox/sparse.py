"""Sparse COO input T = [data, idxs, dims] on the GPU (TensorFox.py:256-280, 457-500, 577-581,
627-640; Compression.py:122-137, 192-237).

The stored entries are uploaded once: data as FP64, coordinates packed on the device into one
int32 column per sorted mode (tfx_sparse_pack, which also rejects coordinates outside dims, the
reference's 'Got too big index' exit).  The MLSVD forms each mode's dense Gram matrix
T_(l) T_(l)^T from the entries (tfx_sparse_gram), takes its top eigenvectors with the dense
path's eigensolver started from the reference's Gaussian sketch, truncates by energy, and forms
the dense core: the first (longest) mode from the entries (tfx_sparse_ttm0), the others as dense
mode products.  dGN / ALS then run on the core as for dense input; there is no refinement, and
the final error is over the stored entries only (tfx_sparse_residual).

Reference behaviour kept:
- n_components = int(max(L, log10(I_l^2 L)) * min(R, I_l)) per mode (Compression.py:218-221):
  the sparse MLSVD keeps more candidate components than the dense one before truncation, and
  the sketch drawn from the global NumPy generator has n_components + 10 columns;
- sigma = sqrt of the Gram's eigenvalues (Compression.py:222);
- a negative tol_mlsvd becomes 1e-16 (with the reference's message), replacing a list value
  (TensorFox.py:491-498);
- options.refine is set to False (TensorFox.py:579-581).
"""

import math
import sys

import numpy as np
import torch

from . import _lib
from .compression import MAX_MODE_DIM, SUBSPACE_ITERS, _pick_tol, clean_compression

# the dense first-mode product of the sparse core has k_0 * prod_{l>0} I_l FP64 entries
MAX_TTM0_ENTRIES = 8_000_000_000
# eigensolver limits (tfx_eig_top): start block p <= 512 columns; a full eigensolve for I <= p
MAX_SKETCH = 512


def is_sparse(T):
    return type(T) == list


def coo_arrays(T):
    """(data FP64 1-D, coordinates uint64 nnz x L, dims int array) as the reference casts them
    (TensorFox.py:476-480)."""
    data, idxs, dims = T
    if isinstance(data, torch.Tensor):
        data = data.detach().to('cpu', torch.float64).numpy()
    if isinstance(idxs, torch.Tensor):
        idxs = idxs.detach().cpu().numpy()
    data = np.asarray(data, dtype=np.float64).ravel()
    dims = np.array([int(d) for d in dims])
    a = np.asarray(idxs)
    # integer arrays wrap like the reference's uint64 cast (a negative coordinate becomes too big)
    a = a.astype(np.uint64) if a.dtype.kind in 'iu' else np.array(a, dtype=np.uint64)
    if a.size == 0:
        a = a.reshape(0, len(dims))
    if a.ndim != 2 or a.shape != (len(data), len(dims)):
        raise ValueError('sparse tensor: idxs must have shape (nnz, L) = (%d, %d), got %s'
                         % (len(data), len(dims), a.shape))
    return data, a, dims


class SparseTensor:
    """The stored entries on the device, coordinates in the sorted mode order."""

    def __init__(self, dev, data, idxs, dims, ordering):
        self.nnz = int(len(data))
        self.L = len(dims)
        self.dims = [int(dims[o]) for o in ordering]
        self.data = dev.upload(data, torch.float64)
        coords = torch.from_numpy(np.ascontiguousarray(idxs).view(np.int64)).to(dev.dev)
        self.idx = dev.empty(max(self.nnz, 1) * self.L, torch.int32)
        bad = dev.zeros(1, torch.int32)
        dev.sparse_pack(coords, ordering, self.dims, self.idx, bad)
        if int(bad.cpu()[0]):
            sys.exit('Got too big index when tried to generate the unfold of sparse tensor.')
        stats = dev.zeros(2, torch.float64)
        dev.tensor_stats(self.data, stats)
        self.Tsize = float(math.sqrt(stats.cpu().numpy()[1]))


def report(data, dims, display):
    """The reference's 'Sparse tensor detected' lines (TensorFox.py:481-484)."""
    if display != 0:
        print('Sparse tensor detected')
        print('    nnz =', len(data))
        print('    Sparsity level =', round(100 * (1 - len(data) / int(np.prod(dims))), 12), '%')


def fix_tol(options):
    """TensorFox.py:488-498: a negative tol_mlsvd (or first entry of a list) becomes 1e-16."""
    tol = options.tol_mlsvd[0] if type(options.tol_mlsvd) == list else options.tol_mlsvd
    if tol < 0:
        print('Using class MLSVD with tol_mlsvd = 1e-16.')
        options.tol_mlsvd = 1e-16


def n_components(I, R, L, svd_n):
    if svd_n is not None:
        return int(svd_n)
    return int(max(L, math.log10(I * I * L)) * min(R, I))


def check_supported(dims, R, svd_n=None):
    """Limits of this path that the reference does not have, raised before any upload."""
    L = len(dims)
    if L > _lib.TFX_MAXL:
        raise NotImplementedError('tensors of order %d are not on the MI355X path (orders 2-%d)'
                                  % (L, _lib.TFX_MAXL))
    for l, I in enumerate(sorted((int(d) for d in dims), reverse=True)):
        if I > MAX_MODE_DIM:
            raise NotImplementedError('mode size %d above %d: the sparse MLSVD forms the I x I Gram '
                                      'matrix of every mode' % (I, MAX_MODE_DIM))
        p = n_components(I, R, L, svd_n) + 10
        if (I > p and p > MAX_SKETCH) or (I <= p and I > MAX_SKETCH):
            raise NotImplementedError('sparse MLSVD of a mode of size %d with %d sketch columns is '
                                      'above the eigensolver limit of %d' % (I, p, MAX_SKETCH))


def mlsvd(dev, X, R, options):
    """Returns (S, S_dims, U list, sigmas list) as compression.mlsvd does for dense input."""
    dims, L = X.dims, X.L
    tol_mlsvd = _pick_tol(options.tol_mlsvd, L)
    rng = np.random.mtrand._rand
    U_list, sig_list = [], []
    for l in range(L):
        I = dims[l]
        nc = n_components(I, R, L, options.svd_n_components)
        p = nc + 10
        draw = rng.normal(size=(I, p))         # randomized_range_finder (Compression.py:379-382)
        k = min(nc, I)
        if options.display != 0:
            print('    Compressing unfolding mode', l + 1)
        G = dev.empty((I, I), torch.float64)
        dev.sparse_gram(X.nnz, dims, X.idx, X.data, l, G)
        U = dev.empty((k, I), torch.float64)   # column-major I x k
        lam = dev.empty(k, torch.float64)
        if I <= p:
            dev.eig_top(G, I, None, p, k, 0, U, lam)
        else:
            Q0 = dev.upload(np.asfortranarray(draw).ravel(order='F'), torch.float64)
            dev.eig_top(G, I, Q0, p, k, SUBSPACE_ITERS, U, lam)
        sigma = np.sqrt(np.maximum(lam.cpu().numpy(), 0.0))
        dim = clean_compression(sigma, tol_mlsvd, L)
        U_list.append(U[:dim].contiguous())
        sig_list.append(sigma[:dim])
    if options.display != 0:
        print('    Computing sparse multilinear multiplication')
    k0 = U_list[0].shape[0]
    cur = [k0] + dims[1:]
    if k0 * int(np.prod(dims[1:])) > MAX_TTM0_ENTRIES:
        raise NotImplementedError('sparse core: first-mode product of %d entries above %d'
                                  % (k0 * int(np.prod(dims[1:])), MAX_TTM0_ENTRIES))
    S = dev.empty(int(np.prod(cur)), torch.float64)
    dev.sparse_ttm0(X.nnz, dims, X.idx, X.data, U_list[0], k0, S)
    for l in range(1, L):
        k = U_list[l].shape[0]
        new = list(cur)
        new[l] = k
        Snew = dev.empty(int(np.prod(new)), torch.float64)
        dev.ttm(S, cur, l, U_list[l], k, Snew, None)
        S, cur = Snew, new
    td = options.trunc_dims
    if type(td) == list:
        view = S.view(*cur)
        for l in range(L):
            view = view.narrow(l, 0, min(td[l], cur[l]))
            U_list[l] = U_list[l][:min(td[l], U_list[l].shape[0])].contiguous()
        S = view.contiguous().view(-1)
        cur = list(view.shape)
    return S, cur, U_list, sig_list


def residual_sq(dev, X, F, R):
    """sum over the stored entries of (data - model)^2 (Critical.py:25-53, squared), F in the
    sorted mode order."""
    out = dev.zeros(1, torch.float64)
    dev.sparse_residual(X.nnz, X.dims, X.idx, X.data, F, R, out)
    return out
