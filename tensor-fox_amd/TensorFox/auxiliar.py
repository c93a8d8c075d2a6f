"""Options, input validation, dimension sorting, the output class and tensor generators.

Mirrors the reference's Auxiliar module for the dense dGN path (Auxiliar.py:24-220, 242-362,
530-572).  Two options are added for the MI355X path: `precision` ('mixed' or 'fp64'; default
chosen from the input dtype, see tensorfox.cpd) and `device` (GPU index, default 0).
"""

import sys

import numpy as np

_DEFAULTS = dict(
    maxiter=200, tol=1e-16, tol_step=1e-6, tol_improv=1e-6, tol_grad=1e-6, tol_jump=10,
    method='dGN', inner_method='cg', cg_maxiter=100, cg_factor=1, cg_tol=1e-16,
    initialization='random', trunc_dims=0, mlsvd_method='seq', tol_mlsvd=1e-16, init_damp=1,
    refine=False, symm=False, factors_norm=0, trials=3, display=0, epochs=1, gpu=False,
    mkl_dot=True, svd_n_components=None, precision=None, device=0)


class Options:
    """Filled option set (the reference builds an ad-hoc `temp_options` class)."""

    def __repr__(self):
        return 'Options(%s)' % ', '.join('%s=%r' % (k, getattr(self, k)) for k in sorted(vars(self)))


def make_options(options=False):
    """Defaults, then every attribute the user object defines (attribute presence, as in
    Auxiliar.py:288-353).  `gpu=True` without an explicit `mlsvd_method` selects the reference's
    'gpu' MLSVD name (Auxiliar.py:359-360), which this package maps to its own Gram method."""
    out = Options()
    for k, v in _DEFAULTS.items():
        setattr(out, k, v)
    out.bi_method_parameters = ['als', 500, 1e-6]
    names = dir(options)
    for k in _DEFAULTS:
        if k in names:
            setattr(out, k, getattr(options, k))
    if 'bi_method' in names:
        out.bi_method_parameters[0] = options.bi_method
        out.bi_method_parameters[1] = {'cg': 1, 'cg_static': 100, 'als': 500}.get(
            options.bi_method, out.bi_method_parameters[1])
    if 'bi_method_maxiter' in names:
        out.bi_method_parameters[1] = options.bi_method_maxiter
    if 'bi_method_tol' in names:
        out.bi_method_parameters[2] = options.bi_method_tol
    if out.gpu and 'mlsvd_method' not in names:
        out.mlsvd_method = 'gpu'
    return out


def consistency(R, dims, options):
    """Same checks and SystemExit messages as Auxiliar.py:24-78, without the 1 s sleep."""
    L = len(dims)
    if L > 12:
        sys.exit('Tensor Fox does not work with tensors of order higher than 12.')
    for l in range(L):
        if dims[l] == 1:
            sys.exit('At least one dimension is equal to 1. This situation not supported by Tensor Fox.')
    if (type(R) != int) or (R < 1):
        sys.exit('Rank must be a positive integer.')
    if L == 3:
        m, n, p = int(dims[0]), int(dims[1]), int(dims[2])
        if R > min(m * n, m * p, n * p):
            sys.exit('Rank must satisfy 1 <= rank <= min(m*n, m*p, n*p) = ' + str(min(m * n, m * p, n * p)) + '.')
    if L > 3 and R == 1:
        sys.exit('Rank must be greater than 1 for tensor with order greater than 3.')
    if options.symm:
        for i in range(L):
            for j in range(L):
                if dims[i] != dims[j]:
                    sys.exit('Symmetric tensors must have equal dimensions.')
    if options.method not in ('dGN', 'als', 'ttcpd'):
        sys.exit("Wrong method name. Must be 'dGN', 'als' or 'ttcpd'.")
    if options.display > 2:
        print('-> With display > 2 Tensor Fox will compute the error of the MLSVD. For large tensors '
              'this routine is very clostly.', file=sys.stderr)


def sort_dims(dims):
    """Permutation putting the dimensions in decreasing order (Auxiliar.py:93-115: argsort(-dims)).
    Returns (sorted dims, ordering)."""
    d = np.array(dims)
    ordering = np.argsort(-d)
    return tuple(int(x) for x in d[ordering]), ordering


def unsort_dims(factors, ordering):
    """Put the factors back in the original mode order (Auxiliar.py:118-128)."""
    new = [None] * len(factors)
    for l in range(len(factors)):
        new[ordering[l]] = factors[l]
    return new


_STOP_MSG = {
    0: '0 - Relative error is small enough.',
    1: '1 - Steps are small enough.',
    2: '2 - Improvement in the relative error is small enough.',
    3: '3 - Gradient is small enough.',
    4: '4 - Average error increased.',
    5: '5 - Limit of iterations was reached.',
    6: '6 - dGN diverged.',
    7: '7 - Average improvement is too small compared to the average error.',
    8: '8 - No refinement was performed.',
}


class Output:
    """Same fields as the reference's output class (Auxiliar.py:161-216), plus `timings`
    (seconds per stage) and `precision`."""

    def __init__(self, num_steps, rel_error, step_sizes, errors, improv, gradients, stop, options,
                 timings=None, precision=None):
        self.num_steps = num_steps
        self.rel_error = rel_error
        self.accuracy = max(0, 100 * (1 - rel_error))
        self.step_sizes = step_sizes
        self.errors = errors
        self.improv = improv
        self.gradients = gradients
        self.stop = stop
        self.options = options
        self.timings = timings or {}
        self.precision = precision

    def stop_msg(self):
        print()
        print('Main stop:')
        if self.stop[0] in _STOP_MSG and self.stop[0] != 8:
            print(_STOP_MSG[self.stop[0]])
        print()
        print('Refinement stop:')
        if self.stop[1] in _STOP_MSG:
            print(_STOP_MSG[self.stop[1]])
        return ''


# ----------------------------------------------------------------------------------------------
# generators
# ----------------------------------------------------------------------------------------------

def cpd2tens(factors):
    """Dense [[F_1, ..., F_L]] in C order (host NumPy, for generators)."""
    L = len(factors)
    letters = 'abcdefghijkl'[:L]
    spec = ','.join(c + 'z' for c in letters) + '->' + letters
    return np.einsum(spec, *factors, optimize=True)


def gen_rand_tensor(dims, R, noise=0):
    """Random rank-R tensor with N(0, 1) factors and absolute Gaussian noise, drawing from the
    global legacy RNG like Auxiliar.py:530-572."""
    F = [np.random.randn(d, R) for d in dims]
    T = cpd2tens(F)
    if noise != 0:
        return T, T + noise * np.random.randn(*dims), F
    return T, F


def add_relative_noise(T0, nu, rs):
    """T0 + nu ||T0|| E / ||E|| with E ~ N(0, 1) from the RandomState rs."""
    E = rs.standard_normal(T0.shape)
    return T0 + nu * np.linalg.norm(T0) * E / np.linalg.norm(E)


def make_lowrank(dims, R, nu, seed=0, col_scale=False, collinear=None, dtype=np.float64):
    """Seeded synthetic CPD problem: factors N(0, 1) per mode (legacy RandomState(seed)),
    optionally the first factor's columns scaled by 10^U(-1, 1), optionally two columns of the
    first factor made collinear with cosine `collinear`, then relative noise nu.
    Returns (T, factors)."""
    rs = np.random.RandomState(seed)
    F = [rs.standard_normal((d, R)) for d in dims]
    if col_scale:
        u = rs.uniform(-1, 1, R)
        F[0] = F[0] * (10.0 ** u)[None, :]
    if collinear is not None:
        a1, a2 = F[0][:, 0], F[0][:, 1]
        n2 = np.linalg.norm(a2)
        u1 = a1 / np.linalg.norm(a1)
        q = a2 - (a2 @ u1) * u1
        q /= np.linalg.norm(q)
        c = float(collinear)
        F[0][:, 1] = c * n2 * u1 + np.sqrt(1 - c * c) * n2 * q
    T0 = cpd2tens(F)
    T = add_relative_noise(T0, nu, rs) if nu else T0
    return T.astype(dtype), F


# ----------------------------------------------------------------------------------------------
# sparse COO tensors [data, idxs, dims]
# ----------------------------------------------------------------------------------------------

def gen_rand_sparse_tensor(dims, R, nnz):
    """Sparse random tensor from sparse factors, drawing from the global legacy RNG in the
    order of Auxiliar.py:575-647 (per mode: randn(nnz) values, randint rows, randint columns), so
    a seeded call returns the reference's tensor.  As there:
    - a factor's (row, column) pairs are made unique and sorted, and the first len(unique) values
      are kept in that order;
    - the stored coordinates are the set of all combinations of the factors' row indexes (set
      order, as the reference iterates it);
    - an entry is sum_r prod over the modes that HAVE an entry (i_l, r) of that entry (a mode
      without one contributes 1, not 0).
    Returns (data, idxs, dims, factors) with factors as dicts {(row, col): value}."""
    import itertools
    L = len(dims)
    factors = []
    for l in range(L):
        vals = np.random.randn(nnz)
        rows = np.random.randint(0, dims[l], size=nnz)
        cols = np.random.randint(0, R, size=nnz)
        pairs = np.unique(np.stack([rows, cols], axis=1), axis=0)
        vals = vals[:pairs.shape[0]]
        factors.append({tuple(p): v for p, v in zip(pairs, vals)})
    rows_per_mode = [[key[0] for key in factors[l].keys()] for l in range(L)]
    coords = list(set(itertools.product(*rows_per_mode)))
    data = []
    for c in coords:
        acc = 0
        for r in range(R):
            p = 1
            for l in range(L):
                v = factors[l].get((c[l], r))
                if v is not None:
                    p *= v
            acc += p
        data.append(acc)
    return np.array(data), np.array(coords), np.array(dims), factors


def sparse2dense(data, idxs, dims):
    """Dense FP64 array of a COO tensor, repeated coordinates summed (Conversion.py:144-170)."""
    T = np.zeros([int(d) for d in dims])
    idx = np.asarray(idxs).astype(np.int64).reshape(-1, len(dims))
    np.add.at(T, tuple(idx.T), np.asarray(data, dtype=np.float64))
    return T


def sparse2txt(idxs, data, filename, start=1):
    """One line per entry: the coordinates shifted by start + 1, then the value (Auxiliar.py:
    650-673 adds `start` and then 1 again; kept, so files match the reference's)."""
    with open(filename, 'w') as fh:
        for i, d in zip(idxs, data):
            i = np.asarray(i) + start
            fh.write(''.join('%d ' % x for x in (i + 1)) + str(d) + '\n')
