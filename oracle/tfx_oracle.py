"""
FP64 NumPy restatement of Tensor Fox's dense CPD path (test infrastructure only).

This module is the ORACLE: a from-scratch CPU restatement, in float64 NumPy, of
the algorithm behind the reference's ``cpd(T, R, options)`` for dense tensors
(MLSVD compression -> random init -> damped Gauss-Newton with preconditioned CG
-> uncompression -> optional refinement -> final error).  Only ``tests/``, the
``smoke()`` check and the CPU-baseline leg of ``bench.py`` may import it, and
only as the checker / CPU baseline.  The product (the HIP path under
``tensor-fox_amd/``) never imports it.

Every function cites the reference function it follows (paths relative to the
reference checkout).  Behaviour that parity depends on is reproduced on
purpose, including these quirks of the reference:

* ``regularization`` scales by max|P1| of the LAST mode, not the running max
  (GaussNewton.py:594-600).
* ``x`` is never re-synchronised with the equalized factors
  (GaussNewton.py:285-289, Conversion.py:39-47).
* CG solves the split-preconditioned system with damping ``damp*Gamma`` and
  returns ``M*y`` (GaussNewton.py:349-359, 379).
* In ``cg`` the first search direction aliases the residual array
  (``P = residual_cg``, GaussNewton.py:349-350), so the in-place residual
  update makes the second direction ``(1+beta)*r1`` (GaussNewton.py:388-392).
* dGN returns the BEST factors; with refinement the reported ``rel_error`` is
  the LAST refinement error (GaussNewton.py:178-180, Auxiliar.py:148-149).
* Main-stage errors are relative to ||S|| (the core), GaussNewton.py:106.

Randomness uses NumPy's global legacy RNG in the reference's order:
range-finder Gaussians per mode (Compression.py:379-382), ``randn`` init per
mode (Initialization.py:62), one ``randint`` per dGN iteration for the CG cap
(GaussNewton.py:267).  The zero-cleaning draw of ``clean_mode_l`` runs under
Numba's private RNG in the reference (Initialization.py:375-386); here it uses
a private ``RandomState`` so that it never disturbs the global stream either.

Deviation: ``consistency`` does not sleep for one second
(Auxiliar.py:76).

Parity status per function is recorded in the tests that pin it
(tests/test_oracle_*.py); every function here has at least one pin that is not
a retyping of its own formula (closed forms, brute-force einsum, an explicit
Jacobian, finite differences, textbook CG on tiny systems, the reference's
doctest thresholds and notebook outputs).
"""

import sys
from copy import deepcopy

import numpy as np
from numpy.linalg import norm
from scipy.linalg import qr, svd

_CLEAN_RNG = np.random.RandomState(12345)


# ----------------------------------------------------------------------------------------------
# Options / consistency / output (Auxiliar.py)
# ----------------------------------------------------------------------------------------------

_DEFAULTS = dict(
    maxiter=200, tol=1e-16, tol_step=1e-6, tol_improv=1e-6, tol_grad=1e-6, tol_jump=10,
    method='dGN', inner_method='cg', cg_maxiter=100, cg_factor=1, cg_tol=1e-16,
    bi_method_parameters=None, initialization='random', trunc_dims=0, mlsvd_method='seq',
    tol_mlsvd=1e-16, init_damp=1, refine=False, symm=False, factors_norm=0, trials=3,
    display=0, epochs=1, gpu=False, mkl_dot=True, svd_n_components=None)


class _Options:
    pass


def make_options(options):
    """Defaults + attribute-presence overrides.  Follows Auxiliar.py:242-362."""
    out = _Options()
    for k, v in _DEFAULTS.items():
        setattr(out, k, v)
    out.bi_method_parameters = ['als', 500, 1e-6]
    names = dir(options)
    for k in _DEFAULTS:
        if k != 'bi_method_parameters' and k in names:
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
    """Input validation with SystemExit messages.  Follows Auxiliar.py:24-78 (no sleep)."""
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


def sort_dims(T):
    """Descending-dimension permutation as a view.  Follows Auxiliar.py:93-115 (dense branch)."""
    dims = np.array(T.shape)
    L = len(dims)
    ordering = np.argsort(-dims)
    return np.moveaxis(T, ordering, np.arange(-L, 0)), ordering


def unsort_dims(factors, ordering):
    """Inverse of sort_dims on the factor list.  Follows Auxiliar.py:118-128."""
    new = [None] * len(factors)
    for l in range(len(factors)):
        new[ordering[l]] = factors[l]
    return new


class Output:
    """Fields of the reference output class.  Follows Auxiliar.py:131-220."""

    def __init__(self, num_steps, rel_error, step_sizes, errors, improv, gradients, stop, options):
        self.num_steps = num_steps
        self.rel_error = rel_error
        self.accuracy = max(0, 100 * (1 - rel_error))
        self.step_sizes = step_sizes
        self.errors = errors
        self.improv = improv
        self.gradients = gradients
        self.stop = stop
        self.options = options


# ----------------------------------------------------------------------------------------------
# Multilinear algebra and conversions (Conversion.py, Critical.py, MultilinearAlgebra.py)
# ----------------------------------------------------------------------------------------------

def unfold(T, mode):
    """Kolda mode-`mode` unfolding (1-based): column index = remaining indices, first remaining
    fastest.  Follows Critical.py:123-150 (order 3) and the generated orders up to 12."""
    n = mode - 1
    return np.reshape(np.moveaxis(T, n, 0), (T.shape[n], -1), order='F')


def foldback(Tl, mode, dims):
    """Inverse of unfold.  Follows Conversion.py:305-315 / Critical.py:1293-1365."""
    n = mode - 1
    rest = [dims[m] for m in range(len(dims)) if m != n]
    X = np.reshape(Tl, [dims[n]] + rest, order='F')
    return np.ascontiguousarray(np.moveaxis(X, 0, n))


def khatri_rao(A, B):
    """M[i*b1 + k, r] = A[i, r] * B[k, r].  Follows MultilinearAlgebra.py:248-273."""
    a1, R = A.shape
    b1 = B.shape[0]
    return (A[:, None, :] * B[None, :, :]).reshape(a1 * b1, R)


def cpd2unfold1(factors):
    """First unfolding of [[F_1..F_L]].  Follows Conversion.py:117-141."""
    M = factors[1]
    for l in range(2, len(factors)):
        M = khatri_rao(factors[l], M)
    return factors[0] @ M.T


def cpd2tens(factors):
    """Dense tensor of [[F_1..F_L]].  Follows Conversion.py:52-79."""
    dims = [f.shape[0] for f in factors]
    return foldback(cpd2unfold1(factors), 1, dims)


def fastnorm(A, B):
    """Frobenius norm of A - B.  Follows Critical.py:14-22."""
    return float(np.sqrt(np.sum((A - B) ** 2)))


def equalize(factors, R):
    """Per rank-one term, give every mode the norm (prod ||f_l^r||)^(1/L).
    Follows Conversion.py:361-387 (in place)."""
    L = len(factors)
    for r in range(R):
        norm_r = np.array([norm(factors[l][:, r]) for l in range(L)])
        if np.prod(norm_r) != 0.0:
            num = np.prod(norm_r) ** (1 / L)
            for l in range(L):
                factors[l][:, r] = (num / norm_r[l]) * factors[l][:, r]
    return factors


def transform(factors, symm, factors_norm):
    """Symmetric averaging and fixed factor norm.  Follows Conversion.py:412-444, including the
    in-place accumulation into factors[0]."""
    L = len(factors)
    if symm:
        s = factors[0]
        for l in range(1, L):
            s += factors[l]
        factors[0] = s / L
        for l in range(1, L):
            factors[l] = factors[0]
    if factors_norm > 0:
        for l in range(L):
            factors[l] = factors_norm * (1 / norm(factors[l])) * factors[l]
    return factors


def x2cpd(x, factors, eq=True):
    """Write the flat point x (column-major per factor) into the factors, then equalize.
    Follows Conversion.py:21-49 (in place)."""
    R = factors[0].shape[1]
    s = 0
    for l in range(len(factors)):
        dim = factors[l].shape[0]
        for r in range(R):
            factors[l][:, r] = x[s:s + dim]
            s += dim
    if eq:
        factors = equalize(factors, R)
    return factors


def multilin_mult(UT, T1, dims):
    """(U_1^T, ..., U_L^T) * T from the first unfolding.  Follows MultilinearAlgebra.py:72-105."""
    L = len(dims)
    S_dims = list(dims)
    unf = T1
    for l in range(L):
        unf2 = UT[l] @ unf
        S_dims[l] = UT[l].shape[0]
        S = foldback(unf2, l + 1, S_dims)
        if l < L - 1:
            unf = unfold(S, l + 2)
        else:
            return S


# ----------------------------------------------------------------------------------------------
# MLSVD (Compression.py)
# ----------------------------------------------------------------------------------------------

def randomized_range_finder(A, size, n_iter):
    """Gaussian sketch, n_iter un-normalised power iterations, economic QR.
    Follows Compression.py:337-397 (draws from the global legacy RNG)."""
    Q = np.random.mtrand._rand.normal(size=(A.shape[1], size))
    for _ in range(n_iter):
        Q = A @ Q
        Q = A.T @ Q
    Q, _ = qr(A @ Q, mode='economic')
    return Q


def svd_flip(u, v, u_based_decision=True):
    """Sign convention: largest-|.| loading positive.  Follows Compression.py:521-564."""
    if u_based_decision:
        idx = np.argmax(np.abs(u), axis=0)
        signs = np.sign(u[idx, range(u.shape[1])])
    else:
        idx = np.argmax(np.abs(v), axis=1)
        signs = np.sign(v[range(v.shape[0]), idx])
    u *= signs
    v *= signs[:, np.newaxis]
    return u, v


def randomized_svd(M, n_components, n_oversamples=10, n_iter=2):
    """Halko et al. Alg. 4.3/4.4 with transpose='auto' and flip_sign=True.
    Follows Compression.py:400-518."""
    n_random = n_components + n_oversamples
    n_samples, n_features = M.shape
    transpose = n_samples < n_features
    if transpose:
        M = M.T
    Q = randomized_range_finder(M, n_random, n_iter)
    B = Q.T @ M
    Uhat, s, V = svd(B, full_matrices=False)
    U = Q @ Uhat
    if not transpose:
        U, V = svd_flip(U, V)
    else:
        U, V = svd_flip(U, V, u_based_decision=False)
    if transpose:
        return V[:n_components, :].T, s[:n_components], U[:, :n_components].T
    return U[:, :n_components], s[:n_components], V[:n_components, :]


def clean_compression(U, sigma, Vt, tol_mlsvd, L):
    """Smallest i with sum(sigma[i:]^2)/sum(sigma^2) < tol_mlsvd/L.  Follows Compression.py:261-301."""
    eps = tol_mlsvd / L
    for i in range(1, len(sigma)):
        if np.sum(sigma[i:] ** 2) / np.sum(sigma ** 2) < eps:
            sigma = sigma[:i]
            U = U[:, :i]
            Vt = Vt[:i, :]
            break
    return U, sigma, Vt, sigma.size


def compute_svd(Tl, dims, R, tol_mlsvd, L, l, svd_n_components):
    """Truncated rSVD of one unfolding.  Follows Compression.py:192-237 (dense CPU branch);
    note n_components uses the ORIGINAL dims (Compression.py:197)."""
    low_rank = min(R, dims[l])
    n_components = low_rank if svd_n_components is None else svd_n_components
    Ul, sigma_l, Vlt = randomized_svd(Tl, n_components, n_oversamples=10, n_iter=2)
    return clean_compression(Ul, sigma_l, Vlt, tol_mlsvd, L)


def _pick_tol(tol_mlsvd, L):
    if type(tol_mlsvd) == list:
        return tol_mlsvd[0] if L > 3 else tol_mlsvd[1]
    return tol_mlsvd


def mlsvd(T, Tsize, R, options):
    """Truncated MLSVD: 'seq' (ST-HOSVD) or 'classic'.  Follows Compression.py:48-189 (dense)."""
    options = make_options(options)
    dims = T.shape
    L = len(dims)
    tol_mlsvd = _pick_tol(options.tol_mlsvd, L)
    trunc_dims = options.trunc_dims
    U, sigmas = [], []
    if tol_mlsvd == -1:
        return np.array(T), [np.identity(dims[l]) for l in range(L)], unfold(T, 1), \
            [np.ones(dims[l]) for l in range(L)]
    if options.mlsvd_method == 'seq':
        S_dims = list(dims)
        S = np.asarray(T)
        T1 = None
        for l in range(L):
            Sl = unfold(S, l + 1)
            if l == 0:
                T1 = np.ascontiguousarray(Sl)
            Ul, sl, Vlt, dim = compute_svd(Sl, dims, R, tol_mlsvd, L, l, options.svd_n_components)
            U.append(Ul)
            sigmas.append(sl)
            Sl = (Vlt.T * sl).T
            S_dims[l] = dim
            S = foldback(Sl, l + 1, S_dims)
    elif options.mlsvd_method == 'classic':
        T1 = None
        for l in range(L):
            Tl = unfold(T, l + 1)
            if l == 0:
                T1 = np.ascontiguousarray(Tl)
            Ul, sl, Vlt, dim = compute_svd(Tl, dims, R, tol_mlsvd, L, l, options.svd_n_components)
            U.append(Ul)
            sigmas.append(sl)
        S = multilin_mult([u.T for u in U], T1, dims)
    else:
        raise ValueError("mlsvd_method must be 'seq' or 'classic' on the dense path")
    if type(trunc_dims) == list:
        slices = []
        for l in range(L):
            slices.append(slice(0, trunc_dims[l]))
            U[l] = U[l][:, :min(trunc_dims[l], U[l].shape[1])]
        S = S[tuple(slices)]
    return S, U, T1, sigmas


# ----------------------------------------------------------------------------------------------
# Initialization (Initialization.py)
# ----------------------------------------------------------------------------------------------

def clean_zeros(init_factors, dims, R):
    """Replace exact zeros by 1e-4*randn*||F||.  Follows Initialization.py:361-386 (private RNG)."""
    for l in range(len(dims)):
        F = init_factors[l]
        for i in range(dims[l]):
            for r in range(R):
                if F[i, r] == 0.0:
                    F[i, r] = 1e-4 * _CLEAN_RNG.randn() * norm(F)
    return init_factors


def starting_point(S, U, R, ordering, options):
    """'random' or user-list initialization.  Follows Initialization.py:20-94."""
    init = options.initialization
    dims = S.shape
    L = len(dims)
    if type(init) == list:
        return [U[l].T @ init[ordering[l]] for l in range(L)]
    if init == 'random':
        F = [np.random.randn(dims[l], R) for l in range(L)]
    elif init == 'smart_random':
        F = smart_random(S, dims, R)
    elif init == 'smart':
        F = smart(S, dims, R)
    else:
        sys.exit('Error with init parameter initialization.')
    F = clean_zeros(F, dims, R)
    F = equalize(F, R)
    return transform(F, options.symm, options.factors_norm)


def smart_sample(S, dims, R):
    """R core coordinates from the triangular index distribution; the rank-one term r is
    S[c_r] e_{c_r}.  Follows Initialization.py:144-191 (linear search over the cumulative sums,
    as assign_values does)."""
    L = len(dims)
    F = [np.zeros((dims[l], R)) for l in range(L)]
    arr = [dims[l] * np.ones(dims[l]) - np.arange(dims[l]) for l in range(L)]
    high = [np.sum(arr[l]) for l in range(L)]
    C = [np.random.randint(high[l], size=R) for l in range(L)]
    for r in range(R):
        idx = []
        for l in range(L):
            for i in range(dims[l]):
                if np.sum(arr[l][0:i]) <= C[l][r] < np.sum(arr[l][0:i + 1]):
                    F[l][i, r] = 1
                    idx.append(i)
                    break
        F[0][idx[0], r] = S[tuple(idx)]
    return F


def smart_random(S, dims, R):
    """Best of 1 + int(sqrt(prod(dims))) smart_sample draws by |S1 - S1_init| / |S| (the full
    unfolding is formed for every sample).  Follows Initialization.py:97-141."""
    samples = 1 + int(np.sqrt(int(np.prod(dims))))
    S1 = unfold(S, 1)
    Ssize = norm(S)
    best_error, best = np.inf, None
    for _ in range(samples):
        F = smart_sample(S, dims, R)
        e = norm(S1 - cpd2unfold1(F)) / Ssize
        if e < best_error:
            best_error, best = e, [f.copy() for f in F]
    return best


def join_rank1(idx):
    """Follows Initialization.py:258-324 literally (list.remove on the working copy)."""
    R = len(idx)
    L = len(idx[0])
    count = 0
    tmp = idx.copy()
    joined = [[] for _ in range(L * R)]
    i = 0
    for l in range(L):
        indexes = [j for j in range(L) if j != l]
        r = 0
        while r < len(tmp):
            term = tuple(tmp[r][j] for j in indexes)
            rr = r + 1
            while rr < len(tmp):
                if tuple(tmp[rr][j] for j in indexes) == term:
                    joined[i].append(tmp[rr])
                    tmp.remove(tmp[rr])
                    count += 1
                rr += 1
            if len(joined[i]) > 0:
                joined[i].append(tmp[r])
                tmp.remove(tmp[r])
            i += 1
            r += 1
    return count, [joined[i] for i in range(L * R) if len(joined[i]) > 0]


def find_pivot(x):
    """Follows Initialization.py:327-358."""
    if len(x) == 1:
        return 0
    for l in range(len(x[0])):
        if x[0][l] != x[1][l]:
            return l


def smart(S, dims, R):
    """Largest-magnitude entries merged into rank-one terms.  Follows Initialization.py:194-255,
    with the reference's repeated argmax over |S| (O(|S|^2): small cores only)."""
    L = len(dims)
    S_cp = np.abs(S).copy()
    idx = []
    for i in range(S.size):
        idx.append(tuple(int(v) for v in np.unravel_index(np.argmax(S_cp), S_cp.shape)))
        S_cp[idx[i]] = -np.inf
    final_idx = idx[:R].copy()
    idx = idx[R:]
    previous_count = 0
    for _ in range(100):
        count, joined = join_rank1(final_idx)
        if count - previous_count <= 0:
            break
        final_idx = final_idx + idx[:count - previous_count]
        idx = idx[count - previous_count:]
        previous_count = count
    rank1_idx = [[] for _ in range(R)]
    r = 0
    for x in joined:
        for y in x:
            rank1_idx[r].append(y)
            final_idx.remove(y)
        r += 1
    for x in final_idx:
        rank1_idx[r].append(x)
        r += 1
    F = [np.zeros((dims[l], R)) for l in range(L)]
    for r in range(R):
        mode = find_pivot(rank1_idx[r])
        for x in rank1_idx[r]:
            for l in range(L):
                if l == mode:
                    F[l][x[l], r] += S[x]
                else:
                    F[l][x[l], r] = 1
    return F


# ----------------------------------------------------------------------------------------------
# Damped Gauss-Newton (GaussNewton.py)
# ----------------------------------------------------------------------------------------------

def gramians(factors):
    """Gr_l = F_l^T F_l, P2[l][m] = prod_{k != l,m} Gr_k, P1[l] = prod_{k != l} Gr_k.
    Follows GaussNewton.py:520-548."""
    L = len(factors)
    R = factors[0].shape[1]
    Gr = np.array([f.T @ f for f in factors])
    P2 = np.ones((L, L, R, R))
    for l in range(L):
        for m in range(L):
            if l != m:
                for k in range(L):
                    if k != l and k != m:
                        P2[l, m] *= Gr[k]
    P1 = np.ones((L, R, R))
    for l in range(L):
        for k in range(L):
            if k != l:
                P1[l] *= Gr[k]
    return Gr, P1, P2


def _split(v, dims, R):
    out, s = [], 0
    for d in dims:
        out.append(v[s:s + R * d].reshape(R, d).T)   # I_l x R view of a column-major block
        s += R * d
    return out


def matvec(factors, v):
    """(J^T J) v for the CPD Jacobian: result_l = V_l P1_l + F_l sum_{m != l} P2[l][m] o (V_m^T F_m).
    Follows GaussNewton.py:342-346, 551-581 (A[m] = V_m^T F_m, B[l] = V_l P1[l])."""
    _, P1, P2 = gramians(factors)
    return _matvec_P(factors, P1, P2, v)


def regularization(P1, dims):
    """gamma[l, r] = max|P1_{L-1}| * sqrt|P1_l[r, r]| (the last mode's max, a reference quirk);
    Gamma repeats gamma[l, r] over the I_l rows of column r.  Follows GaussNewton.py:584-606."""
    L, R = P1.shape[0], P1.shape[1]
    tmp = np.max(np.abs(P1[L - 1]))
    gamma = tmp * np.sqrt(np.abs(np.array([np.diag(P1[l]) for l in range(L)])))
    Gamma = np.concatenate([np.repeat(gamma[l], dims[l]) for l in range(L)])
    return Gamma, gamma


def precond(Gamma, gamma, damp, dims):
    """M = 1/sqrt(gamma^2 + damp^2 Gamma^2).  Follows GaussNewton.py:609-627."""
    g = np.concatenate([np.repeat(gamma[l], dims[l]) for l in range(len(dims))])
    return 1 / np.sqrt(g ** 2 + damp ** 2 * Gamma ** 2)


def compute_grad(T, factors, P1):
    """grad f = vec(F_l P1_l - T_(l) KR_l) with KR_l the Khatri-Rao product of the other factors in
    reversed mode order.  Follows GaussNewton.py:401-427."""
    L = len(factors)
    g = []
    for l in range(L):
        itr = [m for m in reversed(range(L)) if m != l]
        M = factors[itr[0]]
        for ll in range(L - 2):
            M = khatri_rao(M, factors[itr[ll + 1]])
        N = unfold(T, l + 1) @ M
        g.append((factors[l] @ P1[l] - N).T.ravel())
    return np.concatenate(g)


def mttkrp(T, factors, l):
    """T_(l) times the Khatri-Rao product of the other factors (reversed order), the MTTKRP
    inside compute_grad (GaussNewton.py:411-423)."""
    L = len(factors)
    itr = [m for m in reversed(range(L)) if m != l]
    M = factors[itr[0]]
    for ll in range(L - 2):
        M = khatri_rao(M, factors[itr[ll + 1]])
    return unfold(T, l + 1) @ M


def cg(T, factors, damp, maxiter, tol):
    """Preconditioned CG on (J^T J + damp Gamma) y = -grad f, returning M*y.
    Follows GaussNewton.py:315-359 and cg_iterations GaussNewton.py:362-398, including the
    aliasing of P and residual_cg in the first iteration."""
    _, P1, _ = gramians(factors)
    grad = -compute_grad(T, factors, P1)
    return cg_from_grad(factors, grad, damp, maxiter, tol)


def cg_from_grad(factors, grad, damp, maxiter, tol):
    """The body of cg (GaussNewton.py:324-359) for a given right-hand side grad."""
    R = factors[0].shape[1]
    dims = [f.shape[0] for f in factors]
    n = R * sum(dims)
    if n > 100:
        maxiter = min(maxiter, n)
    else:
        tol = 0
    Gr, P1, P2 = gramians(factors)
    Gamma, gamma = regularization(P1, dims)
    M = precond(Gamma, gamma, damp, dims)
    y = np.zeros(n)
    JT_J_grad = _matvec_P(factors, P1, P2, grad)
    r = M * grad
    P = r                      # alias, as in the reference
    rn = float(r @ r)
    if rn == 0.0:
        rn = 1e-6
    itn = 0
    for itn in range(maxiter):
        Q = M * P
        z = _matvec_P(factors, P1, P2, Q) + damp * Gamma * Q
        z = M * z
        den = float(P @ z)
        if den == 0.0:
            den = 1e-6
        alpha = rn / den
        y += alpha * P
        r -= alpha * z         # also updates P while it still aliases r
        rn_new = float(r @ r)
        beta = rn_new / rn
        rn = rn_new
        P = r + beta * P
        if rn <= tol:
            break
    return M * y, grad, JT_J_grad, itn + 1, rn


def _matvec_P(factors, P1, P2, v):
    L = len(factors)
    R = factors[0].shape[1]
    dims = [f.shape[0] for f in factors]
    V = _split(v, dims, R)
    A = [V[m].T @ factors[m] for m in range(L)]
    z = []
    for l in range(L):
        tmp = np.zeros((R, R))
        for m in range(L):
            if m != l:
                tmp += P2[l, m] * A[m]
        z.append((factors[l] @ tmp + V[l] @ P1[l]).T.ravel())
    return np.concatenate(z)


def hessian(factors):
    """Explicit Gauss-Newton matrix J^T J built from an explicit Jacobian (test helper; plays the
    role of GaussNewton.py:666-695 for tiny sizes)."""
    J = jacobian(factors)
    return J.T @ J


def jacobian(factors):
    """Explicit Jacobian of vec_C([[F]]) w.r.t. x = concat_l vec_F(F_l), by brute force."""
    L = len(factors)
    R = factors[0].shape[1]
    dims = [f.shape[0] for f in factors]
    N = int(np.prod(dims))
    cols = []
    for l in range(L):
        for r in range(R):
            for i in range(dims[l]):
                vecs = [f[:, r] for f in factors]
                e = np.zeros(dims[l])
                e[i] = 1.0
                vecs[l] = e
                t = vecs[0]
                for m in range(1, L):
                    t = np.multiply.outer(t, vecs[m])
                cols.append(t.reshape(N))
    return np.array(cols).T


def update_gain_ratio(damp, old_error, error, Tsize, old_x, x, grad):
    """Follows GaussNewton.py:472-487."""
    num = (Tsize ** 2) * (old_error ** 2 - error ** 2)
    den = float((x - old_x) @ (grad + damp * (x - old_x)))
    if den != 0.0:
        return num / den
    if num != 0.0:
        return np.sign(num) * np.inf
    return -np.inf


def update_damp(damp, init_damp, gain_ratio, it):
    """Follows GaussNewton.py:490-503."""
    if type(init_damp) == list:
        return init_damp[it]
    if gain_ratio < 0.25:
        return 3 * damp
    if gain_ratio > 0.75:
        return damp / 3
    return damp


def update_delta(delta, gain_ratio, step_size):
    """Follows GaussNewton.py:506-517."""
    if gain_ratio > 0.75:
        return max(delta, 2 * step_size)
    sigma = 0.25 / (1 + np.exp(-14 * (gain_ratio - 0.75))) + 0.75
    return min(sigma * delta, delta)


def dogleg(y, grad, JT_J_grad, delta):
    """Powell dogleg step.  Follows GaussNewton.py:766-795."""
    alpha = norm(grad) ** 2 / float(grad @ JT_J_grad)
    y_sd = alpha * grad
    c1 = norm(y_sd)
    c2 = float((y - y_sd) @ y_sd)
    c3 = norm(y - y_sd) ** 2
    c4 = delta ** 2 - c1 ** 2
    m = c2 ** 2 + c3 * c4
    if norm(y) < delta:
        return y
    if c1 >= delta:
        return (delta / c1) * y_sd
    if m < 0:
        return y
    c5 = np.sqrt(m)
    beta = (-c2 + c5) / c3 if c2 <= 0 else c4 / (c2 + c5)
    return y_sd + beta * (y - y_sd)


def residual_error(T, factors, Tsize):
    """||T - [[F]]|| / Tsize through the first unfolding.  Follows GaussNewton.py:298-299."""
    return fastnorm(unfold(T, 1), cpd2unfold1(factors)) / Tsize


def compute_dogleg_steps(T, Tsize, factors, grad, JT_J_grad, x, y, error, damp, symm, factors_norm):
    """Up to 11 dogleg trials after an error jump.  Follows GaussNewton.py:713-763."""
    count = 0
    best_x, best_y, best_error = x.copy(), y.copy(), error
    best_factors = deepcopy(factors)
    gain_ratio = 1
    delta = 1
    while gain_ratio > 0:
        old_x, old_error = x, error
        y = dogleg(y, grad, JT_J_grad, delta)
        x = x + y
        factors = x2cpd(x, factors, eq=False)
        factors = transform(factors, symm, factors_norm)
        error = residual_error(T, factors, Tsize)
        gain_ratio = update_gain_ratio(damp, old_error, error, Tsize, old_x, x, grad)
        if error < old_error:
            best_x, best_y, best_error = x.copy(), y.copy(), error
            best_factors = deepcopy(factors)
        delta = update_delta(delta, gain_ratio, norm(x - old_x))
        count += 1
        if count > 10:
            break
    return best_factors, best_x, best_y, best_error


def hessian_blocks(factors, P1, P2):
    """J^T J assembled by blocks: H_ll = P1_l (x) I, H_{l,m} = (P2_{l,m} (x) 1) o [F_l[i, b] F_m[j, a]]
    at ((a, i), (b, j)).  Follows hessian / compute_blocks (GaussNewton.py:666-710)."""
    L = len(factors)
    R = factors[0].shape[1]
    dims = [f.shape[0] for f in factors]
    off = np.concatenate([[0], np.cumsum([R * d for d in dims])])
    H = np.zeros((off[-1], off[-1]))
    for l in range(L):
        H[off[l]:off[l + 1], off[l]:off[l + 1]] = np.kron(P1[l], np.identity(dims[l]))
        vec_m = None
        for m in range(l):
            tmp1 = np.kron(P2[l, m], np.ones((dims[l], dims[m])))
            vec_m = factors[m].flatten('F')[None, :]
            tmp2 = np.zeros((R * dims[l], R * dims[m]))
            for r in range(R):
                tmp = factors[l][:, r][:, None] @ vec_m
                for rr in range(R):
                    tmp2[dims[l] * rr:dims[l] * (rr + 1), dims[m] * r:dims[m] * (r + 1)] = \
                        tmp[:, dims[m] * rr:dims[m] * (rr + 1)]
            H[off[l]:off[l + 1], off[m]:off[m + 1]] = tmp1 * tmp2
            H[off[m]:off[m + 1], off[l]:off[l + 1]] = (tmp1 * tmp2).T
    return H


def direct(T, factors, damp):
    """QR solve of M^2 (H + damp diag(Gamma)) y = M^2 grad (y = 0 on LinAlgError); returns
    (y, grad, '-', |H y - grad|).  Follows GaussNewton.py:630-663."""
    from numpy.linalg import LinAlgError, solve
    dims = [f.shape[0] for f in factors]
    _, P1, P2 = gramians(factors)
    Gamma, gamma = regularization(P1, dims)
    M = precond(Gamma, gamma, damp, dims)
    grad = -compute_grad(T, factors, P1)
    H = hessian_blocks(factors, P1, P2)
    Hd = H + damp * np.diag(Gamma)
    MHd = ((Hd.T) * (M ** 2)).T
    Mgrad = (M ** 2) * grad
    try:
        q, r = qr(MHd)
        y = solve(r, q.T @ Mgrad)
    except LinAlgError:
        y = np.zeros(grad.size)
    return y, grad, '-', norm(H @ y - grad)


def compute_step(T, Tsize, factors, x, inner, it, old_error, L, fixed=None):
    """One dGN step.  Follows GaussNewton.py:254-312 (inner methods 'cg', 'cg_static', 'als',
    'direct'; the dogleg only after a cg step, :298-302)."""
    damp, inner_method, cg_maxiter, cg_factor, cg_tol, tol_jump, symm, factors_norm = inner
    if type(inner_method) == list:
        inner_method = inner_method[it]
    if inner_method == 'cg':
        cg_maxiter = 1 + (L - 2) * int(cg_factor * np.random.randint(1 + it ** 0.4, 2 + it ** 0.9))
    elif inner_method not in ('cg_static', 'als', 'direct'):
        sys.exit("Wrong inner method name. Must be 'cg', 'cg_static', 'als' or 'direct'.")
    if inner_method == 'als':
        # GaussNewton.py:273-276: one ALS sweep replaces the step; y is zeroed, the gradient is
        # reported as [nan] and the dogleg is never tried
        factors = als_iteration([unfold(T, l + 1) for l in range(L)], factors)
        x = np.concatenate([f.flatten('F') for f in factors])
        y = np.zeros(x.size)
        grad, JT_J_grad, itn, rn = np.array([np.nan]), None, '-', np.nan
    elif inner_method == 'direct':
        y, grad, itn, rn = direct(T, factors, damp)
        JT_J_grad = None
    else:
        y, grad, JT_J_grad, itn, rn = cg(T, factors, damp, cg_maxiter, cg_tol)
    x = x + y
    factors = x2cpd(x, factors)
    factors = transform(factors, symm, factors_norm)
    if fixed is not None and L == 3:
        factors[fixed[0]] = deepcopy(fixed[1])
    error = residual_error(T, factors, Tsize)
    if inner_method == 'als':
        return factors, x, y, grad, itn, Tsize * error, error
    if it > 3 and inner_method in ('cg', 'cg_static') and error > tol_jump * old_error:
        x = x - y
        factors, x, y, error = compute_dogleg_steps(T, Tsize, factors, grad, JT_J_grad, x, y, error,
                                                    damp, symm, factors_norm)
    return factors, x, y, grad, itn, rn, error


def _split_fixed(factors):
    """A factor given as [F, l] is held fixed (bicpd).  Follows GaussNewton.py:94-102."""
    fixed = None
    factors = list(factors)
    for l in range(len(factors)):
        if type(factors[l]) == list:
            fixed = (l, deepcopy(factors[l][0]))
            factors[l] = factors[l][0]
    return factors, fixed


def dGN(T, factors, R, options, trace=None):
    """Damped Gauss-Newton.  Follows GaussNewton.py:29-251, including a fixed mode (a factor
    given as [F, l], :94-102, restored after every update, :292-295).  If `trace` is a list,
    per-iteration (error, damp, gain_ratio, itn) tuples are appended to it."""
    factors, fixed = _split_fixed(factors)
    init_damp = options.init_damp
    maxiter = options.maxiter
    tol, tol_step, tol_improv, tol_grad = options.tol, options.tol_step, options.tol_improv, options.tol_grad
    L = len(factors)
    dims = T.shape
    Tsize = norm(T)
    error = 1
    best_error = np.inf
    stop = 5
    damp = init_damp[0] if type(init_damp) == list else init_damp * np.mean(np.abs(T))
    const = 1 + int(maxiter / 10)
    if R * sum(dims) <= 100:
        tol = tol_step = tol_improv = tol_grad = 0
    x = np.concatenate([f.flatten('F') for f in factors])
    step_sizes = np.zeros(maxiter)
    errors = np.zeros(maxiter)
    improv = np.zeros(maxiter)
    gradients = np.zeros(maxiter)
    best_factors = deepcopy(factors)
    it = 0
    for it in range(maxiter):
        old_x, old_error = x, error
        inner = (damp, options.inner_method, options.cg_maxiter, options.cg_factor, options.cg_tol,
                 options.tol_jump, options.symm, options.factors_norm)
        factors, x, y, grad, itn, rn, error = compute_step(T, Tsize, factors, x, inner, it, old_error, L,
                                                           fixed)
        gain_ratio = update_gain_ratio(damp, old_error, error, Tsize, old_x, x, grad)
        if trace is not None:
            trace.append((error, damp, gain_ratio, itn))
        damp = update_damp(damp, init_damp, gain_ratio, it)
        if error < best_error:
            best_error = error
            best_factors = deepcopy(factors)
        errors[it] = error
        step_sizes[it] = norm(x - old_x) / norm(old_x)
        gradients[it] = norm(grad, np.inf)
        improv[it] = errors[it] if it == 0 else abs(errors[it] - errors[it - 1])
        if it > 1:
            if errors[it] < tol:
                stop = 0
                break
            if step_sizes[it] < tol_step:
                stop = 1
                break
            if improv[it] < tol_improv:
                stop = 2
                break
            if gradients[it] < tol_grad:
                stop = 3
                break
            if it > 2 * const and it % const == 0:
                mean1 = np.mean(errors[it - 2 * const: it - const])
                mean2 = np.mean(errors[it - const: it])
                if mean1 - mean2 <= tol_improv:
                    stop = 4
                    break
                mean3 = np.mean(improv[it - const: it])
                if mean3 < 1e-3 * mean2:
                    stop = 7
                    break
            if error > max(1, Tsize ** 2) / (1e-16 + tol):
                stop = 6
                break
    n = it + 1
    return best_factors, step_sizes[:n], errors[:n], improv[:n], gradients[:n], stop


# ----------------------------------------------------------------------------------------------
# Alternating least squares (Alternating_Least_Squares.py)
# ----------------------------------------------------------------------------------------------

def als_iteration(Tl, factors, fix_mode=-1):
    """One Gauss-Seidel sweep over the modes: F_l <- T_(l) pinv(KR_l^T), with KR_l the
    Khatri-Rao product of the other factors in reversed mode order, each mode using the modes
    already updated in this sweep.  `factors` is modified in place and returned.  Follows
    Alternating_Least_Squares.py:206-264; with a fixed mode (third order, bicpd) that mode is
    skipped and the other two are swept in the same order (:233-262)."""
    L = len(factors)
    for l in range(L):
        if l == fix_mode:
            continue
        itr = [m for m in reversed(range(L)) if m != l]
        M = factors[itr[0]]
        for ll in range(L - 2):
            M = khatri_rao(M, factors[itr[ll + 1]])
        factors[l] = Tl[l] @ np.linalg.pinv(M.T)
    return factors


def als(T, factors, R, options):
    """ALS iterations with the reference's stopping rules.  Follows
    Alternating_Least_Squares.py:18-203 (a factor given as [F, l] is held fixed, :60-67,
    119-124): the step size is the absolute
    ||x - old_x|| of the unbalanced point taken before `transform`, the 'gradient' is
    |old_error - error| / step size, and the LAST factors are returned, not the best."""
    factors, fixed = _split_fixed(factors)
    fix_mode = fixed[0] if fixed is not None else -1
    maxiter = options.maxiter
    tol, tol_step, tol_improv, tol_grad = options.tol, options.tol_step, options.tol_improv, options.tol_grad
    L = len(factors)
    Tsize = norm(T)
    error = 1
    stop = 5
    const = 1 + int(maxiter / 10)
    x = np.concatenate([f.flatten('F') for f in factors])
    step_sizes = np.empty(maxiter)
    errors = np.empty(maxiter)
    improv = np.empty(maxiter)
    gradients = np.empty(maxiter)
    Tl = [unfold(T, l + 1) for l in range(L)]
    it = 0
    for it in range(maxiter):
        old_x, old_error = x, error
        factors = als_iteration(Tl, factors, fix_mode)
        x = np.concatenate([f.flatten('F') for f in factors])
        factors = transform(factors, options.symm, options.factors_norm)
        if fixed is not None and L == 3:
            factors[fix_mode] = fixed[1].copy()
        error = fastnorm(Tl[0], cpd2unfold1(factors)) / Tsize
        step_sizes[it] = norm(x - old_x)
        errors[it] = error
        with np.errstate(divide='ignore', invalid='ignore'):
            gradients[it] = np.float64(np.abs(old_error - error)) / step_sizes[it]
        improv[it] = errors[it] if it == 0 else np.abs(errors[it - 1] - errors[it])
        if it > 1:
            if errors[it] < tol:
                stop = 0
                break
            if step_sizes[it] < tol_step:
                stop = 1
                break
            if improv[it] < tol_improv:
                stop = 2
                break
            if gradients[it] < tol_grad:
                stop = 3
                break
            if it > 2 * const and it % const == 0:
                mean1 = np.mean(errors[it - 2 * const: it - const])
                mean2 = np.mean(errors[it - const: it])
                if mean1 - mean2 <= tol_improv:
                    stop = 4
                    break
                mean3 = np.mean(improv[it - const: it])
                if mean3 < 1e-3 * mean2:
                    stop = 7
                    break
            if error > max(1, Tsize ** 2) / (1e-16 + tol):
                stop = 6
                break
    n = it + 1
    return factors, step_sizes[:n], errors[:n], improv[:n], gradients[:n], stop


# ----------------------------------------------------------------------------------------------
# Drivers (TensorFox.py)
# ----------------------------------------------------------------------------------------------

# ----------------------------------------------------------------------------------------------
# Sparse COO input T = [data, idxs, dims] (TensorFox.py:256-280, Compression.py:122-137)
# ----------------------------------------------------------------------------------------------

def sparse_sort_dims(T):
    """Descending-dimension permutation of the coordinate columns.  Follows Auxiliar.py:99-106."""
    data, idxs, dims = T
    idxs = np.array(idxs, dtype=np.uint64)
    dims = np.array(dims)
    ordering = np.argsort(-dims)
    return [data, idxs[:, ordering], tuple(dims[ordering])], ordering


def sparse_unfold(data, idxs, dims, mode):
    """CSR mode-`mode` unfolding (1-based); column = remaining coordinates, first remaining
    fastest.  Follows Conversion.py:255-302."""
    from scipy.sparse import coo_matrix
    L = len(dims)
    dims = [int(d) for d in dims]
    rest = [l for l in range(L) if l != mode - 1]
    K = np.zeros(L, dtype=np.uint64)
    s = 1
    for l in rest:
        K[l] = s
        s *= dims[l]
    idxs = np.asarray(idxs, dtype=np.uint64)
    rows = idxs[:, mode - 1]
    cols = np.sum(K * idxs, axis=1, dtype=np.uint64)
    return coo_matrix((np.asarray(data, dtype=np.float64), (rows, cols)),
                      shape=(dims[mode - 1], s)).tocsr()


def sparse_gram(data, idxs, dims, mode):
    """Dense T_(mode) T_(mode)^T.  Follows Compression.py:211-215 (sparse_dot_mkl / sparse_dot)."""
    Tl = sparse_unfold(data, idxs, dims, mode)
    return np.asarray((Tl @ Tl.T).toarray(), dtype=np.float64)


def sparse_multilin_mult(UT, data, idxs, dims):
    """S = (U_1^T, ..., U_L^T) * T = sum_k data_k (x)_l U_l^T[:, idxs[k, l]] (dense core).
    Follows MultilinearAlgebra.py:145-175 / Critical.py:2704-2721 (orders 3-12)."""
    L = len(dims)
    idxs = np.asarray(idxs).astype(np.int64)
    nnz = len(data)
    M = np.asarray(data, dtype=np.float64)[None, :]
    for l in range(L):
        Ul = UT[l][:, idxs[:, l]]
        M = (M[:, None, :] * Ul[None, :, :]).reshape(-1, nnz)
    return M.sum(axis=1).reshape([u.shape[0] for u in UT])


def sparse_fastnorm(data, idxs, dims, factors):
    """|data - (model at the stored coordinates)|, zeros not counted.  Follows Critical.py:25-53."""
    idxs = np.asarray(idxs).astype(np.int64)
    P = np.ones((len(data), factors[0].shape[1]))
    for l in range(len(factors)):
        P = P * factors[l][idxs[:, l], :]
    return float(norm(np.asarray(data, dtype=np.float64) - P.sum(axis=1)))


def sparse_mlsvd(T, R, options):
    """Per mode: rSVD of the dense Gram of the sparse unfolding with
    n_components = int(max(L, log10(I_l^2 L)) * min(R, I_l)), sigma = sqrt(s), energy
    truncation; then the sparse multilinear product.  Follows Compression.py:122-137, 192-237
    (sparse CPU branch) and the trunc_dims step :171-177."""
    data, idxs, dims = T
    dims = np.array(dims)
    L = len(dims)
    tol_mlsvd = _pick_tol(options.tol_mlsvd, L)
    U, sigmas = [], []
    for l in range(L):
        G = sparse_gram(data, idxs, dims, l + 1)
        low_rank = min(R, dims[l])
        if options.svd_n_components is None:
            n_components = int(max(L, np.log10(dims[l] * dims[l] * L)) * low_rank)
        else:
            n_components = options.svd_n_components
        Ul, sl, Vlt = randomized_svd(G, n_components, n_oversamples=10, n_iter=2)
        sl = np.sqrt(sl)
        Ul, sl, Vlt, dim = clean_compression(Ul, sl, Vlt, tol_mlsvd, L)
        U.append(Ul)
        sigmas.append(sl)
    S = sparse_multilin_mult([u.T for u in U], data, idxs, dims)
    if type(options.trunc_dims) == list:
        slices = []
        for l in range(L):
            slices.append(slice(0, options.trunc_dims[l]))
            U[l] = U[l][:, :min(options.trunc_dims[l], U[l].shape[1])]
        S = S[tuple(slices)]
    return S, U, sigmas


def sparse_tricpd(T, R, options, stages=None):
    """tricpd for T = [data, idxs, dims]: sparse MLSVD, dGN/ALS on the dense core, U*W, no
    refinement, error over the stored entries.  Follows TensorFox.py:457-500, 577-581,
    627-640 and Auxiliar.py:150-159."""
    data_orig = np.array(T[0], dtype=np.float64)
    idxs_orig = np.array(T[1], dtype=np.uint64)
    dims_orig = np.array(T[2])
    L = len(dims_orig)
    tol = options.tol_mlsvd[0] if type(options.tol_mlsvd) == list else options.tol_mlsvd
    T, ordering = sparse_sort_dims([data_orig, idxs_orig, dims_orig])
    Tsize = norm(T[0])
    if tol < 0:
        options.tol_mlsvd = 1e-16
    S, U, sigmas = sparse_mlsvd(T, R, options)
    if options.symm:
        R_min = min(S.shape)
        S = S[tuple(slice(R_min) for _ in range(L))]
        U = [u[:, :R_min] for u in U]
    init = starting_point(S, U, R, ordering, options)
    if stages is not None:
        stages.update(S=S, U=U, init=deepcopy(init), sigmas=sigmas, ordering=ordering)
    solver = als if options.method == 'als' else dGN
    factors, ss_m, e_m, im_m, g_m, stop_m = solver(S, init, R, options)
    factors = [U[l] @ factors[l] for l in range(L)]
    options.refine = False
    factors = unsort_dims(factors, ordering)
    rel_error = sparse_fastnorm(data_orig, idxs_orig, dims_orig, factors) / Tsize
    z = np.array([0])
    out = Output(ss_m.size, rel_error, [ss_m, z], [e_m, z], [im_m, z], [g_m, z], [stop_m, 8],
                 options)
    return factors, out


def sparse2dense(data, idxs, dims):
    """Dense array of a COO tensor (duplicates summed).  Follows Conversion.py:144-170."""
    T = np.zeros([int(d) for d in dims])
    np.add.at(T, tuple(np.asarray(idxs).astype(np.int64).T), np.asarray(data, dtype=np.float64))
    return T


def tricpd(T, R, options, stages=None):
    """sort -> mlsvd -> init -> dGN or ALS -> U*W -> optional refinement -> final error -> unsort.
    Follows TensorFox.py:424-655 (dense, methods 'dGN' and 'als', :563-568, :599-604).  `stages`, if a dict, receives
    the intermediate results (core, U, init factors) for tests."""
    L = T.ndim
    T, ordering = sort_dims(T)
    Tsize = norm(T)
    S, U, T1, sigmas = mlsvd(T, Tsize, R, options)
    if options.symm:
        R_min = min(S.shape)
        S = S[tuple(slice(R_min) for _ in range(L))]
        U = [u[:, :R_min] for u in U]
    init = starting_point(S, U, R, ordering, options)
    if stages is not None:
        stages.update(S=S, U=U, init=deepcopy(init), sigmas=sigmas, ordering=ordering)
    solver = als if options.method == 'als' else dGN
    factors, ss_m, e_m, im_m, g_m, stop_m = solver(S, init, R, options)
    factors = [U[l] @ factors[l] for l in range(L)]
    if stages is not None:
        stages.update(main_factors=deepcopy(factors))
    if options.refine:
        factors, ss_r, e_r, im_r, g_r, stop_r = solver(np.asarray(T), factors, R, options)
    else:
        ss_r, e_r, im_r, g_r, stop_r = np.array([0]), np.array([0]), np.array([0]), np.array([0]), 8
    T1_approx = cpd2unfold1(factors)
    factors = unsort_dims(factors, ordering)
    if options.refine:
        num_steps = ss_m.size + ss_r.size
        rel_error = e_r[-1]
    else:
        num_steps = ss_m.size
        rel_error = fastnorm(T1, T1_approx) / Tsize
    out = Output(num_steps, rel_error, [ss_m, ss_r], [e_m, e_r], [im_m, im_r], [g_m, g_r],
                 [stop_m, stop_r], options)
    return factors, out


# ----------------------------------------------------------------------------------------------
# Tensor-train CPD, method 'ttcpd' (TensorFox.py:252-421, 658-817, Auxiliar.py:365-527)
# ----------------------------------------------------------------------------------------------

def tt_core(V, dims, r1, r2, l):
    """One core of the CPD tensor train: Fortran-order reshape to (r1 I_l) x rest, truncated SVD
    by scikit-learn's randomized_svd with n_iter=0 and all min(shape) components (the sketch is
    drawn from the global generator, random_state=None), rank r2.  Follows Auxiliar.py:365-383."""
    from sklearn.utils.extmath import randomized_svd
    V = V.reshape(r1 * dims[l], int(np.prod(dims[l + 1:])), order='F')
    low_rank = min(V.shape[0], V.shape[1])
    U, S, V = randomized_svd(V, low_rank, n_iter=0)
    U = U[:, :r2]
    V = (np.diag(S) @ V)[:r2, :]
    g = U.reshape(dims[l], r2, order='F') if r1 == 1 else U.reshape(r1, dims[l], r2, order='F')
    return V, g


def cpdtt(T, R):
    """Tensor train dims[0] x R -> R x dims[1] x R -> ... -> R x dims[L-1].  Follows
    TensorFox.py:1037-1051."""
    dims = np.array(T.shape)
    L = len(dims)
    G = []
    r1, r2 = 1, R
    V = T
    for l in range(L - 1):
        V, g = tt_core(V, dims, r1, r2, l)
        r1, r2 = R, R
        G.append(g)
    G.append(V)
    return G


def inflate(T, R, dims):
    """Modes shorter than R padded to R with 1e-12 randn entries (global generator).  Follows
    Conversion.py:460-480."""
    L = len(dims)
    new_dims = [max(int(d), R) for d in dims]
    new_T = 1e-12 * np.random.randn(*new_dims)
    new_T[tuple(slice(int(d)) for d in dims)] = T
    return new_T


def deflate(factors, orig_dims, inflate_status):
    """Follows Conversion.py:483-495."""
    if not inflate_status:
        return factors
    return [factors[l][:orig_dims[l], :] for l in range(len(orig_dims))]


def bicpd(T, R, fixed_factor, options):
    """Third-order CPD with one factor fixed: MLSVD (no mode sorting), random start with the
    fixed mode's start U_l^T F_fixed, ALS (bi_method_parameters[0], default) or dGN with that
    mode held, uncompression of the other two, error with the fixed factor.  Follows
    TensorFox.py:658-817."""
    bi_method = options.bi_method_parameters[0]
    Tsize = norm(T)
    ordering = [0, 1, 2]
    consistency(R, T.shape, options)
    S, U, T1, sigmas = mlsvd(T, Tsize, R, options)
    if options.symm:
        R_min = min(S.shape)
        S = S[:R_min, :R_min, :R_min]
        U = [u[:, :R_min] for u in U]
    X, Y, Z = starting_point(S, U, R, ordering, options)
    mode = fixed_factor[1]
    init = [X, Y, Z]
    init[mode] = [U[mode].T @ fixed_factor[0], mode]
    solver = als if bi_method == 'als' else dGN
    factors, ss, e, im, g, stop = solver(S, init, R, options)
    X, Y, Z = factors
    out = [U[0] @ X, U[1] @ Y, U[2] @ Z]
    out[mode] = factors[mode]
    full = list(out)
    full[mode] = fixed_factor[0]
    rel_error = fastnorm(T1, cpd2unfold1(full)) / Tsize
    z = np.array([0])
    output = Output(ss.size, rel_error, [ss, z], [e, z], [im, z], [g, z], [stop, 5], options)
    return out[0], out[1], out[2], output


def cpd_cores(G, max_trials, epochs, R, display, options):
    """CPDs of the train's cores: tricpd of G[1] (best of max_trials, early exit below 1e-4),
    then forward (and on odd epochs backward) sweeps of bicpd with the shared factor fixed to
    pinv of the neighbour's.  Follows Auxiliar.py:424-527."""
    L = len(G)
    if epochs % 2 == 0:
        epochs += 1
    cpd_list = list(range(L - 2))
    outputs = list(range(L - 2))
    best_error = np.inf
    for _ in range(max_trials):
        factors, output = tricpd(G[1], R, options)
        X, Y, Z = factors
        if output.rel_error < best_error:
            best_output, best_error = output, output.rel_error
            best_X, best_Y, best_Z = X, Y, Z
            if best_error < 1e-4:
                break
    outputs[0] = best_output
    cpd_list[0] = [best_X, best_Y, best_Z]
    low, upp = 2, L - 2
    for epoch in range(epochs):
        if epoch % 2 == 0:
            for l in range(low, L - 1):
                best_error = np.inf
                fixed_X = np.linalg.pinv(best_Z.T)
                for _ in range(max_trials):
                    X, Y, Z, output = bicpd(G[l], R, [fixed_X, 0], options)
                    if output.rel_error < best_error:
                        best_output, best_error = output, output.rel_error
                        best_X, best_Y, best_Z = fixed_X, Y, Z
                        if best_error < 1e-4:
                            break
                if epoch == epochs - 1:
                    outputs[l - 1] = best_output
                    cpd_list[l - 1] = [fixed_X, best_Y, best_Z]
        else:
            if L == 3:
                low, upp = 1, L - 1
            for l in reversed(range(1, upp)):
                best_error = np.inf
                fixed_Z = np.linalg.pinv(best_X.T)
                for _ in range(max_trials):
                    X, Y, Z, output = bicpd(G[l], R, [fixed_Z, 2], options)
                    if output.rel_error < best_error:
                        best_output, best_error = output, output.rel_error
                        best_X, best_Y, best_Z = X, Y, fixed_Z
                        if best_error < 1e-4:
                            break
                if epoch == epochs - 2:
                    outputs[l - 1] = best_output
                    cpd_list[l - 1] = [best_X, best_Y, fixed_Z]
    return cpd_list, outputs, best_Z


def highcpd(T, R, options):
    """CPD of T from the CPDs of its train cores.  Follows TensorFox.py:368-421."""
    L = T.ndim
    options.refine = False
    G = cpdtt(T, R)
    cpd_list, outputs, best_Z = cpd_cores(G, options.trials, options.epochs, R, options.display, options)
    factors = [G[0] @ cpd_list[0][0]]
    for l in range(L - 2):
        factors.append(cpd_list[l][1])
    factors.append(G[-1].T @ best_Z)
    return equalize(factors, R), outputs


class FinalOutputs:
    """Follows make_final_outputs (Auxiliar.py:223-239)."""

    def __init__(self, num_steps, rel_error, accuracy, outputs, options):
        self.num_steps = num_steps
        self.rel_error = rel_error
        self.accuracy = accuracy
        self.cpd_output = outputs
        self.options = options


def ttcpd(T, R, options, stages=None):
    """method 'ttcpd' of cpd: sort, MLSVD, inflate the core to R where shorter, highcpd, deflate,
    U_l W_l, error.  Follows TensorFox.py:256-365 (dense and sparse input)."""
    sparse_in = type(T) == list
    if sparse_in:
        data_orig = np.array(T[0], dtype=np.float64)
        idxs_orig = np.array(T[1], dtype=np.uint64)
        dims_orig = np.array(T[2])
        tol = options.tol_mlsvd[0] if type(options.tol_mlsvd) == list else options.tol_mlsvd
        Ts, ordering = sparse_sort_dims([data_orig, idxs_orig, dims_orig])
        Tsize = norm(Ts[0])
        if tol < 0:
            options.tol_mlsvd = 1e-16
        S, U, _ = sparse_mlsvd(Ts, R, options)
    else:
        Ts, ordering = sort_dims(np.asarray(T, dtype=np.float64))
        Tsize = norm(Ts)
        S, U, T1, _ = mlsvd(Ts, Tsize, R, options)
    L = S.ndim
    S_orig_dims = S.shape
    inflate_status = R > min(S_orig_dims)
    if inflate_status:
        S = inflate(S, R, S_orig_dims)
    options.trunc_dims = 0
    factors, outputs = highcpd(S, R, options)
    factors = deflate(factors, S_orig_dims, inflate_status)
    factors = [U[l] @ factors[l] for l in range(L)]
    if stages is not None:
        stages.update(S=S, U=U, ordering=ordering)
    if sparse_in:
        factors = unsort_dims(factors, ordering)
        rel_error = sparse_fastnorm(data_orig, idxs_orig, dims_orig, factors) / Tsize
    else:
        rel_error = fastnorm(T1, cpd2unfold1(factors)) / Tsize
        factors = unsort_dims(factors, ordering)
    num_steps = sum(o.num_steps for o in outputs)
    return factors, FinalOutputs(num_steps, rel_error, max(0, 100 * (1 - rel_error)), outputs, options)


def cpd(T, R, options=False, stages=None):
    """Drop-in dense CPD.  Follows TensorFox.py:69-365 (dGN and ALS branches, dense input)."""
    dims = T[2] if type(T) == list else T.shape
    options = make_options(options)
    consistency(R, dims, options)
    if options.method == 'ttcpd':
        return ttcpd(T, R, options, stages=stages)
    if type(T) == list:
        return sparse_tricpd(T, R, options, stages=stages)
    return tricpd(np.asarray(T, dtype=np.float64), R, options, stages=stages)


# ----------------------------------------------------------------------------------------------
# Generators
# ----------------------------------------------------------------------------------------------

def gen_rand_tensor(dims, R, noise=0):
    """Random rank-R tensor with N(0,1) factors and ABSOLUTE noise.  Follows Auxiliar.py:530-572."""
    F = [np.random.randn(d, R) for d in dims]
    T = cpd2tens(F)
    if noise != 0:
        return T, T + noise * np.random.randn(*dims), F
    return T, F


def cpd_tensor(factors):
    """Dense [[F]] in C order via einsum (brute force, used to pin cpd2tens)."""
    L = len(factors)
    letters = 'abcdefghijkl'[:L]
    spec = ','.join(ch + 'z' for ch in letters) + '->' + letters
    return np.einsum(spec, *factors)
