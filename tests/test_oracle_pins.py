"""Pins of the FP64 oracle that do not depend on the reference's code: closed forms, brute-force
einsum, an explicit Jacobian, central finite differences, mathematical properties, the reference's
doctest thresholds (tests/simple_tensor_cpd_options.py, tests/condition_number.py) and recorded
notebook outputs (tutorial/1, 6, 7, 8)."""

import itertools

import numpy as np
import pytest

import oracle.tfx_oracle as o


def _rand_factors(dims, R, seed):
    rs = np.random.RandomState(seed)
    return [rs.randn(d, R) for d in dims]


@pytest.mark.parametrize('dims', [(3, 4, 5), (2, 3, 4, 5)])
def test_unfold_index_map_and_roundtrip(dims):
    T = np.random.RandomState(0).randn(*dims)
    L = len(dims)
    for n in range(L):
        Tn = o.unfold(T, n + 1)
        rest = [m for m in range(L) if m != n]
        for idx in itertools.product(*[range(d) for d in dims]):
            j, stride = 0, 1
            for m in rest:            # first remaining index fastest (Kolda)
                j += idx[m] * stride
                stride *= dims[m]
            assert Tn[idx[n], j] == T[idx]
        assert np.array_equal(o.foldback(Tn, n + 1, dims), T)


@pytest.mark.parametrize('dims', [(3, 4, 5), (2, 3, 4, 3)])
def test_cpd2tens_and_khatri_rao_vs_einsum(dims):
    F = _rand_factors(dims, 4, 1)
    assert np.allclose(o.cpd2tens(F), o.cpd_tensor(F), atol=1e-12)
    A, B = F[0], F[1]
    KR = o.khatri_rao(A, B)
    for i in range(A.shape[0]):
        for k in range(B.shape[0]):
            assert np.allclose(KR[i * B.shape[0] + k], A[i] * B[k])


@pytest.mark.parametrize('dims', [(4, 5, 6), (3, 4, 5, 3)])
def test_mttkrp_vs_einsum(dims):
    F = _rand_factors(dims, 3, 2)
    T = np.random.RandomState(3).randn(*dims)
    L = len(dims)
    letters = 'abcd'[:L]
    for l in range(L):
        spec = letters + ',' + ','.join(letters[m] + 'r' for m in range(L) if m != l) + '->' + letters[l] + 'r'
        ref = np.einsum(spec, T, *[F[m] for m in range(L) if m != l])
        assert np.allclose(o.mttkrp(T, F, l), ref, atol=1e-12)


@pytest.mark.parametrize('dims', [(3, 4, 5), (2, 3, 4, 3)])
def test_gradient_vs_central_differences(dims):
    """f(x) = 1/2 ||T - [[F]]||^2 is quadratic in every single entry, so central differences
    are exact up to roundoff."""
    R = 3
    F = _rand_factors(dims, R, 4)
    T = np.random.RandomState(5).randn(*dims)
    _, P1, _ = o.gramians(F)
    g = o.compute_grad(T, F, P1)

    def f(Fs):
        return 0.5 * np.sum((T - o.cpd_tensor(Fs)) ** 2)

    h = 1e-3
    fd = []
    for l in range(len(dims)):
        for r in range(R):
            for i in range(dims[l]):
                Fp = [x.copy() for x in F]
                Fm = [x.copy() for x in F]
                Fp[l][i, r] += h
                Fm[l][i, r] -= h
                fd.append((f(Fp) - f(Fm)) / (2 * h))
    assert np.allclose(g, np.array(fd), rtol=1e-8, atol=1e-8)


def test_gramians_closed_form():
    F = _rand_factors((3, 4, 5, 6), 3, 6)
    Gr, P1, P2 = o.gramians(F)
    for l in range(4):
        assert np.allclose(Gr[l], F[l].T @ F[l])
        others = [m for m in range(4) if m != l]
        assert np.allclose(P1[l], Gr[others[0]] * Gr[others[1]] * Gr[others[2]])
        for m in others:
            rest = [k for k in others if k != m]
            assert np.allclose(P2[l, m], Gr[rest[0]] * Gr[rest[1]])


@pytest.mark.parametrize('dims', [(3, 4, 5), (2, 3, 2, 3)])
def test_matvec_vs_explicit_jacobian(dims):
    F = _rand_factors(dims, 3, 7)
    J = o.jacobian(F)
    # the Jacobian is the exact derivative of vec([[F]]): check one column by differences
    x = np.concatenate([f.flatten('F') for f in F])
    v = np.random.RandomState(8).randn(x.size)
    assert np.allclose(o.matvec(F, v), J.T @ (J @ v), rtol=1e-10, atol=1e-10)
    h = 1e-6
    Fp = o.x2cpd(x + h * np.eye(x.size)[4], [f.copy() for f in F], eq=False)
    col = (o.cpd_tensor(Fp) - o.cpd_tensor(F)).ravel() / h
    assert np.allclose(col, J[:, 4], atol=1e-5)


def test_regularization_uses_last_mode_max():
    F = _rand_factors((3, 4, 5), 2, 9)
    F[0] *= 10.0            # makes max|P1_l| differ between modes
    _, P1, _ = o.gramians(F)
    Gamma, gamma = o.regularization(P1, [3, 4, 5])
    last = np.abs(P1[2]).max()
    assert last != np.abs(P1).max()
    assert np.allclose(gamma, last * np.sqrt(np.abs(np.array([np.diag(p) for p in P1]))))
    assert np.allclose(Gamma[:3], gamma[0, 0]) and np.allclose(Gamma[3:6], gamma[0, 1])
    M = o.precond(Gamma, gamma, 0.5, [3, 4, 5])
    assert np.allclose(M, 1 / (Gamma * np.sqrt(1 + 0.25)))


def test_cg_single_iteration_is_steepest_descent_and_converges():
    dims, R = (3, 4, 5), 2
    F = _rand_factors(dims, R, 10)
    T = np.random.RandomState(11).randn(*dims)
    _, P1, _ = o.gramians(F)
    Gamma, gamma = o.regularization(P1, list(dims))
    damp = 0.2
    M = o.precond(Gamma, gamma, damp, list(dims))
    g = -o.compute_grad(T, F, P1)
    A = M[:, None] * (o.hessian(F) + damp * np.diag(Gamma)) * M[None, :]
    r0 = M * g
    y1, *_ = o.cg(T, F, damp, 1, 0.0)
    assert np.allclose(y1, M * ((r0 @ r0) / (r0 @ A @ r0)) * r0, rtol=1e-10)
    # n = 24 <= 100 unknowns: tol is zeroed and maxiter is not capped, so many iterations reach
    # the solution of the split-preconditioned system
    y, *_ = o.cg(T, F, damp, 400, 0.0)
    assert np.allclose(y, M * np.linalg.solve(A, r0), rtol=1e-6, atol=1e-9)


def test_dogleg_properties():
    rs = np.random.RandomState(12)
    g = rs.randn(10)
    B = rs.randn(10, 10)
    H = B @ B.T + np.eye(10)
    y = np.linalg.solve(H, g)
    Hg = H @ g
    assert np.array_equal(o.dogleg(y, g, Hg, 10 * np.linalg.norm(y)), y)
    alpha = (g @ g) / (g @ Hg)
    tiny = 0.1 * alpha * np.linalg.norm(g)
    assert np.allclose(o.dogleg(y, g, Hg, tiny), tiny / np.linalg.norm(alpha * g) * alpha * g)
    mid = 0.5 * (alpha * np.linalg.norm(g) + np.linalg.norm(y))
    ydl = o.dogleg(y, g, Hg, mid)
    if alpha * np.linalg.norm(g) < mid < np.linalg.norm(y):
        assert np.linalg.norm(ydl) == pytest.approx(mid, rel=1e-10)


def test_gain_damp_delta_closed_forms():
    x0, x1, g = np.zeros(3), np.array([1.0, 0, 0]), np.array([2.0, 0, 0])
    rho = o.update_gain_ratio(0.5, 0.3, 0.2, 2.0, x0, x1, g)
    assert rho == pytest.approx(4 * (0.09 - 0.04) / (2.0 + 0.5))
    assert o.update_gain_ratio(0.5, 0.3, 0.3, 2.0, x0, x0, g) == -np.inf
    assert o.update_damp(1.0, 1, 0.1, 0) == 3.0
    assert o.update_damp(1.0, 1, 0.9, 0) == pytest.approx(1 / 3)
    assert o.update_damp(1.0, 1, 0.5, 0) == 1.0
    assert o.update_damp(1.0, [7.0, 8.0], 0.9, 1) == 8.0
    assert o.update_delta(1.0, 0.9, 2.0) == 4.0
    assert o.update_delta(1.0, 0.75, 2.0) == pytest.approx(0.875)


def test_equalize_and_x2cpd():
    F = _rand_factors((3, 4, 5), 3, 13)
    T0 = o.cpd_tensor(F)
    E = o.equalize([f.copy() for f in F], 3)
    assert np.allclose(o.cpd_tensor(E), T0)
    norms = np.array([[np.linalg.norm(E[l][:, r]) for r in range(3)] for l in range(3)])
    assert np.allclose(norms, norms[0][None, :])
    x = np.concatenate([f.flatten('F') for f in F])
    G = o.x2cpd(x, [np.zeros_like(f) for f in F], eq=False)
    for a, b in zip(F, G):
        assert np.array_equal(a, b)


def test_randomized_svd_on_exact_low_rank():
    rs = np.random.RandomState(14)
    M = rs.randn(30, 5) @ rs.randn(5, 80)
    np.random.seed(0)
    U, s, Vt = o.randomized_svd(M, 5)
    s_ref = np.linalg.svd(M, compute_uv=False)[:5]
    assert np.allclose(s, s_ref, rtol=1e-12)
    assert np.allclose(U @ np.diag(s) @ Vt, M, atol=1e-10)


def test_clean_compression_closed_form():
    s = np.array([10.0, 1.0, 1e-9, 1e-10])
    U, sig, Vt, dim = o.clean_compression(np.eye(4), s, np.eye(4), 1e-16, 3)
    # sum(s[2:]^2)/sum(s^2) ~ 1.0e-20 < 1e-16/3 but sum(s[1:]^2)/sum(s^2) ~ 9.9e-3 is not
    assert dim == 2 and U.shape == (4, 2)


def test_swimmer_multilinear_rank_and_compression(swimmer):
    """tutorial/8 cell 7: multilinear rank (23, 15, 13); tutorial/7 cell 4: sorted swimmer
    compresses from (256, 32, 32) to (13, 23, 15)."""
    T, ordering = o.sort_dims(swimmer)
    assert T.shape == (256, 32, 32)
    np.random.seed(0)
    S, U, T1, sig = o.mlsvd(T, np.linalg.norm(T), 50, o.make_options(False))
    assert S.shape == (13, 23, 15)
    assert np.linalg.norm(T) == pytest.approx(502.85584415416713, rel=1e-14)
    Tc = o.multilin_mult(U, o.unfold(S, 1), S.shape)
    assert np.linalg.norm(T - Tc) / np.linalg.norm(T) < 1e-9


def _t222():
    T = np.zeros((2, 2, 2))
    s = 0
    for k in range(2):
        for i in range(2):
            for j in range(2):
                T[i, j, k] = s
                s += 1
    return T


def test_doctest_simple_tensor_thresholds():
    """tests/simple_tensor_cpd_options.py:17-52 and tutorial/1 cell 5 (||T|| = sqrt(140))."""
    T = _t222()
    assert np.linalg.norm(T) == pytest.approx(11.832159566199232, rel=1e-15)
    best = 1.0
    for seed in range(10):
        np.random.seed(seed)
        _, out = o.cpd(T, 3)
        best = min(best, out.rel_error)
    assert best < 1e-14

    class opt:
        pass
    opt.initialization = [np.ones((2, 3))] * 3
    _, out = o.cpd(T, 3, opt)
    assert out.rel_error < 0.14
    opt.refine = True
    _, out = o.cpd(T, 3, opt)
    assert out.rel_error < 0.14
    opt.inner_method, opt.cg_maxiter, opt.cg_tol, opt.refine = 'cg_static', 3, 1e-7, False
    _, out = o.cpd(T, 3, opt)
    assert out.rel_error < 0.14


def test_best_rank1_error_known_answer():
    """tutorial/6 cell 10 records 0.13080484499593875 as the rank-1 error of the 2x2x2 tensor.
    Without refinement, R = 1 truncates the MLSVD to a 1x1x1 core, so that number is the error of
    the rank-(1,1,1) sequentially truncated HOSVD, checked here with exact SVDs.  With
    refinement the dGN must reach the global optimum, found independently by HOPM."""
    T = _t222()
    X, U = T.copy(), []
    for l in range(3):                  # sequentially truncated HOSVD with exact SVDs
        u = np.linalg.svd(o.unfold(X, l + 1))[0][:, :1]
        X = np.moveaxis(np.tensordot(u.T, np.moveaxis(X, l, 0), axes=1), 0, l)
        U.append(u)
    trunc = np.linalg.norm(T - np.einsum('abc,ia,jb,kc->ijk', X, *U)) / np.linalg.norm(T)
    assert trunc == pytest.approx(0.13080484499593875, abs=1e-14)
    np.random.seed(0)
    _, out = o.cpd(T, 1)
    assert out.rel_error == pytest.approx(0.13080484499593875, abs=1e-14)
    a, b, c = np.ones(2), np.ones(2), np.ones(2)
    for _ in range(2000):
        a = np.einsum('ijk,j,k->i', T, b, c); a /= np.linalg.norm(a)
        b = np.einsum('ijk,i,k->j', T, a, c); b /= np.linalg.norm(b)
        c = np.einsum('ijk,i,j->k', T, a, b)
    hopm = np.linalg.norm(T - np.einsum('i,j,k->ijk', a, b, c)) / np.linalg.norm(T)
    assert hopm < trunc

    class opt:
        refine = True
    np.random.seed(0)
    _, out = o.cpd(T, 1, opt)
    assert out.rel_error == pytest.approx(hopm, abs=1e-9)


def test_condition_number_doctest_recovery():
    """tests/condition_number.py:7-25 and tutorial/10: a noiseless [20, 30, 40] rank-5 tensor is
    recovered (best of 10 tries far below the doctest's needs)."""
    np.random.seed(0)
    T, F = o.gen_rand_tensor([20, 30, 40], 5)
    best, best_f = 1.0, None
    for _ in range(3):
        f, out = o.cpd(T, 5)
        if out.rel_error < best:
            best, best_f = out.rel_error, f
    assert best < 1e-10
    # congruence of the recovered factors with the true ones (up to permutation and scaling)
    C = np.ones((5, 5))
    for l in range(3):
        A = F[l] / np.linalg.norm(F[l], axis=0)
        B = best_f[l] / np.linalg.norm(best_f[l], axis=0)
        C *= np.abs(A.T @ B)
    assert np.all(C.max(axis=1) > 0.9999)


def test_consistency_messages():
    opt = o.make_options(False)
    with pytest.raises(SystemExit, match='positive integer'):
        o.consistency(2.0, (3, 3, 3), opt)
    with pytest.raises(SystemExit, match='equal to 1'):
        o.consistency(2, (3, 1, 3), opt)
    with pytest.raises(SystemExit, match='min\\(m\\*n, m\\*p, n\\*p\\) = 4'):
        o.consistency(5, (2, 2, 3), opt)
    with pytest.raises(SystemExit, match='greater than 1'):
        o.consistency(1, (3, 3, 3, 3), opt)


def test_make_options_defaults():
    opt = o.make_options(False)
    assert (opt.maxiter, opt.tol, opt.tol_step, opt.tol_improv, opt.tol_grad, opt.tol_jump) == \
        (200, 1e-16, 1e-6, 1e-6, 1e-6, 10)
    assert (opt.method, opt.inner_method, opt.cg_maxiter, opt.cg_factor, opt.cg_tol) == \
        ('dGN', 'cg', 100, 1, 1e-16)
    assert (opt.initialization, opt.mlsvd_method, opt.tol_mlsvd, opt.init_damp, opt.refine) == \
        ('random', 'seq', 1e-16, 1, False)

    class u:
        gpu = True
    assert o.make_options(u).mlsvd_method == 'gpu'


def test_als_sweep_is_sequential_least_squares():
    """Each mode update of an ALS sweep is the least-squares solution min ||T_(l) - F KR_l^T||
    given the modes already updated (lstsq on the explicit Khatri-Rao product)."""
    rs = np.random.RandomState(4)
    dims, R = (5, 6, 7), 3
    T = rs.randn(*dims)
    F = [rs.randn(d, R) for d in dims]
    G = [f.copy() for f in F]
    out = o.als_iteration([o.unfold(T, l + 1) for l in range(3)], F)
    for l in range(3):
        other = [m for m in reversed(range(3)) if m != l]
        KR = o.khatri_rao(G[other[0]], G[other[1]])
        sol = np.linalg.lstsq(KR, o.unfold(T, l + 1).T, rcond=None)[0].T
        assert np.allclose(out[l], sol, atol=1e-11)
        G[l] = sol


def test_als_recovers_exact_low_rank():
    rs = np.random.RandomState(5)
    F = [rs.randn(d, 3) for d in (8, 9, 10)]
    T = o.cpd2tens(F)

    class opt:
        method = 'als'
        maxiter = 400
        tol_step = tol_improv = tol_grad = 0
    np.random.seed(1)
    _, out = o.cpd(T, 3, opt)
    assert out.rel_error < 1e-9


def test_transform_symmetric_shared_array_closed_form():
    """After one symmetric transform without factors_norm the reference's factor list holds L
    references to one array; x2cpd then leaves only the last mode's block in it and transform
    doubles it L - 1 times before dividing by L (Conversion.py:21-49, 412-444)."""
    rs = np.random.RandomState(6)
    L, I, R = 3, 4, 2
    F = o.transform([rs.randn(I, R) for _ in range(L)], True, 0)
    assert all(F[l] is F[0] for l in range(L))
    x = rs.randn(L * I * R)
    last = x[(L - 1) * I * R:].reshape(R, I).T
    G = o.x2cpd(x, F, eq=False)
    G = o.transform(G, True, 0)
    assert np.allclose(G[0], last * 2 ** (L - 1) / L, rtol=1e-15, atol=0)
    assert all(G[l] is G[0] for l in range(L))
    # with factors_norm the arrays are separate, so the symmetric transform is the plain mean
    H = o.transform([rs.randn(I, R) for _ in range(L)], True, 1.5)
    assert all(H[l] is not H[0] for l in range(1, L))
    blocks = [x[l * I * R:(l + 1) * I * R].reshape(R, I).T for l in range(L)]
    H = o.transform(o.x2cpd(x, H, eq=False), True, 1.5)
    mean = sum(blocks) / L
    assert np.allclose(H[1], 1.5 * mean / np.linalg.norm(mean), rtol=1e-14)
