"""Sparse COO input and the 'smart' / 'smart_random' starting points, on the CPU: the oracle's
restatements (and the product's host-side pieces: the generator, the index selection of the
starting points) against the reference's own outputs (tests/golden/make_golden_sparse_init.py)
and against what the mathematics fixes (the dense Gram, dense multilinear product and masked
norm of the same tensor)."""

import numpy as np
import pytest

import oracle.tfx_oracle as o
from tests.golden.cases import make_opts

GENS = ['g4', 'g3', 'g3b']


def _gen(z, name):
    meta = z[name + '_meta']
    return z[name + '_data'], z[name + '_idxs'], [int(d) for d in meta[3:]], int(meta[0]), int(meta[1]), int(meta[2])


@pytest.mark.parametrize('name', GENS)
def test_product_generator_reproduces_reference(sgolden, name):
    """Same draws, same set order, same missing-entry-counts-as-one rule (Auxiliar.py:575-647)."""
    import TensorFox as tfx
    data, idxs, dims, R, nnz, seed = _gen(sgolden, name)
    np.random.seed(seed)
    d2, i2, dims2, _ = tfx.gen_rand_sparse_tensor(dims, R, nnz)
    assert np.array_equal(i2.astype(np.int64), idxs)
    assert np.array_equal(d2, data)
    assert list(dims2) == dims


@pytest.mark.parametrize('name', GENS)
def test_sparse_gram_matches_reference_and_dense(sgolden, name):
    data, idxs, dims, R, nnz, seed = _gen(sgolden, name)
    T = o.sparse2dense(data, idxs, dims)
    for l in range(len(dims)):
        G = o.sparse_gram(data, idxs, dims, l + 1)
        ref = sgolden['%s_gram%d' % (name, l)]
        assert np.abs(G - ref).max() <= 1e-12 * np.abs(ref).max()
        Tl = o.unfold(T, l + 1)
        assert np.abs(G - Tl @ Tl.T).max() <= 1e-12 * np.abs(ref).max()


@pytest.mark.parametrize('name', GENS)
def test_sparse_multilin_mult_and_fastnorm(sgolden, name):
    data, idxs, dims, R, nnz, seed = _gen(sgolden, name)
    L = len(dims)
    U = [sgolden['%s_U%d' % (name, l)] for l in range(L)]
    S = o.sparse_multilin_mult([u.T for u in U], data, idxs, dims)
    ref = sgolden[name + '_core']
    assert np.abs(S - ref).max() <= 1e-12 * np.abs(ref).max()
    T = o.sparse2dense(data, idxs, dims)
    dense = o.multilin_mult([u.T for u in U], o.unfold(T, 1), dims)
    assert np.abs(S - dense).max() <= 1e-12 * np.abs(ref).max()
    F = [sgolden['%s_F%d' % (name, l)] for l in range(L)]
    e = o.sparse_fastnorm(data, idxs, dims, F)
    assert e == pytest.approx(float(sgolden[name + '_fastnorm'][0]), rel=1e-13)
    model = o.cpd2tens(F)
    mask = tuple(idxs.astype(np.int64).T)
    assert e == pytest.approx(float(np.linalg.norm(T[mask] - model[mask])), rel=1e-12)


def test_sparse_unfold_rejects_out_of_range():
    with pytest.raises(ValueError):
        o.sparse_unfold(np.ones(2), np.array([[0, 1], [3, 0]]), (3, 2), 1)


RUNS = [('s4', 'g4', {}), ('s3', 'g3', {}), ('s3b', 'g3b', {'maxiter': 30}), ('s3t', 'g3', {'tol_mlsvd': -1}),
        ('s4r', 'g4', None), ('s3r', 'g3', None), ('s3br', 'g3b', {'maxiter': 30})]


def run_options(sgolden, name, kw):
    kw = dict(kw or {})
    if name + '_trunc' in sgolden.files:
        kw['trunc_dims'] = [int(v) for v in sgolden[name + '_trunc']]
    return kw


@pytest.mark.parametrize('name,g,kw', RUNS)
def test_sparse_cpd_trajectory_matches_reference(sgolden, name, g, kw):
    kw = run_options(sgolden, name, kw)
    data, idxs, dims, _, _, _ = _gen(sgolden, g)
    R, seed = [int(v) for v in sgolden[name + '_meta']]
    np.random.seed(seed)
    f, out = o.cpd([data, idxs.astype(np.uint64), dims], R, make_opts(kw))
    ref = sgolden[name + '_errors']
    e = np.asarray(out.errors[0])
    # these tensors are exactly of low rank: once the error is below 1e-6 the iterates are decided
    # by roundoff (the tail reaches 1e-9 .. 1e-15 in a different number of steps), so the history
    # is compared up to there and the final error to the 1e-6 bar
    k = int(np.argmax(ref < 1e-6)) if (ref < 1e-6).any() else ref.size
    assert e.size >= k
    assert np.abs(e[:k] - ref[:k]).max() <= 1e-8 * max(1.0, float(ref[:k].max()))
    assert abs(out.rel_error - float(sgolden[name + '_rel_error'][0])) <= 1e-6
    assert out.stop[1] == 8 and int(sgolden[name + '_stop'][1]) == 8
    T = o.sparse2dense(data, idxs, dims)
    mask = tuple(idxs.astype(np.int64).T)
    assert out.rel_error == pytest.approx(
        float(np.linalg.norm(T[mask] - o.cpd2tens(f)[mask]) / np.linalg.norm(data)), rel=1e-9, abs=1e-15)


CORES = ['c1', 'c2', 'c3']


@pytest.mark.parametrize('name', CORES)
@pytest.mark.parametrize('R', [2, 3, 5])
def test_smart_init_matches_reference(sgolden, name, R):
    from TensorFox import initialization as ini
    S = sgolden[name]
    L = S.ndim
    ref = [sgolden['%s_smart_R%d_%d' % (name, R, l)] for l in range(L)]
    for F in (o.smart(S, S.shape, R), ini.smart(S, S.shape, R)):
        for l in range(L):
            assert np.array_equal(F[l], ref[l])


@pytest.mark.parametrize('name', CORES)
@pytest.mark.parametrize('R', [2, 3, 5])
def test_smart_random_init_matches_reference(sgolden, name, R):
    """The product's O(R) error of each sample picks the same sample as the reference's full
    |S1 - S1_init|, and the draws follow the global generator in the same order."""
    from TensorFox import initialization as ini
    S = sgolden[name]
    L = S.ndim
    ref = [sgolden['%s_smartrand_R%d_%d' % (name, R, l)] for l in range(L)]
    np.random.seed(R)
    Fo = o.smart_random(S, S.shape, R)
    st_o = np.random.get_state()[1].copy()
    np.random.seed(R)
    Fp = ini.smart_random(S, S.shape, R)
    st_p = np.random.get_state()[1].copy()
    assert np.array_equal(st_o, st_p)
    for l in range(L):
        assert np.array_equal(Fo[l], ref[l])
        assert np.array_equal(Fp[l], ref[l])


def test_smart_random_error_formula_brute_force():
    """|S - sum_r S[c_r] e_{c_r}| from the drawn entries only equals the dense norm, repeats
    included (a coordinate drawn m times contributes m S_c)."""
    rs = np.random.RandomState(0)
    S = rs.randn(4, 3, 5)
    coords = np.array([[0, 0, 0], [1, 2, 3], [0, 0, 0], [3, 1, 4], [0, 0, 0]])
    dense = S.copy()
    for c in coords:
        dense[tuple(c)] -= S[tuple(c)]
    flat = np.ravel_multi_index(tuple(coords.T), S.shape)
    u, m = np.unique(flat, return_counts=True)
    v = S.ravel()[u]
    err2 = (np.sum(S * S) - v @ v) + np.sum(((1 - m) * v) ** 2)
    assert np.sqrt(err2) == pytest.approx(np.linalg.norm(dense), rel=1e-13)


def test_hessian_blocks_equal_explicit_jacobian():
    """The block assembly of GaussNewton.py:666-710 against J^T J from an explicit Jacobian."""
    rs = np.random.RandomState(4)
    for dims, R in [((4, 3, 5), 2), ((3, 4, 2, 3), 3)]:
        F = [rs.randn(d, R) for d in dims]
        _, P1, P2 = o.gramians(F)
        H = o.hessian_blocks(F, P1, P2)
        ref = o.hessian(F)
        assert np.abs(H - ref).max() <= 1e-12 * np.abs(ref).max()


DIRECT = ['d1', 'd2', 'd3']


@pytest.mark.parametrize('name', DIRECT)
def test_direct_inner_solver_trajectory_matches_reference(sgolden, name):
    kw = {'d1': {'inner_method': 'direct', 'maxiter': 40},
          'd2': {'inner_method': 'direct', 'refine': True, 'maxiter': 30},
          'd3': {'inner_method': ['direct', 'cg'] * 15, 'maxiter': 30}}[name]
    T = sgolden[name + '_T']
    R, seed = [int(v) for v in sgolden[name + '_meta']]
    np.random.seed(seed)
    f, out = o.cpd(T, R, make_opts(kw))
    for st in (0, 1):
        ref = sgolden['%s_errors%d' % (name, st)]
        e = np.asarray(out.errors[st])
        # the core errors fall to 1e-9 and below, where the iterates are decided by roundoff
        k = int(np.argmax(ref < 1e-6)) if (ref < 1e-6).any() else ref.size
        assert e.size >= k
        assert np.allclose(e[:k], ref[:k], rtol=1e-8, atol=1e-11)
    assert abs(out.rel_error - float(sgolden[name + '_rel_error'][0])) <= 1e-8


TT = {'tt1': {'method': 'ttcpd'}, 'tt2': {'method': 'ttcpd'}, 'tt3': {'method': 'ttcpd'},
      'tt4': {'method': 'ttcpd', 'epochs': 2}, 'tt5': {'method': 'ttcpd', 'bi_method': 'cg'}}


@pytest.mark.parametrize('name', sorted(TT))
def test_ttcpd_matches_reference(sgolden, name):
    """Tensor-train CPD: the train (scikit-learn's randomized SVD of each reshaped core), one
    tricpd and the fixed-factor bicpd sweeps (ALS by default, dGN with bi_method 'cg'), inflation
    of a core shorter than R (tt3), three epochs (tt4)."""
    T = sgolden[name + '_T']
    R, seed = [int(v) for v in sgolden[name + '_meta']]
    np.random.seed(seed)
    st = {}
    f, out = o.cpd(T, R, make_opts(TT[name]), stages=st)
    assert np.array_equal(st['S'], sgolden[name + '_Sinf'])
    i = 0
    while '%s_G%d' % (name, i) in sgolden.files:
        i += 1
    i = 0
    # tt3: the first core's first tricpd trial stalls (0.024) after a roundoff-decided number of
    # iterations, each of which draws a CG cap, so the next trial starts from other draws; the
    # trials converge to the same point, so only the result is compared there
    while name != 'tt3' and '%s_core%d_errors' % (name, i) in sgolden.files:
        ref = sgolden['%s_core%d_errors' % (name, i)]
        e = np.asarray(out.cpd_output[i].errors[0])
        k = int(np.argmax(ref < 1e-6)) if (ref < 1e-6).any() else ref.size
        assert e.size >= k
        assert np.allclose(e[:k], ref[:k], rtol=1e-7, atol=1e-11), i
        i += 1
    if name == 'tt3':
        # R = 4 above the dims 3: the train of the inflated core does not fit T (the reference's
        # own docstring limits highcpd to R below every dimension) and its result follows the
        # shifted draws; both end near 0.18
        assert abs(out.rel_error - float(sgolden[name + '_rel_error'][0])) <= 1e-3
        return
    assert i == len(out.cpd_output)
    assert abs(out.rel_error - float(sgolden[name + '_rel_error'][0])) <= 1e-7
    ref_model = o.cpd2tens([sgolden['%s_f%d' % (name, l)] for l in range(T.ndim)])
    assert np.linalg.norm(o.cpd2tens(f) - ref_model) <= 1e-6 * np.linalg.norm(ref_model)
