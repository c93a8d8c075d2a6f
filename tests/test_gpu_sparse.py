"""GPU parity of the sparse COO path T = [data, idxs, dims] (TensorFox.py:256-280, 457-640) and of
the 'smart' / 'smart_random' starting points (Initialization.py:97-358), through the C-ABI.

Kernels against the oracle on the same entries: the sparse Gram of every mode (sorted-key pair
accumulation with FP64 atomics: equal to the oracle's SciPy product to 1e-12 relative), the
sparse core (first mode from the entries, the rest dense mode products) and the error over the
stored entries.  End to end against the reference's seeded trajectories
(tests/golden/sparse_init_vectors.npz) with the core fixed to the unfoldings' exact ranks: the
main-stage history until the error falls below 1e-6 (after that these exactly low-rank tensors
converge at a roundoff-decided pace), the final error within 1e-6; convergence with the default
truncation (which keeps a roundoff-decided number of null directions); and the reference's own
doctest thresholds (best of several runs below 1e-4, and
below 1e-3 for orders 3 up to this path's largest order 8 with n = 3, nnz = 3)."""

import numpy as np
import pytest

import oracle.tfx_oracle as o
from tests.golden.cases import make_opts

pytestmark = pytest.mark.gpu


def _gen(z, name):
    meta = z[name + '_meta']
    return z[name + '_data'], z[name + '_idxs'], [int(d) for d in meta[3:]]


def _random_coo(seed, dims, nnz, dup=0):
    rs = np.random.RandomState(seed)
    idx = np.stack([rs.randint(0, d, nnz) for d in dims], axis=1)
    if dup:
        idx = np.concatenate([idx, idx[:dup]], axis=0)     # repeated coordinates
    data = rs.randn(idx.shape[0])
    return data, idx.astype(np.int64), list(dims)


def _packed(data, idxs, dims):
    import TensorFox as tfx
    from TensorFox import sparse as sp
    from TensorFox.auxiliar import sort_dims
    dev = tfx.tensorfox.get_device(0)
    d, i, dm = sp.coo_arrays([data, idxs, dims])
    sdims, ordering = sort_dims(tuple(dm))
    print('sparse pack', dims, 'nnz', len(d), flush=True)
    return dev, sp.SparseTensor(dev, d, i, dm, ordering), ordering


CASES = [('golden', 'g4'), ('golden', 'g3'), ('golden', 'g3b'),
         ('rand', (300, 200, 100, 40000, 500, 1)), ('rand', (64, 64, 8, 9, 3000, 0, 2))]


def _case(sgolden, kind, spec):
    if kind == 'golden':
        return _gen(sgolden, spec)
    if len(spec) == 6:
        a, b, c, nnz, dup, seed = spec
        return _random_coo(seed, (a, b, c), nnz, dup)
    a, b, c, d, nnz, dup, seed = spec
    return _random_coo(seed, (a, b, c, d), nnz, dup)


@pytest.mark.parametrize('kind,spec', CASES)
def test_sparse_gram_matches_oracle(sgolden, kind, spec):
    import torch
    data, idxs, dims = _case(sgolden, kind, spec)
    dev, X, ordering = _packed(data, idxs, dims)
    sidx = idxs[:, ordering]
    for l in range(X.L):
        I = X.dims[l]
        G = dev.empty((I, I), torch.float64)
        dev.sparse_gram(X.nnz, X.dims, X.idx, X.data, l, G)
        ref = o.sparse_gram(data, sidx, X.dims, l + 1)
        got = G.cpu().numpy()
        assert np.abs(got - ref).max() <= 1e-12 * max(1.0, np.abs(ref).max()), l


@pytest.mark.parametrize('kind,spec', [CASES[2], CASES[4]])
def test_sparse_gram_and_first_product_are_reproducible(sgolden, kind, spec):
    """The Gram and the first-mode product sum in a fixed order: two calls give the same bits,
    and the Gram is exactly symmetric.  The truncation after them compares roundoff-sized
    singular values (the golden g3b case keeps a core size that roundoff decides), so anything
    less makes whole runs differ from one call to the next."""
    import torch
    data, idxs, dims = _case(sgolden, kind, spec)
    dev, X, ordering = _packed(data, idxs, dims)
    for l in range(X.L):
        I = X.dims[l]
        G1, G2 = dev.empty((I, I), torch.float64), dev.empty((I, I), torch.float64)
        dev.sparse_gram(X.nnz, X.dims, X.idx, X.data, l, G1)
        dev.sparse_gram(X.nnz, X.dims, X.idx, X.data, l, G2)
        g1 = G1.cpu().numpy()
        assert np.array_equal(g1, G2.cpu().numpy()), l
        assert np.array_equal(g1, g1.T), l
    k = min(X.dims[0], 5)
    U = dev.upload(np.random.RandomState(4).randn(X.dims[0] * k), torch.float64)
    n = k * int(np.prod(X.dims[1:]))
    S1, S2 = dev.empty(n, torch.float64), dev.empty(n, torch.float64)
    dev.sparse_ttm0(X.nnz, X.dims, X.idx, X.data, U, k, S1)
    dev.sparse_ttm0(X.nnz, X.dims, X.idx, X.data, U, k, S2)
    assert np.array_equal(S1.cpu().numpy(), S2.cpu().numpy())


@pytest.mark.parametrize('kind,spec', CASES)
def test_sparse_core_and_residual_match_oracle(sgolden, kind, spec):
    import torch
    data, idxs, dims = _case(sgolden, kind, spec)
    dev, X, ordering = _packed(data, idxs, dims)
    sidx = idxs[:, ordering]
    rs = np.random.RandomState(3)
    ks = [min(d, 5) for d in X.dims]
    U = [np.linalg.qr(rs.randn(d, k))[0] for d, k in zip(X.dims, ks)]
    Ud = [dev.upload(np.asfortranarray(u).ravel(order='F'), torch.float64) for u in U]
    cur = [ks[0]] + X.dims[1:]
    S = dev.empty(int(np.prod(cur)), torch.float64)
    dev.sparse_ttm0(X.nnz, X.dims, X.idx, X.data, Ud[0], ks[0], S)
    for l in range(1, X.L):
        new = list(cur)
        new[l] = ks[l]
        Sn = dev.empty(int(np.prod(new)), torch.float64)
        dev.ttm(S, cur, l, Ud[l], ks[l], Sn, None)
        S, cur = Sn, new
    ref = o.sparse_multilin_mult([u.T for u in U], data, sidx, X.dims)
    got = S.cpu().numpy().reshape(cur)
    assert np.abs(got - ref).max() <= 1e-12 * max(1.0, np.abs(ref).max())
    R = 3
    F = [rs.randn(d, R) for d in X.dims]
    Fd = dev.upload(np.concatenate([f.flatten('F') for f in F]), torch.float64)
    from TensorFox import sparse as sp
    e2 = float(sp.residual_sq(dev, X, Fd, R).cpu().numpy()[0])
    e_ref = o.sparse_fastnorm(data, sidx, X.dims, F)
    assert np.sqrt(e2) == pytest.approx(e_ref, rel=1e-12)


def test_sparse_out_of_range_coordinate_exits_like_reference():
    import TensorFox as tfx
    data = np.ones(3)
    idxs = np.array([[0, 1, 2], [1, 1, 1], [2, 5, 0]])      # 5 >= dims[1]
    with pytest.raises(SystemExit, match='Got too big index'):
        tfx.cpd([data, idxs, [3, 4, 3]], 2)


def _ranks(sgolden, g):
    data, idxs, dims = _gen(sgolden, g)
    order = np.argsort(-np.array(dims))
    sd = [dims[i] for i in order]
    out = []
    for l in range(len(dims)):
        w = np.linalg.eigvalsh(o.sparse_gram(data, idxs[:, order], sd, l + 1))
        out.append(int(np.sum(w > 1e-9 * w.max())))
    return out


@pytest.mark.parametrize('name,g,kw', [('s4r', 'g4', {}), ('s3r', 'g3', {}), ('s3br', 'g3b', {'maxiter': 30})])
def test_sparse_cpd_matches_reference(sgolden, name, g, kw):
    """Core size fixed to the exact rank of every unfolding (trunc_dims): the core is then a
    property of the tensor and the trajectory is the reference's."""
    import TensorFox as tfx
    data, idxs, dims = _gen(sgolden, g)
    R, seed = [int(v) for v in sgolden[name + '_meta']]
    kw = dict(kw, trunc_dims=[int(v) for v in sgolden[name + '_trunc']])
    np.random.seed(seed)
    print('sparse cpd', name, dims, R, kw, flush=True)
    f, out = tfx.cpd([data, idxs, dims], R, make_opts(kw))
    ref = sgolden[name + '_errors']
    e = np.asarray(out.errors[0])
    k = int(np.argmax(ref < 1e-6)) if (ref < 1e-6).any() else ref.size
    assert e.size >= k
    assert np.abs(e[:k] - ref[:k]).max() <= 1e-6
    assert abs(out.rel_error - float(sgolden[name + '_rel_error'][0])) <= 1e-6
    assert out.stop[1] == 8 and out.num_steps == e.size
    T = o.sparse2dense(data, idxs, dims)
    mask = tuple(idxs.astype(np.int64).T)
    direct = np.linalg.norm(T[mask] - o.cpd2tens(f)[mask]) / np.linalg.norm(data)
    assert out.rel_error == pytest.approx(direct, rel=1e-6, abs=1e-13)


@pytest.mark.parametrize('name,g,kw', [('s4', 'g4', {}), ('s3', 'g3', {}), ('s3b', 'g3b', {'maxiter': 30}),
                                       ('s3t', 'g3', {'tol_mlsvd': -1})])
def test_sparse_cpd_default_truncation_converges(sgolden, name, g, kw):
    """Default truncation.  The unfoldings of these tensors have rank below the sketch size and
    rows without entries; the randomized SVD gives the null directions roundoff-sized singular
    values, and how many of them clean_compression keeps is decided by roundoff (the reference
    keeps cores such as 6x5x5x7 for ranks 5x5x5x5; the device's subspace iteration likewise).
    What the input fixes: every mode keeps at least its rank, and the decomposition converges as
    the reference's does (which ends below 1e-6 here)."""
    import TensorFox as tfx
    data, idxs, dims = _gen(sgolden, g)
    R, seed = [int(v) for v in sgolden[name + '_meta']]
    np.random.seed(seed)
    print('sparse cpd', name, dims, R, kw, flush=True)
    f, out = tfx.cpd([data, idxs, dims], R, make_opts(kw))
    assert float(sgolden[name + '_rel_error'][0]) < 1e-6
    assert all(c >= r for c, r in zip(out.compressed_dims, _ranks(sgolden, g)))
    assert out.rel_error < 1e-5
    assert out.stop[1] == 8
    T = o.sparse2dense(data, idxs, dims)
    mask = tuple(idxs.astype(np.int64).T)
    direct = np.linalg.norm(T[mask] - o.cpd2tens(f)[mask]) / np.linalg.norm(data)
    assert out.rel_error == pytest.approx(direct, rel=1e-6, abs=1e-13)


def test_sparse_cpd_matches_oracle_midsize():
    """A larger seeded generator tensor, GPU and oracle from the same seed, the core fixed to the
    unfoldings' exact ranks (the default truncation keeps a roundoff-decided number of null
    directions, see above)."""
    import TensorFox as tfx
    np.random.seed(21)
    data, idxs, dims, _ = tfx.gen_rand_sparse_tensor([60, 50, 40], 4, 30)
    print('sparse cpd 60x50x40 nnz', len(data), flush=True)
    ranks = []
    for l in range(3):
        w = np.linalg.eigvalsh(o.sparse_gram(data, idxs, list(dims), l + 1))
        ranks.append(int(np.sum(w > 1e-9 * w.max())))
    kw = {'trunc_dims': ranks}
    np.random.seed(4)
    fo, oo = o.cpd([data, idxs.astype(np.uint64), list(dims)], 4, make_opts(kw))
    np.random.seed(4)
    fg, og = tfx.cpd([data, idxs, list(dims)], 4, make_opts(kw))
    ref = np.asarray(oo.errors[0])
    k = int(np.argmax(ref < 1e-6)) if (ref < 1e-6).any() else ref.size
    assert np.abs(np.asarray(og.errors[0])[:k] - ref[:k]).max() <= 1e-6
    assert abs(og.rel_error - oo.rel_error) <= 1e-6


def test_sparse_reference_doctest_thresholds():
    """tests/sparse_tensors.py of the reference: best of 10 runs on [10]*4, R=5, nnz=10 below
    1e-4; best of 5 runs per order below 1e-3 for n=3, nnz=3, R=5 (orders 3-8 here)."""
    import TensorFox as tfx
    np.random.seed(0)
    data, idxs, dims, _ = tfx.gen_rand_sparse_tensor([10, 10, 10, 10], 5, 10)

    class options:
        method = 'dGN'
    best = 1.0
    for _ in range(10):
        _, out = tfx.cpd([data, idxs, dims], 5, options)
        best = min(best, out.rel_error)
    assert best < 1e-4
    for L in range(3, 9):
        best = 1.0
        for _ in range(5):
            data, idxs, dims, _ = tfx.gen_rand_sparse_tensor(L * [3], 5, 3)
            print('sparse doctest order', L, 'nnz', len(data), flush=True)
            _, out = tfx.cpd([data, idxs, dims], 5)
            best = min(best, out.rel_error)
        assert best < 1e-3, L


@pytest.mark.parametrize('init', ['smart', 'smart_random'])
@pytest.mark.parametrize('refine', [False, True])
def test_smart_initializations_match_oracle(init, refine):
    """Dense cpd with the MLSVD-based starts: the product reads the core back and selects the
    same entries as the oracle (pinned to the reference on the CPU); both clean-up generators
    start from the same state, so the trajectories are the oracle's."""
    import TensorFox as tfx
    import TensorFox.tensorfox as tf_mod
    T, _ = tfx.make_lowrank((30, 25, 20), 6, 1e-6, seed=8)
    opts = {'initialization': init, 'refine': refine}
    o._CLEAN_RNG = np.random.RandomState(77)
    np.random.seed(3)
    fo, oo = o.cpd(T, 6, make_opts(opts))
    tf_mod._CLEAN_RNG = np.random.RandomState(77)
    np.random.seed(3)
    print('cpd init', init, 'refine', refine, flush=True)
    fg, og = tfx.cpd(T, 6, make_opts(opts))
    e_o, e_g = np.asarray(oo.errors[0]), np.asarray(og.errors[0])
    k = min(15, e_o.size)
    assert np.abs(e_g[:k] - e_o[:k]).max() <= 1e-6
    assert abs(og.rel_error - oo.rel_error) <= 1e-6
