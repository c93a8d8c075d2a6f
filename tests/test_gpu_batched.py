"""Batched CPDs (one launch per stage for the whole batch) against the oracle run member by
member with the same seed, and against the single-problem GPU path.

Parity criterion per member, as for `cpd`: final relative error within 1e-6 absolute of the
oracle's, and factors congruent (> 0.999) when the oracle's run converged.  The batched MLSVD
computes exact eigenvectors (full Jacobi) where the oracle runs the reference's randomized SVD,
so trajectories are not compared iteration by iteration."""

import numpy as np
import pytest

import oracle.tfx_oracle as o
from tests.golden.cases import make_opts
from tests.test_gpu_cpd import congruence

pytestmark = pytest.mark.gpu


def _batch(n, lo, hi, rlo, rhi, L, nu, seed, dtype=np.float64):
    rs = np.random.RandomState(seed)
    tensors, ranks, truth = [], [], []
    for m in range(n):
        dims = tuple(int(v) for v in rs.randint(lo, hi + 1, L))
        R = int(rs.randint(rlo, rhi + 1))
        F = [rs.standard_normal((d, R)) for d in dims]
        T0 = o.cpd_tensor(F)
        E = rs.standard_normal(dims)
        T = T0 + nu * np.linalg.norm(T0) * E / np.linalg.norm(E)
        tensors.append(T.astype(dtype))
        ranks.append(R)
        truth.append(F)
    return tensors, ranks, truth


@pytest.mark.parametrize('L,dtype', [(3, np.float64), (3, np.float32), (2, np.float64), (4, np.float64)])
def test_batched_matches_oracle(L, dtype):
    import TensorFox as tfx
    print('batched vs oracle L=%d %s' % (L, np.dtype(dtype).name), flush=True)
    lo, hi = (8, 16) if L < 4 else (5, 8)
    tensors, ranks, truth = _batch(10, lo, hi, 2, 5, L, 1e-4, 100 + L, dtype)
    kw = dict(maxiter=60)
    seeds = list(range(1000, 1000 + len(tensors)))
    res = tfx.cpd_batched(tensors, ranks, make_opts(kw), seeds)
    for m, (T, R) in enumerate(zip(tensors, ranks)):
        np.random.seed(seeds[m])
        fo, oo = o.cpd(np.asarray(T, dtype=np.float64), R, make_opts(kw))
        fg, og = res[m]
        assert abs(og.rel_error - oo.rel_error) < 1e-6, (m, T.shape, R, og.rel_error, oo.rel_error)
        if oo.rel_error < 1e-3:
            assert congruence(fg, fo) > 0.999, m
        assert og.stop[1] == 8 and og.num_steps == len(og.errors[0])


def test_batched_matches_single_gpu_path():
    """Same seeds through tfx.cpd one problem at a time: same compressed dims, errors within
    1e-6 (the single path finds its MLSVD bases by subspace iteration)."""
    import TensorFox as tfx
    print('batched vs single', flush=True)
    tensors, ranks, _ = _batch(6, 8, 20, 2, 6, 3, 1e-4, 7)
    seeds = [5, 6, 7, 8, 9, 10]
    res = tfx.cpd_batched(tensors, ranks, make_opts({}), seeds)
    for m, (T, R) in enumerate(zip(tensors, ranks)):
        np.random.seed(seeds[m])
        _, os_ = tfx.cpd(T, R)
        fg, og = res[m]
        assert og.compressed_dims == os_.compressed_dims
        assert abs(og.rel_error - os_.rel_error) < 1e-6


def test_batched_graph_mode_equals_persistent():
    """One captured launch per dGN iteration (replayed as a HIP graph) gives the same results
    as the single persistent launch: the state carried between launches is exact."""
    import TensorFox as tfx
    print('batched graph vs persistent', flush=True)
    tensors, ranks, _ = _batch(16, 8, 24, 2, 8, 3, 1e-4, 21, np.float32)
    seeds = list(range(16))
    kw = make_opts(dict(maxiter=40))
    a = tfx.cpd_batched(tensors, ranks, kw, seeds, config={'mode': 'persistent', 'order': 'size'})
    b = tfx.cpd_batched(tensors, ranks, kw, seeds, config={'mode': 'graph', 'order': 'natural'})
    for (fa, oa), (fb, ob) in zip(a, b):
        assert oa.rel_error == ob.rel_error
        assert oa.num_steps == ob.num_steps
        assert all(np.array_equal(x, y) for x, y in zip(fa, fb))


def test_batched_small_problems_zero_tolerances():
    """Members with R * sum(core dims) <= 100 zero every tolerance (GaussNewton.py:117-121), so
    they can only stop on the averaged criteria (4, 7), divergence (6) or maxiter (5)."""
    import TensorFox as tfx
    print('batched small', flush=True)
    tensors, ranks, _ = _batch(4, 8, 10, 2, 2, 3, 1e-4, 3)
    seeds = [1, 2, 3, 4]
    kw = dict(maxiter=25)
    res = tfx.cpd_batched(tensors, ranks, make_opts(kw), seeds)
    for m, (f, out) in enumerate(res):
        assert out.stop[0] in (4, 5, 6, 7)
        np.random.seed(seeds[m])
        _, oo = o.cpd(np.asarray(tensors[m]), ranks[m], make_opts(kw))
        assert abs(out.rel_error - oo.rel_error) < 1e-6


def test_batched_rejects_unsupported():
    import TensorFox as tfx
    T = np.random.RandomState(0).randn(70, 8, 8)
    with pytest.raises(NotImplementedError):
        tfx.cpd_batched([T], [3])
    with pytest.raises(NotImplementedError):
        tfx.cpd_batched([T[:10]], [3], make_opts(dict(refine=True)))


def test_batched_4096_members_sampled_against_oracle():
    """The full batch: 4096 FP32 tensors, dims U{8..64}, rank U{2..16}, noise 1e-4, member m
    seeded with m.  Size-independent checks over every member (finite errors, histories of the
    reported length, stop codes 0-7) and oracle parity on sampled members, the largest included."""
    import TensorFox as tfx
    from bench import make_batch
    print('batched 4096', flush=True)
    tensors, ranks = make_batch(4096)
    seeds = list(range(4096))
    res = tfx.cpd_batched(tensors, ranks, None, seeds)
    errs = np.array([out.rel_error for _, out in res])
    assert np.all(np.isfinite(errs))
    assert np.median(errs) < 2e-4
    for f, out in res:
        assert out.stop[0] in range(8) and len(out.errors[0]) == out.num_steps
    sizes = [np.prod(T.shape) * R for T, R in zip(tensors, ranks)]
    sample = sorted(set([int(np.argmax(sizes))] + list(range(0, 4096, 683))))
    for m in sample:
        np.random.seed(seeds[m])
        fo, oo = o.cpd(tensors[m].astype(np.float64), ranks[m])
        assert abs(res[m][1].rel_error - oo.rel_error) < 1e-6, (m, tensors[m].shape, ranks[m])
