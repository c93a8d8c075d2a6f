"""End-to-end parity of the GPU `cpd` with the FP64 oracle (which reproduces the reference's
seeded trajectories, see test_oracle_golden.py) and with the reference's doctest thresholds.

Parity criterion (per problem, same seed for both sides): final relative error within 1e-6
absolute of the oracle's, and factors equal up to permutation and scaling with congruence above
0.999.  The GPU path computes the MLSVD bases by subspace iteration on Gram matrices instead of a
randomized SVD, so per-iteration trajectories are not compared."""

import numpy as np
import pytest
from scipy.optimize import linear_sum_assignment

import oracle.tfx_oracle as o
from tests.golden.cases import make_opts

pytestmark = pytest.mark.gpu


def congruence(A, B):
    """Min over matched rank-one terms of prod_l |cos(a_l^r, b_l^s)| after the best matching."""
    R = A[0].shape[1]
    C = np.ones((R, R))
    for a, b in zip(A, B):
        an = a / np.linalg.norm(a, axis=0)
        bn = b / np.linalg.norm(b, axis=0)
        C *= np.abs(an.T @ bn)
    r, c = linear_sum_assignment(-C)
    return C[r, c].min()


def _both(T, R, kw, seed, precision=None):
    import TensorFox as tfx
    np.random.seed(seed)
    fo, oo = o.cpd(np.asarray(T, dtype=np.float64), R, make_opts(kw))
    kw2 = dict(kw)
    if precision:
        kw2['precision'] = precision
    np.random.seed(seed)
    fg, og = tfx.cpd(T, R, make_opts(kw2))
    return fo, oo, fg, og


WELL_POSED = [
    ('c567_refine', dict(refine=True), 5),
    ('cbig_refine', dict(refine=True), 6),
    ('c4_refine', dict(refine=True, maxiter=60), 8),
]


@pytest.mark.parametrize('name,kw,seed', WELL_POSED)
def test_golden_problems_fp64(golden, name, kw, seed):
    print('cpd fp64', name, flush=True)
    T = golden[name + '_T']
    R = int(golden[name + '_R'])
    fo, oo, fg, og = _both(T, R, kw, seed)
    assert abs(og.rel_error - float(golden[name + '_rel_error'])) < 1e-6
    assert abs(og.rel_error - oo.rel_error) < 1e-6
    assert congruence(fg, fo) > 0.999
    assert og.stop[1] in range(8)


SEEDED = [('fp64', (100, 100, 100), 10, 1e-6, 0), ('fp64', (30, 40, 50), 6, 1e-4, 1),
          ('fp64', (24, 20, 16, 12), 5, 1e-5, 2),
          # the mixed form is specified for FP32 tensors with noise >= 1e-5 relative
          ('mixed', (100, 100, 100), 10, 1e-5, 0), ('mixed', (30, 40, 50), 6, 1e-4, 1),
          ('mixed', (24, 20, 16, 12), 5, 1e-5, 2)]


@pytest.mark.parametrize('precision,dims,R,nu,seed', SEEDED)
def test_seeded_lowrank(dims, R, nu, seed, precision):
    import TensorFox as tfx
    print('cpd', dims, R, precision, flush=True)
    T, Ftrue = tfx.make_lowrank(dims, R, nu, seed=seed)
    if precision == 'mixed':
        T = T.astype(np.float32)
    fo, oo, fg, og = _both(T, R, dict(refine=True), seed, precision)
    assert abs(og.rel_error - oo.rel_error) < 1e-6, (og.rel_error, oo.rel_error)
    assert abs(og.final_rel_error - oo.rel_error) < 1e-6
    assert congruence(fg, fo) > 0.999
    assert congruence(fg, Ftrue) > 0.999


def _t222():
    T = np.zeros((2, 2, 2))
    s = 0
    for k in range(2):
        for i in range(2):
            for j in range(2):
                T[i, j, k] = s
                s += 1
    return T


def test_doctest_simple_tensor_thresholds_gpu():
    """tests/simple_tensor_cpd_options.py:17-52 on the GPU path (FP64)."""
    import TensorFox as tfx
    T = _t222()
    best = 1.0
    for seed in range(10):
        np.random.seed(seed)
        _, out = tfx.cpd(T, 3)
        best = min(best, out.rel_error)
    assert best < 1e-14

    class opt:
        pass
    opt.initialization = [np.ones((2, 3))] * 3
    _, out = tfx.cpd(T, 3, opt)
    assert out.rel_error < 0.14
    opt.refine = True
    _, out = tfx.cpd(T, 3, opt)
    assert out.rel_error < 0.14
    opt.inner_method, opt.cg_maxiter, opt.cg_tol, opt.refine = 'cg_static', 3, 1e-7, False
    _, out = tfx.cpd(T, 3, opt)
    assert out.rel_error < 0.14


def test_rank1_known_answer_gpu():
    """tutorial/6 cell 10: 0.13080484499593875 (rank-1, no refinement)."""
    import TensorFox as tfx
    np.random.seed(0)
    _, out = tfx.cpd(_t222(), 1)
    assert out.rel_error == pytest.approx(0.13080484499593875, abs=1e-12)


def test_swimmer_compression_gpu(swimmer):
    """tutorial/7 cell 4: the sorted swimmer compresses from (256, 32, 32) to (13, 23, 15)."""
    import TensorFox as tfx

    class opt:
        maxiter = 3
    np.random.seed(0)
    _, out = tfx.cpd(swimmer, 50, opt)
    assert out.compressed_dims == (13, 23, 15)


def test_scheduled_damping_trajectory_matches_oracle(golden):
    """With a scheduled damping list and inner-method list (GaussNewton.py:110-111, 262-263,
    495-496) the control flow does not depend on roundoff, so the GPU FP64 trajectory follows the
    oracle's (which reproduces the reference's) iteration by iteration."""
    import TensorFox as tfx
    name, kw, seed = ('cjump_sched', dict(init_damp=[0.5] * 25, maxiter=25,
                                          inner_method=['cg'] * 10 + ['cg_static'] * 15), 10)
    T = golden[name + '_T']
    np.random.seed(seed)
    _, og = tfx.cpd(T, 3, make_opts(kw))
    assert og.num_steps == int(golden[name + '_num_steps'])
    assert np.allclose(og.errors[0], golden[name + '_errors0'], rtol=1e-9, atol=1e-14)
    assert og.rel_error == pytest.approx(float(golden[name + '_rel_error']), rel=1e-9)


def test_dogleg_path_runs():
    """tol_jump < 1 forces the dogleg fallback (GaussNewton.py:303-307, 713-763).  After a dogleg
    with no improving trial the reference continues from x - y with the rejected factors, and the
    next gain ratio divides by a roundoff-sized x - old_x, so the trajectory from there on is
    decided by summation order; only the mechanics are checked here (the dogleg step itself is
    checked against the oracle in test_clib_cpu.py)."""
    import TensorFox as tfx
    T, _ = tfx.make_lowrank((6, 5, 7), 3, 1e-2, seed=24)
    np.random.seed(9)
    _, og = tfx.cpd(T, 3, make_opts(dict(tol_jump=0.9, maxiter=30)))
    assert og.dogleg_calls > 0
    assert np.isfinite(og.rel_error) and og.rel_error < 1.0


@pytest.mark.parametrize('precision', ['fp64', 'mixed'])
def test_order4_collinear_truncated_vs_oracle(precision):
    """Order-4 problem with a collinear pair (cos 0.9) and trunc_dims, against the oracle."""
    import TensorFox as tfx
    print('cpd order4 collinear', precision, flush=True)
    dims, R = (20, 18, 16, 14), 5
    T, Ftrue = tfx.make_lowrank(dims, R, 1e-5, seed=3, collinear=0.9)
    if precision == 'mixed':
        T = T.astype(np.float32)
    fo, oo, fg, og = _both(T, R, dict(trunc_dims=[R] * 4), 3, precision)
    assert og.compressed_dims == (R,) * 4
    assert abs(og.rel_error - oo.rel_error) < 1e-6, (og.rel_error, oo.rel_error)
    assert congruence(fg, fo) > 0.999


def _fp64_rel_residual(T, F):
    """||T - [[F]]|| / ||T|| on the host in FP64, for an order-4 tensor (row blocks of mode 1)."""
    A, B, C, D = F
    CD = (C[:, None, :] * D[None, :, :]).reshape(-1, A.shape[1])
    T2 = T.reshape(A.shape[0] * B.shape[0], -1)
    num, den = 0.0, 0.0
    for i in range(A.shape[0]):
        AB = A[i][None, :] * B
        blk = T2[i * B.shape[0]:(i + 1) * B.shape[0]].astype(np.float64)
        num += float(np.sum((blk - AB @ CD.T) ** 2))
        den += float(np.sum(blk ** 2))
    return np.sqrt(num / den)


def test_order4_96_rank24_collinear_mixed():
    """96^4 FP32 tensor, rank 24, first two columns of the first factor with cosine 0.9, relative
    noise 1e-5; trunc_dims [24]*4, dGN on the 24^4 core in mixed precision, FP64 residual on the
    full tensor.  At this size the oracle is not run: the reported error must equal an FP64 host
    recomputation from the returned factors, reach the noise floor, agree with the FP64 GPU form,
    and the factors must match the generating ones."""
    import TensorFox as tfx
    print('cpd 96^4 R=24 collinear mixed', flush=True)
    dims, R, nu = (96, 96, 96, 96), 24, 1e-5
    T, Ftrue = tfx.make_lowrank(dims, R, nu, seed=0, collinear=0.9, dtype=np.float32)
    res = {}
    for precision in ('mixed', 'fp64'):
        np.random.seed(0)
        f, out = tfx.cpd(T, R, make_opts(dict(trunc_dims=[24] * 4, precision=precision)))
        res[precision] = (f, out)
    f, out = res['mixed']
    assert out.compressed_dims == (24,) * 4
    host = _fp64_rel_residual(T, f)
    assert abs(out.rel_error - host) < 1e-9 * max(1.0, host / 1e-5), (out.rel_error, host)
    assert out.rel_error < 2 * nu
    assert abs(out.rel_error - res['fp64'][1].rel_error) < 1e-6
    assert congruence(f, Ftrue) > 0.999
