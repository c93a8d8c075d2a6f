"""GPU parity of the direct inner solver (inner_method 'direct', GaussNewton.py:630-710) and the
rank / stats / foxit drivers (TensorFox.py:820-1188).

tfx_dgn_direct against the oracle's solve of (H + damp diag(Gamma)) y = grad (the Gauss-Newton
matrix assembled by the reference's block formula, itself pinned to J^T J of an explicit
Jacobian on the CPU), to 1e-9 relative; end to end against the reference's seeded trajectories
(tests/golden/sparse_init_vectors.npz: direct alone, with refinement, alternated with cg), the
history compared while the error is above 1e-6 (below, these small cores converge at a
roundoff-decided pace), the final error to 1e-6."""

import os

import numpy as np
import pytest

import oracle.tfx_oracle as o
from tests.golden.cases import make_opts

pytestmark = pytest.mark.gpu
os.environ.setdefault('MPLBACKEND', 'Agg')


@pytest.mark.parametrize('dims,R,damp', [((5, 4, 3), 2, 0.3), ((12, 10, 8), 5, 1e-3), ((6, 5, 4, 3), 3, 2.0),
                                         ((40, 30, 20), 10, 0.05), ((70, 60, 50), 11, 0.01)])
def test_direct_solve_matches_oracle(dims, R, damp):
    import torch
    import TensorFox as tfx
    dev = tfx.tensorfox.get_device(0)
    rs = np.random.RandomState(len(dims) * 7 + R)
    F = [rs.randn(d, R) for d in dims]
    T = o.cpd2tens([f + 0.1 * rs.randn(*f.shape) for f in F])
    print('direct solve', dims, R, damp, flush=True)
    _, P1, P2 = o.gramians(F)
    Gamma, gamma = o.regularization(P1, dims)
    grad = -o.compute_grad(T, F, P1)
    H = o.hessian_blocks(F, P1, P2)
    y_ref = np.linalg.solve(H + damp * np.diag(Gamma), grad)
    n = R * sum(dims)
    Fd = dev.upload(np.concatenate([f.flatten('F') for f in F]), torch.float64)
    Gr, P1d, P2d, gd = (dev.empty((len(dims), R, R), torch.float64), dev.empty((len(dims), R, R), torch.float64),
                        dev.empty((len(dims), len(dims), R, R), torch.float64), dev.empty((len(dims), R), torch.float64))
    dev.gramians(dims, R, Fd, Gr, P1d, P2d, gd)
    g = dev.upload(grad, torch.float64)
    y = dev.empty(n, torch.float64)
    scal = dev.zeros(2, torch.float64)
    dev.dgn_direct(dims, R, Fd, P1d, P2d, gd, damp, g, y, scal)
    yg = y.cpu().numpy()
    assert np.abs(yg - y_ref).max() <= 1e-9 * np.abs(y_ref).max()
    rn = np.linalg.norm(H @ yg - grad)
    assert np.sqrt(scal.cpu().numpy()[1]) == pytest.approx(rn, rel=1e-6, abs=1e-12 * np.linalg.norm(grad))
    # the reference's QR route on the row-scaled system gives the same step
    y_qr, _, _, _ = o.direct(T, F, damp)
    assert np.abs(yg - y_qr).max() <= 1e-8 * np.abs(y_qr).max()


def test_direct_rejects_oversized_system():
    import torch
    import TensorFox as tfx
    from TensorFox._lib import TfxError
    dev = tfx.tensorfox.get_device(0)
    dims, R = (900, 800, 700), 2          # n = 4800 > 4096
    n = R * sum(dims)
    z = dev.zeros(n, torch.float64)
    small = dev.zeros(3 * 3 * R * R, torch.float64)
    with pytest.raises(TfxError, match='direct solver limited'):
        dev.dgn_direct(dims, R, z, small, small, small, 1.0, z, z, dev.zeros(2, torch.float64))


DIRECT = {'d1': {'inner_method': 'direct', 'maxiter': 40},
          'd2': {'inner_method': 'direct', 'refine': True, 'maxiter': 30},
          'd3': {'inner_method': ['direct', 'cg'] * 15, 'maxiter': 30}}


@pytest.mark.parametrize('name', sorted(DIRECT))
def test_direct_cpd_matches_reference(sgolden, name):
    import TensorFox as tfx
    T = sgolden[name + '_T']
    R, seed = [int(v) for v in sgolden[name + '_meta']]
    np.random.seed(seed)
    print('cpd direct', name, T.shape, R, flush=True)
    f, out = tfx.cpd(T, R, make_opts(dict(DIRECT[name], precision='fp64')))
    for st in (0, 1):
        ref = sgolden['%s_errors%d' % (name, st)]
        e = np.asarray(out.errors[st])
        k = int(np.argmax(ref < 1e-6)) if (ref < 1e-6).any() else ref.size
        assert e.size >= k
        if k:
            assert np.abs(e[:k] - ref[:k]).max() <= 1e-6
    assert abs(out.rel_error - float(sgolden[name + '_rel_error'][0])) <= 1e-6


def test_rank_driver_finds_rank():
    import TensorFox as tfx
    T, _ = tfx.make_lowrank((8, 7, 6), 3, 0.0, seed=1)
    np.random.seed(0)
    print('rank driver', flush=True)
    r, errs = tfx.rank(T, plot=False, trials=2)
    # ranks 1 and 2 leave a large error, rank 3 fits; the search stops at 3, or one rank later
    # when the rank-3 error is above 1e-14 (then the smaller of the two errors decides)
    assert errs[0] > 1e-2 and errs[1] > 1e-3 and errs[2] < 1e-6
    assert r in (3, 4) and errs.size in (3, 4, 5)


def test_stats_and_foxit_drivers():
    import TensorFox as tfx
    T, _ = tfx.make_lowrank((9, 8, 7), 2, 1e-6, seed=2)
    np.random.seed(1)
    print('stats driver', flush=True)
    times, steps, errors = tfx.stats(T, 2, num_samples=3)
    assert times.shape == steps.shape == errors.shape == (3,)
    assert (times > 0).all() and (errors < 1e-4).all()
    for refine in (False, True):
        print('foxit refine', refine, flush=True)

        class options:
            pass
        options.refine = refine
        f, out = tfx.foxit(T, 2, options, bestof=2)
        assert out.rel_error < 1e-4 and len(f) == 3
