"""GPU parity for the option paths whose reference trajectories are stored in
tests/golden/reference_vectors.npz: symmetric factors (shared-array and factors_norm forms,
Conversion.py:412-444), method 'als' (Alternating_Least_Squares.py:18-264), inner_method 'als'
(GaussNewton.py:273-276), mlsvd_method 'classic' (Compression.py:159-170) and the dogleg path
(GaussNewton.py:713-795).

Bar: final relative error within 1e-6 absolute of the reference's, congruence with the
reference's factors above 0.999 where the problem is well posed, and the main-stage error history
equal to the reference's over the iterations that roundoff does not decide (tolerance written in
each test).  The GPU MLSVD takes its bases from Gram eigenpairs instead of a randomized SVD, so
histories agree to the accuracy of the compressed core, not bitwise."""

import numpy as np
import pytest

import oracle.tfx_oracle as o
from tests.golden.cases import CASE_OPTIONS, make_opts
from tests.test_gpu_cpd import congruence

pytestmark = pytest.mark.gpu


def _run(golden, name, **extra):
    import TensorFox as tfx
    print('cpd golden', name, extra, flush=True)
    T = golden[name + '_T']
    R = int(golden[name + '_R'])
    kw = dict(CASE_OPTIONS[name])
    kw.update(extra)
    np.random.seed(int(golden[name + '_seed']))
    return tfx.cpd(T, R, make_opts(kw))


def _ref_factors(golden, name):
    L = golden[name + '_T'].ndim
    return [golden['%s_F%d' % (name, l)] for l in range(L)]


@pytest.mark.parametrize('name,k', [('csymm', 40), ('csymm_refine', 30), ('csymm_norm', 17)])
def test_symmetric_factors_match_reference(golden, name, k):
    """symm=True: after the first symmetric transform the reference's factor list shares one
    array, so each later update keeps only the last mode's block, doubled L - 1 times and divided
    by L; with factors_norm > 0 the arrays stay separate and the transform is the mean.  The whole
    main-stage history (40, 30 and 17 iterations) is compared."""
    f, out = _run(golden, name)
    e_ref = golden[name + '_errors0']
    assert out.errors[0].shape == e_ref.shape
    assert np.allclose(out.errors[0][:k], e_ref[:k], rtol=1e-7, atol=1e-12)
    assert abs(out.rel_error - float(golden[name + '_rel_error'])) < 1e-6
    ref = _ref_factors(golden, name)
    for l in range(3):
        assert np.abs(f[l] - ref[l]).max() <= 1e-5 * np.abs(ref[l]).max()
    assert congruence(f, ref) > 0.999


@pytest.mark.parametrize('name', ['cals', 'cals_refine', 'c4_als'])
def test_als_matches_reference(golden, name):
    """method='als': sweeps of F_l <- N_l pinv(P1_l) on the core, then (refine) on the full
    tensor; the reference returns the last, not the best, factors."""
    f, out = _run(golden, name)
    ref_err = float(golden[name + '_rel_error'])
    assert abs(out.rel_error - ref_err) < 1e-6, (out.rel_error, ref_err)
    assert list(out.stop) == [int(s) for s in golden[name + '_stop']]
    e_ref = golden[name + '_errors0']
    k = min(10, e_ref.size)
    assert np.allclose(out.errors[0][:k], e_ref[:k], rtol=1e-6, atol=1e-12)
    if ref_err < 1e-3:
        assert congruence(f, _ref_factors(golden, name)) > 0.999


def test_inner_als_matches_reference(golden):
    """inner_method list with 'als' sweeps between CG steps (GaussNewton.py:273-276): no CG cap
    draw, the gradient entry is nan and the damping is left as it is on those iterations."""
    name = 'cinner_als'
    f, out = _run(golden, name)
    e_ref = golden[name + '_errors0']
    assert out.errors[0].shape == e_ref.shape
    assert np.allclose(out.errors[0][:10], e_ref[:10], rtol=1e-6, atol=1e-12)
    g = out.gradients[0]
    assert np.isnan(g[4:10]).all() and np.isfinite(g[:4]).all()
    assert abs(out.rel_error - float(golden[name + '_rel_error'])) < 1e-6
    assert congruence(f, _ref_factors(golden, name)) > 0.999


def test_als_sweep_kernel_matches_oracle():
    """One GPU ALS sweep (gramians + single-mode MTTKRP + tfx_als_solve per mode) against the
    oracle's als_iteration (a restatement pinned by the reference's own sweep vectors), FP64,
    orders 3 and 4."""
    import torch
    from TensorFox.als import als_sweep
    from TensorFox.device import get_device
    from TensorFox.gauss_newton import Problem
    dev = get_device(0)
    for dims, R in (((7, 6, 5), 3), ((6, 5, 4, 3), 4), ((40, 30, 20), 12)):
        print('als sweep', dims, R, flush=True)
        rs = np.random.RandomState(sum(dims))
        T = rs.randn(*dims)
        F = [rs.randn(d, R) for d in dims]
        ref = o.als_iteration([o.unfold(T, l + 1) for l in range(len(dims))], [f.copy() for f in F])
        prob = Problem(dev, torch.from_numpy(T).to(dev.dev).view(-1), dims, R, torch.float64, 0)
        Fd = dev.upload(np.concatenate([f.flatten('F') for f in F]), torch.float64)
        als_sweep(prob, Fd)
        got = Fd.cpu().numpy()
        off = 0
        for l, d in enumerate(dims):
            g = got[off:off + R * d].reshape(R, d).T
            off += R * d
            assert np.abs(g - ref[l]).max() <= 1e-10 * np.abs(ref[l]).max(), (dims, l)
        assert int(prob.als_status.cpu()[0]) == 0


def test_als_singular_normal_matrix_takes_pseudo_inverse():
    """Two identical rank-one terms make every P1_l singular: the reference's pinv gives the
    minimum-norm solution, which splits the weight equally between the two columns."""
    import torch
    from TensorFox.device import get_device
    from TensorFox.gauss_newton import Problem
    dev = get_device(0)
    dims, R = (5, 4, 3), 2
    rs = np.random.RandomState(3)
    a = [rs.randn(d) for d in dims]
    T = np.einsum('i,j,k->ijk', *a) + 0.01 * rs.randn(*dims)
    F = [np.stack([v, v], axis=1) for v in (rs.randn(d) for d in dims)]
    prob = Problem(dev, torch.from_numpy(T).to(dev.dev).view(-1), dims, R, torch.float64, 0)
    Fd = dev.upload(np.concatenate([f.flatten('F') for f in F]), torch.float64)
    dev.gramians(dims, R, Fd, prob.Gr, prob.P1, prob.P2, prob.gamma)
    dev.mttkrp(prob.T, dims, Fd, R, prob.N, mask=1)
    dev.als_solve(dims, R, 0, prob.P1, prob.N, Fd, prob.als_status)
    got = Fd.cpu().numpy()[:R * dims[0]].reshape(R, dims[0]).T
    KR = o.khatri_rao(F[2], F[1])
    ref = o.unfold(T, 1) @ np.linalg.pinv(KR.T)
    assert int(prob.als_status.cpu()[0]) == 1
    assert np.abs(got - ref).max() <= 1e-9 * np.abs(ref).max()


def test_classic_mlsvd_matches_reference(golden):
    """mlsvd_method='classic': HOSVD of every unfolding of T, then the multilinear product with
    the bases (Compression.py:159-170, MultilinearAlgebra.py:72-105)."""
    name = 'cbig_classic'
    f, out = _run(golden, name)
    assert abs(out.rel_error - float(golden[name + '_rel_error'])) < 1e-6
    assert out.compressed_dims == (4, 4, 4)
    assert congruence(f, _ref_factors(golden, name)) > 0.999


@pytest.mark.parametrize('precision', ['fp64', 'mixed'])
def test_classic_mlsvd_seeded_vs_oracle(precision):
    import TensorFox as tfx
    print('cpd classic', precision, flush=True)
    dims, R = (30, 24, 20), 5
    T, Ftrue = tfx.make_lowrank(dims, R, 1e-5, seed=7)
    if precision == 'mixed':
        T = T.astype(np.float32)
    np.random.seed(7)
    fo, oo = o.cpd(np.asarray(T, dtype=np.float64), R, make_opts(dict(mlsvd_method='classic', refine=True)))
    np.random.seed(7)
    fg, og = tfx.cpd(T, R, make_opts(dict(mlsvd_method='classic', refine=True, precision=precision)))
    assert abs(og.rel_error - oo.rel_error) < 1e-6
    assert congruence(fg, fo) > 0.999 and congruence(fg, Ftrue) > 0.999


def test_dogleg_trajectory_matches_reference_until_roundoff_decides(golden):
    """tol_jump=0.9 forces dogleg trials (GaussNewton.py:303-307, 713-795).  The history is
    compared up to the first iteration after a dogleg whose trials all failed: from there the
    reference divides by a roundoff-sized x - old_x, so the trajectory is decided by summation
    order (the GPU reaches the same error level, checked at the end)."""
    name = 'cjump_dogleg'
    f, out = _run(golden, name)
    assert out.dogleg_calls > 0
    e_ref = golden[name + '_errors0']
    e = out.errors[0]
    k = 0
    while k < min(e.size, e_ref.size) and abs(e[k] - e_ref[k]) <= 1e-7 * max(1.0, e_ref[k]):
        k += 1
    # iterations 0-4 never try the dogleg (it > 3) and the first trials are deterministic
    assert k >= 6, (k, e[:8], e_ref[:8])
    assert out.rel_error < 2 * float(golden[name + '_rel_error']) + 1e-3
