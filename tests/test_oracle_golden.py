"""Pins the FP64 oracle against vectors produced by the reference itself (tests/golden/make_golden.py):
CG and one ALS sweep on fixed inputs, seeded randomized SVD and MLSVD, and seventeen seeded cpd
trajectories covering defaults, user init, refinement, cg_static, classic MLSVD, order 4, the
dogleg path, scheduled damping / inner methods, symmetric factors (shared-array and factors_norm
forms), method 'als' and inner_method 'als'."""

import numpy as np
import pytest

import oracle.tfx_oracle as o
from tests.golden.cases import CASE_OPTIONS, ROUNDOFF_STOP_TIES, make_opts


@pytest.mark.parametrize('maxiter', [1, 2, 3, 6, 40])
def test_cg_matches_reference(golden, maxiter):
    T = golden['cg_T']
    F = [golden['cg_F%d' % l].copy() for l in range(3)]
    y, grad, jtjg, itn, rn = o.cg(T, F, 0.37, maxiter, 0.0)
    ref = golden['cg_y_%d' % maxiter]
    assert np.abs(y - ref).max() <= 1e-13 * np.abs(ref).max()
    assert itn == int(golden['cg_itn_%d' % maxiter])
    assert rn == pytest.approx(float(golden['cg_rn_%d' % maxiter]), rel=1e-9, abs=1e-25)
    assert np.allclose(grad, golden['cg_grad'], rtol=0, atol=1e-12)
    assert np.allclose(jtjg, golden['cg_jtjg'], rtol=1e-13, atol=1e-11)


def test_cg_aliasing_quirk_is_what_reference_does(golden):
    """Textbook PCG (no aliasing) differs from the reference at maxiter=2; the oracle does not."""
    T = golden['cg_T']
    F = [golden['cg_F%d' % l].copy() for l in range(3)]
    dims = [f.shape[0] for f in F]
    _, P1, P2 = o.gramians(F)
    Gamma, gamma = o.regularization(P1, dims)
    M = o.precond(Gamma, gamma, 0.37, dims)
    g = -o.compute_grad(T, F, P1)
    H = o.hessian(F) + 0.37 * np.diag(Gamma)
    A = (M[:, None] * H) * M[None, :]
    b = M * g
    x = np.zeros_like(b)
    r = b.copy()
    p = r.copy()
    for _ in range(2):
        Ap = A @ p
        a = (r @ r) / (p @ Ap)
        x += a * p
        r_new = r - a * Ap
        p = r_new + (r_new @ r_new) / (r @ r) * p
        r = r_new
    textbook = M * x
    ref = golden['cg_y_2']
    assert np.abs(textbook - ref).max() > 1e-6 * np.abs(ref).max()


def test_randomized_svd_matches_reference(golden):
    M = golden['rsvd_M']
    np.random.seed(5)
    U, s, Vt = o.randomized_svd(M, 4)
    assert np.allclose(U, golden['rsvd_wide_U'], atol=1e-13)
    assert np.allclose(s, golden['rsvd_wide_s'], rtol=1e-13)
    assert np.allclose(Vt, golden['rsvd_wide_Vt'], atol=1e-13)
    np.random.seed(6)
    U, s, Vt = o.randomized_svd(M.T.copy(), 4)
    assert np.allclose(U, golden['rsvd_tall_U'], atol=1e-13)
    assert np.allclose(s, golden['rsvd_tall_s'], rtol=1e-13)
    assert np.allclose(Vt, golden['rsvd_tall_Vt'], atol=1e-13)


@pytest.mark.parametrize('method', ['seq', 'classic'])
def test_mlsvd_matches_reference(golden, method):
    T = golden['mlsvd_T']
    np.random.seed(9)
    S, U, T1, sig = o.mlsvd(T, np.linalg.norm(T), 4, o.make_options(make_opts(dict(mlsvd_method=method))))
    assert np.allclose(S, golden['mlsvd_%s_S' % method], atol=1e-12)
    for l in range(3):
        assert np.allclose(U[l], golden['mlsvd_%s_U%d' % (method, l)], atol=1e-12)
        assert np.allclose(sig[l], golden['mlsvd_%s_sigma%d' % (method, l)], rtol=1e-12)


@pytest.mark.parametrize('name', sorted(CASE_OPTIONS))
def test_cpd_trajectory_matches_reference(golden, name):
    T = golden[name + '_T']
    R = int(golden[name + '_R'])
    np.random.seed(int(golden[name + '_seed']))
    factors, out = o.cpd(T, R, make_opts(CASE_OPTIONS[name]))
    assert out.num_steps == int(golden[name + '_num_steps'])
    ref_stop = [int(s) for s in golden[name + '_stop']]
    for st in range(2):
        if (name, st) in ROUNDOFF_STOP_TIES:
            assert out.stop[st] in (4, 7) and ref_stop[st] in (4, 7)
        else:
            assert out.stop[st] == ref_stop[st]
    assert out.rel_error == pytest.approx(float(golden[name + '_rel_error']), rel=1e-8, abs=1e-15)
    for l, f in enumerate(factors):
        ref = golden['%s_F%d' % (name, l)]
        assert np.abs(f - ref).max() <= 1e-9 * max(1.0, np.abs(ref).max())
    for st in range(2):
        e_ref = golden['%s_errors%d' % (name, st)]
        e = np.asarray(out.errors[st], dtype=float)
        assert e.shape == e_ref.shape
        # The first iterations are deterministic up to roundoff; once the error stagnates at the
        # noise floor, roundoff decides the oscillation, so only the level is compared there.
        k = min(10, e.size)
        assert np.allclose(e[:k], e_ref[:k], rtol=1e-9, atol=1e-15)
        assert np.allclose(e, e_ref, rtol=0.5, atol=1e-14)


def test_als_iteration_matches_reference(golden):
    """One ALS sweep on fixed inputs (Alternating_Least_Squares.py:206-231)."""
    T = golden['als_T']
    F = [golden['als_F%d' % l].copy() for l in range(3)]
    out = o.als_iteration([o.unfold(T, l + 1) for l in range(3)], F)
    for l in range(3):
        ref = golden['als_out%d' % l]
        assert np.abs(out[l] - ref).max() <= 1e-12 * np.abs(ref).max()


def test_c2_oracle_fixture_is_chaotic_beyond_the_horizon():
    """tests/golden/c2_oracle.json (oracle runs of the 400^3 rank-50 column-scaled problem, seed
    0): the runs on T * (1 + 1e-14 E) agree with the unperturbed run to 1e-8 over the first 16
    main-stage iterations, and end at final errors more than 1e-3 apart."""
    import json
    import os
    with open(os.path.join(os.path.dirname(__file__), 'golden', 'c2_oracle.json')) as fh:
        fix = json.load(fh)
    base = np.asarray(fix['colscale_seed0']['errors_main'])
    finals = []
    for r in fix['colscale_seed0_perturbed']:
        e = np.asarray(r['errors_main'])
        assert np.abs(e[:16] - base[:16]).max() <= 1e-8
        finals.append(r['rel_error'])
    assert max(finals) - min(finals) > 1e-3
    # every column-scaled seed stalls; three of the eight unscaled seeds reach the noise floor
    sweep = fix['sweep']
    assert all(r['rel_error'] > 1e-3 for r in sweep if r['col_scale'])
    assert sorted(r['seed'] for r in sweep if not r['col_scale'] and r['rel_error'] < 2e-5) == [2, 5, 7]
