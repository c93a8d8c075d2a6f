"""
Writes tests/golden/reference_vectors.npz: input/output vectors of the reference's own functions
for what closed forms, brute force and finite differences cannot pin -- the seeded dGN trajectory
(GaussNewton.py:29-251, including the CG aliasing and the dogleg path), the seeded randomized SVD
(Compression.py:337-518) and seeded MLSVD outputs (Compression.py:48-189).

Run once, in the build container only (the reference checkout is not available elsewhere):

    python tests/golden/make_golden.py /root/reference

The reference imports Numba and IPython, which are not installed here.  They are replaced by
private stand-ins in sys.modules for the duration of this script: ``njit`` returns the plain
Python function and ``prange`` is ``range``.  The reference therefore runs as plain Python, with
its own NumPy/SciPy calls unchanged.  Only small sizes are used.  The 1 s sleep inside
``consistency`` (Auxiliar.py:76) is part of every cpd call and makes this script slow on purpose.
"""

import os
import sys
import types
import warnings

import numpy as np


def _install_stand_ins():
    numba = types.ModuleType('numba')

    def njit(*args, **kwargs):
        if len(args) == 1 and callable(args[0]) and not kwargs:
            return args[0]
        return lambda f: f

    numba.njit = njit
    numba.jit = njit
    numba.prange = range
    errors = types.ModuleType('numba.errors')
    for name in ('NumbaDeprecationWarning', 'NumbaPendingDeprecationWarning', 'NumbaPerformanceWarning'):
        setattr(errors, name, type(name, (Warning,), {}))
    numba.errors = errors
    sys.modules['numba'] = numba
    sys.modules['numba.errors'] = errors
    ipy = types.ModuleType('IPython')
    ipyd = types.ModuleType('IPython.display')
    ipy.display = ipyd
    sys.modules['IPython'] = ipy
    sys.modules['IPython.display'] = ipyd


def _opts(tfx, **kw):
    class options:
        pass
    for k, v in kw.items():
        setattr(options, k, v)
    return tfx.make_options(options)


def main(ref_root):
    _install_stand_ins()
    sys.path.insert(0, os.path.join(ref_root, 'modules'))
    warnings.simplefilter('ignore')
    import TensorFox as tfx
    import TensorFox.GaussNewton as gn
    import TensorFox.Compression as cmpr
    import TensorFox.Conversion as cnv

    out = {}

    # --- CG on fixed inputs (pins the P/residual aliasing and the returned M*y) -------------
    rs = np.random.RandomState(7)
    dims, R = (4, 5, 6), 3
    T = rs.randn(*dims)
    F = [rs.randn(d, R) for d in dims]
    out['cg_T'] = T
    for l in range(3):
        out['cg_F%d' % l] = F[l].copy()
    Tl = [cnv.unfold_C(T, l + 1) for l in range(3)]
    for maxiter in (1, 2, 3, 6, 40):
        data = gn.prepare_data(dims, R)
        y = np.zeros(R * sum(dims))
        yy, grad, jtjg, itn, rn = gn.cg(Tl, [f.copy() for f in F], data, y, 0.37, maxiter, 0.0)
        out['cg_y_%d' % maxiter] = yy
        out['cg_itn_%d' % maxiter] = np.array(itn)
        out['cg_rn_%d' % maxiter] = np.array(rn)
    out['cg_grad'] = grad
    out['cg_jtjg'] = jtjg

    # --- randomized SVD, wide and tall ----------------------------------------------------
    rs = np.random.RandomState(11)
    M = rs.randn(8, 3) @ rs.randn(3, 30) + 1e-3 * rs.randn(8, 30)
    out['rsvd_M'] = M
    np.random.seed(5)
    U, s, Vt = cmpr.randomized_svd(M, 4, False, n_oversamples=10, n_iter=2)
    out['rsvd_wide_U'], out['rsvd_wide_s'], out['rsvd_wide_Vt'] = U, s, Vt
    np.random.seed(6)
    U, s, Vt = cmpr.randomized_svd(M.T.copy(), 4, False, n_oversamples=10, n_iter=2)
    out['rsvd_tall_U'], out['rsvd_tall_s'], out['rsvd_tall_Vt'] = U, s, Vt

    # --- MLSVD, seq and classic ------------------------------------------------------------
    rs = np.random.RandomState(13)
    Ft = [rs.randn(d, 3) for d in (9, 7, 6)]
    T = cnv.cpd2tens(Ft) + 1e-4 * rs.randn(9, 7, 6)
    out['mlsvd_T'] = T
    for method in ('seq', 'classic'):
        np.random.seed(9)
        o = _opts(tfx, mlsvd_method=method)
        S, U, T1, sig = cmpr.mlsvd(T, np.linalg.norm(T), 4, o)
        out['mlsvd_%s_S' % method] = S
        for l in range(3):
            out['mlsvd_%s_U%d' % (method, l)] = U[l]
            out['mlsvd_%s_sigma%d' % (method, l)] = sig[l]

    # --- one ALS sweep on fixed inputs (Alternating_Least_Squares.py:229-253) -----------------
    import TensorFox.Alternating_Least_Squares as alsq
    rs = np.random.RandomState(31)
    dims, R = (4, 5, 6), 3
    T = rs.randn(*dims)
    F = [rs.randn(d, R) for d in dims]
    out['als_T'] = T
    for l in range(3):
        out['als_F%d' % l] = F[l].copy()
    Fo = alsq.als_iteration([cnv.unfold(T, l + 1) for l in range(3)], [f.copy() for f in F], -1)
    for l in range(3):
        out['als_out%d' % l] = Fo[l]

    # --- whole cpd trajectories -------------------------------------------------------------
    T222 = np.zeros((2, 2, 2))
    s = 0
    for k in range(2):
        for i in range(2):
            for j in range(2):
                T222[i, j, k] = s
                s += 1
    rs = np.random.RandomState(21)
    T567 = cnv.cpd2tens([rs.randn(d, 2) for d in (5, 6, 7)]) + 1e-3 * rs.randn(5, 6, 7)
    rs = np.random.RandomState(22)
    T_big = cnv.cpd2tens([rs.randn(d, 4) for d in (12, 10, 14)]) + 1e-5 * rs.randn(12, 10, 14)
    rs = np.random.RandomState(23)
    T4 = cnv.cpd2tens([rs.randn(d, 3) for d in (5, 4, 6, 3)]) + 1e-4 * rs.randn(5, 4, 6, 3)
    rs = np.random.RandomState(24)
    T_jump = cnv.cpd2tens([rs.randn(d, 3) for d in (6, 5, 7)]) + 1e-2 * rs.randn(6, 5, 7)
    rs = np.random.RandomState(25)
    a = rs.randn(6, 2)
    E = rs.randn(6, 6, 6)
    E = sum(np.transpose(E, p) for p in ((0, 1, 2), (0, 2, 1), (1, 0, 2), (1, 2, 0), (2, 0, 1), (2, 1, 0)))
    T_symm = cnv.cpd2tens([a, a, a]) + 1e-3 * E

    cases = [
        ('c222_default', T222, 3, dict(), 1),
        ('c222_ones', T222, 3, dict(initialization=[np.ones((2, 3))] * 3), 2),
        ('c222_refine', T222, 3, dict(refine=True), 3),
        ('c222_cgstatic', T222, 3, dict(inner_method='cg_static', cg_maxiter=3, cg_tol=1e-7), 4),
        ('c567_refine', T567, 2, dict(refine=True), 5),
        ('cbig_refine', T_big, 4, dict(refine=True), 6),
        ('cbig_classic', T_big, 4, dict(mlsvd_method='classic', maxiter=40), 7),
        ('c4_refine', T4, 3, dict(refine=True, maxiter=60), 8),
        ('cjump_dogleg', T_jump, 3, dict(tol_jump=0.9, maxiter=30), 9),
        ('cjump_sched', T_jump, 3, dict(init_damp=[0.5] * 25, maxiter=25,
                                          inner_method=['cg'] * 10 + ['cg_static'] * 15), 10),
        ('csymm', T_symm, 2, dict(symm=True, maxiter=40), 11),
        ('csymm_refine', T_symm, 2, dict(symm=True, refine=True, maxiter=30), 12),
        ('csymm_norm', T_symm, 2, dict(symm=True, factors_norm=2.0, maxiter=30), 13),
        ('cals', T_big, 4, dict(method='als', maxiter=60), 14),
        ('cals_refine', T567, 2, dict(method='als', refine=True, maxiter=50), 15),
        ('c4_als', T4, 3, dict(method='als', maxiter=40), 16),
        ('cinner_als', T_big, 4, dict(maxiter=20, inner_method=['cg'] * 4 + ['als'] * 6 + ['cg'] * 10), 17),
    ]
    names = []
    for name, T, R, kw, seed in cases:
        o = _opts(tfx, **{k: (list(v) if isinstance(v, list) else v) for k, v in kw.items()})
        np.random.seed(seed)
        factors, output = tfx.cpd(T, R, o)
        names.append(name)
        out[name + '_T'] = T
        out[name + '_R'] = np.array(R)
        out[name + '_seed'] = np.array(seed)
        for l, f in enumerate(factors):
            out['%s_F%d' % (name, l)] = f
        out[name + '_rel_error'] = np.array(output.rel_error)
        out[name + '_num_steps'] = np.array(output.num_steps)
        out[name + '_stop'] = np.array(output.stop)
        for st in range(2):
            out['%s_errors%d' % (name, st)] = np.asarray(output.errors[st], dtype=float)
            out['%s_step_sizes%d' % (name, st)] = np.asarray(output.step_sizes[st], dtype=float)
            out['%s_improv%d' % (name, st)] = np.asarray(output.improv[st], dtype=float)
            out['%s_gradients%d' % (name, st)] = np.asarray(output.gradients[st], dtype=float)
        print(name, output.rel_error, output.num_steps, output.stop, flush=True)
    out['case_names'] = np.array(names)
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), 'reference_vectors.npz')
    np.savez_compressed(path, **out)
    print('wrote', path)


if __name__ == '__main__':
    main(sys.argv[1] if len(sys.argv) > 1 else '/root/reference')
