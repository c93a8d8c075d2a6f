"""Options of the seeded cpd cases stored in reference_vectors.npz (see make_golden.py)."""

import numpy as np

CASE_OPTIONS = {
    'c222_default': {},
    'c222_ones': dict(initialization=[np.ones((2, 3))] * 3),
    'c222_refine': dict(refine=True),
    'c222_cgstatic': dict(inner_method='cg_static', cg_maxiter=3, cg_tol=1e-7),
    'c567_refine': dict(refine=True),
    'cbig_refine': dict(refine=True),
    'cbig_classic': dict(mlsvd_method='classic', maxiter=40),
    'c4_refine': dict(refine=True, maxiter=60),
    'cjump_dogleg': dict(tol_jump=0.9, maxiter=30),
    'cjump_sched': dict(init_damp=[0.5] * 25, maxiter=25,
                        inner_method=['cg'] * 10 + ['cg_static'] * 15),
    'csymm': dict(symm=True, maxiter=40),
    'csymm_refine': dict(symm=True, refine=True, maxiter=30),
    'csymm_norm': dict(symm=True, factors_norm=2.0, maxiter=30),
    'cals': dict(method='als', maxiter=60),
    'cals_refine': dict(method='als', refine=True, maxiter=50),
    'c4_als': dict(method='als', maxiter=40),
    'cinner_als': dict(maxiter=20, inner_method=['cg'] * 4 + ['als'] * 6 + ['cg'] * 10),
}

# Stop codes decided by a roundoff-level tie (mean1 - mean2 vs a zeroed tol_improv on a
# stagnated error sequence, GaussNewton.py:229-238): (case, stage) pairs where 4 and 7 are both
# acceptable.
ROUNDOFF_STOP_TIES = {('c4_refine', 1)}


def make_opts(kw):
    class options:
        pass
    for k, v in kw.items():
        setattr(options, k, list(v) if isinstance(v, list) else v)
    return options
