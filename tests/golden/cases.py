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
