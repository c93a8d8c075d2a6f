"""A sweep of seeded low-rank problems through the GPU `cpd`, each against the FP64 oracle with the
same seed.

The problems cover:
- orders 2 to 5;
- odd mode sizes;
- a rank above one mode size (the core keeps min(R, I) there, Compression.py:192-237);
- last dimensions with K % 8 == 4 (the split16 MTTKRP's half-masked column groups) and
  K % 4 != 0 (its FMA fallback);
- both precision forms;
- runs with and without refinement, and the classic MLSVD.

Parity bar, as in test_gpu_cpd.py: final relative error within 1e-6 absolute of the oracle's,
and (order >= 3, where the CPD is essentially unique) the generating factors recovered with
congruence above 0.999."""

import numpy as np
import pytest

import oracle.tfx_oracle as o
from tests.golden.cases import make_opts
from tests.test_gpu_cpd import congruence

pytestmark = pytest.mark.gpu

SWEEP = [
    # (dims, R, noise, seed, precision, options)
    ((40, 30), 5, 1e-5, 11, 'fp64', dict(refine=True)),
    ((13, 17, 19), 4, 1e-5, 12, 'fp64', dict(refine=True)),
    ((64, 9, 33), 7, 1e-5, 13, 'fp64', dict()),
    ((8, 50, 50), 10, 1e-5, 14, 'fp64', dict(refine=True)),
    ((10, 12, 14, 9), 3, 1e-5, 15, 'fp64', dict(refine=True)),
    ((6, 7, 5, 8, 6), 3, 1e-5, 16, 'fp64', dict()),
    ((30, 30, 30), 5, 1e-5, 17, 'fp64', dict(mlsvd_method='classic', refine=True)),
    ((100, 36, 20), 12, 1e-5, 18, 'mixed', dict(refine=True)),
    ((50, 36, 100), 17, 1e-5, 19, 'mixed', dict(refine=True)),
    ((60, 44, 30), 6, 1e-4, 20, 'mixed', dict(refine=True)),
    ((40, 41, 42), 8, 1e-5, 21, 'mixed', dict(refine=True)),
    ((16, 20, 24, 28), 4, 1e-5, 22, 'mixed', dict(refine=True)),
    ((70, 60), 6, 1e-5, 23, 'mixed', dict(refine=True)),
    ((30, 30, 30), 5, 1e-5, 24, 'mixed', dict(mlsvd_method='classic', refine=True)),
]


@pytest.mark.parametrize('dims,R,nu,seed,precision,kw', SWEEP)
def test_sweep_matches_oracle(dims, R, nu, seed, precision, kw):
    import TensorFox as tfx
    print('sweep', dims, R, precision, kw, flush=True)
    T, Ftrue = tfx.make_lowrank(dims, R, nu, seed=seed)
    if precision == 'mixed':
        T = T.astype(np.float32)
    np.random.seed(seed)
    fo, oo = o.cpd(np.asarray(T, dtype=np.float64), R, make_opts(kw))
    kw2 = dict(kw)
    kw2['precision'] = precision
    np.random.seed(seed)
    fg, og = tfx.cpd(T, R, make_opts(kw2))
    assert abs(og.rel_error - oo.rel_error) < 1e-6, (og.rel_error, oo.rel_error)
    assert og.rel_error < 10 * nu
    if len(dims) >= 3:
        assert congruence(fg, Ftrue) > 0.999
        assert congruence(fg, fo) > 0.999
