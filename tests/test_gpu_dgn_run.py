"""tfx_dgn_run (the dGN loop as one library call) against the package's Python loop over the same
entry points: the same kernels in the same order with the same host arithmetic, so the
histories, stop codes, dogleg counts, returned factors and the global NumPy generator's state
afterwards must be bitwise equal.  Cases: defaults, refinement, a scheduled damping with mixed
'cg' / 'cg_static' inner methods, the dogleg path, shared-array symmetric factors, an order-4
tensor and the mixed-precision form (FP32 tensor)."""

import numpy as np
import pytest

from tests.golden.cases import CASE_OPTIONS, make_opts

pytestmark = pytest.mark.gpu


def _both(T, R, kw, seed):
    import TensorFox as tfx
    import TensorFox.gauss_newton as gn
    res = []
    for off in (True, False):
        gn.NATIVE_LOOP_OFF = off
        np.random.seed(seed)
        try:
            f, out = tfx.cpd(T, R, make_opts(kw))
        finally:
            gn.NATIVE_LOOP_OFF = False
        res.append((f, out, np.random.get_state()[1].copy(), np.random.get_state()[2]))
    return res


def _same(a, b, rtol=0.0):
    """Bitwise equal (rtol 0), or within rtol where a kernel's own summation order varies between
    runs (the split16 MTTKRP adds its brick partials with LDS atomics)."""
    (fa, oa, ka, pa), (fb, ob, kb, pb) = a, b
    close = (lambda x, y: np.array_equal(x, y, equal_nan=True)) if rtol == 0 else \
        (lambda x, y: np.allclose(x, y, rtol=rtol, atol=0, equal_nan=True))
    for st in (0, 1):
        # with a varying summation order the refinement's steps and gradients of a converged
        # problem move relatively more than its errors; those are compared in the FP64 cases
        fields = ('errors', 'step_sizes', 'improv', 'gradients') if rtol == 0 or st == 0 else ('errors',)
        for field in fields:
            x, y = np.asarray(getattr(oa, field)[st]), np.asarray(getattr(ob, field)[st])
            assert x.shape == y.shape and close(x, y), (field, st)
    assert list(oa.stop) == list(ob.stop)
    assert oa.dogleg_calls == ob.dogleg_calls
    assert close(np.array([oa.rel_error]), np.array([ob.rel_error]))
    for x, y in zip(fa, fb):
        assert close(x, y) if rtol == 0 else np.abs(x - y).max() <= 1e-6 * np.abs(y).max()
    assert np.array_equal(ka, kb) and pa == pb


@pytest.mark.parametrize('name', ['c567_refine', 'cjump_sched', 'cjump_dogleg', 'csymm', 'c4_refine'])
def test_native_loop_equals_python_loop_golden_cases(golden, name):
    print('dgn_run vs python loop', name, flush=True)
    T = golden[name + '_T']
    R = int(golden[name + '_R'])
    res = _both(T, R, CASE_OPTIONS[name], int(golden[name + '_seed']))
    _same(*res)
    if name == 'cjump_dogleg':
        assert res[0][1].dogleg_calls > 0


@pytest.mark.parametrize('precision', ['fp64', 'mixed'])
def test_native_loop_equals_python_loop_lowrank(precision):
    import TensorFox as tfx
    T, _ = tfx.make_lowrank((60, 50, 40), 6, 1e-5, seed=3, col_scale=True,
                            dtype=np.float32 if precision == 'mixed' else np.float64)
    print('dgn_run vs python loop lowrank', precision, flush=True)
    _same(*_both(T, 6, {'refine': True, 'precision': precision}, 11),
          rtol=0.0 if precision == 'fp64' else 1e-9)
