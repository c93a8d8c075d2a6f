"""CPU-side checks of the product: the C-ABI library loads and exports every symbol declared in
include/tfxhip.h (no compute call is made without a GPU), and the host logic of the GPU driver
agrees with the oracle."""

import os
import re

import numpy as np
import pytest

import oracle.tfx_oracle as o

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared_symbols():
    text = open(os.path.join(ROOT, 'include', 'tfxhip.h')).read()
    return sorted(set(re.findall(r'^(?:int|const char\*)\s+(tfx_\w+)\s*\(', text, flags=re.M)))


def test_header_declares_entry_points():
    syms = _declared_symbols()
    assert 'tfx_mttkrp' in syms and 'tfx_cg' in syms and len(syms) >= 15


def test_library_loads_and_exports_every_declared_symbol():
    from TensorFox import _lib
    lib = _lib.load()
    for name in _declared_symbols():
        assert hasattr(lib, name), name
    # the ctypes prototypes cover every declared compute symbol
    for name in _declared_symbols():
        if name != 'tfx_last_error':
            assert (name in _lib.SIGNATURES or name in _lib.HOST_SIGNATURES
                    or name in _lib.HOST_I64_SIGNATURES), name


def test_dgn_run_workspace_size():
    """tfx_dgn_run_work (host-only): 18 aligned vectors of R * sum(dims) doubles plus the Gram
    blocks, and -1 for invalid shapes."""
    from TensorFox import _lib
    lib = _lib.load()
    dims = (50, 40, 30)
    need = lib.tfx_dgn_run_work(3, _lib.dims_arg(dims), 7)
    n = 7 * sum(dims)
    assert need >= 18 * n + 3 * 3 * 49
    assert lib.tfx_dgn_run_work(1, _lib.dims_arg(dims), 7) == -1
    assert lib.tfx_dgn_run_work(3, _lib.dims_arg(dims), 0) == -1


def test_make_options_matches_oracle():
    import TensorFox as tfx

    class u:
        maxiter = 7
        refine = True
        inner_method = ['cg', 'cg_static']
        bi_method = 'cg'
    a, b = tfx.make_options(u), o.make_options(u)
    for k in vars(b):
        assert getattr(a, k) == getattr(b, k), k


def test_consistency_messages_match_oracle():
    import TensorFox as tfx
    opt = tfx.make_options(False)
    for R, dims in [(2.0, (3, 3, 3)), (0, (3, 3, 3)), (2, (3, 1, 3)), (5, (2, 2, 3)), (1, (3, 3, 3, 3))]:
        with pytest.raises(SystemExit) as e1:
            tfx.consistency(R, dims, opt)
        with pytest.raises(SystemExit) as e2:
            o.consistency(R, dims, o.make_options(False))
        assert str(e1.value) == str(e2.value)


def test_sort_dims_matches_oracle():
    import TensorFox as tfx
    T = np.zeros((3, 7, 5, 7))
    d, ordering = tfx.sort_dims(T.shape)
    Ts, ordering2 = o.sort_dims(T)
    assert d == Ts.shape and list(ordering) == list(ordering2)


@pytest.mark.parametrize('seed', range(6))
def test_dogleg_coefficients_match_oracle(seed):
    from TensorFox.gauss_newton import _dogleg_coeffs
    rs = np.random.RandomState(seed)
    g = rs.randn(12)
    B = rs.randn(12, 12)
    H = B @ B.T + 0.1 * np.eye(12)
    y = np.linalg.solve(H, g) * rs.uniform(0.5, 1.5)
    Jg = H @ g
    for delta in [1e-3, 0.1, 0.5, 1.0, 3.0, 100.0]:
        a, b = _dogleg_coeffs(np.array([y @ y, y @ g, g @ g, g @ Jg]), delta)
        assert np.allclose(a * y + b * g, o.dogleg(y, g, Jg, delta), rtol=1e-10, atol=1e-12)


def test_clean_compression_matches_oracle():
    from TensorFox.compression import clean_compression
    for s in [np.array([10.0, 1.0, 1e-9, 1e-10]), np.array([3.0, 2.0, 1.0]), np.array([1.0, 1e-20])]:
        _, sig, _, dim = o.clean_compression(np.eye(len(s)), s, np.eye(len(s)), 1e-16, 3)
        assert clean_compression(s, 1e-16, 3) == dim


def test_generators_are_seeded():
    import TensorFox as tfx
    T1, F1 = tfx.make_lowrank((6, 5, 4), 3, 1e-3, seed=3, col_scale=True)
    T2, F2 = tfx.make_lowrank((6, 5, 4), 3, 1e-3, seed=3, col_scale=True)
    assert np.array_equal(T1, T2)
    T0 = tfx.cpd2tens(F1)
    assert np.linalg.norm(T1 - T0) / np.linalg.norm(T0) == pytest.approx(1e-3, rel=1e-9)
    T, F = tfx.make_lowrank((8, 5, 4, 3), 3, 0, seed=1, collinear=0.9)
    a1, a2 = F[0][:, 0], F[0][:, 1]
    assert a1 @ a2 / np.linalg.norm(a1) / np.linalg.norm(a2) == pytest.approx(0.9, abs=1e-12)


def test_product_does_not_import_oracle():
    pkg = os.path.join(ROOT, 'tensor-fox_amd', 'TensorFox')
    for f in os.listdir(pkg):
        if f.endswith('.py'):
            src = open(os.path.join(pkg, f)).read()
            assert 'oracle' not in src, f


def test_batched_draws_follow_the_oracle_stream():
    """The batched driver replays a member's random draws up front (MLSVD sketches, starting
    factors, the CG caps of all maxiter iterations).  With every iteration run (tolerances 0),
    the oracle's cpd consumes exactly the same stream: the generator states must agree, and the
    starting point must be the oracle's."""
    import TensorFox as tfx
    from TensorFox.batched import _draw_member
    from tests.golden.cases import make_opts
    rs = np.random.RandomState(4)
    dims, R = (12, 10, 9), 3
    F = [rs.randn(d, R) for d in dims]
    T = o.cpd_tensor(F) + 1e-3 * rs.randn(*dims)
    kw = dict(maxiter=6, tol=0, tol_step=0, tol_improv=0, tol_grad=0)
    np.random.seed(11)
    o.cpd(T, R, make_opts(kw))
    s1 = np.random.get_state()
    x, caps = _draw_member(11, dims, (3, 3, 3), R, tfx.make_options(make_opts(kw)), 6)
    s2 = np.random.get_state()
    assert s1[2] == s2[2] and np.array_equal(s1[1], s2[1])
    assert caps.shape == (6,) and caps[0] == 2      # 1 + (L - 2) * randint(1, 2)
    assert x.shape == (R * 9,)


def test_host_equalize_bitwise_equals_oracle():
    """The host equalization of starting points (vectorised scaling) gives bitwise the oracle's
    per-column result (Conversion.py:361-387), zero columns included."""
    from TensorFox.tensorfox import _equalize
    rs = np.random.RandomState(2)
    for trial in range(400):
        L, R = rs.randint(2, 5), rs.randint(1, 17)
        F = [rs.randn(rs.randint(1, 65), R) * 10 ** rs.uniform(-3, 3) for _ in range(L)]
        if trial % 7 == 0:
            F[1][:, R - 1] = 0
        A = o.equalize([f.copy() for f in F], R)
        B = _equalize([f.copy() for f in F], R)
        assert all(np.array_equal(a, b) for a, b in zip(A, B))


def test_native_batch_draws_match_numpy_stream():
    """tfx_batch_draws (host C++, MT19937 + polar Gaussians + masked bounded integers) against
    NumPy's legacy global generator driven by the Python replay: identical CG caps (drawn after
    every Gaussian, so any drift in the stream would show) and the same equalized start point up
    to rounding (the equalization sums differ from BLAS's)."""
    import ctypes
    import TensorFox as tfx
    from TensorFox import _lib
    from TensorFox.batched import _cap_bounds, _draw_member
    lib = _lib.load()
    rs = np.random.RandomState(9)
    nb, maxiter = 40, 60
    L = 3
    dims = np.sort(rs.randint(8, 65, (nb, L)), axis=1)[:, ::-1].astype(np.int32).copy()
    ranks = rs.randint(2, 17, nb).astype(np.int32)
    kd = np.ones((nb, 4), dtype=np.int32)
    for m in range(nb):
        kd[m, :L] = np.minimum(ranks[m], dims[m])
    kd[3, 1] = max(1, kd[3, 1] - 1)                  # a truncated mode changes later sketch shapes
    seeds = (rs.randint(0, 2 ** 31, nb)).astype(np.uint32)
    nx = np.array([ranks[m] * kd[m, :L].sum() for m in range(nb)], dtype=np.int64)
    xoff = np.concatenate([[0], np.cumsum(nx)[:-1]]).astype(np.int64)
    x0 = np.zeros(int(nx.sum()))
    caps = np.zeros((nb, maxiter), dtype=np.int32)
    hz = np.zeros(nb, dtype=np.int32)
    low, high = _cap_bounds(maxiter)
    rng = (high - 1 - low).astype(np.int64)
    p = lambda a: a.ctypes.data_as(ctypes.c_void_p)
    rc = lib.tfx_batch_draws(nb, L, p(seeds), p(dims), p(kd), p(ranks), maxiter, p(low), p(rng),
                             1.0, 0, 100, 0.0, p(xoff), p(x0), p(caps), p(hz), 4)
    assert rc == 0
    opts = tfx.make_options(False)
    for m in range(nb):
        x, c = _draw_member(int(seeds[m]), tuple(dims[m]), kd[m, :L], int(ranks[m]), opts, maxiter)
        assert np.array_equal(c, caps[m]), m
        got = x0[xoff[m]:xoff[m] + nx[m]]
        assert np.abs(got - x).max() <= 1e-13 * np.abs(x).max(), m
    assert not hz.any()


def test_mttkrp_planner_routes_large_last_dims_to_fma():
    """The split16 kernel addresses each slab T[o] (J x K) with 32-bit byte offsets of rows up to
    J + 32: shapes with (J + 32) K 4 > 2^31 - 16 (or K % 4 != 0) go to the FMA path; FP64 tensors
    use the FP64 matrix cores unless an unfolding has 2^31 columns or more."""
    from TensorFox import _lib
    lib = _lib.load()
    S16, FMA, F64 = _lib.TFX_PATH_SPLIT16_MFMA, _lib.TFX_PATH_FMA, _lib.TFX_PATH_F64_MFMA

    def plan(dims, t=_lib.TFX_F32, f=_lib.TFX_F64, prec=_lib.TFX_PREC_SPLIT16, mask=7):
        return lib.tfx_mttkrp_plan(t, f, prec, len(dims), _lib.dims_arg(dims), 50, mask)
    assert plan((400, 400, 400)) == S16
    assert plan((4, 4, 14913080)) == S16                # (4 + 32) K 4 = 2^31 - 128
    assert plan((4, 4, 14913084)) == FMA                # (4 + 32) K 4 > 2^31 - 16
    assert plan((2, 100000, 5000)) == S16
    assert plan((2, 110000, 5000)) == FMA
    assert plan((400, 400, 402)) == FMA                 # K % 4 != 0
    assert plan((30, 20, 10), prec=_lib.TFX_PREC_NATIVE) == FMA
    assert plan((30, 20, 10), t=_lib.TFX_F64, prec=_lib.TFX_PREC_NATIVE) == F64
    assert plan((2, 2, 2 ** 31), t=_lib.TFX_F64, prec=_lib.TFX_PREC_NATIVE, mask=1) == FMA
    assert plan((30, 20, 10), t=_lib.TFX_F64, prec=_lib.TFX_PREC_SPLIT16) < 0


def test_unsupported_shapes_fail_before_upload():
    """Orders above TFX_MAXL and modes above the MLSVD's Gram size pass the reference's checks
    but raise NotImplementedError before the tensor reaches the device."""
    import TensorFox as tfx
    with pytest.raises(NotImplementedError, match='order 9'):
        tfx.cpd(np.ones((2,) * 9), 2)
    with pytest.raises(NotImplementedError, match='8192'):
        tfx.cpd(np.ones((8193, 2, 2)), 2)
