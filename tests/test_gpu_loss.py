"""CPDLoss (0.5 ||T - [[F]]||^2 with a hand-written backward) on the GPU.

Checked against the oracle's loss and its gradient compute_grad (GaussNewton.py:401-427),
against central finite differences with step 1e-3 in FP64 (f is quadratic in each single factor
entry, so central differences are exact up to roundoff), and with torch.autograd.gradcheck.
The 256^3 rank-32 FP32 configuration runs the bf16x3 MFMA MTTKRP in the backward: its gradient
agrees with the FP64 oracle to ~1e-5 of the gradient's scale."""

import numpy as np
import pytest
import torch

import oracle.tfx_oracle as o

pytestmark = pytest.mark.gpu


def _vec(factors):
    return np.concatenate([f.T.ravel() for f in factors])


def _central_differences(tfx, Td, Fd, entries, h=1e-3):
    out = []
    with torch.no_grad():
        for (l, i, r) in entries:
            Fp = [f.detach().clone() for f in Fd]
            Fm = [f.detach().clone() for f in Fd]
            Fp[l][i, r] += h
            Fm[l][i, r] -= h
            out.append((tfx.cpd_loss(Td, *Fp).item() - tfx.cpd_loss(Td, *Fm).item()) / (2 * h))
    return out


@pytest.mark.parametrize('dims,R', [((8, 9, 10), 3), ((5, 6, 4, 7), 4), ((12, 16), 5)])
def test_loss_and_gradient_fp64(dims, R):
    import TensorFox as tfx
    print('cpdloss fp64', dims, R, flush=True)
    rs = np.random.RandomState(31)
    F = [rs.randn(d, R) for d in dims]
    T = o.cpd_tensor([f + 0.3 * rs.randn(*f.shape) for f in F])
    Td = torch.tensor(T, device='cuda')
    Fd = [torch.tensor(f, device='cuda', requires_grad=True) for f in F]
    loss = tfx.cpd_loss(Td, *Fd)
    ref = 0.5 * np.sum((T - o.cpd_tensor(F)) ** 2)
    assert abs(loss.item() - ref) / ref < 1e-12
    loss.backward()
    _, P1, _ = o.gramians(F)
    g = o.compute_grad(T, F, P1)
    got = _vec([f.grad.cpu().numpy() for f in Fd])
    assert np.abs(got - g).max() / np.abs(g).max() < 1e-11
    entries = [(0, 1, 0), (len(dims) - 1, dims[-1] - 1, R - 1), (1, 2, R // 2)]
    fd = _central_differences(tfx, Td, Fd, entries)
    for (l, i, r), v in zip(entries, fd):
        assert v == pytest.approx(Fd[l].grad[i, r].item(), rel=1e-7, abs=1e-9 * np.abs(g).max())


def test_gradcheck():
    import TensorFox as tfx
    rs = np.random.RandomState(5)
    dims, R = (4, 5, 6), 2
    Td = torch.tensor(rs.randn(*dims), device='cuda')
    Fd = tuple(torch.tensor(rs.randn(d, R), device='cuda', requires_grad=True) for d in dims)
    assert torch.autograd.gradcheck(lambda *F: tfx.CPDLoss.apply(Td, *F), Fd, eps=1e-6, atol=1e-7)


def test_loss_rejects_bad_input():
    import TensorFox as tfx
    Td = torch.zeros((3, 4, 5), device='cuda')
    with pytest.raises(ValueError):
        tfx.cpd_loss(Td, torch.zeros((3, 2), device='cuda'), torch.zeros((4, 2), device='cuda'))
    with pytest.raises(ValueError):
        tfx.cpd_loss(Td.clone().requires_grad_(), *[torch.zeros((d, 2), device='cuda') for d in (3, 4, 5)])


def test_cpdloss_256_rank32_fp32():
    """Dense 256^3 FP32 target from seeded rank-32 factors plus 1e-3 relative noise; factors
    A, B, C (256 x 32 FP32) as leaf tensors."""
    import TensorFox as tfx
    print('cpdloss 256^3 R=32', flush=True)
    dims, R = (256, 256, 256), 32
    T32, Ft = tfx.make_lowrank(dims, R, 1e-3, seed=0, dtype=np.float32)
    rs = np.random.RandomState(1)
    F = [(f + 0.1 * rs.standard_normal(f.shape)).astype(np.float32) for f in Ft]
    Td = torch.tensor(T32, device='cuda')
    Fd = [torch.tensor(f, device='cuda', requires_grad=True) for f in F]
    loss = tfx.cpd_loss(Td, *Fd)
    loss.backward()
    F64 = [f.astype(np.float64) for f in F]
    T64 = T32.astype(np.float64)
    ref = 0.5 * np.sum((T64 - o.cpd_tensor(F64)) ** 2)
    # FP64 accumulation; the only rounding is the FP32 result
    assert abs(loss.item() - ref) / ref < 1e-6
    _, P1, _ = o.gramians(F64)
    g = o.compute_grad(T64, F64, P1)
    got = _vec([f.grad.cpu().double().numpy() for f in Fd])
    assert np.abs(got - g).max() / np.abs(g).max() < 1e-4
    # FP64 path on the same data: central differences (step 1e-3) against its backward
    Td64 = torch.tensor(T64, device='cuda')
    Fd64 = [torch.tensor(f, device='cuda', requires_grad=True) for f in F64]
    tfx.cpd_loss(Td64, *Fd64).backward()
    g64 = _vec([f.grad.cpu().numpy() for f in Fd64])
    assert np.abs(g64 - g).max() / np.abs(g).max() < 1e-10
    entries = [(0, 0, 0), (1, 255, 31), (2, 100, 7)]
    fd = _central_differences(tfx, Td64, Fd64, entries)
    for (l, i, r), v in zip(entries, fd):
        assert v == pytest.approx(Fd64[l].grad[i, r].item(), rel=1e-6, abs=1e-9 * np.abs(g).max())
