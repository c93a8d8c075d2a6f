"""Differentiable CPD residual loss on the MI355X kernels.

    loss = CPDLoss.apply(T, A, B, C)        # 0.5 * ||T - [[A, B, C]]||_F^2, any order >= 2

Forward: one pass over T (`tfx_cpd_residual_mixed`, FP64 accumulation).  Backward: the
gradient of Tensor Fox's error function, grad_l = F_l P1_l - T_(l) KR_l with P1_l the Hadamard
product of the other factors' Gram matrices (GaussNewton.py:401-427, 520-548): one fused pass over
T for every mode's MTTKRP (`tfx_mttkrp_mixed`; the bf16x3 MFMA form for an FP32 tensor whose last
dimension is a multiple of 4), the Grams (`tfx_gramians`) and `tfx_neg_grad`.  The Khatri-Rao
products are formed inside the kernels, tile by tile, and never stored.  Factors enter the
kernels as FP64 whatever their dtype; the loss and the gradients are returned in the factors'
dtype.  The target T is a constant (a target that requires grad is rejected).
"""

import torch

from . import _lib
from .device import get_device


def _flat64(factors):
    """concat_l vec(F_l) (column-major per factor), FP64 -- the kernels' factor layout (one
    converting copy per factor into its slice)."""
    n = sum(f.numel() for f in factors)
    x = torch.empty(n, dtype=torch.float64, device=factors[0].device)
    s = 0
    for f in factors:
        d, R = f.shape
        x[s:s + R * d].view(R, d).copy_(f.detach().t())
        s += R * d
    return x


class CPDLoss(torch.autograd.Function):
    @staticmethod
    def forward(ctx, T, *factors):
        if T.requires_grad:
            raise ValueError('CPDLoss treats the target tensor as a constant; pass T.detach()')
        if len(factors) != T.dim():
            raise ValueError('CPDLoss needs one factor per mode (%d), got %d' % (T.dim(), len(factors)))
        R = int(factors[0].shape[1])
        for l, f in enumerate(factors):
            if f.dim() != 2 or f.shape[0] != T.shape[l] or f.shape[1] != R:
                raise ValueError('factor %d has shape %s, expected (%d, %d)' % (l, tuple(f.shape), T.shape[l], R))
            if f.device != T.device:
                raise ValueError('factors and target must be on the same device')
        if T.device.type != 'cuda':
            raise ValueError('CPDLoss runs on the GPU: move T and the factors to a ROCm device')
        if T.dtype not in (torch.float32, torch.float64):
            raise TypeError('T must be float32 or float64')
        dev = get_device(T.device.index or 0)
        dims = tuple(int(d) for d in T.shape)
        Tc = T.contiguous()
        x = _flat64(factors)
        out = torch.empty(1, dtype=torch.float64, device=T.device)
        dev.residual_sq(Tc.view(-1), dims, x, R, out)
        ctx.save_for_backward(Tc, x)
        ctx.fdtypes = [f.dtype for f in factors]
        return (0.5 * out[0]).to(factors[0].dtype)

    @staticmethod
    def backward(ctx, g):
        Tc, x = ctx.saved_tensors
        dev = get_device(Tc.device.index or 0)
        dims = tuple(int(d) for d in Tc.shape)
        L, R = len(dims), x.numel() // sum(dims)
        f64 = torch.float64
        Gr, P1, P2 = dev.empty((L, R, R), f64), dev.empty((L, R, R), f64), dev.empty((L, L, R, R), f64)
        gamma = dev.empty((L, R), f64)
        dev.gramians(dims, R, x, Gr, P1, P2, gamma)
        N = dev.empty(x.numel(), f64)
        prec = _lib.TFX_PREC_SPLIT16 if Tc.dtype == torch.float32 else _lib.TFX_PREC_NATIVE
        dev.mttkrp(Tc.view(-1), dims, x, R, N, prec=prec)
        ng = dev.empty(x.numel(), f64)
        dev.neg_grad(dims, R, x, P1, N, ng)            # N - F P1 = minus the gradient
        # gradient = -g (N - F P1), scaled in FP64 and converted once for the whole vector when
        # the factors share a dtype; factor l's gradient is the (I_l, R) transpose of its block
        gall = ng.mul_(-g.to(f64))
        same = all(dt == ctx.fdtypes[0] for dt in ctx.fdtypes)
        if same and ctx.fdtypes[0] != f64:
            gall = gall.to(ctx.fdtypes[0])
        grads, s = [], 0
        for l, d in enumerate(dims):
            gl = gall[s:s + R * d].view(R, d).t()
            if gl.dtype != ctx.fdtypes[l]:
                gl = gl.to(ctx.fdtypes[l])
            grads.append(gl if ctx.needs_input_grad[l + 1] else None)
            s += R * d
        return (None, *grads)


def cpd_loss(T, *factors):
    """0.5 * ||T - [[factors]]||_F^2 with the hand-written backward of CPDLoss."""
    return CPDLoss.apply(T, *factors)
