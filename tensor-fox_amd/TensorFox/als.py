"""Alternating least squares on the GPU (Alternating_Least_Squares.py:18-264).

One sweep updates the modes in order, each from the modes already updated in the sweep:
F_l <- T_(l) pinv(KR_l^T) = N_l pinv(P1_l), with N_l the MTTKRP of mode l (one pass over the
tensor, `tfx_mttkrp`) and P1_l the Hadamard product of the other factors' Grams
(`tfx_gramians`); the R x R (pseudo-)inverse and the product run in `tfx_als_solve`.  The host
keeps the reference's scalar bookkeeping (step size, error, stopping codes), reading a few FP64
scalars per iteration.
"""

import numpy as np
import torch

from . import _lib
from .display import print_header, print_row


def restore_fixed(prob, F, fixed):
    """Put the held factor back into the flat factors F (bicpd's fixed mode,
    Alternating_Least_Squares.py:119-124, GaussNewton.py:292-295)."""
    if fixed is None or len(prob.dims) != 3:
        return
    mode, block = fixed
    off = prob.R * sum(prob.dims[:mode])
    F[off:off + block.numel()].copy_(block)


def als_sweep(prob, F, fix_mode=-1):
    """One ALS sweep over the flat factors F (device, in place): Alternating_Least_Squares.py:
    206-264; a fixed mode (bicpd) is skipped, the others keep their order."""
    dev, dims, R = prob.dev, prob.dims, prob.R
    for l in range(len(dims)):
        if l == fix_mode:
            continue
        dev.gramians(dims, R, F, prob.Gr, prob.P1, prob.P2, prob.gamma)
        dev.mttkrp(prob.T, dims, F, R, prob.N, mask=1 << l, prec=prob.prec)
        dev.als_solve(dims, R, l, prob.P1, prob.N, F, prob.als_status)


def als(prob, F0, options, Tsize, fixed=None):
    """ALS from the flat factors F0 (device).  Returns (last flat factors, step_sizes, errors,
    improv, gradients, stop) like Alternating_Least_Squares.py:18-203: the step size is the
    absolute ||x - old_x|| of the sweep's factors before `transform`, the gradient entry is
    |old_error - error| / step size, and the factors of the LAST iteration are returned.
    fixed: (mode, flat block) of a factor held fixed (bicpd), restored after every sweep."""
    dev, dims, R = prob.dev, prob.dims, prob.R
    maxiter = options.maxiter
    tol, tol_step, tol_improv, tol_grad = options.tol, options.tol_step, options.tol_improv, options.tol_grad
    symm, factors_norm = options.symm, options.factors_norm
    display = options.display
    error = 1.0
    stop = 5
    const = 1 + int(maxiter / 10)
    x_old, x_new = prob.xa, prob.xb
    x_old.copy_(F0)
    F = prob.F
    F.copy_(F0)
    zero = torch.zeros_like(F)
    d = torch.empty_like(F)
    step_sizes = np.empty(maxiter)
    errors = np.empty(maxiter)
    improv = np.empty(maxiter)
    gradients = np.empty(maxiter)
    sc = prob.scal
    if display > 1:
        print_header(display, als=True)
    it = 0
    for it in range(maxiter):
        old_error = error
        als_sweep(prob, F, fixed[0] if fixed is not None else -1)
        x_new.copy_(F)
        if symm or factors_norm > 0:
            # the sweep assigns a new array to every mode, so the symmetric transform averages
            # separate arrays (Conversion.py:432-444)
            dev.dgn_update(dims, R, x_new, zero, 0.0, zero, 0.0, False, 1 if symm else 0,
                           factors_norm, d, F, sc[8:12])
        restore_fixed(prob, F, fixed)
        dev.residual_sq(prob.T, dims, F, R, sc[6:7])
        dev.axpby(1.0, x_new, -1.0, x_old, d)
        dev.dogleg_dots(d, d, d, sc[0:4])
        h = sc[:8].cpu().numpy()
        error = float(np.sqrt(h[6])) / Tsize
        x_old, x_new = x_new, x_old
        step_sizes[it] = np.sqrt(h[0])
        errors[it] = error
        with np.errstate(divide='ignore', invalid='ignore'):
            gradients[it] = np.float64(abs(old_error - error)) / step_sizes[it]
        improv[it] = errors[it] if it == 0 else abs(errors[it - 1] - errors[it])
        if display > 1:
            print_row(display, it, errors[it], step_sizes[it], improv[it], gradients[it])
        if it > 1:
            if errors[it] < tol:
                stop = 0
                break
            if step_sizes[it] < tol_step:
                stop = 1
                break
            if improv[it] < tol_improv:
                stop = 2
                break
            if gradients[it] < tol_grad:
                stop = 3
                break
            if it > 2 * const and it % const == 0:
                mean1 = np.mean(errors[it - 2 * const: it - const])
                mean2 = np.mean(errors[it - const: it])
                if mean1 - mean2 <= tol_improv:
                    stop = 4
                    break
                mean3 = np.mean(improv[it - const: it])
                if mean3 < 1e-3 * mean2:
                    stop = 7
                    break
            if error > max(1, Tsize ** 2) / (1e-16 + tol):
                stop = 6
                break
    k = it + 1
    return F.clone(), step_sizes[:k], errors[:k], improv[:k], gradients[:k], stop
