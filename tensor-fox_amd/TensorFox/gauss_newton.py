"""Damped Gauss-Newton on the GPU (GaussNewton.py:29-251 and compute_step :254-312), with the
inner methods 'cg', 'cg_static', 'als' and 'direct'.

Per iteration the device runs: Gram/Hadamard products and regularization, the fused MTTKRP,
minus the gradient, the persistent preconditioned CG, the point update with equalization, and
the residual.  The host keeps only the scalar control flow of the reference (CG cap draw from the
global NumPy RNG, gain ratio, damping, best point, stopping codes), reading one small vector of
FP64 scalars per iteration.  For the CG inner methods (the default) the same loop runs inside the
library as one call (tfx_dgn_run), without interpreter overhead between the kernel launches.
"""

import sys

import numpy as np
import torch

from . import _lib
from .als import als_sweep, restore_fixed
from .display import print_header, print_row

_S_CG, _S_UPD, _S_ERR = 0, 2, 6     # offsets in the per-iteration scalar vector
# the loop below instead of tfx_dgn_run for the CG inner methods (tests compare the two)
NATIVE_LOOP_OFF = False


class Problem:
    """Device buffers of one dGN run on a tensor with dims `dims`, rank R, working dtype `wdt`."""

    def __init__(self, dev, T, dims, R, wdt, prec):
        self.dev, self.T, self.dims, self.R, self.wdt, self.prec = dev, T, tuple(dims), R, wdt, prec
        L = len(dims)
        n = R * sum(dims)
        self.n = n
        e = dev.empty
        self.Gr = e((L, R, R), wdt)
        self.P1 = e((L, R, R), wdt)
        self.P2 = e((L, L, R, R), wdt)
        self.gamma = e((L, R), wdt)
        self.N = e(n, wdt)
        self.grad = e(n, wdt)
        self.y = e(n, wdt)
        self.jtjg = e(n, wdt)
        self.xa = e(n, wdt)
        self.xb = e(n, wdt)
        self.F = e(n, wdt)
        self.best_F = e(n, wdt)
        self.scal = dev.zeros(16, torch.float64)
        self.als_status = dev.zeros(1, torch.int32)
        self.dogleg_calls = 0
        self.tmp_scal = dev.zeros(16, torch.float64)

    def error_sq(self, F, out):
        self.dev.residual_sq(self.T, self.dims, F, self.R, out)


def _mttkrp_prec(prob):
    return prob.prec


def _gain_ratio(damp, old_error, error, Tsize, gain_den):
    """update_gain_ratio (GaussNewton.py:472-487) with the denominator computed on the device."""
    num = (Tsize ** 2) * (old_error ** 2 - error ** 2)
    if gain_den != 0.0:
        return num / gain_den
    if num != 0.0:
        return np.sign(num) * np.inf
    return -np.inf


def _update_damp(damp, init_damp, gain_ratio, it):
    """GaussNewton.py:490-503."""
    if type(init_damp) == list:
        return init_damp[it]
    if gain_ratio < 0.25:
        return 3 * damp
    if gain_ratio > 0.75:
        return damp / 3
    return damp


def _update_delta(delta, gain_ratio, step_size):
    """GaussNewton.py:506-517."""
    if gain_ratio > 0.75:
        return max(delta, 2 * step_size)
    with np.errstate(over='ignore'):
        sigma = 0.25 / (1 + np.exp(-14 * (gain_ratio - 0.75))) + 0.75
    return min(sigma * delta, delta)


def _dogleg_coeffs(dots, delta):
    """dogleg (GaussNewton.py:766-795) expressed as y_dl = a*y + b*grad from the FP64 inner
    products [y.y, y.g, g.g, g.Jg]."""
    yy, yg, gg, gjg = dots
    alpha = gg / gjg
    c1 = abs(alpha) * np.sqrt(gg)
    c2 = alpha * yg - alpha * alpha * gg
    c3 = yy - 2 * alpha * yg + alpha * alpha * gg
    c4 = delta ** 2 - c1 ** 2
    m = c2 ** 2 + c3 * c4
    if np.sqrt(yy) < delta:
        return 1.0, 0.0
    if c1 >= delta:
        return 0.0, (delta / c1) * alpha
    if m < 0:
        return 1.0, 0.0
    c5 = np.sqrt(m)
    beta = (-c2 + c5) / c3 if c2 <= 0 else c4 / (c2 + c5)
    return beta, alpha * (1 - beta)


def _dogleg_steps(prob, x, y, error, Tsize, damp, symm, factors_norm):
    """compute_dogleg_steps (GaussNewton.py:713-763) with device vectors.  `x` already had y
    removed; `symm` is the update's transform form (_symm_mode).  Returns (best_F, best_x,
    best_y, best_error) as device tensors / float."""
    dev = prob.dev
    best_x, best_y, best_error = x.clone(), y.clone(), error
    best_F = prob.F.clone()
    gain_ratio, delta, count = 1.0, 1.0, 0
    sc = prob.tmp_scal
    while gain_ratio > 0:
        old_error = error
        dev.dogleg_dots(y, prob.grad, prob.jtjg, sc[0:4])
        a, b = _dogleg_coeffs(sc[0:4].cpu().numpy(), delta)
        y_new = torch.empty_like(y)
        dev.axpby(a, y, b, prob.grad, y_new)
        x_new = torch.empty_like(x)
        dev.dgn_update(prob.dims, prob.R, x, y_new, 1.0, prob.grad, damp, False, symm, factors_norm,
                       x_new, prob.F, sc[4:8])
        prob.error_sq(prob.F, sc[8:9])
        h = sc[4:9].cpu().numpy()
        error = float(np.sqrt(h[4])) / Tsize
        gain_ratio = _gain_ratio(damp, old_error, error, Tsize, float(h[0]))
        if error < old_error:
            best_x, best_y, best_error = x_new.clone(), y_new.clone(), error
            best_F = prob.F.clone()
        delta = _update_delta(delta, gain_ratio, float(np.sqrt(h[1])))
        x, y = x_new, y_new
        count += 1
        if count > 10:
            break
    return best_F, best_x, best_y, best_error


def _stop_code(it, errors, step_sizes, improv, gradients, tol, tol_step, tol_improv, tol_grad,
               const, error, Tsize):
    """Stopping conditions of GaussNewton.py:215-242 (None: continue)."""
    if it <= 1:
        return None
    if errors[it] < tol:
        return 0
    if step_sizes[it] < tol_step:
        return 1
    if improv[it] < tol_improv:
        return 2
    if gradients[it] < tol_grad:
        return 3
    if it > 2 * const and it % const == 0:
        mean1 = np.mean(errors[it - 2 * const: it - const])
        mean2 = np.mean(errors[it - const: it])
        if mean1 - mean2 <= tol_improv:
            return 4
        mean3 = np.mean(improv[it - const: it])
        if mean3 < 1e-3 * mean2:
            return 7
    if error > max(1, Tsize ** 2) / (1e-16 + tol):
        return 6
    return None


def _symm_mode(symm, factors_norm, aliased):
    """Which form of the reference's symmetric transform an update reproduces: 0 none; 1 the
    average of separate factor arrays; 2 the factor list whose entries all reference one array,
    as left by an earlier symmetric transform without factors_norm (Conversion.py:432-438)."""
    if not symm:
        return 0
    if factors_norm > 0:
        return 1
    return 2 if aliased else 1


def _native_inner(options, trace, fixed, prob):
    """The per-iteration inner methods when the whole loop can run as one library call
    (tfx_dgn_run): CG inner methods only, FP64 factors, no per-iteration display, trace or fixed
    mode; None otherwise."""
    if trace is not None or fixed is not None or options.display > 1 or prob.wdt != torch.float64:
        return None
    maxiter = options.maxiter
    inner = options.inner_method
    inners = inner if type(inner) == list else [inner] * maxiter
    if len(inners) < maxiter or any(i not in ('cg', 'cg_static') for i in inners[:maxiter]):
        return None
    if type(options.init_damp) == list and len(options.init_damp) < maxiter:
        return None
    return np.array([0 if i == 'cg' else 1 for i in inners[:maxiter]], dtype=np.int8)


def _dgn_native(prob, F0, options, Tsize, mean_abs, aliased, inner):
    """dGN as one tfx_dgn_run call: the same kernels, arguments and host arithmetic as the loop
    below, with the CG caps drawn from the global NumPy generator's MT19937 state inside the
    library (the state is read before and written back after, advanced by one draw per 'cg'
    iteration run, as the reference advances it)."""
    import ctypes
    dev, dims, R = prob.dev, prob.dims, prob.R
    maxiter = options.maxiter
    init_damp = options.init_damp
    if type(init_damp) == list:
        damp0 = init_damp[0]
        sched = np.ascontiguousarray(np.asarray(init_damp[:maxiter], dtype=np.float64))
    else:
        damp0 = init_damp * mean_abs
        sched = None
    opts = _lib.DgnOpts(int(maxiter), float(options.tol), float(options.tol_step),
                        float(options.tol_improv), float(options.tol_grad), float(options.tol_jump),
                        float(options.cg_tol), float(options.cg_factor), int(options.cg_maxiter),
                        float(damp0), int(bool(options.symm)), int(bool(aliased)),
                        float(options.factors_norm))
    st = np.random.get_state()
    key = np.ascontiguousarray(st[1], dtype=np.uint32).copy()
    pos = ctypes.c_int(int(st[2]))
    need = int(dev.lib.tfx_dgn_run_work(len(dims), _lib.dims_arg(dims), int(R)))
    if getattr(prob, 'run_work', None) is None or prob.run_work.numel() < need:
        prob.run_work = dev.empty(need, torch.float64)
    F_best = dev.empty(prob.n, torch.float64)
    hist = np.zeros(4 * maxiter)
    info = (ctypes.c_int * 3)()
    try:
        dev.dgn_run(prob.T, prob.prec, dims, R, F0, opts, inner, sched, key, pos, Tsize,
                    prob.run_work, F_best, hist, info)
    finally:
        np.random.set_state((st[0], key, pos.value, st[3], st[4]))
    k, stop = int(info[0]), int(info[1])
    prob.dogleg_calls += int(info[2])
    m = maxiter
    return (F_best, hist[0:k].copy(), hist[m:m + k].copy(), hist[2 * m:2 * m + k].copy(),
            hist[3 * m:3 * m + k].copy(), stop)


def dGN(prob, F0, options, Tsize, mean_abs, trace=None, aliased=False, fixed=None):
    """Damped Gauss-Newton from the flat factors F0 (device, working dtype).  Returns
    (best flat factors, step_sizes, errors, improv, gradients, stop) like GaussNewton.py:29-251.
    `aliased`: the starting factors came out of a symmetric transform (random initialization
    with symm, Initialization.py:74-83), so the reference's factor list already holds one shared
    array.  `fixed`: (mode, flat block) of a factor held fixed (bicpd, GaussNewton.py:94-102),
    put back after every update (:292-295)."""
    native = _native_inner(options, trace, fixed, prob)
    if native is not None and not NATIVE_LOOP_OFF:
        return _dgn_native(prob, F0, options, Tsize, mean_abs, aliased, native)
    dev = prob.dev
    dims, R = prob.dims, prob.R
    L = len(dims)
    init_damp = options.init_damp
    maxiter = options.maxiter
    tol, tol_step, tol_improv, tol_grad = options.tol, options.tol_step, options.tol_improv, options.tol_grad
    tol_jump = options.tol_jump
    symm, factors_norm = options.symm, options.factors_norm
    display = options.display
    error = 1
    best_error = np.inf
    stop = 5
    damp = init_damp[0] if type(init_damp) == list else init_damp * mean_abs
    const = 1 + int(maxiter / 10)
    if R * sum(dims) <= 100:
        tol = tol_step = tol_improv = tol_grad = 0
    x = prob.xa
    x.copy_(F0)
    prob.F.copy_(F0)
    prob.best_F.copy_(F0)
    x_spare = prob.xb
    step_sizes = np.zeros(maxiter)
    errors = np.zeros(maxiter)
    improv = np.zeros(maxiter)
    gradients = np.zeros(maxiter)
    if display > 1:
        print_header(display)
    it = 0
    for it in range(maxiter):
        old_error = error
        inner = options.inner_method[it] if type(options.inner_method) == list else options.inner_method
        if inner == 'cg':
            cap = 1 + (L - 2) * int(options.cg_factor * np.random.randint(1 + it ** 0.4, 2 + it ** 0.9))
        elif inner == 'cg_static':
            cap = options.cg_maxiter
        elif inner == 'direct':
            if prob.wdt != torch.float64:
                raise NotImplementedError("inner_method 'direct' runs in FP64 only")
        elif inner != 'als':
            sys.exit("Wrong inner method name. Must be 'cg', 'cg_static', 'als' or 'direct'.")
        if inner == 'als':
            # --- compute_step: one ALS sweep replaces the step (GaussNewton.py:273-276); y is
            # zeroed, x is the swept factors, then x2cpd + equalize + transform as below -------
            als_sweep(prob, prob.F, fixed[0] if fixed is not None else -1)
            prob.y.zero_()
            x_als = torch.empty_like(x)
            x_als.copy_(prob.F)
            dev.dgn_update(dims, R, x_als, prob.y, 1.0, prob.y, damp, True,
                           _symm_mode(symm, factors_norm, False), factors_norm, x_spare, prob.F,
                           prob.scal[_S_UPD:_S_UPD + 4])
            aliased = True
            restore_fixed(prob, prob.F, fixed)
            prob.error_sq(prob.F, prob.scal[_S_ERR:_S_ERR + 1])
            d = torch.empty_like(x)
            dev.axpby(1.0, x_spare, -1.0, x, d)
            dev.dogleg_dots(d, x, x, prob.tmp_scal[0:4])
            h = prob.scal[:8].cpu().numpy()
            hs = prob.tmp_scal[0:4].cpu().numpy()
            x, x_spare = x_spare, x
            error = float(np.sqrt(h[_S_ERR])) / Tsize
            gain_ratio = np.nan      # the reference's gradient is [nan] here (GaussNewton.py:309)
            if trace is not None:
                trace.append((error, damp, gain_ratio, 0))
            damp = _update_damp(damp, init_damp, gain_ratio, it)
            if error < best_error:
                best_error = error
                prob.best_F.copy_(prob.F)
            errors[it] = error
            step_sizes[it] = np.sqrt(hs[0]) / np.sqrt(hs[2])
            gradients[it] = np.nan
            improv[it] = errors[it] if it == 0 else abs(errors[it] - errors[it - 1])
            if display > 1:
                print_row(display, it, errors[it], step_sizes[it], improv[it], gradients[it],
                          Tsize * error, '-')
            code = _stop_code(it, errors, step_sizes, improv, gradients, tol, tol_step, tol_improv,
                              tol_grad, const, error, Tsize)
            if code is not None:
                stop = code
                break
            continue
        # --- compute_step: cg ------------------------------------------------------------------
        dev.gramians(dims, R, prob.F, prob.Gr, prob.P1, prob.P2, prob.gamma)
        dev.mttkrp_neg_grad(prob.T, dims, prob.F, R, prob.P1, prob.N, prob.grad, prec=_mttkrp_prec(prob))
        if inner == 'direct':
            # explicit J^T J and a Cholesky solve (GaussNewton.py:630-710)
            dev.dgn_direct(dims, R, prob.F, prob.P1, prob.P2, prob.gamma, damp, prob.grad, prob.y,
                           prob.scal[_S_CG:_S_CG + 2])
        else:
            dev.cg(dims, R, prob.F, prob.P1, prob.P2, prob.gamma, damp, prob.grad, cap,
                   options.cg_tol, prob.y, prob.jtjg, prob.scal[_S_CG:_S_CG + 2])
        # --- x = x + y; factors = transform(equalize(x2cpd(x))); error ----------------------------
        dev.dgn_update(dims, R, x, prob.y, 1.0, prob.grad, damp, True,
                       _symm_mode(symm, factors_norm, aliased), factors_norm, x_spare, prob.F,
                       prob.scal[_S_UPD:_S_UPD + 4])
        aliased = True
        restore_fixed(prob, prob.F, fixed)
        prob.error_sq(prob.F, prob.scal[_S_ERR:_S_ERR + 1])
        h = prob.scal[:8].cpu().numpy()
        old_x = x
        x, x_spare = x_spare, x
        error = float(np.sqrt(h[_S_ERR])) / Tsize
        gain_den = float(h[_S_UPD])
        dx2 = float(h[_S_UPD + 1])
        x2 = float(h[_S_UPD + 2])
        gmax = float(h[_S_UPD + 3])
        if it > 3 and inner != 'direct' and error > tol_jump * old_error:
            prob.dogleg_calls += 1
            x_back = torch.empty_like(x)
            dev.axpby(1.0, x, -1.0, prob.y, x_back)
            bF, bx, by, error = _dogleg_steps(prob, x_back, prob.y.clone(), error, Tsize, damp,
                                              _symm_mode(symm, factors_norm, True), factors_norm)
            prob.F.copy_(bF)
            x.copy_(bx)
            prob.y.copy_(by)
            # x - old_x statistics for the gain ratio / step size of this iteration
            d = torch.empty_like(x)
            dev.axpby(1.0, x, -1.0, old_x, d)
            scratch_x, scratch_F = torch.empty_like(x), torch.empty_like(x)
            dev.dgn_update(dims, R, old_x, d, 1.0, prob.grad, damp, False, False, 0.0, scratch_x,
                           scratch_F, prob.tmp_scal[0:4])
            hs = prob.tmp_scal[0:4].cpu().numpy()
            gain_den, dx2 = float(hs[0]), float(hs[1])
        gain_ratio = _gain_ratio(damp, old_error, error, Tsize, gain_den)
        if trace is not None:
            trace.append((error, damp, gain_ratio, int(h[_S_CG])))
        damp = _update_damp(damp, init_damp, gain_ratio, it)
        if error < best_error:
            best_error = error
            prob.best_F.copy_(prob.F)
        errors[it] = error
        step_sizes[it] = np.sqrt(dx2) / np.sqrt(x2)
        gradients[it] = gmax
        improv[it] = errors[it] if it == 0 else abs(errors[it] - errors[it - 1])
        if display > 1:
            print_row(display, it, errors[it], step_sizes[it], improv[it], gradients[it],
                      float(h[_S_CG + 1]), '-' if inner == 'direct' else int(h[_S_CG]))
        code = _stop_code(it, errors, step_sizes, improv, gradients, tol, tol_step, tol_improv,
                          tol_grad, const, error, Tsize)
        if code is not None:
            stop = code
            break
    k = it + 1
    return prob.best_F.clone(), step_sizes[:k], errors[:k], improv[:k], gradients[:k], stop
