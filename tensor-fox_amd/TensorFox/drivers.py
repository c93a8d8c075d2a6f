"""Drivers over repeated cpd calls (TensorFox.py:820-1188): rank, stats and foxit.

Each call of `cpd` inside them is the GPU pipeline; the drivers only decide which calls to make
and report the results, as the reference's do.  Plots use matplotlib like the reference (with a
non-interactive backend `plt.show()` draws nothing); `plot=False` in `rank` skips them.

Deviation: foxit with refine=True plots the refinement step sizes from `step_sizes[1]`; the
reference reads `best_outputs.step_size[1]` (TensorFox.py:1166), an attribute its output class
does not have, and stops with AttributeError there.
"""

import sys
import time
from copy import deepcopy
from decimal import Decimal

import numpy as np

from .auxiliar import make_options


def _dims_of(T):
    if type(T) == list:
        return tuple(int(d) for d in T[2])
    return tuple(T.shape)


def rank(T, options=False, plot=True, trials=3):
    """Search for the rank of T by the error of the best of `trials` CPDs for R = 1, 2, ...
    (TensorFox.py:820-948).  Stops when the error is below 1e-14, when it stops improving by
    1e-5 (with error below 0.5), or when its improvement drops by a factor 100 (error below 0.5).
    Returns (rank, errors per rank)."""
    from .tensorfox import cpd
    dims = _dims_of(T)
    L = len(dims)
    options = make_options(options)
    if L > 3:
        if options.method == 'dGN':
            Rmin = 1
            Rmax = int(np.prod(np.sort(np.array(dims))[1:]))
        else:
            Rmin, Rmax = 2, int(np.min(dims))
    else:
        m, n, p = dims
        Rmin, Rmax = 1, min(m * n, m * p, n * p)
    err = np.empty(Rmax)
    print('Start searching for rank')
    print('Stops at R =', Rmax, ' or less')
    print('-----------------------------')
    final_rank = final_error = None
    r = 0
    for r in range(1, Rmax + 1):
        sys.stdout.write('\rTesting R = ' + str(r))
        best = np.inf
        for _ in range(trials):
            if L > 3 and r == 1:
                e = 1
            else:
                _, out = cpd(T, r, options)
                e = out.rel_error
            best = min(best, e)
        err[r - 1] = best
        if best < 1e-14:
            final_rank, final_error = r, err[r - 1]
            break
        if r > Rmin and np.abs(err[r - 1] - err[r - 2]) < 1e-5 and best < 0.5:
            final_rank = int(np.nanargmin(err[0:r])) + 1
            final_error = err[final_rank - 1]
            break
        if r > 2:
            prev = np.abs(err[r - 2] - err[r - 3])
            cur = np.abs(err[r - 1] - err[r - 2])
            with np.errstate(divide='ignore', invalid='ignore'):
                ratio = prev / cur
            if ratio > 100 and best < 0.5:
                final_rank, final_error = r - 1, err[r - 2]
                break
    err = err[0:r]
    if final_rank is None:
        final_rank = int(np.nanargmin(err[0:r])) + 1
        final_error = err[final_rank - 1]
    print('\nrank(T) =', final_rank)
    print('|T - T_approx|/|T| =', final_error)
    if plot:
        import matplotlib.pyplot as plt
        plt.figure(figsize=[14, 4])
        plt.plot(range(1, r + 1), err, color='blue')
        plt.plot(range(1, r + 1), err, 's', color='blue')
        plt.plot(final_rank, final_error, marker='s', color='red')
        plt.xlabel('Rank')
        plt.ylabel('Relative error')
        plt.yscale('log')
        plt.xticks(range(1, r + 1, int((r + 1) / 20)) if r > 20 else range(1, r + 1))
        plt.grid()
        plt.show()
    return int(final_rank), err


def stats(T, R, options=False, num_samples=100):
    """num_samples CPDs from random starts: seconds, steps and relative error of each, with
    their histograms (TensorFox.py:951-1034).  Returns (times, steps, errors)."""
    from .tensorfox import cpd
    options = make_options(options)
    times, steps, errors = np.empty(num_samples), np.empty(num_samples), np.empty(num_samples)
    for trial in range(1, num_samples + 1):
        t0 = time.time()
        _, out = cpd(T, R, options)
        times[trial - 1] = time.time() - t0
        steps[trial - 1] = out.num_steps
        errors[trial - 1] = out.rel_error
        x = 100 * trial // num_samples
        sys.stdout.write('\r[' + x * '=' + (100 - x) * ' ' + '] ' + str(np.round(100 * trial / num_samples, 2)) + '%')
    import matplotlib.pyplot as plt
    for vals, xlabel, title in ((times, 'Seconds', 'Histogram of the total time of each trial'),
                                (steps, 'Number of steps', 'Histogram of the number of steps of each trial'),
                                (np.log10(errors), r'$\log_{10} \|T - \tilde{T}\|/\|T\|$',
                                 'Histogram of the log10 of the relative error of each trial')):
        plt.hist(vals, 50, edgecolor='black')
        plt.xlabel(xlabel)
        plt.ylabel('Quantity')
        plt.title(title)
        plt.show()
    return times, steps, errors


def foxit(T, R, options=False, bestof=1):
    """Best of `bestof` CPDs, its results and parameters printed and its step sizes, errors,
    improvements and gradients plotted (TensorFox.py:1049-1188).  Returns (factors, output)."""
    from .tensorfox import cpd
    options = make_options(options)
    best_error, best_factors, best = np.inf, None, None
    for _ in range(bestof):
        factors, out = cpd(T, R, options)
        if out.rel_error < best_error:
            best_error = out.rel_error
            best_factors, best = deepcopy(factors), deepcopy(out)
    print('Final results')
    print('    Number of steps =', best.num_steps)
    print('    Relative error =', best.rel_error)
    print('    Accuracy = ', float('%.6e' % Decimal(best.accuracy)), '%')
    print()
    print('==========================================================================')
    print()
    print('Parameters used')
    print('    initialization:', options.initialization)
    print('    maximum of iterations:', options.maxiter)
    print('    error tolerance:', options.tol)
    print('    steps size tolerance:', options.tol_step)
    print('    improvement tolerance:', options.tol_improv)
    print('    gradient norm tolerance:', options.tol_grad)
    print('    inner algorithm parameters:')
    if options.method == 'als':
        print('        method: alternating least squares')
    elif options.method == 'ttcpd':
        print('        method: tensor train cpd')
    elif options.inner_method == 'cg_static':
        print('        method: conjugate gradient static')
        print('        cg maximum of iterations:', options.cg_maxiter)
        print('        cg tolerance:', options.cg_tol)
    elif options.inner_method == 'cg':
        print('        method: conjugate gradient dynamic/random')
        print('        cg factor:', options.cg_factor)
        print('        cg tolerance:', options.cg_tol)
    elif options.inner_method == 'direct':
        print('        method: direct solver')
    elif type(options.inner_method) == list:
        print('        method: hybrid strategy')
    print()
    import matplotlib.pyplot as plt
    plt.figure(figsize=[9, 6])
    sz1 = best.step_sizes[0].size
    x1 = np.arange(sz1)
    series = ((best.step_sizes, 'k', 'Step sizes'), (best.errors, 'b', 'Relative errors'),
              (best.improv, 'g', 'Improvements'), (best.gradients, 'r', 'Gradients'))
    for vals, c, name in series:
        plt.plot(x1, vals[0], c + '-', markersize=2, label=name + ' - Main')
        if options.refine:
            x2 = np.arange(sz1 - 1, sz1 + best.step_sizes[1].size - 1)
            plt.plot(x2, vals[1], c + '--', markersize=2, label=name + ' - Refinement')
    plt.xlabel('iteration')
    plt.yscale('log')
    plt.grid()
    plt.legend()
    plt.show()
    return best_factors, best
