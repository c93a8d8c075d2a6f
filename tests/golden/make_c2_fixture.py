"""Writes tests/golden/c2_oracle.json: FP64 oracle results (the NumPy restatement of the reference
algorithm, oracle/tfx_oracle.py) for the dense 400 x 400 x 400 rank-50 problems the GPU tests
and bench.py compare against.  It runs the oracle only, never the reference.

    python tests/golden/make_c2_fixture.py [--procs 4]

Contents:
- "colscale_seed0": the problem with factors N(0, 1), the first factor's columns scaled by
  10^U(-1, 1), relative noise 1e-5, stored FP32, seed 0, options refine=True: the oracle's final
  relative error, stop codes, step counts and main-stage error history;
- "colscale_seed0_perturbed": the same run on T * (1 + 1e-14 E), E ~ N(0, 1) (RandomState(p)),
  for p = 1..8.  These show how far roundoff-sized changes of the input move the reference
  algorithm's result on this problem;
- "sweep": seeds 0-7 with and without the column scaling (final error, stop codes, steps, the
  error of the model against the noise-free tensor, oracle seconds).

The numbers depend on the BLAS summation order only through the chaotic part of stalled runs;
the first iterations agree across thread counts and perturbations (see the perturbed runs).
"""

import argparse
import json
import os
import sys
import time
from concurrent.futures import ProcessPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
OUT = os.path.join(HERE, 'c2_oracle.json')
DIMS, R, NU = (400, 400, 400), 50, 1e-5


def _job(args):
    seed, col_scale, pert = args
    for p in (ROOT, os.path.join(ROOT, 'tensor-fox_amd')):
        if p not in sys.path:
            sys.path.insert(0, p)
    import numpy as np
    import oracle.tfx_oracle as o
    from TensorFox.auxiliar import make_lowrank
    T, F = make_lowrank(DIMS, R, NU, seed=seed, col_scale=col_scale, dtype=np.float32)
    T = T.astype(np.float64)
    if pert:
        T *= 1 + 1e-14 * np.random.RandomState(pert).standard_normal(T.shape)

    class opt:
        refine = True
    np.random.seed(seed)
    t0 = time.perf_counter()
    f, out = o.cpd(T, R, opt)
    t = time.perf_counter() - t0
    del T
    T0 = o.cpd2tens(F)
    model = float(np.linalg.norm(o.cpd2tens(f) - T0) / np.linalg.norm(T0))
    return dict(seed=seed, col_scale=col_scale, perturbation=pert, rel_error=float(out.rel_error),
                stop=[int(s) for s in out.stop], steps=[int(out.errors[0].size), int(out.errors[1].size)],
                model_error_vs_noise_free=model, oracle_seconds=round(t, 2),
                errors_main=[float(e) for e in out.errors[0][:40]])


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument('--procs', type=int, default=4)
    ap.add_argument('--threads', type=int, default=2)
    a = ap.parse_args()
    os.environ['OMP_NUM_THREADS'] = os.environ['OPENBLAS_NUM_THREADS'] = str(a.threads)
    jobs = [(0, True, 0)] + [(0, True, p) for p in range(1, 9)]
    jobs += [(s, cs, 0) for cs in (True, False) for s in range(8) if (s, cs) != (0, True)]
    with ProcessPoolExecutor(a.procs) as ex:
        res = list(ex.map(_job, jobs))
    for r in res:
        print(r['seed'], r['col_scale'], r['perturbation'], r['rel_error'], r['stop'], r['steps'], flush=True)
    base = res[0]
    data = {
        'problem': 'dense %s rank %d, factors N(0,1), relative noise %g, stored FP32, refine=True' % (
            'x'.join(map(str, DIMS)), R, NU),
        'blas_threads': a.threads,
        'colscale_seed0': base,
        'colscale_seed0_perturbed': res[1:9],
        'sweep': [base] + res[9:],
    }
    with open(OUT, 'w') as fh:
        json.dump(data, fh, indent=1)
    print('wrote', OUT)


if __name__ == '__main__':
    main()
