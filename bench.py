"""Benchmark of the dense CPD hot path on one MI355X.

Workload (a "step" = one full `cpd` call: MLSVD -> dGN on the core -> uncompression ->
refinement -> final FP64 error): dense 400 x 400 x 400 FP32 tensor, rank 50, factors N(0, 1),
relative noise 1e-5, seed 2 (an instance on which the reference algorithm converges to the noise
floor); the tensor is resident in HBM before the timed region.  The same problem with the first
factor's columns scaled by 10^U(-1, 1) (seed 0), on which the reference algorithm stalls at a
relative error of about 3e-2, is timed once and reported under "colscale_instance".  The mixed-precision form is timed for `--steps`
steps; the FP64 GPU form runs the same problem once for the speedup; the dominant kernel (the
fused MTTKRP of the refinement) is timed with HIP events on the stream it runs on; the CPU
baseline times the FP64 NumPy restatement of the reference algorithm on a bounded sample.

    python bench.py [--gpus 1] [--steps K] [--warmup W]
"""

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
for p in (ROOT, os.path.join(ROOT, 'tensor-fox_amd')):
    if p not in sys.path:
        sys.path.insert(0, p)

HBM_PEAK_GBS = 8000.0

# HBM bytes per all-mode split16 MTTKRP call measured with rocprofv3 --pmc FETCH_SIZE and
# --pmc WRITE_SIZE (separate runs of `bench.py --kernel-only`), each scaled by the counts of a
# plain 256 MB device copy in the same run (FETCH_SIZE x 2.0, WRITE_SIZE x 1.0):
# reads 262.4 MB + writes 12.2 MB (brick partials) + the partial sums (6.1 MB read).
MEASURED_TRAFFIC = {((400, 400, 400), 50): 280.7e6}


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument('--gpus', type=int, default=1)
    ap.add_argument('--steps', type=int, default=3)
    ap.add_argument('--warmup', type=int, default=1)
    ap.add_argument('--dims', type=int, nargs='+', default=[400, 400, 400])
    ap.add_argument('--rank', type=int, default=50)
    ap.add_argument('--noise', type=float, default=1e-5)
    ap.add_argument('--seed', type=int, default=2)
    ap.add_argument('--no-colscale', action='store_true', help='skip the column-scaled instance')
    ap.add_argument('--no-fp64', action='store_true', help='skip the FP64 GPU form')
    ap.add_argument('--no-cpu', action='store_true', help='skip the CPU baseline')
    ap.add_argument('--no-loss', action='store_true', help='skip the CPDLoss forward/backward timing')
    ap.add_argument('--no-order4', action='store_true', help='skip the order-4 96^4 problem')
    ap.add_argument('--no-batched', action='store_true', help='skip the 4096-member batch')
    ap.add_argument('--batch-size', type=int, default=4096)
    ap.add_argument('--no-pipeline', action='store_true',
                    help='skip the 400^3 pipeline (and its FP64, column-scaled and CPU legs): time only '
                         'the other sections, e.g. to profile one of them')
    ap.add_argument('--cpu-sample-iters', type=int, default=1)
    ap.add_argument('--kernel-reps', type=int, default=20)
    ap.add_argument('--kernel-only', action='store_true',
                    help='time only the fused MTTKRP on a random resident tensor of the same shape, '
                         'and a plain device copy of the tensor (the reference for HBM counters)')
    ap.add_argument('--compression-only', action='store_true',
                    help='time only the MLSVD mode Gram and mode product on a random resident tensor '
                         'of the same shape')
    return ap.parse_args()


class Opt:
    refine = True


def time_cpd(tfx, Tdev, R, precision, seed):
    import numpy as np
    import torch
    o = Opt()
    o.precision = precision
    torch.cuda.synchronize()
    np.random.seed(seed)
    t0 = time.perf_counter()
    factors, out = tfx.cpd(Tdev, R, o)
    torch.cuda.synchronize()
    return time.perf_counter() - t0, factors, out


def mttkrp_kernel_time(dev, Tdev, dims, R, prec, reps):
    """Average duration of one full MTTKRP call (all modes) measured with HIP events recorded on
    the stream the kernels are launched on (torch's current stream)."""
    import numpy as np
    import torch
    rs = np.random.RandomState(1)
    F = dev.upload(rs.standard_normal(R * sum(dims)), torch.float64)
    out = dev.zeros(R * sum(dims), torch.float64)
    for _ in range(3):
        dev.mttkrp(Tdev, dims, F, R, out, prec=prec)
    st = torch.cuda.current_stream()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record(st)
    for _ in range(reps):
        dev.mttkrp(Tdev, dims, F, R, out, prec=prec)
    e1.record(st)
    e1.synchronize()
    return e0.elapsed_time(e1) / 1e3 / reps


def kernel_only(args):
    """MTTKRP timing alone: FP32 tensor of the workload's shape (random values), FP64 factors,
    split16 precision; then `reps` plain copies of the tensor (torch's copy kernel), whose counters
    give the bytes a kernel moving exactly 2|T| reports."""
    import numpy as np
    import torch
    from TensorFox import _lib
    from TensorFox.device import get_device
    dims, R = tuple(args.dims), args.rank
    dev = get_device(0)
    g = torch.Generator(device=dev.dev)
    g.manual_seed(0)
    Tdev = torch.randn(dims, generator=g, device=dev.dev, dtype=torch.float32)
    kt = mttkrp_kernel_time(dev, Tdev, dims, R, _lib.TFX_PREC_SPLIT16, args.kernel_reps)
    N = int(np.prod(dims))
    alg_bytes = 4 * N + 8 * R * sum(dims) * 2
    dst = torch.empty_like(Tdev)
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(args.kernel_reps):
        dst.copy_(Tdev)
    e1.record()
    e1.synchronize()
    ct = e0.elapsed_time(e1) / 1e3 / args.kernel_reps
    res = {'kernel': 'tfx_mttkrp split16', 'dims': list(dims), 'rank': R,
           'avg_launch_s': kt, 'alg_bytes': alg_bytes, 'achieved_GBs': round(alg_bytes / kt / 1e9, 1),
           'frac': round(alg_bytes / kt / 1e9 / HBM_PEAK_GBS, 4),
           'copy_avg_s': ct, 'copy_GBs': round(2 * 4 * N / ct / 1e9, 1)}
    print(json.dumps(res), flush=True)


BF16_MFMA_PEAK_TFLOPS = 2500.0    # MI355X dense bf16 MFMA
FP32_MFMA_PEAK_TFLOPS = 157.3     # MI355X dense FP32 MFMA


def _event_time(fn, reps, warm=3):
    import torch
    for _ in range(warm):
        fn()
    st = torch.cuda.current_stream()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record(st)
    for _ in range(reps):
        fn()
    e1.record(st)
    e1.synchronize()
    return e0.elapsed_time(e1) / 1e3 / reps


def compression_time(dev, Tdev, dims, R, reps):
    """The MLSVD's two passes over the FP32 tensor (mode 0, as in the first ST-HOSVD step), timed
    with HIP events on the stream the kernels run on:
    - the bf16x3 MFMA mode Gram: useful flops I (I + 1) K (the distinct entries of the symmetric
      product, K = N / I columns); the MFMA pipe executes 3x that (hi*hi + hi*lo + lo*hi);
    - the FP32-MFMA mode product with fused squared row norms, k = min(R, I) rows: HBM bound,
      4 N + 8 k K bytes (tensor read, FP64 result written)."""
    import numpy as np
    import torch
    from TensorFox import _lib
    I = dims[0]
    N = int(np.prod(dims))
    K = N // I
    k = min(R, I)
    G = dev.empty((I, I), torch.float64)
    rs = np.random.RandomState(3)
    U = dev.upload(np.linalg.qr(rs.standard_normal((I, k)))[0].ravel(order='F'), torch.float64)
    Y = dev.empty(k * K, torch.float64)
    sig2 = dev.empty(k, torch.float64)
    sp = _lib.TFX_PREC_SPLIT16
    tg = _event_time(lambda: dev.gram_mode(Tdev, dims, 0, G, prec=sp), reps)
    tt = _event_time(lambda: dev.ttm(Tdev, dims, 0, U, k, Y, sig2, prec=sp), reps)
    useful = I * (I + 1) * K / tg / 1e12
    ttm_bytes = 4 * N + 8 * k * K
    return {
        'gram': {'kernel': 'tfx_gram_mode split16 mode 0 (k_gram_mfma<float> + k_gram_reduce)',
                 'bound': 'mfma', 'avg_launch_s': tg, 'useful_TFLOPs': round(useful, 1),
                 'achieved': round(3 * useful, 1), 'peak': BF16_MFMA_PEAK_TFLOPS, 'unit': 'TFLOP/s',
                 'frac': round(3 * useful / BF16_MFMA_PEAK_TFLOPS, 4),
                 'input_read_GBs': round(4 * N / tg / 1e9, 1)},
        'ttm': {'kernel': 'tfx_ttm split16 mode 0, k=%d (k_ttm_prep + k_ttm_mfma<float> + k_ttm_sig2)' % k,
                'bound': 'hbm', 'avg_launch_s': tt, 'achieved': round(ttm_bytes / tt / 1e9, 1),
                'peak': HBM_PEAK_GBS, 'unit': 'GB/s', 'frac': round(ttm_bytes / tt / 1e9 / HBM_PEAK_GBS, 4),
                'fp32_mfma_TFLOPs': round(2 * k * I * K / tt / 1e12, 1),
                'fp32_mfma_frac': round(2 * k * I * K / tt / 1e12 / FP32_MFMA_PEAK_TFLOPS, 4)},
    }


def compression_only(args):
    """compression_time alone on a random FP32 tensor of the workload's shape (for counters)."""
    import torch
    from TensorFox.device import get_device
    dims = tuple(args.dims)
    dev = get_device(0)
    g = torch.Generator(device=dev.dev)
    g.manual_seed(0)
    Tdev = torch.randn(dims, generator=g, device=dev.dev, dtype=torch.float32)
    print(json.dumps(compression_time(dev, Tdev, dims, args.rank, args.kernel_reps)), flush=True)


def cpdloss_time(tfx, dev, reps):
    """CPDLoss on a 256^3 FP32 target from seeded rank-32 factors plus 1e-3 relative noise, with
    A, B, C (256 x 32 FP32) as leaf tensors: forward (one FP64-accumulated pass over T) and
    forward + backward (adds the all-mode bf16x3 MFMA MTTKRP pass), timed with HIP events."""
    import numpy as np
    import torch
    dims, R = (256, 256, 256), 32
    T, Ft = tfx.make_lowrank(dims, R, 1e-3, seed=0, dtype=np.float32)
    rs = np.random.RandomState(1)
    Td = torch.from_numpy(T).to(dev.dev)
    Fd = [torch.tensor(f + 0.1 * rs.standard_normal(f.shape), dtype=torch.float32, device=dev.dev,
                       requires_grad=True) for f in Ft]
    st = torch.cuda.current_stream()

    def run(backward):
        for f in Fd:
            f.grad = None
        loss = tfx.cpd_loss(Td, *Fd)
        if backward:
            loss.backward()
        return loss

    res = {'workload': 'CPDLoss 256x256x256 FP32, rank 32, noise 1e-3, factors as FP32 leaf tensors'}
    for name, bw in (('forward_s', False), ('forward_backward_s', True)):
        for _ in range(3):
            run(bw)
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(st)
        for _ in range(reps):
            run(bw)
        e1.record(st)
        e1.synchronize()
        res[name] = e0.elapsed_time(e1) / 1e3 / reps
    res['tensor_passes_GBs'] = round(2 * 4 * T.size / res['forward_backward_s'] / 1e9, 1)
    return res


def order4_time(tfx, dev):
    """96^4 FP32 tensor, rank 24, first two columns of the first factor with cosine 0.9, relative
    noise 1e-5, trunc_dims [24]*4, no refinement (the error is the FP64 residual on the full
    tensor): the mixed form (warm) and the FP64 form."""
    import numpy as np
    import torch
    dims, R = (96, 96, 96, 96), 24
    T, _ = tfx.make_lowrank(dims, R, 1e-5, seed=0, collinear=0.9, dtype=np.float32)
    Td = torch.from_numpy(T).to(dev.dev)
    del T

    class O:
        trunc_dims = [24] * 4
    res = {'workload': 'dense CPD 96^4 FP32 rank 24, collinear pair (cos 0.9), noise 1e-5, '
                       'trunc_dims [24]*4, seed 0'}
    for prec in ('mixed', 'mixed', 'fp64'):
        o = O()
        o.precision = prec
        X = Td if prec == 'mixed' else Td.to(torch.float64)
        torch.cuda.synchronize()
        np.random.seed(0)
        t0 = time.perf_counter()
        _, out = tfx.cpd(X, R, o)
        torch.cuda.synchronize()
        res[prec + '_seconds'] = round(time.perf_counter() - t0, 5)
        res[prec + '_rel_error'] = out.rel_error
        res[prec + '_steps'] = int(out.num_steps)
        res[prec + '_stage_seconds'] = {a: round(b, 5) for a, b in out.timings.items()}
        del X
    return res


def make_batch(nb, seed=0):
    """nb dense 3rd-order FP32 tensors, each mode size uniform in 8..64, rank uniform in 2..16,
    factors N(0, 1), relative noise 1e-4 (legacy RandomState(seed) draws)."""
    import numpy as np
    rs = np.random.RandomState(seed)
    dims = rs.randint(8, 65, (nb, 3))
    ranks = rs.randint(2, 17, nb)
    tensors = []
    for m in range(nb):
        I, J, K = (int(v) for v in dims[m])
        R = int(ranks[m])
        A, B, C = (rs.standard_normal((d, R)) for d in (I, J, K))
        T0 = ((A[:, None, :] * B[None, :, :]).reshape(I * J, R) @ C.T).reshape(I, J, K)
        E = rs.standard_normal((I, J, K))
        T = T0 + 1e-4 * np.linalg.norm(T0) * E / np.linalg.norm(E)
        tensors.append(T.astype(np.float32))
    return tensors, [int(r) for r in ranks]


def batched_time(tfx, nb, steps):
    """The batch resident on the device, autotuned launch configuration, then `steps` timed
    solves of the whole batch (member m seeded with m); plus the oracle on a bounded sample of
    members on the host cores."""
    import numpy as np
    import oracle.tfx_oracle as o
    tensors, ranks = make_batch(nb)
    seeds = list(range(nb))
    best, tune_times = tfx.autotune_batched(tensors, ranks, None, seeds, sample=256)
    prob = tfx.BatchedCPD(tensors, ranks, None, config=best)
    prob.run(seeds)
    times, res = [], None
    for _ in range(max(1, steps)):
        t0 = time.perf_counter()
        res = prob.run(seeds)
        times.append(time.perf_counter() - t0)
    t = float(np.mean(times))
    errs = np.array([out.rel_error for _, out in res])
    nsteps = np.array([out.num_steps for _, out in res])
    out = {'workload': '%d dense FP32 tensors, dims U{8..64}^3, rank U{2..16}, noise 1e-4, '
                       'default options (MLSVD, dGN on the core), member m seeded with m' % nb,
           'seconds': round(t, 5), 'members_per_s': round(nb / t, 1),
           'stage_seconds': {a: round(b, 5) for a, b in res[0][1].timings.items()},
           'config': best, 'tune_seconds': {a: round(b, 4) for a, b in tune_times.items()},
           'rel_error_median': float(np.median(errs)),
           'rel_error_below_1e-3': float(np.mean(errs < 1e-3)),
           'dgn_steps_mean': float(nsteps.mean())}
    # oracle on a bounded sample (every 512th member), scaled to the batch
    idx = list(range(0, nb, max(1, nb // 8)))
    t0 = time.perf_counter()
    for i in idx:
        np.random.seed(seeds[i])
        o.cpd(tensors[i].astype(np.float64), ranks[i])
    tc = time.perf_counter() - t0
    out['cpu_baseline'] = {'value': round(tc / len(idx) * nb, 2), 'unit': 's', 'cores': os.cpu_count(),
                           'kind': 'port',
                           'sample': 'oracle (FP64 NumPy restatement of the reference) on %d members '
                                     '(%.2f s), scaled to %d members' % (len(idx), tc, nb)}
    return out


def cpu_baseline(T32, R, sample_iters, seed):
    """FP64 NumPy restatement of the reference algorithm (the oracle) on the host cores: the
    MLSVD, `sample_iters` main-stage and `sample_iters` refinement dGN iterations are timed, then
    scaled to the iteration counts of a full run given by the caller."""
    import numpy as np
    import oracle.tfx_oracle as o
    T = T32.astype(np.float64)

    class opt:
        refine = True
        maxiter = sample_iters
    np.random.seed(seed)
    t0 = time.perf_counter()
    Ts, ordering = o.sort_dims(T)
    S, U, T1, sig = o.mlsvd(Ts, np.linalg.norm(Ts), R, o.make_options(opt))
    t1 = time.perf_counter()
    init = o.starting_point(S, U, R, ordering, o.make_options(opt))
    F, *_ = o.dGN(S, init, R, o.make_options(opt))
    t2 = time.perf_counter()
    F = [U[l] @ F[l] for l in range(3)]
    o.dGN(np.asarray(Ts), F, R, o.make_options(opt))
    t3 = time.perf_counter()
    return dict(mlsvd=t1 - t0, main_iter=(t2 - t1) / sample_iters, refine_iter=(t3 - t2) / sample_iters)


def main():
    args = parse()
    if args.gpus != 1:
        print('bench.py: only one GPU is measured; the CPD hot path here is a single-GPU kernel '
              'set (sharding it over several GPUs is not implemented). Use --gpus 1.',
              file=sys.stderr)
        sys.exit(2)
    if args.kernel_only:
        kernel_only(args)
        return
    if args.compression_only:
        compression_only(args)
        return
    import numpy as np
    import torch
    import TensorFox as tfx
    from TensorFox import _lib
    from TensorFox.device import get_device

    dev = get_device(0)
    result = {'metric': 'time to rel.err 1e-6, 400^3 R=50: mixed vs FP64 GPU dGN; MTTKRP GB/s',
              'value': None, 'unit': 's', 'n_gpus': 1, 'steps': args.steps, 'warmup': args.warmup,
              'higher_is_better': False, 'scaling': 'weak', 'vs_baseline': None}
    if not args.no_pipeline:
        dims, R, nu = tuple(args.dims), args.rank, args.noise
        log('generating', dims, R, nu)
        T, Ftrue = tfx.make_lowrank(dims, R, nu, seed=args.seed, dtype=np.float32)
        Tdev = torch.from_numpy(T).to(dev.dev)
        torch.cuda.synchronize()
        T0 = tfx.cpd2tens(Ftrue)
        T0n = float(np.linalg.norm(T0))

        for w in range(args.warmup):
            t, _, out = time_cpd(tfx, Tdev, R, 'mixed', args.seed)
            log('warmup %d: %.3f s rel_error %.3e' % (w, t, out.rel_error))
        times, outs, facs = [], [], []
        torch.cuda.synchronize()
        for k in range(args.steps):
            t, f, out = time_cpd(tfx, Tdev, R, 'mixed', args.seed)
            times.append(t)
            outs.append(out)
            facs.append(f)
            log('step %d: %.3f s rel_error %.6e final %.6e steps %s stop %s timings %s' % (
                k, t, out.rel_error, out.final_rel_error, out.num_steps, out.stop,
                {a: round(b, 4) for a, b in out.timings.items()}))
        ms = 1e3 * float(np.mean(times))
        out = outs[-1]
        model_err = float(np.linalg.norm(tfx.cpd2tens(facs[-1]) - T0)) / T0n

        result = {
            'metric': 'time to rel.err 1e-6, 400^3 R=50: mixed vs FP64 GPU dGN; MTTKRP GB/s',
            'value': round(ms / 1e3, 5),
            'unit': 's',
            'n_gpus': 1,
            'steps': args.steps,
            'warmup': args.warmup,
            'ms_per_step': round(ms, 3),
            'higher_is_better': False,
            'scaling': 'weak',
            'vs_baseline': None,
            'dtype': 'f32+f64',
            'data': 'synthetic (seeded low-rank + 1e-5 relative noise)',
            'config': {'workload': 'dense CPD %s rank %d, noise %g, seed %d, refine, mixed precision' % (
                           'x'.join(map(str, dims)), R, nu, args.seed),
                       'dims': list(dims), 'rank': R, 'noise': nu, 'seed': args.seed, 'refine': True,
                       'precision': 'mixed'},
            'rel_error': out.rel_error,
            'final_rel_error': out.final_rel_error,
            'model_error_vs_noise_free': model_err,
            'num_steps': int(out.num_steps),
            'stage_seconds': {a: round(b, 5) for a, b in out.timings.items()},
        }

        # dominant kernel: fused MTTKRP over the full tensor
        kt = mttkrp_kernel_time(dev, Tdev, dims, R, _lib.TFX_PREC_SPLIT16, args.kernel_reps)
        N = int(np.prod(dims))
        alg_bytes = 4 * N + 8 * R * sum(dims) * 2
        gbs = alg_bytes / kt / 1e9
        result['roofline'] = {'bound': 'hbm', 'achieved': round(gbs, 1), 'peak': HBM_PEAK_GBS, 'unit': 'GB/s',
                              'frac': round(gbs / HBM_PEAK_GBS, 4),
                              'traffic': MEASURED_TRAFFIC.get((dims, R)),
                              'kernel': 'tfx_mttkrp split16 (all modes, FP32 tensor, bf16x3 MFMA)', 'avg_launch_s': kt}
        result['compression_roofline'] = compression_time(dev, Tdev, dims, R, args.kernel_reps)
        log('compression:', result['compression_roofline'])

        if not args.no_fp64:
            T64 = Tdev.to(torch.float64)          # the FP64 form's input, resident before timing
            time_cpd(tfx, T64, R, 'fp64', args.seed)
            t64, f64, out64 = time_cpd(tfx, T64, R, 'fp64', args.seed)
            del T64
            log('fp64: %.3f s rel_error %.6e timings %s' % (t64, out64.rel_error, out64.timings))
            result['fp64_seconds'] = round(t64, 5)
            result['fp64_rel_error'] = out64.rel_error
            result['speedup_mixed_vs_fp64'] = round(t64 / (ms / 1e3), 3)
            result['abs_err_diff_vs_fp64'] = abs(out.rel_error - out64.rel_error)

        if not args.no_colscale:
            Tc, Fc = tfx.make_lowrank(dims, R, nu, seed=0, col_scale=True, dtype=np.float32)
            Tcd = torch.from_numpy(Tc).to(dev.dev)
            tm, _, om = time_cpd(tfx, Tcd, R, 'mixed', 0)
            res = {'workload': 'same, first factor columns x 10^U(-1,1), seed 0', 'mixed_seconds': round(tm, 5),
                   'mixed_rel_error': om.rel_error}
            if not args.no_fp64:
                t6, _, o6 = time_cpd(tfx, Tcd.to(torch.float64), R, 'fp64', 0)
                res.update(fp64_seconds=round(t6, 5), fp64_rel_error=o6.rel_error)
            log('colscale instance:', res)
            result['colscale_instance'] = res
            del Tcd

    if not args.no_loss:
        result['cpdloss'] = cpdloss_time(tfx, dev, args.kernel_reps)
        log('cpdloss:', result['cpdloss'])

    if not args.no_batched:
        result['batched'] = batched_time(tfx, args.batch_size, args.steps)
        log('batched:', {k: v for k, v in result['batched'].items() if k != 'tune_seconds'})

    if not args.no_order4:
        result['order4'] = order4_time(tfx, dev)
        log('order4:', result['order4'])

    if not args.no_cpu and not args.no_pipeline:
        cores = os.cpu_count()
        log('cpu baseline on %d cores' % cores)
        c = cpu_baseline(T, R, args.cpu_sample_iters, args.seed)
        n_main = int(out.errors[0].size)
        n_ref = int(out.errors[1].size)
        est = c['mlsvd'] + n_main * c['main_iter'] + n_ref * c['refine_iter']
        result['cpu_baseline'] = {
            'value': round(est, 3), 'unit': 's', 'cores': cores, 'kind': 'port',
            'sample': 'oracle (FP64 NumPy restatement of the reference) on the same tensor: MLSVD + '
                      '%d main + %d refinement dGN iterations timed (%.2f s, %.3f s/it, %.3f s/it), '
                      'scaled to the %d + %d iterations of the GPU run' % (
                          args.cpu_sample_iters, args.cpu_sample_iters, c['mlsvd'], c['main_iter'],
                          c['refine_iter'], n_main, n_ref)}
    print(json.dumps(result), flush=True)


if __name__ == '__main__':
    main()
