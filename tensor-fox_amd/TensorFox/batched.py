"""Many small dense CPDs solved together on one MI355X.

    results = cpd_batched(tensors, ranks, options, seeds)
    factors, output = results[m]          # as cpd(tensors[m], ranks[m], options) returns them

Member m follows the reference's `np.random.seed(seeds[m]); cpd(tensors[m], ranks[m], options)`
(TensorFox.py:69 -> tricpd :424-655 with default options): modes sorted by size, MLSVD 'seq'
compression, random start on the core, dGN with CG on the core, uncompression, relative error
of the result on the full tensor.  The reference has no batched call; its users loop over `cpd`.

Device work (libtfxhip, batch.hip): one launch computes every member's MLSVD (one workgroup per
member); one launch runs every member's whole dGN loop (one wavefront per member, the CG on
FP64 MFMA with its operands in LDS) -- or, in graph mode, one captured launch per dGN iteration
replayed as one HIP graph; one launch uncompresses and measures the error.  The host only
replays each member's random draws: the MLSVD sketches (drawn and discarded, in the reference's
shapes, because the device computes exact eigenvectors), the starting factors and the CG caps
of every iteration (GaussNewton.py:263-266), all from the member's seed -- with seeds, in native
code on host threads (tfx_batch_draws, the same stream as NumPy's legacy generator); without
seeds, from NumPy's global generator member after member (pre-drawing every iteration's cap, so
the global stream then differs from a loop of reference calls).

Supported options: method 'dGN', inner_method 'cg' or 'cg_static', initialization 'random',
mlsvd_method 'seq', no trunc_dims / svd_n_components, no refinement, no symm; order 2-4,
dims <= 64, rank <= 16.  Anything else raises NotImplementedError.
"""

import ctypes
import gc
import json
import os
import time

import numpy as np
import torch

from . import _lib
from .auxiliar import Output, consistency, make_options, sort_dims, unsort_dims
from .compression import _pick_tol
from .device import _p, get_device
from .tensorfox import _clean_zeros, _equalize, _transform, _flat

BM_STRIDE, RS_STRIDE, NV = 16, 16, 14
RS_SSIZE, RS_DONE, RS_STEPS, RS_STOP, RS_DOGLEG, RS_E2, RS_T2 = 0, 7, 9, 6, 10, 11, 12
V_BEST = 3
MAX_DIM, MAX_RANK, MAX_ORDER = 64, 16, 4

DEFAULT_CONFIG = {'mode': 'persistent', 'order': 'size'}
_TUNE_ENV = 'TFX_BATCH_TUNE_CACHE'


def _check_options(o):
    bad = []
    if o.method != 'dGN':
        bad.append("method '%s'" % o.method)
    if type(o.inner_method) == list or o.inner_method not in ('cg', 'cg_static'):
        bad.append('inner_method %r' % (o.inner_method,))
    if o.initialization != 'random':
        bad.append('initialization %r' % (o.initialization,))
    if o.mlsvd_method != 'seq':
        bad.append("mlsvd_method '%s'" % o.mlsvd_method)
    if type(o.trunc_dims) == list or o.svd_n_components is not None:
        bad.append('trunc_dims / svd_n_components')
    if o.refine:
        bad.append('refine')
    if type(o.init_damp) == list:
        bad.append('scheduled init_damp')
    if o.symm:
        bad.append('symm')
    if bad:
        raise NotImplementedError('cpd_batched supports the default pipeline only; unsupported: '
                                  + ', '.join(bad))


class _Plan:
    """Host-side layout of a batch: member records and buffer sizes."""

    def __init__(self, dims_sorted, ranks, maxiter, order_rule):
        nb = len(dims_sorted)
        self.nb = nb
        self.L = len(dims_sorted[0])
        meta = np.zeros((nb, BM_STRIDE), dtype=np.int64)
        t = u = sc = x = f = h = w = 0
        self.nx = []
        for m, (d, R) in enumerate(zip(dims_sorted, ranks)):
            L = len(d)
            kmax = [min(R, di) for di in d]
            nx = R * sum(kmax)
            wsz = 1
            for l in range(L):
                sz = int(np.prod(kmax[:l])) * kmax[l] * int(np.prod(d[l + 1:]))
                wsz = max(wsz, sz)
            meta[m, 0], meta[m, 1] = L, R
            meta[m, 2:2 + L] = d
            meta[m, 2 + L:6] = 1
            meta[m, 6], meta[m, 7], meta[m, 8] = t, u, sc
            meta[m, 9], meta[m, 10], meta[m, 11], meta[m, 12] = x, nx, f, h
            meta[m, 13], meta[m, 14] = w, wsz
            t += int(np.prod(d))
            u += sum(di * k for di, k in zip(d, kmax))
            sc += int(np.prod(kmax))
            x += NV * nx
            f += R * sum(d)
            h += 4 * maxiter
            w += 2 * wsz
            self.nx.append(nx)
        self.meta = meta
        self.sizes = dict(T=t, U=u, S=sc, X=x, F=f, H=h, W=w)
        if order_rule == 'size':
            cost = np.array([np.prod(d) * R for d, R in zip(dims_sorted, ranks)], dtype=np.float64)
            self.order = np.argsort(-cost, kind='stable').astype(np.int32)
        else:
            self.order = np.arange(nb, dtype=np.int32)


def _cap_bounds(maxiter):
    """Integer bounds of the CG cap draws randint(1 + it**0.4, 2 + it**0.9), it = 0..maxiter-1,
    with the powers taken on Python floats as the reference's loop takes them (randint truncates
    float bounds)."""
    low = np.array([int(1 + it ** 0.4) for it in range(maxiter)], dtype=np.int64)
    high = np.array([int(2 + it ** 0.9) for it in range(maxiter)], dtype=np.int64)
    return low, high


def _draw_member(seed, dims, kd, R, options, maxiter):
    """The member's random draws in the reference's order: MLSVD sketches (Compression.py:
    337-409 draws one Gaussian sketch per mode), starting factors (Initialization.py:20-94) and
    the CG caps of every iteration (GaussNewton.py:263-266)."""
    if seed is not None:
        np.random.seed(int(seed))
    L = len(dims)
    S_dims = list(dims)
    for l in range(L):
        n_comp = min(R, dims[l])
        n_rest = int(np.prod(S_dims)) // S_dims[l]
        transpose = S_dims[l] < n_rest
        np.random.mtrand._rand.normal(size=(S_dims[l] if transpose else n_rest, n_comp + 10))
        S_dims[l] = int(kd[l])
    F = [np.random.randn(int(kd[l]), R) for l in range(L)]
    F = _clean_zeros(F, R)
    F = _equalize(F, R)
    F = _transform(F, options.symm, options.factors_norm)
    if options.inner_method == 'cg':
        low, high = _cap_bounds(maxiter)
        draws = np.random.randint(low, high)
        caps = 1 + (L - 2) * np.trunc(options.cg_factor * draws).astype(np.int64)
    else:
        caps = np.full(maxiter, int(options.cg_maxiter), dtype=np.int64)
    return _flat(F), caps.astype(np.int32)


def _tune_cache_path():
    p = os.environ.get(_TUNE_ENV)
    if p:
        return p
    return os.path.join(os.path.expanduser('~'), '.cache', 'tensorfox_amd', 'batch_tune.json')


def _tune_key(dev, L, dims_sorted, ranks, maxiter):
    nb = len(dims_sorted)
    mean_n = float(np.mean([np.prod(d) for d in dims_sorted]))
    bucket = int(np.round(np.log2(max(mean_n, 1.0))))
    return '%s|L%d|n%d|b%d|R%d|it%d' % (torch.cuda.get_device_name(dev.dev), L,
                                         int(np.round(np.log2(nb))), bucket,
                                         int(np.max(ranks)), maxiter)


def _load_tuned(key):
    try:
        with open(_tune_cache_path()) as f:
            return json.load(f).get(key)
    except (OSError, ValueError):
        return None


def _store_tuned(key, cfg):
    path = _tune_cache_path()
    try:
        os.makedirs(os.path.dirname(path), exist_ok=True)
        try:
            with open(path) as f:
                data = json.load(f)
        except (OSError, ValueError):
            data = {}
        data[key] = cfg
        with open(path, 'w') as f:
            json.dump(data, f, indent=1, sort_keys=True)
    except OSError:
        pass


class BatchedCPD:
    """A batch resident on the device: validated, sorted, packed and uploaded once.  `run`
    solves it (every member from its seed) and returns the (factors, output) pairs."""

    def __init__(self, tensors, ranks, options=False, config=None, device=None):
        t0 = time.perf_counter()
        options = make_options(options)
        _check_options(options)
        tensors = [T.detach().cpu().numpy() if isinstance(T, torch.Tensor) else np.asarray(T)
                   for T in tensors]
        nb = len(tensors)
        if nb < 1 or nb != len(ranks):
            raise ValueError('one rank per tensor expected')
        dims_sorted, orderings = [], []
        for T, R in zip(tensors, ranks):
            consistency(R, tuple(T.shape), options)
            d, ordering = sort_dims(T.shape)
            if len(d) < 2 or len(d) > MAX_ORDER or max(d) > MAX_DIM or R > MAX_RANK:
                raise NotImplementedError('cpd_batched covers orders 2-%d, dims <= %d and rank <= %d; '
                                          'got dims %s, rank %d' % (MAX_ORDER, MAX_DIM, MAX_RANK,
                                                                   tuple(T.shape), R))
            dims_sorted.append(d)
            orderings.append(ordering)
        L = len(dims_sorted[0])
        if any(len(d) != L for d in dims_sorted):
            raise NotImplementedError('cpd_batched needs members of one order')
        self.dev = dev = get_device(options.device if device is None else device)
        precision = options.precision
        if precision is None:
            precision = 'mixed' if all(T.dtype == np.float32 for T in tensors) else 'fp64'
        self.tdt = torch.float32 if precision == 'mixed' else torch.float64
        npdt = np.float32 if precision == 'mixed' else np.float64
        maxiter = int(options.maxiter)
        if config is None:
            config = _load_tuned(_tune_key(dev, L, dims_sorted, ranks, maxiter)) or DEFAULT_CONFIG
        self.config = config
        self.options, self.precision, self.nb, self.L = options, precision, nb, L
        self.ranks = [int(r) for r in ranks]
        self.dims_sorted, self.orderings = dims_sorted, orderings
        plan = self.plan = _Plan(dims_sorted, self.ranks, maxiter, config.get('order', 'size'))
        sz = plan.sizes
        host = np.empty(sz['T'], dtype=npdt)
        for m, (T, ordering) in enumerate(zip(tensors, orderings)):
            A = T
            if any(int(ordering[l]) != l for l in range(L)):
                A = np.moveaxis(A, ordering, np.arange(L))
            o = int(plan.meta[m, 6])
            host[o:o + A.size] = np.ascontiguousarray(A, dtype=npdt).ravel()
        e = dev.empty
        self.Tbuf = torch.from_numpy(host).to(dev.dev)
        self.meta = torch.from_numpy(plan.meta).to(dev.dev)
        self.order = torch.from_numpy(plan.order).to(dev.dev)
        self.kdims = dev.zeros((nb, 4), torch.int32)
        self.U = e(max(sz['U'], 1), torch.float64)
        self.S = e(max(sz['S'], 1), torch.float64)
        self.W = e(max(sz['W'], 1), torch.float64)
        self.work = dev.zeros(max(sz['X'], 1), torch.float64)
        self.hist = dev.zeros(max(sz['H'], 1), torch.float64)
        self.res = dev.zeros(nb * RS_STRIDE, torch.float64)
        self.Ffull = e(max(sz['F'], 1), torch.float64)
        torch.cuda.synchronize()
        self.upload_seconds = time.perf_counter() - t0

    def run(self, seeds=None):
        t_all = time.perf_counter()
        nb, L, plan, options = self.nb, self.L, self.plan, self.options
        if seeds is not None and len(seeds) != nb:
            raise ValueError('one seed per tensor expected')
        maxiter = int(options.maxiter)
        self.hist.zero_()
        self.res.zero_()
        timings = {'upload': self.upload_seconds}
        kd, resh, hh, Fh = _solve(self.dev, plan, options, seeds, self.config, self.precision,
                                  self.tdt, self.Tbuf, self.meta, self.order, self.kdims, self.U,
                                  self.S, self.W, self.work, self.hist, self.res, self.Ffull,
                                  timings)
        t0 = time.perf_counter()
        # Factors and histories are views into the two host arrays this run copied back (each
        # run gets fresh ones): factor l of a member is the (d_l, R) transpose of its flat
        # column-major block, with the values the kernels wrote.  Per-member scalars are
        # converted in bulk.
        results = []
        z = np.array([0])
        steps = resh[:, RS_STEPS].astype(np.int64).tolist()
        errs = (np.sqrt(resh[:, RS_E2]) / np.sqrt(resh[:, RS_T2])).tolist()
        stops = resh[:, RS_STOP].astype(np.int64).tolist()
        doglegs = resh[:, RS_DOGLEG].astype(np.int64).tolist()
        kdl = kd[:, :L].astype(np.int64).tolist()
        fofs = plan.meta[:, 11].astype(np.int64).tolist()
        hofs = plan.meta[:, 12].astype(np.int64).tolist()
        prec = self.precision
        # the cyclic collector is paused while the ~5 nb objects are created (a full collection
        # set off in the middle of the loop costs more than the loop), and restored after
        gc_was_on = gc.isenabled()
        gc.disable()
        try:
            results = self._assemble(results, z, steps, errs, stops, doglegs, kdl, fofs, hofs, prec,
                                     Fh, hh, maxiter, timings)
        finally:
            if gc_was_on:
                gc.enable()
        timings['outputs'] = time.perf_counter() - t0
        timings['total'] = time.perf_counter() - t_all
        return results

    def _assemble(self, results, z, steps, errs, stops, doglegs, kdl, fofs, hofs, prec, Fh, hh,
                  maxiter, timings):
        nb, L, options = self.nb, self.L, self.options
        for m in range(nb):
            R = self.ranks[m]
            d = self.dims_sorted[m]
            fo = fofs[m]
            Fm = []
            for l in range(L):
                n = R * d[l]
                Fm.append(Fh[fo:fo + n].reshape(R, d[l]).T)
                fo += n
            Fm = unsort_dims(Fm, self.orderings[m])
            k = steps[m]
            h0 = hofs[m]
            hm = hh[h0:h0 + 4 * maxiter].reshape(4, maxiter)[:, :k]
            final_err = errs[m]
            o = Output(k, final_err, [hm[1], z], [hm[0], z], [hm[2], z], [hm[3], z],
                       [stops[m], 8], options, timings=timings if m == 0 else {},
                       precision=prec)
            o.final_rel_error = final_err
            o.compressed_dims = tuple(kdl[m])
            o.dogleg_calls = doglegs[m]
            results.append((Fm, o))
        return results


def cpd_batched(tensors, ranks, options=False, seeds=None, config=None, device=None):
    """CPD of every tensor in `tensors` with the rank in `ranks` (Python ints).  Returns a list
    of (factors, output) pairs.  seeds[m] (optional, 0 .. 2**32 - 1) seeds the global NumPy
    generator for member m's draws, as `np.random.seed(seeds[m])` before a call of the
    reference's cpd would.  `config` ({'mode': 'persistent'|'graph', 'order': 'size'|'natural'})
    overrides the tuned launch configuration.  output.timings of the first member holds the
    batch's stage times (seconds; 'total' excludes packing and upload, which are 'upload')."""
    return BatchedCPD(tensors, ranks, options, config, device).run(seeds)


def _lds_bytes(plan, kd, maxiter):
    """Dynamic LDS of the dGN kernel: the largest member's CG operands (Fs, qs, Ts, P1s, P2s
    with row pitch Rp + 1) plus gamma, column norms and two histories."""
    best = 0
    for m in range(plan.nb):
        L, R = int(plan.meta[m, 0]), int(plan.meta[m, 1])
        Rp = (R + 15) // 16 * 16
        R4 = (R + 3) // 4 * 4
        PR = Rp + 1
        nrows = sum((int(k) + 3) // 4 * 4 for k in kd[m, :L]) + 16
        d = 2 * nrows * PR + 2 * L * R4 * PR + (L * (L - 1) // 2) * R * R
        best = max(best, d)
    return 8 * (best + 128 + 2 * maxiter)


def _solve(dev, plan, options, seeds, config, precision, tdt, Tbuf, meta, order, kdims, U, S, W,
           work, hist, res, Ffull, timings):
    nb, L = plan.nb, plan.L
    st = dev.stream
    t0 = time.perf_counter()
    tol = _pick_tol(options.tol_mlsvd, L)
    if tol == -1:
        raise NotImplementedError('cpd_batched needs the MLSVD compression (tol_mlsvd >= 0)')
    dev._call('tfx_batch_mlsvd', _p(Tbuf), _lib.TFX_F32 if tdt == torch.float32 else _lib.TFX_F64,
              int(nb), _p(meta), _p(order), float(tol), _p(kdims), _p(U), _p(S), _p(W), st())
    kd = kdims.cpu().numpy()
    t1 = time.perf_counter()
    timings['mlsvd'] = t1 - t0
    maxiter = int(options.maxiter)
    x0 = np.zeros(plan.sizes['X'])
    caps = np.empty((nb, maxiter), dtype=np.int32)
    redo = range(nb)
    if seeds is not None:
        # every member's stream from its own seed, on host threads (libtfxhip, batch_draws.cpp)
        low, high = _cap_bounds(maxiter)
        rng = (high - 1 - low).astype(np.int64)
        sd = np.asarray(seeds, dtype=np.int64)
        if sd.min() < 0 or sd.max() > 2 ** 32 - 1:
            raise ValueError('seeds must be between 0 and 2**32 - 1')
        sd = sd.astype(np.uint32)
        dims = np.ascontiguousarray(plan.meta[:, 2:2 + L], dtype=np.int32)
        rk = np.ascontiguousarray(plan.meta[:, 1], dtype=np.int32)
        kdc = np.ascontiguousarray(kd, dtype=np.int32)
        xoff = np.ascontiguousarray(plan.meta[:, 9], dtype=np.int64)
        hz = np.zeros(nb, dtype=np.int32)
        cp = lambda a: ctypes.c_void_p(a.ctypes.data)
        rc = dev.lib.tfx_batch_draws(int(nb), int(L), cp(sd), cp(dims), cp(kdc), cp(rk), maxiter,
                                     cp(low), cp(rng), float(options.cg_factor),
                                     int(options.inner_method == 'cg_static'),
                                     int(options.cg_maxiter), float(options.factors_norm),
                                     cp(xoff), cp(x0), cp(caps), cp(hz), 0)
        _lib.check(rc)
        redo = np.nonzero(hz)[0]     # a start entry exactly 0: the reference redraws it
    for m in redo:
        R = int(plan.meta[m, 1])
        d = tuple(int(v) for v in plan.meta[m, 2:2 + L])
        x, c = _draw_member(None if seeds is None else seeds[m], d, kd[m, :L], R, options, maxiter)
        o = int(plan.meta[m, 9])
        x0[o:o + x.size] = x
        caps[m] = c
    work.copy_(torch.from_numpy(x0).to(dev.dev))
    capsd = torch.from_numpy(caps).to(dev.dev)
    bo = _lib.BatchOpts(maxiter=maxiter, symm=int(bool(options.symm)), tol=float(options.tol),
                        tol_step=float(options.tol_step), tol_improv=float(options.tol_improv),
                        tol_grad=float(options.tol_grad), tol_jump=float(options.tol_jump),
                        init_damp=float(options.init_damp), cg_tol=float(options.cg_tol),
                        factors_norm=float(options.factors_norm))
    lds = _lds_bytes(plan, kd, maxiter)
    torch.cuda.synchronize()
    t2 = time.perf_counter()
    timings['init'] = t2 - t1

    def launch(a, b):
        dev._call('tfx_batch_dgn', int(L), int(nb), _p(meta), _p(order), _p(kdims), _p(S), _p(work),
                  _p(hist), _p(res), _p(capsd), ctypes.byref(bo), int(a), int(b), int(lds), st())

    if config.get('mode', 'persistent') == 'graph':
        # one launch per dGN iteration, captured once and replayed as one graph; members that
        # have stopped return at once
        launch(0, 1)
        if maxiter > 1:
            g = torch.cuda.CUDAGraph()
            with torch.cuda.graph(g, capture_error_mode='thread_local'):
                for it in range(1, maxiter):
                    launch(it, it + 1)
            g.replay()
    else:
        launch(0, maxiter)
    torch.cuda.synchronize()
    t3 = time.perf_counter()
    timings['dgn'] = t3 - t2
    dev._call('tfx_batch_finish', _p(Tbuf), _lib.TFX_F32 if tdt == torch.float32 else _lib.TFX_F64,
              int(L), int(nb), _p(meta), _p(order), _p(kdims), _p(U), _p(work), _p(Ffull), _p(res), st())
    # results to fresh page-locked host buffers (the returned factors and histories are views
    # into them), the three copies in flight together
    host = []
    for t in (res, hist, Ffull):
        h = torch.empty(t.shape, dtype=t.dtype, pin_memory=True)
        h.copy_(t, non_blocking=True)
        host.append(h)
    torch.cuda.current_stream().synchronize()
    resh = host[0].numpy().reshape(nb, RS_STRIDE)
    hh = host[1].numpy()
    Fh = host[2].numpy()
    timings['finish'] = time.perf_counter() - t3
    return kd, resh, hh, Fh


def autotune_batched(tensors, ranks, options=False, seeds=None, sample=256, device=None,
                     candidates=None):
    """Times every launch configuration on a sample of the batch (the largest members first,
    then every k-th) and stores the fastest for batches of this signature (device, order,
    batch size, mean member size, max rank, maxiter) in the cache file
    (~/.cache/tensorfox_amd/batch_tune.json, or $TFX_BATCH_TUNE_CACHE).  Returns
    (best config, {config name: seconds})."""
    if candidates is None:
        candidates = [{'mode': 'persistent', 'order': 'size'},
                      {'mode': 'persistent', 'order': 'natural'},
                      {'mode': 'graph', 'order': 'size'}]
    nb = len(tensors)
    idx = list(range(0, nb, max(1, nb // sample)))[:sample]
    st = [tensors[i] for i in idx]
    rk = [ranks[i] for i in idx]
    sd = None if seeds is None else [seeds[i] for i in idx]
    times = {}
    for cfg in candidates:
        cpd_batched(st, rk, options, sd, config=cfg, device=device)     # warm-up
        t0 = time.perf_counter()
        cpd_batched(st, rk, options, sd, config=cfg, device=device)
        times['%s/%s' % (cfg['mode'], cfg['order'])] = time.perf_counter() - t0
    best = candidates[int(np.argmin([times['%s/%s' % (c['mode'], c['order'])] for c in candidates]))]
    opts = make_options(options)
    dev = get_device(opts.device if device is None else device)
    dims_sorted = [sort_dims(T.shape)[0] for T in tensors]
    _store_tuned(_tune_key(dev, len(dims_sorted[0]), dims_sorted, ranks, int(opts.maxiter)), best)
    return best, times
