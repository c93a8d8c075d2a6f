"""
Writes tests/golden/sparse_init_vectors.npz: input/output vectors of the reference's own
functions for the sparse COO path and the MLSVD-based starting points, where closed forms do not
fix the result: the seeded generator gen_rand_sparse_tensor (Auxiliar.py:575-647, set order and
the missing-entry-counts-as-one rule), the sparse Gram / multilinear product / error
(Conversion.py:255-302, MultilinearAlgebra.py:145-175, 524-571, Critical.py:25-53), seeded sparse
cpd trajectories (TensorFox.py:424-655 with T = [data, idxs, dims]), the index selection of
smart / smart_random (Initialization.py:97-358), seeded trajectories of the direct inner
solver (GaussNewton.py:630-710) and seeded tensor-train CPDs (method 'ttcpd').

Run once, in the build container only (the reference checkout is not available elsewhere):

    python tests/golden/make_golden_sparse_init.py /root/reference

The reference is imported with the stand-ins of make_golden.py (Numba and IPython replaced; the
reference's own NumPy/SciPy calls unchanged).  Its sparse Gram goes through the required
dependency sparse_dot_mkl (setup.py: sparse_dot_mkl>=0.7), which is not installed here: a
stand-in ``dot_product_mkl(A, B, copy, dense)`` returns the product A B (dense when asked), which
is that function's published contract, computed with SciPy.  Without it the reference falls back
to MultilinearAlgebra.sparse_dot, whose worker processes only pair rows inside their own row
slice (Critical.py:56-84), so its result depends on the number of CPUs and omits products; that
fallback is not what these vectors record.
"""

import os
import sys
import warnings

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
OUT = os.path.join(HERE, 'sparse_init_vectors.npz')


def main(ref_root):
    sys.path.insert(0, HERE)
    from make_golden import _install_stand_ins, _opts
    _install_stand_ins()
    import types
    mkl = types.ModuleType('sparse_dot_mkl')

    def dot_product_mkl(A, B, copy=False, dense=False):
        C = A @ B
        return np.asarray(C.toarray()) if dense and hasattr(C, 'toarray') else C

    mkl.dot_product_mkl = dot_product_mkl
    sys.modules['sparse_dot_mkl'] = mkl
    sys.path.insert(0, os.path.join(ref_root, 'modules'))
    warnings.simplefilter('ignore')
    import TensorFox as tfx
    import TensorFox.Initialization as ini
    import TensorFox.Auxiliar as aux
    import TensorFox.Compression as cmpr
    import TensorFox.Conversion as cnv
    import TensorFox.MultilinearAlgebra as mla
    import TensorFox.Critical as crt

    out = {}

    # --- generator --------------------------------------------------------------------------
    gens = [('g4', [10, 10, 10, 10], 5, 10, 3), ('g3', [30, 20, 12], 4, 12, 5), ('g3b', [9, 40, 17], 3, 9, 11)]
    for name, dims, R, nnz, seed in gens:
        np.random.seed(seed)
        data, idxs, dims_o, _ = tfx.gen_rand_sparse_tensor(dims, R, nnz)
        out[name + '_data'] = np.asarray(data, dtype=np.float64)
        out[name + '_idxs'] = np.asarray(idxs, dtype=np.int64)
        out[name + '_meta'] = np.array([R, nnz, seed] + list(dims))

    # --- sparse Gram, multilinear product, error on fixed inputs -----------------------------
    for name, dims, R, nnz, seed in gens:
        data, idxs = out[name + '_data'], out[name + '_idxs'].astype(np.uint64)
        L = len(dims)
        for l in range(L):
            G = cmpr.sparse_dot_mkl_call(cnv.sparse_unfold(data, idxs, dims, l + 1), 0)
            out['%s_gram%d' % (name, l)] = np.asarray(G.toarray() if hasattr(G, 'toarray') else G,
                                                      dtype=np.float64)
        rs = np.random.RandomState(seed + 100)
        U = [rs.randn(d, min(d, 4)) for d in dims]
        for l in range(L):
            out['%s_U%d' % (name, l)] = U[l]
        out[name + '_core'] = mla.sparse_multilin_mult([u.T for u in U], data, idxs, dims)
        F = [rs.randn(d, R) for d in dims]
        for l in range(L):
            out['%s_F%d' % (name, l)] = F[l]
        out[name + '_fastnorm'] = np.array([crt.sparse_fastnorm(data, idxs, dims, F)])

    # --- seeded sparse cpd trajectories ------------------------------------------------------
    # Unfoldings of these tensors have zero rows, so their Grams have zero eigenvalues, and the
    # randomized SVD returns roundoff-sized singular values for them that clean_compression may
    # keep: how many null directions the core keeps is decided by roundoff.  The 'r' runs fix the
    # core size to the exact rank of every (sorted-mode) unfolding with trunc_dims, which makes
    # the core a property of the tensor.
    def ranks(g):
        data, idxs = out[g + '_data'], out[g + '_idxs'].astype(np.uint64)
        dims = list(out[g + '_meta'][3:])
        order = np.argsort(-np.array(dims))
        sd = [dims[o] for o in order]
        si = idxs[:, order]
        rk = []
        for l in range(len(dims)):
            G = cmpr.sparse_dot_mkl_call(cnv.sparse_unfold(data, si, sd, l + 1), 0)
            w = np.linalg.eigvalsh(np.asarray(G))
            rk.append(int(np.sum(w > 1e-9 * w.max())))
        return rk

    runs = [('s4', 'g4', 5, 0, {}), ('s3', 'g3', 4, 1, {}), ('s3b', 'g3b', 3, 2, {'maxiter': 30}),
            ('s3t', 'g3', 4, 4, {'tol_mlsvd': -1}),
            ('s4r', 'g4', 5, 0, {'trunc_dims': ranks('g4')}), ('s3r', 'g3', 4, 1, {'trunc_dims': ranks('g3')}),
            ('s3br', 'g3b', 3, 2, {'maxiter': 30, 'trunc_dims': ranks('g3b')})]
    for name, g, R, seed, kw in runs:
        if 'trunc_dims' in kw:
            out[name + '_trunc'] = np.array(kw['trunc_dims'])
        data, idxs = out[g + '_data'], out[g + '_idxs'].astype(np.uint64)
        dims = list(out[g + '_meta'][3:])
        np.random.seed(seed)
        f, o = tfx.cpd([data, idxs, dims], R, _opts(tfx, **kw))
        out[name + '_errors'] = np.asarray(o.errors[0], dtype=np.float64)
        out[name + '_rel_error'] = np.array([o.rel_error])
        out[name + '_stop'] = np.array([o.stop[0], o.stop[1]])
        for l in range(len(dims)):
            out['%s_f%d' % (name, l)] = f[l]
        out[name + '_meta'] = np.array([R, seed])

    # --- smart / smart_random on fixed cores -------------------------------------------------
    rs = np.random.RandomState(5)
    cores = {'c1': rs.randn(4, 5, 3), 'c2': rs.randn(6, 4, 4, 3)}
    S = 1e-3 * rs.randn(5, 4, 3)
    for c, v in [((0, 0, 0), 9.0), ((0, 1, 0), -8.0), ((0, 2, 0), 7.5), ((1, 1, 1), 7.0),
                 ((2, 1, 1), -6.5), ((3, 3, 2), 6.0), ((3, 0, 2), 5.5)]:
        S[c] = v
    cores['c3'] = S
    for name, S in cores.items():
        out[name] = S
        for R in (2, 3, 5):
            try:
                F = ini.smart(S, S.shape, R)
                for l, f in enumerate(F):
                    out['%s_smart_R%d_%d' % (name, R, l)] = f
            except Exception as e:       # the reference's own failure for this R
                out['%s_smart_R%d_error' % (name, R)] = np.array([str(type(e).__name__)])
            np.random.seed(R)
            F = ini.smart_random(S, S.shape, R)
            for l, f in enumerate(F):
                out['%s_smartrand_R%d_%d' % (name, R, l)] = f

    # --- inner_method 'direct' (GaussNewton.py:630-710) on small dense tensors -----------------
    dcases = [('d1', (6, 5, 4), 2, 1e-4, 11, {'inner_method': 'direct', 'maxiter': 40}),
              ('d2', (7, 6, 5), 3, 1e-5, 12, {'inner_method': 'direct', 'refine': True, 'maxiter': 30}),
              ('d3', (6, 6, 5), 3, 1e-5, 13, {'inner_method': ['direct', 'cg'] * 15, 'maxiter': 30})]
    for name, dims, R, noise, seed, kw in dcases:
        rs = np.random.RandomState(seed)
        F = [rs.randn(d, R) for d in dims]
        T = cnv.cpd2tens(F)
        T = T + noise * np.linalg.norm(T) * rs.randn(*dims) / np.sqrt(T.size)
        out[name + '_T'] = T
        np.random.seed(seed)
        f, o = tfx.cpd(T, R, _opts(tfx, **kw))
        out[name + '_errors0'] = np.asarray(o.errors[0], dtype=np.float64)
        out[name + '_errors1'] = np.asarray(o.errors[1], dtype=np.float64)
        out[name + '_rel_error'] = np.array([o.rel_error])
        out[name + '_meta'] = np.array([R, seed])
        for l in range(len(dims)):
            out['%s_f%d' % (name, l)] = f[l]

    # --- method 'ttcpd' (TensorFox.py:252-421, 658-817, Auxiliar.py:365-527) -------------------
    tcases = [('tt1', (10, 9, 8, 7), 3, 1e-6, 21, {'method': 'ttcpd'}),
              ('tt2', (6, 6, 5, 5, 4), 2, 1e-6, 22, {'method': 'ttcpd'}),
              ('tt3', (3, 3, 3, 3), 4, 1e-6, 23, {'method': 'ttcpd'}),
              ('tt4', (9, 8, 7, 6), 3, 1e-6, 24, {'method': 'ttcpd', 'epochs': 2}),
              ('tt5', (8, 8, 7, 6), 3, 1e-6, 25, {'method': 'ttcpd', 'bi_method': 'cg'})]
    for name, dims, R, noise, seed, kw in tcases:
        rs = np.random.RandomState(seed)
        F = [rs.randn(d, R) for d in dims]
        T = cnv.cpd2tens(F)
        T = T + noise * np.linalg.norm(T) * rs.randn(*dims) / np.sqrt(T.size)
        out[name + '_T'] = T
        np.random.seed(seed)
        # the raw options class: cpd runs make_options itself, and a second pass would reset
        # bi_method_parameters (only 'bi_method' sets them, Auxiliar.py:312-320)
        raw = type('options', (), dict(kw))
        f, o = tfx.cpd(T, R, raw)
        out[name + '_rel_error'] = np.array([o.rel_error])
        out[name + '_num_steps'] = np.array([o.num_steps])
        out[name + '_meta'] = np.array([R, seed])
        # the stages of the same call: sorted-mode MLSVD core, inflated core, train cores
        np.random.seed(seed)
        opts = tfx.make_options(raw)
        Ts, _ = aux.sort_dims(T)
        S, U, T1, _ = cmpr.mlsvd(Ts, np.linalg.norm(Ts), R, opts)
        out[name + '_S'] = S
        if R > min(S.shape):
            S = cnv.inflate(S, R, S.shape)
        out[name + '_Sinf'] = S
        for i, g in enumerate(tfx.cpdtt(S, R)):
            out['%s_G%d' % (name, i)] = g
        for i, co in enumerate(o.cpd_output):
            out['%s_core%d_errors' % (name, i)] = np.asarray(co.errors[0], dtype=np.float64)
            out['%s_core%d_rel_error' % (name, i)] = np.array([co.rel_error])
        for l in range(len(dims)):
            out['%s_f%d' % (name, l)] = f[l]

    np.savez_compressed(OUT, **out)
    print('wrote', OUT, len(out), 'arrays')


if __name__ == '__main__':
    main(sys.argv[1] if len(sys.argv) > 1 else '/root/reference')
