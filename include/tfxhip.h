/*
 * tfxhip.h -- C ABI of libtfxhip.so, the MI355X (gfx950) kernels behind the dense CPD path of
 * Tensor Fox (reference: modules/TensorFox/TensorFox.py:69 `cpd` -> :424 `tricpd`).
 *
 * The reference has no native boundary: its `cpd` is Python calling NumPy/SciPy/Numba
 * (modules/TensorFox/__init__.py:6 re-exports it).  Each entry point below replaces the CPU
 * function cited next to it; the Python package TensorFox (tensor-fox_amd/TensorFox) keeps the
 * reference's `cpd(T, R, options)` signature and binds these symbols with ctypes.
 *
 * Conventions
 *  - every call returns 0 on success and a negative TFX_E* code on failure; the message is
 *    available from tfx_last_error(ctx); nothing throws across the ABI;
 *  - all array arguments are device pointers owned by the caller; the context owns only
 *    workspace, grown on demand with hipMalloc;
 *  - every call is asynchronous on `stream`; scalar results are written to device memory;
 *  - one context per host thread / stream;
 *  - tensors are dense C-order (last index fastest) with dims[0..L-1], 2 <= L <= TFX_MAXL;
 *  - factor matrices use the reference's flat point layout x = concat_l vec_F(F_l)
 *    (GaussNewton.py:125, Conversion.py:21-49): F_l[i, r] sits at off_l + r*I_l + i with
 *    off_l = R * sum_{m<l} I_m.  "flat" below means a buffer of n = R*sum(I_l) scalars in that
 *    layout;
 *  - dtype: TFX_F32 or TFX_F64 selects the storage / working type of every floating array of a
 *    call, unless the argument says double.
 */
#ifndef TFXHIP_H
#define TFXHIP_H

#include <stdint.h>
#include <hip/hip_runtime_api.h>

#ifdef __cplusplus
extern "C" {
#endif

#define TFX_MAXL 8

enum { TFX_F32 = 0, TFX_F64 = 1 };

/* arithmetic used for the tensor-touching products (MTTKRP, Gram, TTM) */
enum { TFX_PREC_NATIVE = 0,     /* FMA in the storage type */
       TFX_PREC_SPLIT16 = 1 };  /* FP32 storage split into 16-bit hi+lo MFMA operands, FP32 acc */

enum { TFX_OK = 0, TFX_E_ARG = -1, TFX_E_HIP = -2, TFX_E_NOMEM = -3, TFX_E_UNSUPPORTED = -4 };

typedef struct tfx_ctx tfx_ctx;

/* Context and errors (no reference counterpart: the reference reports errors with sys.exit,
 * Auxiliar.py:33-71, which the Python layer reproduces). */
int tfx_ctx_create(int device, tfx_ctx** out);
int tfx_ctx_destroy(tfx_ctx* ctx);
const char* tfx_last_error(const tfx_ctx* ctx);
int tfx_device_info(tfx_ctx* ctx, int* cu_count, int* arch_major, int* arch_minor);

/* MTTKRP for the modes set in mode_mask, in one pass over T.
 * Replaces compute_grad's Khatri-Rao build + dot(Tl[l], M) (GaussNewton.py:411-423).
 * out_flat receives N_l = T_(l) * KR_l in the flat layout (modes not in the mask are untouched).
 * prec: TFX_PREC_NATIVE, or TFX_PREC_SPLIT16 with dtype TFX_F32. */
int tfx_mttkrp(tfx_ctx* ctx, const void* T, int dtype, int L, const int64_t* dims,
               const void* F_flat, int R, unsigned mode_mask, void* out_flat, int prec,
               hipStream_t stream);

/* As tfx_mttkrp with the tensor in t_dtype and factors / outputs in f_dtype (FP32 tensor with FP64
 * factors is the mixed-precision refinement; TFX_PREC_SPLIT16 requires exactly that pair). */
int tfx_mttkrp_mixed(tfx_ctx* ctx, const void* T, int t_dtype, int L, const int64_t* dims,
                     const void* F_flat, int f_dtype, int R, unsigned mode_mask, void* out_flat,
                     int prec, hipStream_t stream);

/* Kernel family tfx_mttkrp_mixed uses for these arguments (host only, no GPU, no context):
 * TFX_PATH_SPLIT16_MFMA (bf16x3 tiles; needs an FP32 tensor whose last dimension K is a multiple
 * of 4 with 64 K 4 < 2^31 bytes), TFX_PATH_F64_MFMA (FP64 tensors whose selected unfoldings have
 * fewer than 2^31 columns) or TFX_PATH_FMA; negative on bad arguments.  No reference
 * counterpart: the reference always multiplies unfoldings with BLAS (GaussNewton.py:423). */
enum { TFX_PATH_SPLIT16_MFMA = 0, TFX_PATH_FMA = 1, TFX_PATH_F64_MFMA = 2 };
int tfx_mttkrp_plan(int t_dtype, int f_dtype, int prec, int L, const int64_t* dims, int R,
                    unsigned mode_mask);

/* The split16 MTTKRP of tfx_mttkrp_mixed (FP32 tensor, FP64 factors and outputs,
 * TFX_PREC_SPLIT16) without any allocation: the brick partials live in a static device array of
 * TFX_NOALLOC_SCRATCH_MB MiB per GPU (held until the process ends) instead of the context's
 * workspace.  No counterpart in the reference; it is for callers that cannot allocate device
 * memory (code under HIP graph capture, serving loops, harnesses that forbid hipMalloc).
 * Shapes outside the tiled kernel (see tfx_mttkrp_plan) or needing more scratch than the array
 * holds return TFX_E_UNSUPPORTED.  Calls on one GPU share the array: calls that overlap on one
 * GPU (different streams or host threads) race. */
#define TFX_NOALLOC_SCRATCH_MB 64
int tfx_mttkrp_noalloc(tfx_ctx* ctx, const void* T, int L, const int64_t* dims,
                       const void* F_flat, int R, unsigned mode_mask, void* out_flat,
                       hipStream_t stream);

/* ||T - [[F]]||_F^2, accumulated in FP64, written to *out_sqnorm (device).
 * Replaces cpd2unfold1 + fastnorm (Conversion.py:117-141, Critical.py:14-22). */
int tfx_cpd_residual(tfx_ctx* ctx, const void* T, int dtype, int L, const int64_t* dims,
                     const void* F_flat, int R, double* out_sqnorm, hipStream_t stream);

/* As tfx_cpd_residual with the factors in f_dtype and the tensor in t_dtype (FP32 tensor with
 * FP64 factors gives the final FP64 error evaluation of the mixed-precision path). */
int tfx_cpd_residual_mixed(tfx_ctx* ctx, const void* T, int t_dtype, int L, const int64_t* dims,
                           const void* F_flat, int f_dtype, int R, double* out_sqnorm,
                           hipStream_t stream);

/* out[0] = sum |x|, out[1] = sum x^2 over n entries (FP64): ||T|| (TensorFox.py:503) and
 * mean|T| for the initial damping (GaussNewton.py:113). */
int tfx_tensor_stats(tfx_ctx* ctx, const void* X, int dtype, int64_t n, double* out,
                     hipStream_t stream);

/* Gr_l = F_l^T F_l, P1_l = o_{k!=l} Gr_k, P2_{l,m} = o_{k!=l,m} Gr_k (L x L x R x R, row-major),
 * and gamma[l, r] = max|P1_{L-1}| * sqrt|P1_l[r, r]| (L x R).
 * Replaces gramians + regularization (GaussNewton.py:520-548, 584-606). */
int tfx_gramians(tfx_ctx* ctx, int dtype, int L, const int64_t* dims, int R, const void* F_flat,
                 void* Gr, void* P1, void* P2, void* gamma, hipStream_t stream);

/* grad_flat = N_flat - F_l P1_l per mode, i.e. minus the gradient of 1/2||T - [[F]]||^2.
 * Replaces the tail of compute_grad and the sign flip in cg (GaussNewton.py:423-425, 339). */
int tfx_neg_grad(tfx_ctx* ctx, int dtype, int L, const int64_t* dims, int R, const void* F_flat,
                 const void* P1, const void* N_flat, void* grad_flat, hipStream_t stream);

/* tfx_mttkrp_mixed over every mode followed by tfx_neg_grad, with the same results: N_flat =
 * T_(l) KR_l and grad_flat = N_flat - F_l P1_l.  For FP64 tensors on the FP64 matrix-core
 * MTTKRP the gradient is formed in the MTTKRP's final sum (one launch fewer).
 * Replaces compute_grad (GaussNewton.py:401-427) with the sign flip of cg (GaussNewton.py:339). */
int tfx_mttkrp_neg_grad(tfx_ctx* ctx, const void* T, int t_dtype, int L, const int64_t* dims,
                        const void* F_flat, int f_dtype, int R, const void* P1, void* N_flat,
                        void* grad_flat, int prec, hipStream_t stream);

/* Split-preconditioned CG on (J^T J + damp Gamma) y = grad, one persistent workgroup.
 * Replaces cg + cg_iterations + matvec + precond (GaussNewton.py:315-398, 551-581, 609-627),
 * including the aliasing of P and residual_cg in the first iteration (GaussNewton.py:349-392).
 * Writes y_flat = M*y, jtjg_flat = J^T J grad, and scal[0] = iterations done, scal[1] = final
 * residual norm^2 (double).  Inner products and norms are accumulated in FP64. */
int tfx_cg(tfx_ctx* ctx, int dtype, int L, const int64_t* dims, int R, const void* F_flat,
           const void* P1, const void* P2, const void* gamma, double damp, const void* grad_flat,
           int maxiter, double tol, void* y_flat, void* jtjg_flat, double* scal,
           hipStream_t stream);

/* Direct inner solver (FP64): forms H = J^T J (n x n, n = R * sum(dims)) from the flat factors
 * and P1 / P2 of tfx_gramians, solves (H + damp diag(Gamma)) y = grad by Cholesky (y = 0 when a
 * pivot is not positive) and writes scal[0] = 0, scal[1] = ||H y - grad||^2.  n <= 4096.
 * Replaces direct + hessian + compute_blocks (GaussNewton.py:630-710); the reference's QR of the
 * row-scaled system M^2 (H + damp diag(Gamma)) has the same solution. */
int tfx_dgn_direct(tfx_ctx* ctx, int L, const int64_t* dims, int R, const double* F_flat,
                   const double* P1, const double* P2, const double* gamma, double damp,
                   const double* grad_flat, double* y_flat, double* scal, hipStream_t stream);

/* x_new = x + alpha*y; F = x2cpd(x_new) with equalize when eq != 0, then transform(symm,
 * factors_norm).  Replaces x = x + y / x2cpd / equalize / transform (GaussNewton.py:285-289,
 * Conversion.py:21-49, 361-387, 412-444).  symm: 0 no symmetric transform; 1 the average of
 * the L blocks (the reference's transform on separate factor arrays); 2 the reference's
 * transform on a factor list whose entries share one array, as every symmetric transform
 * without factors_norm leaves it (Conversion.py:437-438): the last block alone, equalized with
 * L equal column norms, times 2^(L-1)/L, copied to every mode.  scal (double, 4 entries) receives
 * [ (alpha y).(g + damp alpha y), ||alpha y||^2, ||x||^2, max|g| ] for the gain ratio, step size
 * and gradient norm (GaussNewton.py:174, 184-185, 472-487). */
int tfx_dgn_update(tfx_ctx* ctx, int dtype, int L, const int64_t* dims, int R,
                   const void* x_flat, const void* y_flat, double alpha, const void* grad_flat,
                   double damp, int eq, int symm, double factors_norm, void* x_new_flat,
                   void* F_flat, double* scal, hipStream_t stream);

/* ALS mode update F_l = N_l pinv(P1_l) (flat layouts; P1 as written by tfx_gramians, N_l the
 * MTTKRP block of mode `mode`), which equals the reference's T_(l) pinv(KR_l^T)
 * (Alternating_Least_Squares.py:206-231): pinv(KR^T) = KR (KR^T KR)^+ and KR^T KR = P1_l.
 * P1_l (R <= 80) is inverted by Gauss-Jordan elimination in FP64; when it is singular to working
 * precision the pseudo-inverse comes from its Jacobi eigendecomposition.  *status (device int)
 * receives 0, or 1 when the pseudo-inverse path was taken.  Only the mode's block of F_flat is
 * written. */
int tfx_als_solve(tfx_ctx* ctx, int dtype, int L, const int64_t* dims, int R, int mode,
                  const void* P1, const void* N_flat, void* F_flat, int* status,
                  hipStream_t stream);

/* dots[0..3] = [y.y, y.g, g.g, g.Jg] in FP64 (the inner products of dogleg, GaussNewton.py:766-795). */
int tfx_dogleg_dots(tfx_ctx* ctx, int dtype, int64_t n, const void* y, const void* g,
                    const void* jg, double* dots, hipStream_t stream);

/* out = a*y + b*g (the dogleg step y_dl, GaussNewton.py:781-795). */
int tfx_axpby(tfx_ctx* ctx, int dtype, int64_t n, double a, const void* y, double b,
              const void* g, void* out, hipStream_t stream);

/* G = X_(mode) X_(mode)^T (I_mode x I_mode, FP64 output).
 * Replaces the range finder's Gram-type products on unfold(S, l) (Compression.py:147-150,
 * 389-395); the eigenpairs of G give the left singular pairs of the unfolding. */
int tfx_gram_mode(tfx_ctx* ctx, const void* X, int dtype, int L, const int64_t* dims, int mode,
                  double* G, int prec, hipStream_t stream);

/* Top-k eigenpairs of a symmetric n x n FP64 matrix by block subspace iteration from the
 * n x p start block Q0 (column-major), then Rayleigh-Ritz with a Jacobi eigensolver, all FP64.
 * U (n x k, column-major), lam (k, descending).  Sign: the largest-|.| entry of each column of U
 * is made positive (svd_flip with the V-based decision, Compression.py:510-512, 559-563).
 * Replaces randomized_svd's power iterations, QR and small SVD (Compression.py:337-518). */
int tfx_eig_top(tfx_ctx* ctx, const double* G, int n, const double* Q0, int p, int k, int iters,
                double* U, double* lam, hipStream_t stream);

/* Y = X x_mode U^T: Y[.., s, ..] = sum_a U[a, s] X[.., a, ..]; U is I_mode x k column-major
 * FP64; Y has dims with dims[mode] replaced by k, same dtype as X.  sig2 (k, FP64, optional)
 * receives the squared norms of the rows of Y_(mode).
 * Replaces Sl = diag(sigma) Vlt + foldback (Compression.py:153-156) and multilin_mult
 * (MultilinearAlgebra.py:72-105). */
int tfx_ttm(tfx_ctx* ctx, const void* X, int dtype, int L, const int64_t* dims, int mode,
            const double* U, int k, void* Y, double* sig2, int prec, hipStream_t stream);

/* As tfx_ttm with X in x_dtype and Y in y_dtype (FP32 tensor -> FP64 partial core). */
int tfx_ttm_mixed(tfx_ctx* ctx, const void* X, int x_dtype, int L, const int64_t* dims, int mode,
                  const double* U, int k, void* Y, int y_dtype, double* sig2, int prec,
                  hipStream_t stream);

/* C (m x n, column-major, dtype) = op(A) * B with A (FP64, column-major, m x kk or kk x m when
 * transA) and B (dtype, column-major kk x n).  Used for the uncompression U_l W_l
 * (TensorFox.py:573-574) and the projection of a user initialization U_l^T F_l
 * (Initialization.py:58-59). */
int tfx_gemm_small(tfx_ctx* ctx, int dtype, int transA, int m, int n, int kk, const double* A,
                   const void* B, void* C, hipStream_t stream);

/* Element-type conversion between flat buffers (FP64 <-> FP32), used when a stage hands its
 * factors to a stage of another precision. */
int tfx_convert(tfx_ctx* ctx, int64_t n, const void* src, int src_dtype, void* dst, int dst_dtype,
                hipStream_t stream);

/* ---- the dGN loop as one call --------------------------------------------------------------
 * Options of tfx_dgn_run (the reference's dGN options, GaussNewton.py:29-251).  damp0 is the
 * initial damping (init_damp * mean|T|, or init_damp[0] for a schedule); symm / factors_norm as
 * in tfx_dgn_update, `aliased` when the start point already came out of a symmetric transform
 * (the first update then uses the shared-array form). */
typedef struct tfx_dgn_opts {
  int maxiter;
  double tol, tol_step, tol_improv, tol_grad, tol_jump;
  double cg_tol, cg_factor;
  int cg_maxiter;
  double damp0;
  int symm, aliased;
  double factors_norm;
} tfx_dgn_opts;

/* Doubles of device workspace that tfx_dgn_run needs (host-only, no context). */
int64_t tfx_dgn_run_work(int L, const int64_t* dims, int R);

/* Damped Gauss-Newton with the CG inner methods from the flat FP64 start F0 (GaussNewton.py:
 * 29-251, compute_step :254-312, dogleg :713-795, stopping codes 0-7): every iteration runs
 * tfx_gramians, tfx_mttkrp_neg_grad, tfx_cg, tfx_dgn_update and
 * tfx_cpd_residual_mixed on `stream` and the scalar control flow on the host (one 8-double
 * readback per iteration).  inner[it] = 0 for 'cg' (cap 1 + (L - 2) int(cg_factor randint(1 +
 * it^0.4, 2 + it^0.9)) drawn from the MT19937 state mt_key / *mt_pos of NumPy's legacy global
 * generator, advanced in place, one draw per iteration run), 1 for 'cg_static' (cap cg_maxiter);
 * damp_sched: NULL, or maxiter damping values (an init_damp list).  Tsize = ||T||.  Outputs:
 * F_best (device, the best factors), hist (host, 4 x maxiter: step sizes, errors, improvements,
 * gradient norms), info (host: iterations, stop code, dogleg calls).  work: device, at least
 * tfx_dgn_run_work doubles.  Replaces dGN and compute_step (GaussNewton.py:29-312) for inner
 * methods 'cg' / 'cg_static'. */
int tfx_dgn_run(tfx_ctx* ctx, const void* T, int t_dtype, int prec, int L, const int64_t* dims, int R,
                const double* F0, const tfx_dgn_opts* opts, const int8_t* inner, const double* damp_sched,
                uint32_t* mt_key, int* mt_pos, double Tsize, double* work, int64_t work_doubles,
                double* F_best, double* hist, int* info, hipStream_t stream);

/* ---- sparse COO input T = [data, idxs, dims] (TensorFox.py:256-280) ----------------------
 * Entries are handed over as data (nnz, FP64) and coordinates (nnz x L, uint64, row-major, the
 * caller's mode order).  tfx_sparse_pack turns the coordinates into L columns of nnz int32 in
 * the sorted mode order (idx[l * nnz + e] = coords[e * L + ordering[l]]), checking every one
 * against sorted_dims; *bad (device int) is set to 1 when any is out of range (the entry is
 * then stored as 0).  The other sparse entry points take that packed form and assume it was
 * validated.  Replaces the coordinate permutation of sort_dims (Auxiliar.py:99-106) and the
 * index check of sparse_unfold (Conversion.py:293-297). */
int tfx_sparse_pack(tfx_ctx* ctx, int64_t nnz, int L, const uint64_t* coords, const int32_t* ordering,
                    const int64_t* sorted_dims, int32_t* idx, int* bad, hipStream_t stream);

/* G = T_(mode) T_(mode)^T (dims[mode] x dims[mode], FP64, zeroed here): entries sorted by the
 * key of their other coordinates, then every pair in a run of equal keys adds d_a d_b
 * (duplicate coordinates therefore count as their sum).  Replaces sparse_unfold +
 * sparse_dot_mkl_call / sparse_dot (Conversion.py:255-302, Compression.py:211-215, 240-258,
 * MultilinearAlgebra.py:524-571). */
int tfx_sparse_gram(tfx_ctx* ctx, int64_t nnz, int L, const int64_t* dims, const int32_t* idx,
                    const double* data, int mode, double* G, hipStream_t stream);

/* Y = T x_0 U^T: dense C-order (k, dims[1], ..., dims[L-1]) FP64 (zeroed here), U dims[0] x k
 * column-major FP64.  The first mode of sparse_multilin_mult (MultilinearAlgebra.py:145-175,
 * Critical.py:2704-3096); the remaining modes are dense products (tfx_ttm). */
int tfx_sparse_ttm0(tfx_ctx* ctx, int64_t nnz, int L, const int64_t* dims, const int32_t* idx,
                    const double* data, const double* U, int k, double* Y, hipStream_t stream);

/* out[0] = sum_e (data_e - sum_r prod_l F_l[idx_le, r])^2 over the stored entries only; F is
 * the flat FP64 factor layout (factor l column-major dims[l] x R).  Replaces sparse_fastnorm
 * (Critical.py:25-53), squared. */
int tfx_sparse_residual(tfx_ctx* ctx, int64_t nnz, int L, const int64_t* dims, const int32_t* idx,
                        const double* data, const double* F, int R, double* out, hipStream_t stream);

/* ---- batches of small CPDs --------------------------------------------------------------
 * A batch runs many independent problems of uneven shape and rank, each taken through the
 * reference pipeline with default options (TensorFox.py:424-655: sorted modes, MLSVD 'seq',
 * random start on the core, dGN with CG, uncompression, error on the full tensor).  The
 * reference has no batch call: its users loop over `cpd` (TensorFox.py:69).  Members are
 * described by device int64 records of 16 fields, built by the Python layer
 * (TensorFox/batched.py): [0] order L, [1] rank R (<= 16), [2..5] sorted dims (<= 64),
 * [6] tensor offset (elements), [7] U offset, [8] core offset, [9] work-vector offset,
 * [10] work-vector stride, [11] full-factor offset, [12] history offset, [13] MLSVD scratch
 * offset, [14] scratch half size (all in doubles unless stated).  `order` lists member indices
 * in launch order.  Per-member results and dGN state: 16 doubles each (`res`). */
typedef struct tfx_batch_opts {
  int maxiter;
  int symm;
  double tol, tol_step, tol_improv, tol_grad, tol_jump, init_damp, cg_tol, factors_norm;
} tfx_batch_opts;

/* MLSVD of every member (Compression.py:141-156, 192-301 with the Gram eigensolver of
 * tfx_eig_top): one workgroup per member.  Writes the truncated core dims (nbatch x 4 int32),
 * the bases U_l and the cores. */
int tfx_batch_mlsvd(tfx_ctx* ctx, const void* T, int dtype, int nbatch, const int64_t* meta,
                    const int32_t* order, double tol_mlsvd, int32_t* kdims, double* U, double* S,
                    double* W, hipStream_t stream);

/* dGN iterations [it_begin, it_end) of every member on its core (GaussNewton.py:29-312), one
 * wavefront per member; the member's start point is work vector 0 when it_begin == 0, its
 * state is kept in `res` / `work` between calls.  caps: nbatch x maxiter CG iteration caps.
 * lds_bytes: dynamic LDS per member (the largest member's need). */
int tfx_batch_dgn(tfx_ctx* ctx, int L, int nbatch, const int64_t* meta, const int32_t* order,
                  const int32_t* kdims, const double* S, double* work, double* hist, double* res,
                  const int32_t* caps, const tfx_batch_opts* opts, int it_begin, int it_end,
                  int lds_bytes, hipStream_t stream);

/* Host-side (no GPU, no context): every member's random draws from its seed, as NumPy's legacy
 * global generator produces them for one reference `cpd` call -- the MLSVD sketches
 * (Compression.py:337-409, discarded), the starting factors randn(k_l, R) (Initialization.py:
 * 20-94), equalized (Conversion.py:361-387) and scaled to factors_norm when positive, written
 * flat at x0 + x_off[m], and the CG cap of each iteration from the integer bounds
 * randint(cap_low[it], cap_low[it] + cap_rng[it] + 1) (GaussNewton.py:263-266), or cg_maxiter
 * when cg_static.  dims: nbatch x L sorted dims; kdims: nbatch x 4 core dims.  has_zero[m] is
 * set when a starting entry is exactly 0 (the reference then redraws it; the caller handles
 * that member).  Runs on up to nthreads host threads (0: all cores). */
int tfx_batch_draws(int nbatch, int L, const uint32_t* seeds, const int32_t* dims,
                    const int32_t* kdims, const int32_t* ranks, int maxiter, const int64_t* cap_low,
                    const int64_t* cap_rng, double cg_factor, int cg_static, int cg_maxiter,
                    double factors_norm, const int64_t* x_off, double* x0, int32_t* caps,
                    int32_t* has_zero, int nthreads);

/* Uncompression F_l = U_l W_l (TensorFox.py:573-574) of the best core factors and the FP64
 * squared error / squared norm of every member's full tensor (res fields 11, 12). */
int tfx_batch_finish(tfx_ctx* ctx, const void* T, int dtype, int L, int nbatch,
                     const int64_t* meta, const int32_t* order, const int32_t* kdims,
                     const double* U, const double* work, double* Ffull, double* res,
                     hipStream_t stream);

#ifdef __cplusplus
}
#endif
#endif
