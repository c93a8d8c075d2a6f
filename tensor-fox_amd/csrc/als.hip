// Alternating least squares mode update F_l <- N_l pinv(P1_l) on the GPU.
//
// The reference's ALS sweep (Alternating_Least_Squares.py:206-231) forms the Khatri-Rao product
// KR_l of the other factors and sets F_l = T_(l) pinv(KR_l^T).  Because pinv(KR^T) =
// KR (KR^T KR)^+ and KR^T KR is the Hadamard product of the other factors' Grams (P1_l of
// GaussNewton.py:520-548), the update is N_l P1_l^+ with N_l = T_(l) KR_l the MTTKRP: the
// tensor is read by the MTTKRP kernels and only an R x R matrix is inverted here.
//
// k_als_pinv: one workgroup inverts P1_l in LDS by Gauss-Jordan elimination without pivoting
// (P1_l is symmetric positive semidefinite).  A pivot at or below R * eps * max(diag) means
// P1_l is singular to working precision; the workgroup then takes the pseudo-inverse from a
// Jacobi eigendecomposition, dropping eigenvalues at or below R * eps * lambda_max (the level
// that rounding of the R x R Gram product can resolve; the reference's SVD of KR^T drops
// singular values below max(R, K) * eps * s_max).
// k_als_apply: F_l[i, r] = sum_k N_l[i, k] Pinv[k, r], one thread per output entry.
#include "common.h"
#include "jacobi.h"

namespace {

constexpr int ALS_MAXR = 80;   // 2 R (R + 1) doubles of LDS for the fallback

template <typename S>
__global__ __launch_bounds__(256) void k_als_pinv(const S* __restrict__ P, int R,
                                                   double* __restrict__ Pinv,
                                                   int* __restrict__ status) {
  extern __shared__ __align__(16) double sm[];
  __shared__ double red[4];
  __shared__ double cs[2 * (ALS_MAXR + 2)];
  const int ld = R | 1;            // odd row pitch: column reads hit distinct banks
  double* A = sm;                  // R x ld
  double* V = A + R * ld;          // R x R (fallback eigenvectors, column-major)
  double* colk = V + R * R;        // R
  double* rowk = colk + R;         // R
  const int t = threadIdx.x, nt = blockDim.x;
  double dmax = 0.0;
  for (int e = t; e < R * R; e += nt) {
    const int i = e / R, j = e - (e / R) * R;
    const double v = (double)P[e];
    A[i * ld + j] = v;
    if (i == j) dmax = fabs(v) > dmax ? fabs(v) : dmax;
  }
  dmax = block_max(dmax, red);
  const double tol = (double)R * 2.220446049250313e-16 * dmax;
  bool ok = dmax > 0.0;
  for (int k = 0; k < R && ok; ++k) {
    __syncthreads();
    const double p = A[k * ld + k];
    if (!(p > tol)) {
      ok = false;
      break;
    }
    for (int i = t; i < R; i += nt) {
      colk[i] = A[i * ld + k];
      rowk[i] = A[k * ld + i];
    }
    __syncthreads();
    const double ip = 1.0 / p;
    for (int e = t; e < R * R; e += nt) {
      const int i = e / R, j = e - (e / R) * R;
      double v;
      if (i == k) v = j == k ? ip : rowk[j] * ip;
      else if (j == k) v = -colk[i] * ip;
      else v = A[i * ld + j] - colk[i] * (rowk[j] * ip);
      A[i * ld + j] = v;
    }
  }
  __syncthreads();
  if (ok) {
    for (int e = t; e < R * R; e += nt) {
      const int i = e / R, j = e - (e / R) * R;
      Pinv[e] = A[i * ld + j];
    }
    if (t == 0) *status = 0;
    return;
  }
  // numerically singular: pseudo-inverse from the eigendecomposition
  for (int e = t; e < R * R; e += nt) {
    const int i = e / R, j = e - (e / R) * R;
    A[i * ld + j] = (double)P[e];
  }
  __syncthreads();
  tfxjac::wg_jacobi(A, R, ld, V, cs, red);
  __syncthreads();
  double lmax = 0.0;
  for (int k = t; k < R; k += nt) lmax = A[k * ld + k] > lmax ? A[k * ld + k] : lmax;
  lmax = block_max(lmax, red);
  const double cut = (double)R * 2.220446049250313e-16 * lmax;
  for (int k = t; k < R; k += nt) {
    const double lam = A[k * ld + k];
    colk[k] = lam > cut ? 1.0 / lam : 0.0;
  }
  __syncthreads();
  for (int e = t; e < R * R; e += nt) {
    const int i = e / R, j = e - (e / R) * R;
    double acc = 0.0;
    for (int k = 0; k < R; ++k) acc += V[k * R + i] * colk[k] * V[k * R + j];
    Pinv[e] = acc;
  }
  if (t == 0) *status = 1;
}

template <typename S>
__global__ __launch_bounds__(256) void k_als_apply(const S* __restrict__ N,
                                                    const double* __restrict__ Pinv, int64_t I,
                                                    int R, S* __restrict__ F) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= I * R) return;
  const int64_t r = e / I, i = e - r * I;
  double acc = 0.0;
  for (int k = 0; k < R; ++k) acc += (double)N[(int64_t)k * I + i] * Pinv[(int64_t)k * R + r];
  F[e] = (S)acc;
}

}  // namespace

extern "C" int tfx_als_solve(tfx_ctx* ctx, int dtype, int L, const int64_t* dims, int R, int mode,
                             const void* P1, const void* N_flat, void* F_flat, int* status,
                             hipStream_t stream) {
  if (!ctx || !P1 || !N_flat || !F_flat || !status)
    return tfx_set_error(ctx, TFX_E_ARG, "null argument");
  Shape s;
  int rc = tfx_make_shape(ctx, L, dims, R, &s);
  if (rc) return rc;
  if (mode < 0 || mode >= L) return tfx_set_error(ctx, TFX_E_ARG, "mode %d outside [0, %d)", mode, L);
  if (R > ALS_MAXR)
    return tfx_set_error(ctx, TFX_E_UNSUPPORTED, "ALS update supports rank <= %d (got %d)", ALS_MAXR, R);
  void* ws;
  rc = tfx_workspace(ctx, sizeof(double) * (size_t)R * R, stream, &ws);
  if (rc) return rc;
  double* Pinv = (double*)ws;
  const size_t lds = sizeof(double) * ((size_t)R * (R | 1) + (size_t)R * R + 2 * (size_t)R);
  const int64_t I = s.d[mode];
  const int64_t nout = I * R;
  const unsigned blocks = (unsigned)((nout + 255) / 256);
  if (dtype == TFX_F64) {
    TFX_HIP(ctx, hipFuncSetAttribute((const void*)k_als_pinv<double>,
                                     hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    hipLaunchKernelGGL(k_als_pinv<double>, dim3(1), dim3(256), lds, stream,
                       (const double*)P1 + (size_t)mode * R * R, R, Pinv, status);
    TFX_CHECK_LAUNCH(ctx);
    hipLaunchKernelGGL(k_als_apply<double>, dim3(blocks), dim3(256), 0, stream,
                       (const double*)N_flat + s.off[mode], (const double*)Pinv, I, R,
                       (double*)F_flat + s.off[mode]);
  } else if (dtype == TFX_F32) {
    TFX_HIP(ctx, hipFuncSetAttribute((const void*)k_als_pinv<float>,
                                     hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    hipLaunchKernelGGL(k_als_pinv<float>, dim3(1), dim3(256), lds, stream,
                       (const float*)P1 + (size_t)mode * R * R, R, Pinv, status);
    TFX_CHECK_LAUNCH(ctx);
    hipLaunchKernelGGL(k_als_apply<float>, dim3(blocks), dim3(256), 0, stream,
                       (const float*)N_flat + s.off[mode], (const double*)Pinv, I, R,
                       (float*)F_flat + s.off[mode]);
  } else {
    return tfx_set_error(ctx, TFX_E_ARG, "bad dtype %d", dtype);
  }
  TFX_CHECK_LAUNCH(ctx);
  return TFX_OK;
}
