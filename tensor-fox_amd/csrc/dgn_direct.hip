// Direct inner solver of the damped Gauss-Newton step (inner_method 'direct', GaussNewton.py:
// 630-710): the Gauss-Newton matrix H = J^T J is formed explicitly from the factors and their
// Gram products, and (H + damp diag(Gamma)) y = grad is solved by Cholesky.  The reference
// solves the row-scaled system M^2 (H + damp diag(Gamma)) y = M^2 grad by QR; the scaling does
// not change the solution, and H + damp diag(Gamma) is symmetric positive definite (Gamma > 0),
// so the two agree to the conditioning of the system.  As in the reference this is for small
// problems: n = R * sum(dims) unknowns, an n x n matrix, O(n^3) work in one workgroup.
//
// H in the flat factor layout (index of (mode l, rank a, row i) = off_l + a I_l + i):
//   H[(l,a,i), (l,b,i')] = delta_{i i'} P1_l[a, b]
//   H[(l,a,i), (m,b,j)]  = P2_{l,m}[a, b] F_l[i, b] F_m[j, a]          (l != m)
// (hessian and compute_blocks, GaussNewton.py:666-710).

#include "common.h"

namespace {

constexpr int kThreads = 1024;
constexpr int kMaxN = 4096;

__global__ __launch_bounds__(256) void k_hessian(Shape s, const double* __restrict__ F, const double* __restrict__ P1,
                                                 const double* __restrict__ P2, const double* __restrict__ gamma,
                                                 double damp, double* __restrict__ H, double* __restrict__ A) {
  const int64_t n = s.n, total = n * n;
  const int R = s.R, L = s.L;
  for (int64_t e = blockIdx.x * 256 + threadIdx.x; e < total; e += (int64_t)gridDim.x * 256) {
    const int64_t row = e / n, col = e - row * n;
    int l = 0, m = 0;
    while (l + 1 < L && row >= s.off[l + 1]) ++l;
    while (m + 1 < L && col >= s.off[m + 1]) ++m;
    const int64_t ro = row - s.off[l], co = col - s.off[m];
    const int a = (int)(ro / s.d[l]), b = (int)(co / s.d[m]);
    const int64_t i = ro - (int64_t)a * s.d[l], j = co - (int64_t)b * s.d[m];
    double h;
    if (l == m) {
      h = i == j ? P1[((int64_t)l * R + a) * R + b] : 0.0;
    } else {
      h = P2[(((int64_t)l * L + m) * R + a) * R + b] * F[s.off[l] + (int64_t)b * s.d[l] + i] *
          F[s.off[m] + (int64_t)a * s.d[m] + j];
    }
    H[e] = h;
    A[e] = row == col ? h + damp * gamma[(int64_t)l * R + a] : h;
  }
}

// In-place Cholesky A = C C^T of the row-major n x n matrix (lower triangle of A becomes C),
// left-looking, one column per step; then y = A^{-1} g by the two triangular solves.  A pivot
// that is not positive (the matrix is not numerically positive definite) gives y = 0, as the
// reference's solve failure does (GaussNewton.py:662-667).
__global__ __launch_bounds__(kThreads) void k_chol_solve(double* __restrict__ A, int n, const double* __restrict__ g,
                                                         double* __restrict__ y) {
  __shared__ double rowj[kMaxN];
  __shared__ int fail;
  __shared__ double piv;
  const int t = threadIdx.x;
  if (t == 0) fail = 0;
  __syncthreads();
  for (int j = 0; j < n; ++j) {
    for (int k = t; k < j; k += kThreads) rowj[k] = A[(int64_t)j * n + k];
    __syncthreads();
    // v_i = A[i][j] - sum_{k<j} C[i][k] C[j][k], i >= j
    for (int i = j + t; i < n; i += kThreads) {
      const double* ci = A + (int64_t)i * n;
      double v = ci[j];
      for (int k = 0; k < j; ++k) v -= ci[k] * rowj[k];
      if (i == j) {
        if (!(v > 0.0)) fail = 1;
        piv = v > 0.0 ? sqrt(v) : 1.0;
      }
      A[(int64_t)i * n + j] = v;
    }
    __syncthreads();
    if (fail) break;
    const double p = piv;
    for (int i = j + t; i < n; i += kThreads) {
      double* cij = A + (int64_t)i * n + j;
      *cij = i == j ? p : *cij / p;
    }
    __syncthreads();
  }
  if (fail) {
    for (int i = t; i < n; i += kThreads) y[i] = 0.0;
    return;
  }
  // forward: C z = g, column-oriented (z in y)
  for (int i = t; i < n; i += kThreads) y[i] = g[i];
  __syncthreads();
  for (int j = 0; j < n; ++j) {
    const double zj = y[j] / A[(int64_t)j * n + j];
    __syncthreads();
    if (t == 0) y[j] = zj;
    for (int i = j + 1 + t; i < n; i += kThreads) y[i] -= A[(int64_t)i * n + j] * zj;
    __syncthreads();
  }
  // backward: C^T y = z
  for (int j = n - 1; j >= 0; --j) {
    const double yj = y[j] / A[(int64_t)j * n + j];
    __syncthreads();
    if (t == 0) y[j] = yj;
    for (int i = t; i < j; i += kThreads) y[i] -= A[(int64_t)j * n + i] * yj;
    __syncthreads();
  }
}

// scal[1] = ||H y - g||^2 (one workgroup; H row-major n x n)
__global__ __launch_bounds__(kThreads) void k_direct_residual(const double* __restrict__ H, int n,
                                                              const double* __restrict__ y,
                                                              const double* __restrict__ g,
                                                              double* __restrict__ scal) {
  __shared__ double red[kThreads / 64];
  double acc = 0.0;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  for (int i = wave; i < n; i += kThreads / 64) {
    const double* hi = H + (int64_t)i * n;
    double d = 0.0;
    for (int k = lane; k < n; k += 64) d += hi[k] * y[k];
    d = wave_sum(d);
    const double r = d - g[i];
    if (lane == 0) acc += r * r;
  }
  acc = block_sum(acc, red);
  if (threadIdx.x == 0) {
    scal[0] = 0.0;
    scal[1] = acc;
  }
}

}  // namespace

extern "C" int tfx_dgn_direct(tfx_ctx* ctx, int L, const int64_t* dims, int R, const double* F_flat,
                              const double* P1, const double* P2, const double* gamma, double damp,
                              const double* grad_flat, double* y_flat, double* scal, hipStream_t stream) {
  if (!ctx || !dims || !F_flat || !P1 || !P2 || !gamma || !grad_flat || !y_flat || !scal)
    return tfx_set_error(ctx, TFX_E_ARG, "null argument");
  Shape s;
  int rc = tfx_make_shape(ctx, L, dims, R, &s);
  if (rc) return rc;
  if (s.n > kMaxN)
    return tfx_set_error(ctx, TFX_E_UNSUPPORTED, "direct solver limited to R * sum(dims) <= %d (got %lld)", kMaxN,
                         (long long)s.n);
  const int n = (int)s.n;
  const size_t mat = align256(sizeof(double) * (size_t)n * n);
  void* wsp;
  rc = tfx_workspace(ctx, 2 * mat, stream, &wsp);
  if (rc) return rc;
  double* H = (double*)wsp;
  double* A = (double*)((char*)wsp + mat);
  const int64_t total = (int64_t)n * n;
  const int nb = (int)std::min<int64_t>((total + 255) / 256, (int64_t)ctx->cu_count * 16);
  hipLaunchKernelGGL(k_hessian, dim3(nb), dim3(256), 0, stream, s, F_flat, P1, P2, gamma, damp, H, A);
  TFX_CHECK_LAUNCH(ctx);
  hipLaunchKernelGGL(k_chol_solve, dim3(1), dim3(kThreads), 0, stream, A, n, grad_flat, y_flat);
  TFX_CHECK_LAUNCH(ctx);
  hipLaunchKernelGGL(k_direct_residual, dim3(1), dim3(kThreads), 0, stream, H, n, y_flat, grad_flat, scal);
  TFX_CHECK_LAUNCH(ctx);
  return TFX_OK;
}
