// MLSVD building blocks (Compression.py:48-237): Gram matrix of a mode unfolding, top-k FP64
// eigenpairs by subspace iteration + Rayleigh-Ritz (parallel cyclic Jacobi), the mode product
// Y = X x_mode U^T with the squared row norms of its unfolding, and a small GEMM for the
// uncompression U_l W_l.  The Gram and mode-product kernels here are the FMA forms for FP32 data
// at native precision; FP64 data and the split precision run the MFMA kernels of mlsvd_mfma.hip.
//
// A tensor X with dims d[0..L-1] is viewed around `mode` as (P, I, Q):
//   P = prod_{m < mode} d[m], I = d[mode], Q = prod_{m > mode} d[m],
// element (p, a, q) at (p*I + a)*Q + q.  The unfolding X_(mode) has rows a and columns (p, q).
#include "common.h"
#include "jacobi.h"

namespace {

struct PIQ {
  int64_t P, I, Q;
};

PIQ piq_of(const Shape& s, int mode) {
  PIQ v{1, s.d[mode], 1};
  for (int m = 0; m < mode; ++m) v.P *= s.d[m];
  for (int m = mode + 1; m < s.L; ++m) v.Q *= s.d[m];
  return v;
}

constexpr int GT = 32;   // Gram output tile
constexpr int GK = 32;   // reduction chunk

// G[a][b] += sum over a chunk of (p, q) of X[p, a, q] X[p, b, q]; one workgroup per
// (tile pair with tb >= ta, k-split); FP64 atomics into G (zeroed by the caller).
template <typename S>
__global__ __launch_bounds__(256) void k_gram_mode(const S* __restrict__ X, PIQ v, int ntile,
                                                    int64_t kchunk, double* __restrict__ G) {
  __shared__ S Xa[GK][GT + 1];
  __shared__ S Xb[GK][GT + 1];
  // decode (ta, tb) from the triangular pair index
  int pair = blockIdx.x;
  int ta = 0;
  while (pair >= ntile - ta) {
    pair -= ntile - ta;
    ++ta;
  }
  const int tb = ta + pair;
  const int64_t KT = v.P * v.Q;
  const int64_t kb = (int64_t)blockIdx.y * kchunk;
  const int64_t ke = min(KT, kb + kchunk);
  const int t = threadIdx.x;
  const int ox = t % 16, oy = t / 16;   // each thread: outputs (2 ox + {0,1}, 2 oy + {0,1})
  double acc[2][2] = {{0, 0}, {0, 0}};
  S sacc[2][2] = {{0, 0}, {0, 0}};
  for (int64_t k0 = kb; k0 < ke; k0 += GK) {
    __syncthreads();
    for (int e = t; e < GK * GT; e += 256) {
      int kk, a;
      if (v.Q >= GK) {        // k contiguous: consecutive threads walk k
        kk = e % GK;
        a = e / GK;
      } else {                // a contiguous when Q == 1
        a = e % GT;
        kk = e / GT;
      }
      const int64_t k = k0 + kk;
      S va = S(0), vb = S(0);
      if (k < ke) {
        const int64_t p = k / v.Q, q = k % v.Q;
        const int64_t ra = (int64_t)ta * GT + a, rb = (int64_t)tb * GT + a;
        if (ra < v.I) va = X[(p * v.I + ra) * v.Q + q];
        if (rb < v.I) vb = X[(p * v.I + rb) * v.Q + q];
      }
      Xa[kk][a] = va;
      Xb[kk][a] = vb;
    }
    __syncthreads();
#pragma unroll 8
    for (int kk = 0; kk < GK; ++kk) {
      const S a0 = Xa[kk][2 * oy], a1 = Xa[kk][2 * oy + 1];
      const S b0 = Xb[kk][2 * ox], b1 = Xb[kk][2 * ox + 1];
      sacc[0][0] += a0 * b0;
      sacc[0][1] += a0 * b1;
      sacc[1][0] += a1 * b0;
      sacc[1][1] += a1 * b1;
    }
    // fold into FP64 every chunk so that long reductions keep FP32 partials short
    for (int i = 0; i < 2; ++i)
      for (int j = 0; j < 2; ++j) {
        acc[i][j] += (double)sacc[i][j];
        sacc[i][j] = S(0);
      }
  }
  for (int i = 0; i < 2; ++i)
    for (int j = 0; j < 2; ++j) {
      const int64_t ra = (int64_t)ta * GT + 2 * oy + i, rb = (int64_t)tb * GT + 2 * ox + j;
      if (ra < v.I && rb < v.I) {
        atomicAdd(&G[ra * v.I + rb], acc[i][j]);
        if (ta != tb) atomicAdd(&G[rb * v.I + ra], acc[i][j]);
      }
    }
}

// Y[p, s, q] = sum_a U[a + s*I] X[p, a, q]
template <typename S, typename SY>
__global__ void k_ttm(const S* __restrict__ X, PIQ v, const double* __restrict__ U, int k,
                      SY* __restrict__ Y) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t tot = v.P * k * v.Q;
  if (e >= tot) return;
  const int64_t q = e % v.Q;
  const int64_t sidx = (e / v.Q) % k;
  const int64_t p = e / (v.Q * k);
  const double* u = U + sidx * v.I;
  const S* x = X + p * v.I * v.Q + q;
  double acc = 0.0;
  S sacc = S(0);
  for (int64_t a = 0; a < v.I; ++a) {
    sacc += (S)u[a] * x[a * v.Q];
    if ((a & 31) == 31) {
      acc += (double)sacc;
      sacc = S(0);
    }
  }
  acc += (double)sacc;
  Y[e] = (SY)acc;
}

// sig2[s] = sum_{p, q} Y[p, s, q]^2 (FP64)
template <typename S>
__global__ __launch_bounds__(256) void k_rownorm2(const S* __restrict__ Y, PIQ v,
                                                   double* __restrict__ sig2) {
  __shared__ double red[4];
  const int64_t sidx = blockIdx.x;
  double acc = 0.0;
  for (int64_t e = threadIdx.x; e < v.P * v.Q; e += blockDim.x) {
    const int64_t p = e / v.Q, q = e % v.Q;
    const double y = (double)Y[(p * v.I + sidx) * v.Q + q];
    acc += y * y;
  }
  acc = block_sum(acc, red);
  if (threadIdx.x == 0) sig2[sidx] = acc;
}

// ---------------------------------------------------------------------------------------------
// FP64 eigensolver, one workgroup.
// ---------------------------------------------------------------------------------------------

// C (n x p, col-major) = G (n x n, row-major symmetric) * B (n x p, col-major)
__device__ void wg_symm_mul(const double* G, int n, const double* B, int p, double* C) {
  for (int e = threadIdx.x; e < n * p; e += blockDim.x) {
    const int i = e % n, j = e / n;
    const double* g = G + (int64_t)i * n;
    const double* b = B + (int64_t)j * n;
    double acc = 0.0;
    for (int k = 0; k < n; ++k) acc += g[k] * b[k];
    C[e] = acc;
  }
  __syncthreads();
}

// z_j -= sum_{i<j} (z_i . z_j) z_i over the finished columns, applied twice.  The j inner
// products are computed one wavefront each (no workgroup barrier per product), then subtracted
// in one pass.  `dots` holds >= p doubles.
__device__ void wg_project_twice(double* Z, int n, int j, double* dots) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, nwave = blockDim.x >> 6;
  double* zj = Z + (int64_t)j * n;
  for (int pass = 0; pass < 2 && j > 0; ++pass) {
    for (int i = wave; i < j; i += nwave) {
      const double* zi = Z + (int64_t)i * n;
      double a = 0.0;
      for (int e = lane; e < n; e += 64) a += zi[e] * zj[e];
      a = wave_sum(a);
      if (lane == 0) dots[i] = a;
    }
    __syncthreads();
    for (int e = threadIdx.x; e < n; e += blockDim.x) {
      double v = zj[e];
      for (int i = 0; i < j; ++i) v -= dots[i] * Z[(int64_t)i * n + e];
      zj[e] = v;
    }
    __syncthreads();
  }
}

// Row r minimizing sum_{i<j} Z[i][r]^2 (the coordinate direction least covered by the finished
// columns; the smallest r among equal values), returned in every thread.  For j < n orthonormal
// columns that sum is at most j/n on average, so e_r keeps at least 1 - j/n of its norm.
__device__ int wg_least_covered_row(const double* Z, int n, int j, double* red) {
  __shared__ int ired[16];
  double best = 1e300;
  int arg = n;
  for (int r = threadIdx.x; r < n; r += blockDim.x) {
    double c = 0.0;
    for (int i = 0; i < j; ++i) c += Z[(int64_t)i * n + r] * Z[(int64_t)i * n + r];
    if (c < best) { best = c; arg = r; }
  }
  for (int o = 32; o > 0; o >>= 1) {
    const double ob = __shfl_xor(best, o, 64);
    const int oa = __shfl_xor(arg, o, 64);
    if (ob < best || (ob == best && oa < arg)) { best = ob; arg = oa; }
  }
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63, nw = blockDim.x >> 6;
  __syncthreads();
  if (lane == 0) { red[w] = best; ired[w] = arg; }
  __syncthreads();
  best = red[0];
  arg = ired[0];
  for (int i = 1; i < nw; ++i)
    if (red[i] < best || (red[i] == best && ired[i] < arg)) { best = red[i]; arg = ired[i]; }
  __syncthreads();
  return arg;
}

// Classical Gram-Schmidt applied twice (CGS2) on the p columns of Z (n x p col-major, p < n), in
// place.  A column that the two projections leave with less than 1e-12 of its norm lies in the
// span of the finished columns up to rounding (the block is rank deficient, e.g. power steps on a
// Gram matrix of rank below p, or with rows that are exactly zero): it is replaced by the unit
// vector of the least covered row, projected twice and normalized, so that the block is
// orthonormal whatever its rank, as a Householder QR makes it.
__device__ void wg_orthonormalize(double* Z, int n, int p, double* red, double* dots) {
  for (int j = 0; j < p; ++j) {
    double* zj = Z + (int64_t)j * n;
    double a0 = 0.0;
    for (int e = threadIdx.x; e < n; e += blockDim.x) a0 += zj[e] * zj[e];
    a0 = block_sum(a0, red);
    wg_project_twice(Z, n, j, dots);
    double a = 0.0;
    for (int e = threadIdx.x; e < n; e += blockDim.x) a += zj[e] * zj[e];
    a = block_sum(a, red);
    if (!(a > 1e-24 * a0) && j < n) {
      const int r = wg_least_covered_row(Z, n, j, red);
      for (int e = threadIdx.x; e < n; e += blockDim.x) zj[e] = e == r ? 1.0 : 0.0;
      __syncthreads();
      wg_project_twice(Z, n, j, dots);
      a = 0.0;
      for (int e = threadIdx.x; e < n; e += blockDim.x) a += zj[e] * zj[e];
      a = block_sum(a, red);
    }
    const double nr = sqrt(a);
    const double inv = nr > 0.0 ? 1.0 / nr : 0.0;
    for (int e = threadIdx.x; e < n; e += blockDim.x) zj[e] *= inv;
    __syncthreads();
  }
}

// Orthonormalize the p columns of Z (n x p, col-major, global) in place with CGS2.
__global__ __launch_bounds__(1024) void k_orth(double* Z, int n, int p) {
  __shared__ double red[16];
  __shared__ double dots[512];
  wg_orthonormalize(Z, n, p, red, dots);
}

// Block classical Gram-Schmidt with reorthogonalization (BCGS2) on the p columns of Z (n x p,
// column-major), in place, one workgroup.  Columns go in panels of OPB: a panel is staged in LDS,
// projected twice against all finished columns (each finished column is read once per pass:
// H = Q^T Z_P, then Z_P -= Q H), then orthonormalized within itself by CGS2 on a single
// wavefront (each lane owns fixed rows of the panel, inner products are wave reductions, so no
// workgroup barrier is needed there), and written back.  When projections and in-panel step
// together remove all but a fraction 1e-3 of a column's norm (the block is numerically rank
// deficient there), what is left of that column carries rounding components along the finished
// columns that normalization would magnify: such a panel gets one more projection against the
// finished columns and one more in-panel pass ("twice is enough": Parlett, The Symmetric
// Eigenvalue Problem, sec. 6-9).  The in-panel CGS2 costs p * OPB sequential wave reductions and the
// projections one pass over the finished columns per panel; on 400 x 60 blocks panels of 4
// measured fastest (8: 590 µs, 4: 450 µs, 2: 440 µs per call).
constexpr int OPB = 4;
constexpr int ONU = 8;    // rows per lane of the register-resident in-panel CGS2 (n <= 512)
constexpr int ORTH_MAXN = 2048;   // n * OPB doubles of LDS

__global__ __launch_bounds__(1024) void k_orth_block(double* Z, int n, int p) {
  extern __shared__ __align__(16) unsigned char smem[];
  __shared__ int redo_s;
  __shared__ int collapse_s;   // a column still dependent after the extra round
  __shared__ double red_s[16];
  if (threadIdx.x == 0) collapse_s = 0;
  __syncthreads();
  double* zp = (double*)smem;                 // [OPB][n]
  double* H = zp + (size_t)OPB * n;           // [p][OPB]
  const int t = threadIdx.x, nt = blockDim.x;
  const int lane = t & 63, wave = t >> 6, nwave = nt >> 6;
  for (int j0 = 0; j0 < p; j0 += OPB) {
    const int pb = p - j0 < OPB ? p - j0 : OPB;
    for (int u = t; u < pb * n; u += nt) zp[u] = Z[(int64_t)j0 * n + u];
    __syncthreads();
    for (int round = 0; round < 2; ++round) {
      double pre[OPB];   // wave 0: squared column norms before this round's projections
      if (wave == 0 && n <= 64 * ONU) {
        double zv[OPB][ONU];   // all of the lane's reads in flight together
#pragma unroll
        for (int c = 0; c < OPB; ++c)
#pragma unroll
          for (int u = 0; u < ONU; ++u) {
            const int e = lane + 64 * u;
            zv[c][u] = (c < pb && e < n) ? zp[c * n + e] : 0.0;
          }
#pragma unroll
        for (int c = 0; c < OPB; ++c) {
          double a = 0.0;
#pragma unroll
          for (int u = 0; u < ONU; ++u) a += zv[c][u] * zv[c][u];
          pre[c] = wave_sum(a);
        }
      } else if (wave == 0) {
#pragma unroll
        for (int c = 0; c < OPB; ++c) {
          double a = 0.0;
          if (c < pb)
            for (int e = lane; e < n; e += 64) a += zp[c * n + e] * zp[c * n + e];
          pre[c] = wave_sum(a);
        }
      }
      for (int pass = 0; pass < 2 - round && j0 > 0; ++pass) {
        // H[i][c] = Q_i . zp_c: one wavefront per finished column i
        for (int i = wave; i < j0; i += nwave) {
          const double* qi = Z + (int64_t)i * n;
          double acc[OPB];
#pragma unroll
          for (int c = 0; c < OPB; ++c) acc[c] = 0.0;
#pragma unroll 4
          for (int e = lane; e < n; e += 64) {
            const double q = qi[e];
#pragma unroll
            for (int c = 0; c < OPB; ++c)
              if (c < pb) acc[c] += q * zp[c * n + e];
          }
#pragma unroll
          for (int c = 0; c < OPB; ++c) {
            const double v = wave_sum(acc[c]);
            if (lane == 0 && c < pb) H[i * OPB + c] = v;
          }
        }
        __syncthreads();
        // zp_c -= sum_i Q_i H[i][c] (eight finished columns' loads in flight at a time)
        for (int e = t; e < n; e += nt) {
          double acc[OPB];
#pragma unroll
          for (int c = 0; c < OPB; ++c) acc[c] = c < pb ? zp[c * n + e] : 0.0;
          int i = 0;
          for (; i + 8 <= j0; i += 8) {
            double q[8];
#pragma unroll
            for (int u = 0; u < 8; ++u) q[u] = Z[(int64_t)(i + u) * n + e];
#pragma unroll
            for (int u = 0; u < 8; ++u)
#pragma unroll
              for (int c = 0; c < OPB; ++c) acc[c] -= q[u] * H[(i + u) * OPB + c];
          }
          for (; i < j0; ++i) {
            const double q = Z[(int64_t)i * n + e];
#pragma unroll
            for (int c = 0; c < OPB; ++c) acc[c] -= q * H[i * OPB + c];
          }
#pragma unroll
          for (int c = 0; c < OPB; ++c)
            if (c < pb) zp[c * n + e] = acc[c];
        }
        __syncthreads();
      }
      // CGS2 inside the panel on wavefront 0 (lane-owned rows e = lane + 64 u); for n <= 64 ONU
      // the panel stays in registers for the whole pass (same operations in the same order)
      if (wave == 0 && n <= 64 * ONU) {
        double zr[OPB][ONU];
#pragma unroll
        for (int c = 0; c < OPB; ++c)
#pragma unroll
          for (int u = 0; u < ONU; ++u) {
            const int e = lane + 64 * u;
            zr[c][u] = (c < pb && e < n) ? zp[c * n + e] : 0.0;
          }
        bool redo = false;
#pragma unroll
        for (int c = 0; c < OPB; ++c) {
          if (c >= pb) break;
          for (int pass = 0; pass < 2 && c > 0; ++pass) {
            double d[OPB];
#pragma unroll
            for (int i = 0; i < OPB; ++i) {
              double a = 0.0;
              if (i < c) {
#pragma unroll
                for (int u = 0; u < ONU; ++u) a += zr[i][u] * zr[c][u];
              }
              d[i] = i < c ? wave_sum(a) : 0.0;
            }
#pragma unroll
            for (int u = 0; u < ONU; ++u) {
              double v = zr[c][u];
#pragma unroll
              for (int i = 0; i < OPB; ++i)
                if (i < c) v -= d[i] * zr[i][u];
              zr[c][u] = v;
            }
          }
          double a = 0.0;
#pragma unroll
          for (int u = 0; u < ONU; ++u) a += zr[c][u] * zr[c][u];
          a = wave_sum(a);
          if (a < 1e-6 * pre[c]) redo = true;
          if (!(a >= 1e-6 * pre[c]) && (round == 1 || j0 == 0) && lane == 0) collapse_s = 1;
          const double inv = a > 0.0 ? 1.0 / sqrt(a) : 0.0;
#pragma unroll
          for (int u = 0; u < ONU; ++u) zr[c][u] *= inv;
        }
#pragma unroll
        for (int c = 0; c < OPB; ++c)
#pragma unroll
          for (int u = 0; u < ONU; ++u) {
            const int e = lane + 64 * u;
            if (c < pb && e < n) zp[c * n + e] = zr[c][u];
          }
        if (lane == 0) redo_s = redo && round == 0 && j0 > 0;
      } else if (wave == 0) {
        bool redo = false;
#pragma unroll
        for (int c = 0; c < OPB; ++c) {
          if (c >= pb) break;
          double* zc = zp + c * n;
          for (int pass = 0; pass < 2 && c > 0; ++pass) {
            double d[OPB];
#pragma unroll
            for (int i = 0; i < OPB; ++i) d[i] = 0.0;
            for (int e = lane; e < n; e += 64) {
              const double v = zc[e];
#pragma unroll
              for (int i = 0; i < OPB; ++i)
                if (i < c) d[i] += zp[i * n + e] * v;
            }
#pragma unroll
            for (int i = 0; i < OPB; ++i) d[i] = i < c ? wave_sum(d[i]) : 0.0;
            for (int e = lane; e < n; e += 64) {
              double v = zc[e];
#pragma unroll
              for (int i = 0; i < OPB; ++i)
                if (i < c) v -= d[i] * zp[i * n + e];
              zc[e] = v;
            }
          }
          double a = 0.0;
          for (int e = lane; e < n; e += 64) a += zc[e] * zc[e];
          a = wave_sum(a);
          if (a < 1e-6 * pre[c]) redo = true;
          if (!(a >= 1e-6 * pre[c]) && (round == 1 || j0 == 0) && lane == 0) collapse_s = 1;
          const double inv = a > 0.0 ? 1.0 / sqrt(a) : 0.0;
          for (int e = lane; e < n; e += 64) zc[e] *= inv;
        }
        if (lane == 0) redo_s = redo && round == 0 && j0 > 0;
      }
      __syncthreads();
      if (!redo_s) break;
      __syncthreads();   // every thread has read redo_s before wave 0 may write it again
    }
    for (int u = t; u < pb * n; u += nt) Z[(int64_t)j0 * n + u] = zp[u];
    __syncthreads();
  }
  // rank-deficient block: one column-by-column pass that replaces the dependent columns (the
  // independent ones only move by rounding); H's p * OPB doubles hold the inner products
  if (collapse_s) wg_orthonormalize(Z, n, p, red_s, H);
}

// Jacobi eigensolver of the symmetric m x m matrix H (row-major, global, destroyed), one
// workgroup with H and V staged in LDS when they fit (IN_LDS: a compile-time choice, so that the
// solver's accesses compile to LDS instructions rather than generic ones).  Writes the
// eigenvectors of the k largest eigenvalues, in descending order, to Vk (m x k col-major) and the
// eigenvalues to lam.
template <bool IN_LDS>
__global__ __launch_bounds__(1024) void k_jacobi_top(double* H, int m, int k, double* Vg,
                                                     double* Vk, double* lam) {
  extern __shared__ __align__(16) unsigned char smem[];
  __shared__ double red[16];
  __shared__ __align__(16) double cs[4 * 260];
  __shared__ int order[512];
  // in LDS, H has an odd row stride (no bank conflicts on its column accesses)
  const int ld = IN_LDS ? (m | 1) : m;
  double* Hs = IN_LDS ? (double*)smem : H;
  double* Vs = IN_LDS ? (double*)smem + m * ld : Vg;
  // symmetrized copy (the Rayleigh-Ritz product is symmetric only up to rounding)
  for (int e = threadIdx.x; e < m * m; e += blockDim.x) {
    const int i = e / m, j = e % m;
    if (i <= j) {
      const double a = i == j ? H[e] : 0.5 * (H[i * m + j] + H[j * m + i]);
      Hs[i * ld + j] = a;
      Hs[j * ld + i] = a;
    }
  }
  __syncthreads();
  tfxjac::wg_jacobi(Hs, m, ld, Vs, cs, red);
  for (int i = threadIdx.x; i < m; i += blockDim.x) {
    const double di = Hs[i * ld + i];
    int rank = 0;
    for (int j = 0; j < m; ++j) {
      const double dj = Hs[j * ld + j];
      if (dj > di || (dj == di && j < i)) ++rank;
    }
    order[rank] = i;
  }
  __syncthreads();
  for (int e = threadIdx.x; e < m * k; e += blockDim.x) {
    const int r = e % m, c = e / m;
    Vk[e] = Vs[order[c] * m + r];
  }
  for (int c = threadIdx.x; c < k; c += blockDim.x) lam[c] = Hs[order[c] * ld + order[c]];
}

// Make the largest-|.| entry of every column of U (n x k col-major) positive; one workgroup per
// column.
__global__ __launch_bounds__(256) void k_sign_fix(double* U, int n) {
  __shared__ double bv[256];
  __shared__ int bi[256];
  double* u = U + (int64_t)blockIdx.x * n;
  double best = -1.0;
  int idx = 0;
  for (int r = threadIdx.x; r < n; r += blockDim.x) {
    const double a = fabs(u[r]);
    if (a > best) {
      best = a;
      idx = r;
    }
  }
  bv[threadIdx.x] = best;
  bi[threadIdx.x] = idx;
  __syncthreads();
  for (int o = blockDim.x / 2; o > 0; o >>= 1) {
    if (threadIdx.x < o) {
      const double b2 = bv[threadIdx.x + o];
      const int i2 = bi[threadIdx.x + o];
      if (b2 > bv[threadIdx.x] || (b2 == bv[threadIdx.x] && i2 < bi[threadIdx.x])) {
        bv[threadIdx.x] = b2;
        bi[threadIdx.x] = i2;
      }
    }
    __syncthreads();
  }
  const double sg = u[bi[0]] < 0 ? -1.0 : 1.0;
  for (int r = threadIdx.x; r < n; r += blockDim.x) u[r] *= sg;
}

// C (m x n col-major) = op(A) B; A FP64 col-major (m x kk, or kk x m when transA), B col-major
template <typename S>
__global__ void k_gemm_small(int transA, int m, int n, int kk, const double* __restrict__ A,
                             const S* __restrict__ B, S* __restrict__ C) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= (int64_t)m * n) return;
  const int i = (int)(e % m), j = (int)(e / m);
  double acc = 0.0;
  for (int q = 0; q < kk; ++q) {
    const double a = transA ? A[(int64_t)i * kk + q] : A[(int64_t)q * m + i];
    acc += a * (double)B[(int64_t)j * kk + q];
  }
  C[e] = (S)acc;
}

typedef double f64x4 __attribute__((ext_vector_type(4)));

// C (m x n col-major) = op(A) B in FP64 on the matrix cores: one workgroup per 16 x 16 tile of
// C; its 4 waves take the K-steps of 4 in turn (v_mfma_f64_16x16x4f64), and the 4 partial tiles
// are summed in LDS in a fixed order.  Loads are branch-free (indices clamped, values masked).
__global__ __launch_bounds__(256) void k_gemm_mfma64(int transA, int m, int n, int kk,
                                                     const double* __restrict__ A,
                                                     const double* __restrict__ B,
                                                     double* __restrict__ C) {
  __shared__ double part[4][256];
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63, l16 = lane & 15, lg = lane >> 4;
  const int i = blockIdx.x * 16 + l16, j = blockIdx.y * 16 + l16;
  const bool iok = i < m, jok = j < n;
  const int ic = iok ? i : 0, jc = jok ? j : 0;
  f64x4 acc = {0.0, 0.0, 0.0, 0.0};
  for (int k0 = 4 * w; k0 < kk; k0 += 16) {
    const int k = k0 + lg;
    const bool kok = k < kk;
    const int kc = kok ? k : 0;
    const double av = transA ? A[(int64_t)ic * kk + kc] : A[(int64_t)kc * m + ic];
    const double bv = B[(int64_t)jc * kk + kc];
    acc = __builtin_amdgcn_mfma_f64_16x16x4f64((kok && iok) ? av : 0.0, (kok && jok) ? bv : 0.0,
                                               acc, 0, 0, 0);
  }
  // register r of lane l holds C row 4 r + l / 16, column l % 16 of the tile
#pragma unroll
  for (int r = 0; r < 4; ++r) part[w][(4 * r + lg) * 16 + l16] = acc[r];
  __syncthreads();
  const int e = threadIdx.x, row = e >> 4, col = e & 15;
  const int ci = blockIdx.x * 16 + row, cj = blockIdx.y * 16 + col;
  if (ci < m && cj < n)
    C[(int64_t)cj * m + ci] = ((part[0][e] + part[1][e]) + part[2][e]) + part[3][e];
}

void launch_gemm64(int transA, int m, int n, int kk, const double* A, const double* B, double* C,
                   hipStream_t st) {
  const dim3 grid((unsigned)((m + 15) / 16), (unsigned)((n + 15) / 16));
  hipLaunchKernelGGL(k_gemm_mfma64, grid, dim3(256), 0, st, transA, m, n, kk, A, B, C);
}

template <typename S>
int gram_t(tfx_ctx* ctx, const S* X, const Shape& s, int mode, double* G, hipStream_t st) {
  const PIQ v = piq_of(s, mode);
  TFX_HIP(ctx, hipMemsetAsync(G, 0, sizeof(double) * v.I * v.I, st));
  const int ntile = (int)((v.I + GT - 1) / GT);
  const int npairs = ntile * (ntile + 1) / 2;
  const int64_t KT = v.P * v.Q;
  // enough workgroups to fill the chip: split the reduction
  int64_t want = (int64_t)ctx->cu_count * 8;
  int64_t ksplit = (want + npairs - 1) / npairs;
  int64_t kchunk = (KT + ksplit - 1) / ksplit;
  kchunk = ((kchunk + GK - 1) / GK) * GK;
  if (kchunk < GK) kchunk = GK;
  ksplit = (KT + kchunk - 1) / kchunk;
  hipLaunchKernelGGL(k_gram_mode<S>, dim3(npairs, (unsigned)ksplit), dim3(256), 0, st, X, v,
                     ntile, kchunk, G);
  TFX_CHECK_LAUNCH(ctx);
  return TFX_OK;
}

template <typename S, typename SY>
int ttm_t(tfx_ctx* ctx, const S* X, const Shape& s, int mode, const double* U, int k, SY* Y,
          double* sig2, hipStream_t st) {
  const PIQ v = piq_of(s, mode);
  const int64_t tot = v.P * k * v.Q;
  hipLaunchKernelGGL((k_ttm<S, SY>), dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, st, X, v, U, k, Y);
  TFX_CHECK_LAUNCH(ctx);
  if (sig2) {
    const PIQ vy{v.P, k, v.Q};
    hipLaunchKernelGGL(k_rownorm2<SY>, dim3(k), dim3(256), 0, st, Y, vy, sig2);
    TFX_CHECK_LAUNCH(ctx);
  }
  return TFX_OK;
}

}  // namespace

// MFMA kernels (mlsvd_mfma.hip): bf16x3 Gram and FP32-MFMA product on FP32 data, FP64 MFMA on
// FP64 data.
int tfx_gram_split16(tfx_ctx* ctx, const float* X, const Shape& s, int mode, double* G,
                     hipStream_t st);
int tfx_ttm_split16(tfx_ctx* ctx, const float* X, const Shape& s, int mode, const double* U,
                    int k, double* Y, double* sig2, hipStream_t st);
int tfx_gram_f64mfma(tfx_ctx* ctx, const double* X, const Shape& s, int mode, double* G,
                     hipStream_t st);
int tfx_ttm_f64mfma(tfx_ctx* ctx, const double* X, const Shape& s, int mode, const double* U,
                    int k, double* Y, double* sig2, hipStream_t st);

extern "C" int tfx_gram_mode(tfx_ctx* ctx, const void* X, int dtype, int L, const int64_t* dims,
                             int mode, double* G, int prec, hipStream_t stream) {
  if (!ctx || !X || !G) return tfx_set_error(ctx, TFX_E_ARG, "null argument");
  Shape s;
  int rc = tfx_make_shape(ctx, L, dims, 1, &s);
  if (rc) return rc;
  if (mode < 0 || mode >= L) return tfx_set_error(ctx, TFX_E_ARG, "mode %d out of range", mode);
  if (prec == TFX_PREC_SPLIT16) {
    if (dtype != TFX_F32) return tfx_set_error(ctx, TFX_E_ARG, "split16 Gram needs FP32 data");
    return tfx_gram_split16(ctx, (const float*)X, s, mode, G, stream);
  }
  if (dtype == TFX_F32) return gram_t<float>(ctx, (const float*)X, s, mode, G, stream);
  if (dtype == TFX_F64) return tfx_gram_f64mfma(ctx, (const double*)X, s, mode, G, stream);
  return tfx_set_error(ctx, TFX_E_ARG, "bad dtype %d", dtype);
}

extern "C" int tfx_eig_top(tfx_ctx* ctx, const double* G, int n, const double* Q0, int p, int k,
                           int iters, double* U, double* lam, hipStream_t stream) {
  if (!ctx || !G || !U || !lam) return tfx_set_error(ctx, TFX_E_ARG, "null argument");
  if (n < 1 || k < 1 || k > n) return tfx_set_error(ctx, TFX_E_ARG, "bad sizes n=%d k=%d", n, k);
  const bool direct = n <= p;
  if (!direct && (!Q0 || p < k)) return tfx_set_error(ctx, TFX_E_ARG, "start block needs p >= k columns");
  if (n > 8192 || p > 512) return tfx_set_error(ctx, TFX_E_UNSUPPORTED, "eigensolver limited to n <= 8192, p <= 512");
  const int m = direct ? n : p;
  if (m > 512) return tfx_set_error(ctx, TFX_E_UNSUPPORTED, "Rayleigh-Ritz size %d > 512", m);
  // workspace (doubles): Q n*p | Z n*p | H m*m | V m*m | Vk m*k
  const size_t qz = direct ? 0 : (size_t)n * p;
  const size_t dbl = 2 * qz + 2 * (size_t)m * m + (size_t)m * k + 64;
  void* wsp;
  int rc = tfx_workspace(ctx, sizeof(double) * dbl, stream, &wsp);
  if (rc) return rc;
  double* Q = (double*)wsp;
  double* Z = Q + qz;
  double* H = Z + qz;
  double* V = H + (size_t)m * m;
  double* Vk = V + (size_t)m * m;
  auto gemm = [&](int transA, int mm_, int nn_, int kk_, const double* A_, const double* B_, double* C_) {
    launch_gemm64(transA, mm_, nn_, kk_, A_, B_, C_, stream);
  };
  // orthonormalization: block CGS2 with its panel in LDS when it fits, else column CGS2
  const size_t orth_lds = sizeof(double) * ((size_t)OPB * n + (size_t)p * OPB);
  const bool orth_block = n <= ORTH_MAXN && orth_lds <= 150 * 1024;
  if (orth_block)
    TFX_HIP(ctx, hipFuncSetAttribute((const void*)k_orth_block,
                                     hipFuncAttributeMaxDynamicSharedMemorySize, (int)orth_lds));
  auto orth = [&](double* Zo) {
    if (orth_block)
      hipLaunchKernelGGL(k_orth_block, dim3(1), dim3(1024), orth_lds, stream, Zo, n, p);
    else
      hipLaunchKernelGGL(k_orth, dim3(1), dim3(1024), 0, stream, Zo, n, p);
  };
  if (direct) {
    TFX_HIP(ctx, hipMemcpyAsync(H, G, sizeof(double) * n * n, hipMemcpyDeviceToDevice, stream));
  } else {
    // range finder Q = orth(G^iters Q0) with plain power steps, as the reference's randomized
    // range finder (Compression.py:386-395: un-normalised power iterations, then one QR), except
    // that the block is orthonormalized after every second product as well: direction i then
    // keeps a relative precision of about eps (lambda_1 / lambda_i)^2 in FP64 whatever `iters`
    // (the reference's products on the unfolding give (sigma_1 / sigma_i)^(2 iters + 1)).
    TFX_HIP(ctx, hipMemcpyAsync(Q, Q0, sizeof(double) * n * p, hipMemcpyDeviceToDevice, stream));
    for (int it = 0; it < iters; ++it) {
      gemm(0, n, p, n, G, Q, Z);
      double* t = Q;
      Q = Z;
      Z = t;
      if (it % 2 == 1 && it + 1 < iters) orth(Q);
    }
    orth(Q);
    // Rayleigh-Ritz: H = Q^T (G Q)
    gemm(0, n, p, n, G, Q, Z);
    gemm(1, p, p, n, Q, Z, H);
  }
  TFX_CHECK_LAUNCH(ctx);
  const size_t lds = sizeof(double) * ((size_t)m * (m | 1) + (size_t)m * m);
  if (lds <= 140 * 1024) {
    TFX_HIP(ctx, hipFuncSetAttribute((const void*)k_jacobi_top<true>,
                                     hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    hipLaunchKernelGGL(k_jacobi_top<true>, dim3(1), dim3(1024), lds, stream, H, m, k, V, Vk, lam);
  } else {
    hipLaunchKernelGGL(k_jacobi_top<false>, dim3(1), dim3(1024), 0, stream, H, m, k, V, Vk, lam);
  }
  TFX_CHECK_LAUNCH(ctx);
  if (direct) {
    TFX_HIP(ctx, hipMemcpyAsync(U, Vk, sizeof(double) * n * k, hipMemcpyDeviceToDevice, stream));
  } else {
    gemm(0, n, k, p, Q, Vk, U);
    TFX_CHECK_LAUNCH(ctx);
  }
  hipLaunchKernelGGL(k_sign_fix, dim3(k), dim3(256), 0, stream, U, n);
  TFX_CHECK_LAUNCH(ctx);
  return TFX_OK;
}

extern "C" int tfx_ttm_mixed(tfx_ctx* ctx, const void* X, int x_dtype, int L, const int64_t* dims,
                             int mode, const double* U, int k, void* Y, int y_dtype, double* sig2,
                             int prec, hipStream_t stream) {
  if (!ctx || !X || !U || !Y) return tfx_set_error(ctx, TFX_E_ARG, "null argument");
  Shape s;
  int rc = tfx_make_shape(ctx, L, dims, 1, &s);
  if (rc) return rc;
  if (mode < 0 || mode >= L || k < 1) return tfx_set_error(ctx, TFX_E_ARG, "bad mode/k");
  if (prec == TFX_PREC_SPLIT16) {
    if (x_dtype != TFX_F32 || y_dtype != TFX_F64)
      return tfx_set_error(ctx, TFX_E_ARG, "split16 TTM needs FP32 input and FP64 output");
    return tfx_ttm_split16(ctx, (const float*)X, s, mode, U, k, (double*)Y, sig2, stream);
  }
  if (x_dtype == TFX_F32 && y_dtype == TFX_F32) return ttm_t<float, float>(ctx, (const float*)X, s, mode, U, k, (float*)Y, sig2, stream);
  if (x_dtype == TFX_F32 && y_dtype == TFX_F64) return ttm_t<float, double>(ctx, (const float*)X, s, mode, U, k, (double*)Y, sig2, stream);
  if (x_dtype == TFX_F64 && y_dtype == TFX_F64) return tfx_ttm_f64mfma(ctx, (const double*)X, s, mode, U, k, (double*)Y, sig2, stream);
  return tfx_set_error(ctx, TFX_E_ARG, "unsupported dtype pair %d/%d", x_dtype, y_dtype);
}

extern "C" int tfx_ttm(tfx_ctx* ctx, const void* X, int dtype, int L, const int64_t* dims,
                       int mode, const double* U, int k, void* Y, double* sig2, int prec,
                       hipStream_t stream) {
  return tfx_ttm_mixed(ctx, X, dtype, L, dims, mode, U, k, Y, dtype, sig2, prec, stream);
}

extern "C" int tfx_gemm_small(tfx_ctx* ctx, int dtype, int transA, int m, int n, int kk,
                              const double* A, const void* B, void* C, hipStream_t stream) {
  if (!ctx || !A || !B || !C) return tfx_set_error(ctx, TFX_E_ARG, "null argument");
  if (m < 1 || n < 1 || kk < 1) return tfx_set_error(ctx, TFX_E_ARG, "bad sizes");
  const int64_t tot = (int64_t)m * n;
  const unsigned nb = (unsigned)((tot + 255) / 256);
  if (dtype == TFX_F32)
    hipLaunchKernelGGL(k_gemm_small<float>, dim3(nb), dim3(256), 0, stream, transA, m, n, kk, A,
                       (const float*)B, (float*)C);
  else if (dtype == TFX_F64)
    launch_gemm64(transA, m, n, kk, A, (const double*)B, (double*)C, stream);
  else
    return tfx_set_error(ctx, TFX_E_ARG, "bad dtype %d", dtype);
  TFX_CHECK_LAUNCH(ctx);
  return TFX_OK;
}
