// MTTKRP of a large third-order FP64 tensor as two passes of FP64 matrix-core GEMM
// (v_mfma_f64_16x16x4f64), for the FP64 form's refinement on the full tensor
// (compute_grad, GaussNewton.py:401-427; N_l = T_(l) KR_l).
//
// With T stored C order as an (I0 I1) x I2 matrix T_mat (row = i0 I1 + i1):
//   pass 1  W = T_mat F_2                     (I0 I1) x R, one read of T
//           N_0[i0, r] = sum_i1 W[(i0, i1), r] F_1[i1, r]
//           N_1[i1, r] = sum_i0 W[(i0, i1), r] F_0[i0, r]
//   pass 2  N_2 = T_mat^T KR,  KR[(i0, i1), r] = F_0[i0, r] F_1[i1, r]   one more read of T
// so T is read twice for all three modes (the per-mode kernel reads it three times and forms the
// Khatri-Rao operand of every MFMA step from factor entries), and the Khatri-Rao operand of
// pass 2 is formed once per staged chunk of rows.
//
// Both passes stage 64 x 32 tiles of T and 32 x 64 tiles of the other operand in LDS, with the
// next chunk's loads in registers while the MFMAs run on the current one; each of the four
// waves of a workgroup owns 16 output rows and all 64 ranks of the chunk (four 16 x 16
// accumulators).  Ranks go in chunks of 64 (operands past R are zero).  Pass 2 splits the rows
// over workgroups and sums the partial tiles in a fixed order; W is written row-major with the
// 64 ranks of the chunk contiguous, so the reductions for modes 0 and 1 read whole rows.  No
// atomics: results do not depend on scheduling.
//
// Fragment layouts of v_mfma_f64_16x16x4f64 (lane l): A[l % 16][l / 16], B[l / 16][l % 16],
// result register j = D[4 j + l / 16][l % 16].
#include "common.h"

int tfx_mttkrp_f64_gemm3(tfx_ctx* ctx, const double* T, const Shape& s, const double* F,
                         unsigned mask, double* out, hipStream_t stream);

namespace {

typedef double f64x4 __attribute__((ext_vector_type(4)));

constexpr int GT = 64;    // output rows per workgroup tile
constexpr int GR = 64;    // ranks per chunk
constexpr int KC = 32;    // reduction chunk staged in LDS
constexpr int NT = 256;
constexpr int PER = GT * KC / NT;   // staged values per thread per operand (8)

// pass 1: W[row][r] = sum_k T[row][k] F2[k][rc0 + r] for the 64 rows of the tile
__global__ __launch_bounds__(NT) void k_f64g_w(const double* __restrict__ T, int64_t M, int64_t K,
                                               const double* __restrict__ F2, int R, int rc0,
                                               double* __restrict__ W) {
  __shared__ double As[GT][KC + 1];
  __shared__ double Bs[KC][GR + 1];
  const int t = threadIdx.x, wv = t >> 6, lane = t & 63, li = lane & 15, lk = lane >> 4;
  const int64_t row0 = (int64_t)blockIdx.x * GT;
  f64x4 acc[4];
#pragma unroll
  for (int q = 0; q < 4; ++q) acc[q] = f64x4{0.0, 0.0, 0.0, 0.0};
  double ra[PER], rb[PER];
  auto load = [&](int64_t k0) {
#pragma unroll
    for (int q = 0; q < PER; ++q) {
      const int u = t + NT * q;
      const int rr = u / KC, kk = u % KC;       // A: row rr, column kk (coalesced along k)
      const int64_t row = row0 + rr, k = k0 + kk;
      const bool ok = row < M && k < K;
      ra[q] = T[ok ? row * K + k : 0];
      ra[q] = ok ? ra[q] : 0.0;
      const int r = rc0 + rr;                    // B: rank rr (as the row of u), k = kk
      const bool okb = r < R && k < K;
      rb[q] = F2[okb ? (int64_t)r * K + k : 0];
      rb[q] = okb ? rb[q] : 0.0;
    }
  };
  load(0);
  for (int64_t k0 = 0; k0 < K; k0 += KC) {
#pragma unroll
    for (int q = 0; q < PER; ++q) {
      const int u = t + NT * q;
      As[u / KC][u % KC] = ra[q];
      Bs[u % KC][u / KC] = rb[q];
    }
    __syncthreads();
    if (k0 + KC < K) load(k0 + KC);
#pragma unroll
    for (int ks = 0; ks < KC; ks += 4) {
      const double a = As[16 * wv + li][ks + lk];
#pragma unroll
      for (int q = 0; q < 4; ++q)
        acc[q] = __builtin_amdgcn_mfma_f64_16x16x4f64(a, Bs[ks + lk][16 * q + li], acc[q], 0, 0, 0);
    }
    __syncthreads();
  }
#pragma unroll
  for (int q = 0; q < 4; ++q)
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int64_t row = row0 + 16 * wv + 4 * j + lk;
      if (row < M) W[row * GR + 16 * q + li] = acc[q][j];
    }
}

// FT[i][r] = F[rc0 + r][i] for the rank chunk (zero past R): the factor rows that the
// reductions and the Khatri-Rao operand read, with the 64 ranks of a row contiguous
__global__ __launch_bounds__(256) void k_f64g_ft(const double* __restrict__ F, int64_t I, int R, int rc0,
                                                 double* __restrict__ FT) {
  const int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (e >= I * GR) return;
  const int64_t i = e / GR;
  const int r = (int)(e - i * GR);
  FT[e] = rc0 + r < R ? F[(int64_t)(rc0 + r) * I + i] : 0.0;
}

// modes 0 and 1 from W: one wavefront per output row i (block b < I0: mode 0, else mode 1),
// lane r = rank rc0 + r; a lane reads one entry of each W row it sums (whole rows per wave)
__global__ __launch_bounds__(64) void k_f64g_wred(const double* __restrict__ W, int64_t I0, int64_t I1,
                                                  const double* __restrict__ F0T, const double* __restrict__ F1T,
                                                  int R, int rc0, unsigned mask, int64_t off0, int64_t off1,
                                                  double* __restrict__ out) {
  const int r = threadIdx.x;
  const int rg = rc0 + r;
  const bool rok = rg < R;
  const int64_t b = blockIdx.x;
  double acc = 0.0;
  if (b < I0) {
    if (!(mask & 1u)) return;
    const int64_t i0 = b;
    const double* w = W + (i0 * I1) * GR + r;
    const double* f = F1T + r;
    int64_t i1 = 0;
    for (; i1 + 4 <= I1; i1 += 4) {
      double wv[4], fv[4];
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        wv[u] = w[(i1 + u) * GR];
        fv[u] = f[(i1 + u) * GR];
      }
#pragma unroll
      for (int u = 0; u < 4; ++u) acc += wv[u] * fv[u];
    }
    for (; i1 < I1; ++i1) acc += w[i1 * GR] * f[i1 * GR];
    if (rok) out[off0 + (int64_t)rg * I0 + i0] = acc;
  } else {
    if (!(mask & 2u)) return;
    const int64_t i1 = b - I0;
    const double* w = W + i1 * GR + r;
    const double* f = F0T + r;
    int64_t i0 = 0;
    for (; i0 + 4 <= I0; i0 += 4) {
      double wv[4], fv[4];
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        wv[u] = w[(i0 + u) * I1 * GR];
        fv[u] = f[(i0 + u) * GR];
      }
#pragma unroll
      for (int u = 0; u < 4; ++u) acc += wv[u] * fv[u];
    }
    for (; i0 < I0; ++i0) acc += w[i0 * I1 * GR] * f[i0 * GR];
    if (rok) out[off1 + (int64_t)rg * I1 + i1] = acc;
  }
}

// pass 2: partial N_2 tile for columns c0..c0+63 of T_mat and rows [rb, re) of the split:
// part[split][c][r] = sum_row T[row][c] F_0[i0(row)][r] F_1[i1(row)][r]
__global__ __launch_bounds__(NT) void k_f64g_t(const double* __restrict__ T, int64_t M, int64_t K,
                                               int64_t I1, const double* __restrict__ F0T,
                                               const double* __restrict__ F1T,
                                               int64_t rows_per_split, double* __restrict__ part) {
  __shared__ double Ts[KC][GT + 1];    // [row][c]
  __shared__ double Ks[KC][GR + 1];    // [row][r]
  const int t = threadIdx.x, wv = t >> 6, lane = t & 63, li = lane & 15, lk = lane >> 4;
  const int64_t c0 = (int64_t)blockIdx.x * GT;
  const int64_t rb = (int64_t)blockIdx.y * rows_per_split;
  const int64_t re = rb + rows_per_split < M ? rb + rows_per_split : M;
  f64x4 acc[4];
#pragma unroll
  for (int q = 0; q < 4; ++q) acc[q] = f64x4{0.0, 0.0, 0.0, 0.0};
  double rt[PER], rk[PER];
  // thread t stages rows kk = t / 64 + 4 q (q = 0..7) of a chunk, column / rank cc = t % 64:
  // (i0, i1) of its first row by one division per chunk, then stepped by 4 rows
  const int cc = t % GT, kk0 = t / GT;
  auto load = [&](int64_t r0) {
    int64_t row = r0 + kk0;
    int64_t i0 = row / I1, i1 = row - i0 * I1;
#pragma unroll
    for (int q = 0; q < PER; ++q) {
      const int64_t c = c0 + cc;
      const bool ok = row < re && c < K;          // T: row of the chunk, column cc (coalesced)
      rt[q] = T[ok ? row * K + c : 0];
      rt[q] = ok ? rt[q] : 0.0;
      const bool okr = row < re;                  // KR: factor rows with contiguous ranks
      const double a = F0T[(okr ? i0 : 0) * GR + cc], bq = F1T[(okr ? i1 : 0) * GR + cc];
      rk[q] = okr ? a * bq : 0.0;
      row += NT / GT;
      i1 += NT / GT;
      while (i1 >= I1) {
        i1 -= I1;
        ++i0;
      }
    }
  };
  load(rb);
  for (int64_t r0 = rb; r0 < re; r0 += KC) {
#pragma unroll
    for (int q = 0; q < PER; ++q) {
      const int u = t + NT * q;
      Ts[u / GT][u % GT] = rt[q];
      Ks[u / GT][u % GT] = rk[q];
    }
    __syncthreads();
    if (r0 + KC < re) load(r0 + KC);
#pragma unroll
    for (int ks = 0; ks < KC; ks += 4) {
      const double a = Ts[ks + lk][16 * wv + li];
#pragma unroll
      for (int q = 0; q < 4; ++q)
        acc[q] = __builtin_amdgcn_mfma_f64_16x16x4f64(a, Ks[ks + lk][16 * q + li], acc[q], 0, 0, 0);
    }
    __syncthreads();
  }
  double* p = part + ((int64_t)blockIdx.y * gridDim.x + blockIdx.x) * (GT * GR);
#pragma unroll
  for (int q = 0; q < 4; ++q)
#pragma unroll
    for (int j = 0; j < 4; ++j) p[(16 * wv + 4 * j + lk) * GR + 16 * q + li] = acc[q][j];
}

// N_2[c][r] = sum over splits (in order) of the partial tiles
__global__ __launch_bounds__(256) void k_f64g_tred(const double* __restrict__ part, int64_t K, int nblk,
                                                   int nsplit, int R, int rc0, int64_t off2,
                                                   double* __restrict__ out) {
  const int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x;
  const int64_t c = e / GR;
  const int r = (int)(e - c * GR);
  if (c >= K || rc0 + r >= R) return;
  const int64_t blk = c / GT, cc = c - blk * GT;
  const double* src = part + (blk * GT + cc) * GR + r;
  const int64_t stride = (int64_t)nblk * GT * GR;
  double acc = 0.0;
  int sp = 0;
  for (; sp + 4 <= nsplit; sp += 4) {
    double v[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) v[u] = src[(int64_t)(sp + u) * stride];
#pragma unroll
    for (int u = 0; u < 4; ++u) acc += v[u];
  }
  for (; sp < nsplit; ++sp) acc += src[(int64_t)sp * stride];
  out[off2 + (int64_t)(rc0 + r) * K + c] = acc;
}

}  // namespace

// Third-order FP64 MTTKRP of the modes in `mask` (flat factor / output layouts of Shape s).
// Workspace: W ((I0 I1) x 64 doubles) when mode 0 or 1 is selected, and the pass-2 partials.
int tfx_mttkrp_f64_gemm3(tfx_ctx* ctx, const double* T, const Shape& s, const double* F,
                         unsigned mask, double* out, hipStream_t stream) {
  if (s.L != 3) return tfx_set_error(ctx, TFX_E_UNSUPPORTED, "two-pass FP64 MTTKRP is third order only");
  const int64_t I0 = s.d[0], I1 = s.d[1], I2 = s.d[2];
  const int64_t M = I0 * I1, K = I2;
  const bool need_w = (mask & 3u) != 0, need_t = (mask & 4u) != 0;
  const int nblk = (int)((K + GT - 1) / GT);
  // pass-2 row splits: about 8 workgroups per CU in total, whole KC chunks per split
  int nsplit = (int)std::max<int64_t>(1, (8 * (int64_t)ctx->cu_count + nblk - 1) / nblk);
  int64_t rps = (M + nsplit - 1) / nsplit;
  rps = (rps + KC - 1) / KC * KC;
  nsplit = (int)((M + rps - 1) / rps);
  const size_t wbytes = need_w ? align256(sizeof(double) * (size_t)M * GR) : 0;
  const size_t pbytes = need_t ? align256(sizeof(double) * (size_t)nsplit * nblk * GT * GR) : 0;
  const size_t tbytes = align256(sizeof(double) * (size_t)(I0 + I1) * GR);
  void* wsp;
  int rc = tfx_workspace(ctx, wbytes + pbytes + tbytes + 256, stream, &wsp);
  if (rc) return rc;
  double* W = (double*)wsp;
  double* part = (double*)((char*)wsp + wbytes);
  double* F0T = (double*)((char*)wsp + wbytes + pbytes);
  double* F1T = F0T + I0 * GR;
  const double* F0 = F + s.off[0];
  const double* F1 = F + s.off[1];
  const double* F2 = F + s.off[2];
  for (int rc0 = 0; rc0 < s.R; rc0 += GR) {
    hipLaunchKernelGGL(k_f64g_ft, dim3((unsigned)((I0 * GR + 255) / 256)), dim3(256), 0, stream, F0, I0, s.R, rc0,
                       F0T);
    hipLaunchKernelGGL(k_f64g_ft, dim3((unsigned)((I1 * GR + 255) / 256)), dim3(256), 0, stream, F1, I1, s.R, rc0,
                       F1T);
    TFX_CHECK_LAUNCH(ctx);
    if (need_w) {
      hipLaunchKernelGGL(k_f64g_w, dim3((unsigned)((M + GT - 1) / GT)), dim3(NT), 0, stream, T, M, K, F2, s.R,
                         rc0, W);
      TFX_CHECK_LAUNCH(ctx);
      hipLaunchKernelGGL(k_f64g_wred, dim3((unsigned)(I0 + I1)), dim3(64), 0, stream, W, I0, I1, F0T, F1T, s.R,
                         rc0, mask, s.off[0], s.off[1], out);
      TFX_CHECK_LAUNCH(ctx);
    }
    if (need_t) {
      hipLaunchKernelGGL(k_f64g_t, dim3((unsigned)nblk, (unsigned)nsplit), dim3(NT), 0, stream, T, M, K, I1, F0T,
                         F1T, rps, part);
      TFX_CHECK_LAUNCH(ctx);
      const unsigned nr = (unsigned)((K * GR + 255) / 256);
      hipLaunchKernelGGL(k_f64g_tred, dim3(nr), dim3(256), 0, stream, part, K, nblk, nsplit, s.R, rc0,
                         s.off[2], out);
      TFX_CHECK_LAUNCH(ctx);
    }
  }
  return TFX_OK;
}
