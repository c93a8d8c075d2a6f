// MTTKRP of every mode in one pass over T, and the CPD residual ||T - [[F]]||^2.
//
// T (C order) is read as NR = N / K rows of length K = I_{L-1} (the contiguous last mode).
// For row rho = (i_0, ..., i_{L-2}):
//   W[rho, r] = prod_{m < L-1} F_m[i_m, r]              (Khatri-Rao row of the other modes)
//   Y[rho, r] = sum_k T[rho, k] F_{L-1}[k, r]
//   N_{L-1}[k, r]  += T[rho, k] W[rho, r]
//   N_m[i_m, r]    += Y[rho, r] prod_{m' != m, m' < L-1} F_{m'}[i_{m'}, r]     (m < L-1)
// which together are T_(l) KR_l for every l (compute_grad, GaussNewton.py:411-423), with the
// Khatri-Rao products never stored.  This file holds the FMA path (FP32 or FP64 arithmetic);
// the MFMA split-precision path is in mttkrp_mfma.hip.
#include "common.h"

// FP64 MFMA residual (residual_mfma.hip): 1 when it ran, 0 when the shape is not covered
int residual_mfma_try(tfx_ctx* ctx, const void* T, int t_dtype, const Shape& s, const double* F,
                      double* out, hipStream_t st);

namespace {

constexpr int RB = 32;   // rows per block step
constexpr int KC = 32;   // k per chunk

template <typename S> struct RCOf;
template <> struct RCOf<float> { static constexpr int v = 64; };
template <> struct RCOf<double> { static constexpr int v = 32; };

// Row strides of the row index over modes 0..L-2 (last of them fastest).
__device__ inline void row_to_idx(const Shape& s, int64_t rho, int64_t* idx) {
  for (int m = s.L - 2; m >= 0; --m) {
    idx[m] = rho % s.d[m];
    rho /= s.d[m];
  }
}

// Pass 1: Y (NR x R, row-major, columns rc0..rc0+RC-1) and per-workgroup partials of N_{L-1}
// (accumulated in LDS when NP_LDS, a compile-time choice so that those accesses compile to LDS
// instructions, else in the workgroup's slice of `part`).
template <typename TS, typename S, bool NP_LDS>
__global__ __launch_bounds__(256) void k_mttkrp_rows(const TS* __restrict__ T, Shape s,
                                                      const S* __restrict__ F, int rc0,
                                                      int64_t rows_per_wg, S* __restrict__ Y,
                                                      S* __restrict__ part) {
  constexpr bool np_in_lds = NP_LDS;
  constexpr int RC = RCOf<S>::v;
  constexpr int JN = RC / 8;
  extern __shared__ __align__(16) unsigned char smem[];
  S* Tt = (S*)smem;                 // RB x (KC+1)
  S* Fl = Tt + RB * (KC + 1);       // KC x RC
  S* W = Fl + KC * RC;              // RB x RC
  S* Np_lds = W + RB * RC;          // K x RC (when it fits)

  const int t = threadIdx.x;
  const int L = s.L, R = s.R;
  const int64_t K = s.d[L - 1];
  const int64_t NR = s.N / K;
  const int rce = min(RC, R - rc0);
  const int64_t w = blockIdx.x;
  S* Np = np_in_lds ? Np_lds : part + w * K * RC;

  for (int64_t e = t; e < K * RC; e += 256) Np[e] = S(0);

  const int64_t rho_begin = w * rows_per_wg;
  const int64_t rho_end = min(NR, rho_begin + rows_per_wg);
  const S* Flast = F + s.off[L - 1];

  for (int64_t rho0 = rho_begin; rho0 < rho_end; rho0 += RB) {
    const int nrows = (int)min((int64_t)RB, rho_end - rho0);
    __syncthreads();
    // Khatri-Rao rows of the leading modes
    for (int e = t; e < RB * RC; e += 256) {
      const int rr = e / RC, r = e % RC;
      S v = S(0);
      if (rr < nrows && r < rce) {
        int64_t idx[TFX_MAXL];
        row_to_idx(s, rho0 + rr, idx);
        v = S(1);
        for (int m = 0; m < L - 1; ++m) v *= F[s.off[m] + (int64_t)(rc0 + r) * s.d[m] + idx[m]];
      }
      W[rr * RC + r] = v;
    }
    S yacc[JN];
#pragma unroll
    for (int j = 0; j < JN; ++j) yacc[j] = S(0);
    const int yr = t >> 3, rl = t & 7;

    for (int64_t k0 = 0; k0 < K; k0 += KC) {
      __syncthreads();
      for (int e = t; e < RB * KC; e += 256) {
        const int rr = e / KC, kk = e % KC;
        S v = S(0);
        if (rr < nrows && k0 + kk < K) v = (S)T[(rho0 + rr) * K + k0 + kk];
        Tt[rr * (KC + 1) + kk] = v;
      }
      for (int e = t; e < KC * RC; e += 256) {
        const int r = e / KC, kk = e % KC;
        S v = S(0);
        if (r < rce && k0 + kk < K) v = Flast[(int64_t)(rc0 + r) * K + k0 + kk];
        Fl[kk * RC + r] = v;
      }
      __syncthreads();
      // Y[rho, r] += sum_kk T[rho, kk] F[kk, r]
#pragma unroll 4
      for (int kk = 0; kk < KC; ++kk) {
        const S tv = Tt[yr * (KC + 1) + kk];
#pragma unroll
        for (int j = 0; j < JN; ++j) yacc[j] += tv * Fl[kk * RC + rl + 8 * j];
      }
      // N_last[k, r] += sum_rho T[rho, k] W[rho, r]
      const int kk = t >> 3;
      if (k0 + kk < K) {
        S nacc[JN];
#pragma unroll
        for (int j = 0; j < JN; ++j) nacc[j] = S(0);
        for (int rr = 0; rr < nrows; ++rr) {
          const S tv = Tt[rr * (KC + 1) + kk];
#pragma unroll
          for (int j = 0; j < JN; ++j) nacc[j] += tv * W[rr * RC + rl + 8 * j];
        }
#pragma unroll
        for (int j = 0; j < JN; ++j) Np[(k0 + kk) * RC + rl + 8 * j] += nacc[j];
      }
    }
    if (Y && yr < nrows) {
#pragma unroll
      for (int j = 0; j < JN; ++j) {
        const int r = rl + 8 * j;
        if (r < rce) Y[(rho0 + yr) * R + rc0 + r] = yacc[j];
      }
    }
  }
  if (np_in_lds) {
    __syncthreads();
    S* dst = part + w * K * RC;
    for (int64_t e = t; e < K * RC; e += 256) dst[e] = Np[e];
  }
}

// Pass 2: N_{L-1}[k, rc0 + r] = sum_w part[w][k][r].  One workgroup per k: thread (g, r) sums
// w = g, g + 8, .. (the loads of a step are contiguous over r), then the 8 sums are added in
// order of g.
template <typename S>
__global__ __launch_bounds__(256) void k_mttkrp_reduce_last(Shape s, int rc0, int nwg,
                                                             const S* __restrict__ part,
                                                             S* __restrict__ out) {
  constexpr int RC = RCOf<S>::v;
  constexpr int NG = 256 / RC;
  __shared__ S red[NG][RC];
  const int64_t K = s.d[s.L - 1];
  const int64_t k = blockIdx.x;
  const int r = threadIdx.x % RC, g = threadIdx.x / RC;
  S acc = S(0);
  const S* p = part + k * RC + r;
  for (int w = g; w < nwg; w += NG) acc += p[(int64_t)w * K * RC];
  red[g][r] = acc;
  __syncthreads();
  if (g == 0 && rc0 + r < s.R) {
    S tot = red[0][r];
    for (int q = 1; q < NG; ++q) tot += red[q][r];
    out[s.off[s.L - 1] + (int64_t)(rc0 + r) * K + k] = tot;
  }
}

// Pass 3: N_m for m < L-1 from Y (NR x R, row-major).  One workgroup per (m, i_m); lane = rank
// (chunks of 64), the eight waves split the rows with i_m = i.  The rows are walked as an outer
// index over the row modes other than m and the innermost varying mode qi (the last row mode,
// or the one before it when m is the last), decomposed once per outer step with 32-bit
// arithmetic, around an inner loop over qi; the product of the outer modes' factor entries is
// formed once per outer step.
template <typename S>
__global__ __launch_bounds__(512) void k_mttkrp_rowreduce(Shape s, const S* __restrict__ F,
                                                           const S* __restrict__ Y,
                                                           unsigned mask, S* __restrict__ out) {
  constexpr int NWV = 8;
  __shared__ S red[NWV][64];
  const int L = s.L, R = s.R;
  const int LR = L - 1;                       // row modes 0..LR-1
  int b = blockIdx.x;
  int m = 0;
  while (m < LR && b >= s.d[m]) {
    b -= (int)s.d[m];
    ++m;
  }
  if (m >= LR || !((mask >> m) & 1u)) return;
  const int i = b;
  int rs[TFX_MAXL];                           // row strides over the row modes
  rs[LR - 1] = 1;
  for (int q = LR - 2; q >= 0; --q) rs[q] = rs[q + 1] * (int)s.d[q + 1];
  const int qi = LR == 1 ? -1 : (m == LR - 1 ? LR - 2 : LR - 1);
  int n_outer = 1;
  for (int q = 0; q < LR; ++q)
    if (q != m && q != qi) n_outer *= (int)s.d[q];
  const int ni = qi >= 0 ? (int)s.d[qi] : 1;
  const int si = qi >= 0 ? rs[qi] : 0;
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  for (int r0 = 0; r0 < R; r0 += 64) {
    const int r = r0 + lane;
    const bool rok = r < R;
    const int rc = rok ? r : 0;
    S acc = S(0);
    for (int ot = wv; ot < n_outer; ot += NWV) {
      int rem = ot, base = i * rs[m];
      S pre = S(1);
      for (int q = LR - 1; q >= 0; --q) {
        if (q == m || q == qi) continue;
        const int dq = (int)s.d[q];
        const int iq = rem % dq;
        rem /= dq;
        base += iq * rs[q];
        pre *= F[s.off[q] + (int64_t)rc * dq + iq];
      }
      S inner = S(0);
      if (qi >= 0) {
        // four independent partial sums: four loads of Y and F in flight per step
        const S* fq = F + s.off[qi] + (int64_t)rc * ni;
        const S* yp = Y + (int64_t)base * R + rc;
        const int64_t ys = (int64_t)si * R;
        S p0 = S(0), p1 = S(0), p2 = S(0), p3 = S(0);
        int iq = 0;
        for (; iq + 4 <= ni; iq += 4) {
          p0 += yp[iq * ys] * fq[iq];
          p1 += yp[(iq + 1) * ys] * fq[iq + 1];
          p2 += yp[(iq + 2) * ys] * fq[iq + 2];
          p3 += yp[(iq + 3) * ys] * fq[iq + 3];
        }
        for (; iq < ni; ++iq) p0 += yp[iq * ys] * fq[iq];
        inner = (p0 + p1) + (p2 + p3);
      } else {
        inner = Y[(int64_t)base * R + rc];
      }
      acc += pre * inner;
    }
    __syncthreads();
    red[wv][lane] = rok ? acc : S(0);
    __syncthreads();
    if (wv == 0 && rok) {
      S tot = red[0][lane];
#pragma unroll
      for (int q = 1; q < NWV; ++q) tot += red[q][lane];
      out[s.off[m] + (int64_t)r * s.d[m] + i] = tot;
    }
  }
}

// Residual: per-workgroup partial sums of (T - model)^2 in FP64.
template <typename TS, typename S, int RMAX>
__global__ __launch_bounds__(256) void k_residual_rows(const TS* __restrict__ T, Shape s,
                                                        const S* __restrict__ F,
                                                        int64_t rows_per_wg,
                                                        double* __restrict__ part) {
  extern __shared__ __align__(16) unsigned char smem[];
  constexpr int FS = RMAX + 1;     // padded row pitch: threads of a wave read different k
  S* W = (S*)smem;                 // RB x RMAX
  S* Fl = W + RB * RMAX;           // KC x FS
  __shared__ double red[4];
  const int t = threadIdx.x;
  const int L = s.L, R = s.R;
  const int64_t K = s.d[L - 1];
  const int64_t NR = s.N / K;
  const int64_t w = blockIdx.x;
  const int64_t rho_begin = w * rows_per_wg;
  const int64_t rho_end = min(NR, rho_begin + rows_per_wg);
  const S* Flast = F + s.off[L - 1];
  double acc = 0.0;
  const int kk = t & 31;
  const int rq = t >> 5;  // 0..7, rows rq + 8 q
  for (int64_t rho0 = rho_begin; rho0 < rho_end; rho0 += RB) {
    const int nrows = (int)min((int64_t)RB, rho_end - rho0);
    __syncthreads();
    for (int e = t; e < RB * RMAX; e += 256) {
      const int rr = e / RMAX, r = e % RMAX;
      S v = S(0);
      if (rr < nrows && r < R) {
        int64_t idx[TFX_MAXL];
        row_to_idx(s, rho0 + rr, idx);
        v = S(1);
        for (int m = 0; m < L - 1; ++m) v *= F[s.off[m] + (int64_t)r * s.d[m] + idx[m]];
      }
      W[e] = v;
    }
    for (int64_t k0 = 0; k0 < K; k0 += KC) {
      __syncthreads();
      for (int e = t; e < KC * RMAX; e += 256) {
        const int r = e / KC, k = e % KC;
        S v = S(0);
        if (r < R && k0 + k < K) v = Flast[(int64_t)r * K + k0 + k];
        Fl[k * FS + r] = v;
      }
      __syncthreads();
      if (k0 + kk < K) {
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          const int rr = rq + 8 * q;
          if (rr < nrows) {
            S mdl = S(0);
            for (int r = 0; r < R; ++r) mdl += W[rr * RMAX + r] * Fl[kk * FS + r];
            const double d = (double)T[(rho0 + rr) * K + k0 + kk] - (double)mdl;
            acc += d * d;
          }
        }
      }
    }
  }
  const double tot = block_sum(acc, red);
  if (t == 0) part[w] = tot;
}

__global__ void k_sum_parts(const double* __restrict__ part, int n, double* __restrict__ out) {
  __shared__ double red[4];
  double acc = 0.0;
  for (int i = threadIdx.x; i < n; i += blockDim.x) acc += part[i];
  const double tot = block_sum(acc, red);
  if (threadIdx.x == 0) *out = tot;
}

template <typename TS, typename S>
int mttkrp_native(tfx_ctx* ctx, const TS* T, const Shape& s, const S* F, unsigned mask, S* out,
                  hipStream_t stream) {
  constexpr int RC = RCOf<S>::v;
  const int L = s.L, R = s.R;
  const int64_t K = s.d[L - 1];
  const int64_t NR = s.N / K;
  if (NR > 0x7fffffff)
    return tfx_set_error(ctx, TFX_E_UNSUPPORTED, "MTTKRP (FMA form) needs fewer than 2^31 rows");
  const int64_t nblocks = (NR + RB - 1) / RB;
  int nwg = (int)std::min<int64_t>(nblocks, (int64_t)ctx->cu_count);
  if (nwg < 1) nwg = 1;
  const int64_t rows_per_wg = ((nblocks + nwg - 1) / nwg) * RB;
  nwg = (int)((NR + rows_per_wg - 1) / rows_per_wg);
  const bool need_rows = (mask & ((1u << (L - 1)) - 1u)) != 0;
  const size_t y_bytes = need_rows ? align256(sizeof(S) * NR * R) : 0;
  const size_t p_bytes = align256(sizeof(S) * (size_t)nwg * K * RC);
  void* wsp;
  int rc = tfx_workspace(ctx, y_bytes + p_bytes + 256, stream, &wsp);
  if (rc) return rc;
  S* Y = need_rows ? (S*)wsp : nullptr;
  S* part = (S*)((char*)wsp + y_bytes);

  size_t lds_base = sizeof(S) * (RB * (KC + 1) + KC * RC + RB * RC);
  size_t lds_np = sizeof(S) * (size_t)K * RC;
  const bool np_in_lds = (lds_base + lds_np) <= 150 * 1024;
  size_t lds = lds_base + (np_in_lds ? lds_np : 0);
  const void* kfn = np_in_lds ? (const void*)k_mttkrp_rows<TS, S, true>
                              : (const void*)k_mttkrp_rows<TS, S, false>;
  TFX_HIP(ctx, hipFuncSetAttribute(kfn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
  for (int rc0 = 0; rc0 < R; rc0 += RC) {
    if (np_in_lds)
      hipLaunchKernelGGL((k_mttkrp_rows<TS, S, true>), dim3(nwg), dim3(256), lds, stream, T, s, F,
                         rc0, rows_per_wg, Y, part);
    else
      hipLaunchKernelGGL((k_mttkrp_rows<TS, S, false>), dim3(nwg), dim3(256), lds, stream, T, s, F,
                         rc0, rows_per_wg, Y, part);
    TFX_CHECK_LAUNCH(ctx);
    if ((mask >> (L - 1)) & 1u) {
      hipLaunchKernelGGL(k_mttkrp_reduce_last<S>, dim3((unsigned)K), dim3(256), 0, stream, s, rc0,
                         nwg, part, out);
      TFX_CHECK_LAUNCH(ctx);
    }
  }
  if (need_rows) {
    int64_t nb = 0;
    for (int m = 0; m < L - 1; ++m) nb += s.d[m];
    hipLaunchKernelGGL(k_mttkrp_rowreduce<S>, dim3((unsigned)nb), dim3(512), 0, stream, s, F, Y,
                       mask, out);
    TFX_CHECK_LAUNCH(ctx);
  }
  return TFX_OK;
}

template <typename TS, typename S, int RMAX>
int residual_native(tfx_ctx* ctx, const TS* T, const Shape& s, const S* F, double* out,
                    hipStream_t stream) {
  const int64_t K = s.d[s.L - 1];
  const int64_t NR = s.N / K;
  const int64_t nblocks = (NR + RB - 1) / RB;
  int nwg = (int)std::min<int64_t>(nblocks, (int64_t)ctx->cu_count * 4);
  if (nwg < 1) nwg = 1;
  const int64_t rows_per_wg = ((nblocks + nwg - 1) / nwg) * RB;
  nwg = (int)((NR + rows_per_wg - 1) / rows_per_wg);
  void* wsp;
  int rc = tfx_workspace(ctx, align256(sizeof(double) * nwg), stream, &wsp);
  if (rc) return rc;
  double* part = (double*)wsp;
  size_t lds = sizeof(S) * (RB * RMAX + KC * (RMAX + 1));
  TFX_HIP(ctx, hipFuncSetAttribute((const void*)k_residual_rows<TS, S, RMAX>,
                                   hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
  hipLaunchKernelGGL((k_residual_rows<TS, S, RMAX>), dim3(nwg), dim3(256), lds, stream, T, s, F,
                     rows_per_wg, part);
  TFX_CHECK_LAUNCH(ctx);
  hipLaunchKernelGGL(k_sum_parts, dim3(1), dim3(256), 0, stream, part, nwg, out);
  TFX_CHECK_LAUNCH(ctx);
  return TFX_OK;
}

template <typename TS, typename S>
int residual_dispatch(tfx_ctx* ctx, const void* T, const Shape& s, const void* F, double* out,
                      hipStream_t stream) {
  if (s.R <= 64) return residual_native<TS, S, 64>(ctx, (const TS*)T, s, (const S*)F, out, stream);
  return residual_native<TS, S, 256>(ctx, (const TS*)T, s, (const S*)F, out, stream);
}

__global__ __launch_bounds__(256) void k_stats_f32(const float* __restrict__ x, int64_t n, double* part) {
  __shared__ double red[4];
  double a = 0.0, q = 0.0;
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < n; e += (int64_t)gridDim.x * blockDim.x) {
    const double v = x[e];
    a += fabs(v);
    q += v * v;
  }
  a = block_sum(a, red);
  q = block_sum(q, red);
  if (threadIdx.x == 0) {
    part[2 * blockIdx.x] = a;
    part[2 * blockIdx.x + 1] = q;
  }
}

__global__ __launch_bounds__(256) void k_stats_f64(const double* __restrict__ x, int64_t n, double* part) {
  __shared__ double red[4];
  double a = 0.0, q = 0.0;
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < n; e += (int64_t)gridDim.x * blockDim.x) {
    const double v = x[e];
    a += fabs(v);
    q += v * v;
  }
  a = block_sum(a, red);
  q = block_sum(q, red);
  if (threadIdx.x == 0) {
    part[2 * blockIdx.x] = a;
    part[2 * blockIdx.x + 1] = q;
  }
}

__global__ void k_stats_final(const double* __restrict__ part, int nb, double* out) {
  __shared__ double red[4];
  double a = 0.0, q = 0.0;
  for (int i = threadIdx.x; i < nb; i += blockDim.x) {
    a += part[2 * i];
    q += part[2 * i + 1];
  }
  a = block_sum(a, red);
  q = block_sum(q, red);
  if (threadIdx.x == 0) {
    out[0] = a;
    out[1] = q;
  }
}

}  // namespace

int tfx_mttkrp_split16(tfx_ctx* ctx, const float* T, const Shape& s, const double* F,
                       unsigned mask, double* out, hipStream_t stream);

int tfx_mttkrp_f64mfma(tfx_ctx* ctx, const double* T, const Shape& s, const double* F,
                       unsigned mask, double* out, hipStream_t stream);

int tfx_mttkrp_native_fd(tfx_ctx* ctx, const float* T, const Shape& s, const double* F,
                         unsigned mask, double* out, hipStream_t stream) {
  return mttkrp_native<float, double>(ctx, T, s, F, mask, out, stream);
}

int tfx_mttkrp_split16_fits(const Shape& s);

// Kernel family chosen for a call (host only).
static int mttkrp_path(const Shape& s, int t_dtype, int f_dtype, unsigned mask, int prec) {
  if (prec == TFX_PREC_SPLIT16) {
    if (t_dtype != TFX_F32 || f_dtype != TFX_F64) return TFX_E_ARG;
    return tfx_mttkrp_split16_fits(s) ? TFX_PATH_SPLIT16_MFMA : TFX_PATH_FMA;
  }
  if (t_dtype == TFX_F32 && (f_dtype == TFX_F32 || f_dtype == TFX_F64)) return TFX_PATH_FMA;
  if (t_dtype == TFX_F64 && f_dtype == TFX_F64) {
    // FP64 matrix cores; the FMA form takes unfoldings with 2^31 columns or more
    for (int l = 0; l < s.L; ++l)
      if (((mask >> l) & 1u) && s.N / s.d[l] > 0x7fffffff) return TFX_PATH_FMA;
    return TFX_PATH_F64_MFMA;
  }
  return TFX_E_ARG;
}

extern "C" int tfx_mttkrp_plan(int t_dtype, int f_dtype, int prec, int L, const int64_t* dims,
                               int R, unsigned mode_mask) {
  if (!dims) return TFX_E_ARG;
  Shape s;
  const int rc = tfx_make_shape(nullptr, L, dims, R, &s);
  if (rc) return rc;
  return mttkrp_path(s, t_dtype, f_dtype, mode_mask, prec);
}

extern "C" int tfx_mttkrp_mixed(tfx_ctx* ctx, const void* T, int t_dtype, int L,
                                const int64_t* dims, const void* F_flat, int f_dtype, int R,
                                unsigned mode_mask, void* out_flat, int prec, hipStream_t stream) {
  if (!ctx || !T || !F_flat || !out_flat) return tfx_set_error(ctx, TFX_E_ARG, "null argument");
  Shape s;
  int rc = tfx_make_shape(ctx, L, dims, R, &s);
  if (rc) return rc;
  if (mode_mask == 0) return TFX_OK;
  const int path = mttkrp_path(s, t_dtype, f_dtype, mode_mask, prec);
  if (path < 0)
    return tfx_set_error(ctx, TFX_E_ARG,
                         prec == TFX_PREC_SPLIT16
                             ? "split16 MTTKRP needs an FP32 tensor and FP64 factors"
                             : "unsupported dtype pair %d/%d", t_dtype, f_dtype);
  if (t_dtype == TFX_F32 && f_dtype == TFX_F64) {
    if (path == TFX_PATH_SPLIT16_MFMA)
      return tfx_mttkrp_split16(ctx, (const float*)T, s, (const double*)F_flat, mode_mask,
                                (double*)out_flat, stream);
    return mttkrp_native<float, double>(ctx, (const float*)T, s, (const double*)F_flat, mode_mask,
                                        (double*)out_flat, stream);
  }
  if (t_dtype == TFX_F32)
    return mttkrp_native<float, float>(ctx, (const float*)T, s, (const float*)F_flat, mode_mask,
                                       (float*)out_flat, stream);
  if (path == TFX_PATH_F64_MFMA)
    return tfx_mttkrp_f64mfma(ctx, (const double*)T, s, (const double*)F_flat, mode_mask,
                              (double*)out_flat, stream);
  return mttkrp_native<double, double>(ctx, (const double*)T, s, (const double*)F_flat, mode_mask,
                                       (double*)out_flat, stream);
}

int tfx_mttkrp_f64_uses_gemm3(const Shape& s);
int tfx_mttkrp_f64mfma_g(tfx_ctx* ctx, const double* T, const Shape& s, const double* F,
                         unsigned mask, double* out, const double* P1, double* grad,
                         hipStream_t stream);

extern "C" int tfx_mttkrp_neg_grad(tfx_ctx* ctx, const void* T, int t_dtype, int L,
                                   const int64_t* dims, const void* F_flat, int f_dtype, int R,
                                   const void* P1, void* N_flat, void* grad_flat, int prec,
                                   hipStream_t stream) {
  if (!ctx || !T || !F_flat || !P1 || !N_flat || !grad_flat)
    return tfx_set_error(ctx, TFX_E_ARG, "null argument");
  Shape s;
  int rc = tfx_make_shape(ctx, L, dims, R, &s);
  if (rc) return rc;
  const unsigned all = (1u << L) - 1u;
  if (t_dtype == TFX_F64 && f_dtype == TFX_F64 &&
      mttkrp_path(s, t_dtype, f_dtype, all, prec) == TFX_PATH_F64_MFMA &&
      !tfx_mttkrp_f64_uses_gemm3(s))
    return tfx_mttkrp_f64mfma_g(ctx, (const double*)T, s, (const double*)F_flat, all,
                                (double*)N_flat, (const double*)P1, (double*)grad_flat, stream);
  rc = tfx_mttkrp_mixed(ctx, T, t_dtype, L, dims, F_flat, f_dtype, R, all, N_flat, prec, stream);
  if (rc) return rc;
  return tfx_neg_grad(ctx, f_dtype, L, dims, R, F_flat, P1, N_flat, grad_flat, stream);
}

extern "C" int tfx_mttkrp(tfx_ctx* ctx, const void* T, int dtype, int L, const int64_t* dims,
                          const void* F_flat, int R, unsigned mode_mask, void* out_flat, int prec,
                          hipStream_t stream) {
  return tfx_mttkrp_mixed(ctx, T, dtype, L, dims, F_flat, dtype, R, mode_mask, out_flat, prec, stream);
}

extern "C" int tfx_cpd_residual(tfx_ctx* ctx, const void* T, int dtype, int L, const int64_t* dims,
                                const void* F_flat, int R, double* out_sqnorm, hipStream_t stream) {
  return tfx_cpd_residual_mixed(ctx, T, dtype, L, dims, F_flat, dtype, R, out_sqnorm, stream);
}

extern "C" int tfx_cpd_residual_mixed(tfx_ctx* ctx, const void* T, int t_dtype, int L,
                                      const int64_t* dims, const void* F_flat, int f_dtype, int R,
                                      double* out_sqnorm, hipStream_t stream) {
  if (!ctx || !T || !F_flat || !out_sqnorm) return tfx_set_error(ctx, TFX_E_ARG, "null argument");
  Shape s;
  int rc = tfx_make_shape(ctx, L, dims, R, &s);
  if (rc) return rc;
  if (R > 256) return tfx_set_error(ctx, TFX_E_UNSUPPORTED, "residual supports R <= 256");
  if (f_dtype == TFX_F64) {
    const int m = residual_mfma_try(ctx, T, t_dtype, s, (const double*)F_flat, out_sqnorm, stream);
    if (m != 0) return m < 0 ? m : TFX_OK;
  }
  if (t_dtype == TFX_F32 && f_dtype == TFX_F32) return residual_dispatch<float, float>(ctx, T, s, F_flat, out_sqnorm, stream);
  if (t_dtype == TFX_F32 && f_dtype == TFX_F64) return residual_dispatch<float, double>(ctx, T, s, F_flat, out_sqnorm, stream);
  if (t_dtype == TFX_F64 && f_dtype == TFX_F64) return residual_dispatch<double, double>(ctx, T, s, F_flat, out_sqnorm, stream);
  return tfx_set_error(ctx, TFX_E_ARG, "unsupported dtype pair %d/%d", t_dtype, f_dtype);
}

extern "C" int tfx_tensor_stats(tfx_ctx* ctx, const void* X, int dtype, int64_t n, double* out,
                                hipStream_t stream) {
  if (!ctx || !X || !out) return tfx_set_error(ctx, TFX_E_ARG, "null argument");
  int nb = (int)std::min<int64_t>((n + 255) / 256, (int64_t)ctx->cu_count * 4);
  if (nb < 1) nb = 1;
  void* wsp;
  int rc = tfx_workspace(ctx, sizeof(double) * 2 * nb + 256, stream, &wsp);
  if (rc) return rc;
  double* part = (double*)wsp;
  if (dtype == TFX_F32)
    hipLaunchKernelGGL(k_stats_f32, dim3(nb), dim3(256), 0, stream, (const float*)X, n, part);
  else if (dtype == TFX_F64)
    hipLaunchKernelGGL(k_stats_f64, dim3(nb), dim3(256), 0, stream, (const double*)X, n, part);
  else
    return tfx_set_error(ctx, TFX_E_ARG, "bad dtype %d", dtype);
  TFX_CHECK_LAUNCH(ctx);
  hipLaunchKernelGGL(k_stats_final, dim3(1), dim3(256), 0, stream, part, nb, out);
  TFX_CHECK_LAUNCH(ctx);
  return TFX_OK;
}
