// ||T - [[F]]||^2 with the model on FP64 MFMA (cpd2unfold1 + fastnorm, Conversion.py:117-141,
// Critical.py:14-22), FP64 factors, rank R <= 64, order L = 2..4.
//
// T (C order) is NR = N / K rows of length K = I_{L-1}.  The model of row rho = (i_0..i_{L-2})
// is M[rho, k] = sum_r W[rho, r] C[k, r] with W[rho, r] = prod_{m < L-1} F_m[i_m, r] and
// C = F_{L-1}: a (NR x R) (R x K) product, formed 16 x 16 at a time by v_mfma_f64_16x16x4_f64
// (A = W from registers, B = C from LDS), compared with T in registers and never stored.
// A workgroup owns 64 columns of the last mode and a run of 16-row tiles; its C chunk is staged
// once in LDS.  Accumulation of (T - M)^2 is FP64.  The per-workgroup partials are summed by the
// last workgroup to finish (an arrival counter in the context scratch, reset by that workgroup;
// no workgroup waits for another), 256 threads over the partials in index order, so the result
// does not depend on scheduling.
#include "common.h"

namespace {

typedef double dbl4 __attribute__((ext_vector_type(4)));

constexpr int NT = 256;         // 4 waves
constexpr int KCH = 64;         // columns per workgroup (4 tiles of 16)
constexpr int CP = KCH + 1;     // LDS pitch of the staged C chunk (odd)

__device__ inline dbl4 mfma64(double a, double b, dbl4 c) {
  return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}

// NQ: MFMA k steps per tile, a compile-time count (R rounded up to a multiple of 8, ranks past R
// masked), so the step loop unrolls without branches.
// Sum of the n partials by one workgroup of NT threads (strided accumulation, then block_sum).
__device__ inline void sum_partials(const double* part, int n, double* red, double* out) {
  double acc = 0.0;
  for (int i = threadIdx.x; i < n; i += NT) acc += __builtin_nontemporal_load(part + i);
  const double tot = block_sum(acc, red);
  if (threadIdx.x == 0) *out = tot;
}

template <typename TS, int LM1, int NQ>
__global__ __launch_bounds__(NT) void k_residual_mfma(const TS* __restrict__ T, Shape s,
                                                       const double* __restrict__ F,
                                                       int tiles_per_wave,
                                                       double* __restrict__ part,
                                                       unsigned* __restrict__ counter,
                                                       double* __restrict__ out) {
  __shared__ double Cs[64 * CP];
  __shared__ double red[NT / 64];
  const int R = s.R;
  const int K = (int)s.d[LM1];
  const int64_t NR = s.N / K;
  const int kc0 = blockIdx.x * KCH;
  const int t = threadIdx.x, wave = t >> 6, lane = t & 63, li = lane & 15, lg = lane >> 4;
  const int R4 = 4 * NQ;   // rows of the staged C chunk (ranks past R are zero)
  const double* FL = F + s.off[LM1];
  for (int e = t; e < R4 * KCH; e += NT) {
    const int r = e / KCH, kk = e - r * KCH;
    const int k = kc0 + kk;
    Cs[r * CP + kk] = (r < R && k < K) ? FL[(int64_t)r * K + k] : 0.0;
  }
  __syncthreads();
  int dm[LM1];
  int64_t offm[LM1];
#pragma unroll
  for (int m = 0; m < LM1; ++m) {
    dm[m] = (int)s.d[m];
    offm[m] = s.off[m];
  }
  double acc2 = 0.0;
  const int64_t rt_begin = (int64_t)blockIdx.y * tiles_per_wave * 4;
  for (int it = 0; it < tiles_per_wave; ++it) {
    const int64_t rho0 = (rt_begin + (int64_t)it * 4 + wave) * 16;
    if (rho0 >= NR) break;   // wave-uniform
    // T values of the 4 column tiles, in the accumulator layout: row 4 j + lg, column li
    TS tv[4][4];
#pragma unroll
    for (int ct = 0; ct < 4; ++ct) {
      const int col = kc0 + 16 * ct + li;
      const int colc = col < K ? col : K - 1;
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const int64_t row = rho0 + 4 * j + lg;
        const int64_t rowc = row < NR ? row : NR - 1;
        tv[ct][j] = T[rowc * K + colc];
      }
    }
    // A fragment: W[rho0 + li][4 q + lg]
    const int64_t rho = rho0 + li;
    const bool rok = rho < NR;
    int idx[LM1];
    {
      int64_t rem = rok ? rho : 0;
#pragma unroll
      for (int m = LM1 - 1; m >= 0; --m) {
        const int64_t qd = rem / dm[m];
        idx[m] = (int)(rem - qd * dm[m]);
        rem = qd;
      }
    }
    double a[NQ];
#pragma unroll
    for (int q = 0; q < NQ; ++q) {
      const int r = 4 * q + lg;
      const int rc = r < R ? r : R - 1;
      double w = 1.0;
#pragma unroll
      for (int m = 0; m < LM1; ++m) w *= F[offm[m] + (int64_t)rc * dm[m] + idx[m]];
      a[q] = (rok && r < R) ? w : 0.0;
    }
    dbl4 acc[4];
#pragma unroll
    for (int ct = 0; ct < 4; ++ct) acc[ct] = dbl4{0.0, 0.0, 0.0, 0.0};
    const double* cb = Cs + lg * CP + li;
#pragma unroll
    for (int q = 0; q < NQ; ++q)
#pragma unroll
      for (int ct = 0; ct < 4; ++ct) acc[ct] = mfma64(a[q], cb[4 * q * CP + 16 * ct], acc[ct]);
#pragma unroll
    for (int ct = 0; ct < 4; ++ct) {
      const bool cok = kc0 + 16 * ct + li < K;
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const bool ok = cok && rho0 + 4 * j + lg < NR;
        const double d = (double)tv[ct][j] - acc[ct][j];
        acc2 += ok ? d * d : 0.0;
      }
    }
  }
  const double tot = block_sum(acc2, red);
  __shared__ int is_last;
  const unsigned nwg = gridDim.x * gridDim.y;
  if (t == 0) {
    part[(int64_t)blockIdx.y * gridDim.x + blockIdx.x] = tot;
    __threadfence();
    is_last = atomicAdd(counter, 1u) == nwg - 1;
  }
  __syncthreads();
  if (is_last) {
    __threadfence();
    sum_partials(part, (int)nwg, red, out);
    if (t == 0) *counter = 0u;
  }
}


template <typename TS, int LM1, int NQ>
void launch_nq(const TS* T, const Shape& s, const double* F, int kchunks, int groups, int tpw,
               double* part, unsigned* counter, double* out, hipStream_t st) {
  hipLaunchKernelGGL((k_residual_mfma<TS, LM1, NQ>), dim3((unsigned)kchunks, (unsigned)groups),
                     dim3(NT), 0, st, T, s, F, tpw, part, counter, out);
}

template <typename TS, int LM1>
int launch(tfx_ctx* ctx, const TS* T, const Shape& s, const double* F, double* out,
           hipStream_t st) {
  const int64_t K = s.d[s.L - 1];
  const int64_t NR = s.N / K;
  const int64_t kchunks = (K + KCH - 1) / KCH;
  const int64_t rtiles = (NR + 15) / 16;
  // about 8 workgroups per CU, at least one 16-row tile per wave
  const int64_t target = 8LL * ctx->cu_count;
  int64_t groups = std::max<int64_t>(1, target / kchunks);
  int64_t tpw = std::max<int64_t>(1, (rtiles + 4 * groups - 1) / (4 * groups));
  tpw = std::min<int64_t>(tpw, 64);
  groups = (rtiles + 4 * tpw - 1) / (4 * tpw);
  if (groups > 65535 || kchunks > 65535) return 0;
  const int64_t nparts = groups * kchunks;
  void* wsp;
  int rc = tfx_workspace(ctx, align256(sizeof(double) * nparts), st, &wsp);
  if (rc) return rc;
  double* part = (double*)wsp;
  const int nq8 = ((s.R + 3) / 4 + 1) / 2;   // k steps in pairs: 2 nq8 steps, 8 nq8 >= R
  const int kc = (int)kchunks, gr = (int)groups, tp = (int)tpw;
  unsigned* counter = (unsigned*)(ctx->dscal + TFX_DSCAL_RES_COUNTER);
  switch (nq8) {
    case 1: launch_nq<TS, LM1, 2>(T, s, F, kc, gr, tp, part, counter, out, st); break;
    case 2: launch_nq<TS, LM1, 4>(T, s, F, kc, gr, tp, part, counter, out, st); break;
    case 3: launch_nq<TS, LM1, 6>(T, s, F, kc, gr, tp, part, counter, out, st); break;
    case 4: launch_nq<TS, LM1, 8>(T, s, F, kc, gr, tp, part, counter, out, st); break;
    case 5: launch_nq<TS, LM1, 10>(T, s, F, kc, gr, tp, part, counter, out, st); break;
    case 6: launch_nq<TS, LM1, 12>(T, s, F, kc, gr, tp, part, counter, out, st); break;
    case 7: launch_nq<TS, LM1, 14>(T, s, F, kc, gr, tp, part, counter, out, st); break;
    default: launch_nq<TS, LM1, 16>(T, s, F, kc, gr, tp, part, counter, out, st); break;
  }
  TFX_CHECK_LAUNCH(ctx);
  return 1;
}

}  // namespace

// The FP64 MFMA residual for FP64 factors, R <= 64, L = 2..4: returns 1 when it ran, 0 when the
// shape is not covered (the caller then uses the FMA kernel), or an error code.
int residual_mfma_try(tfx_ctx* ctx, const void* T, int t_dtype, const Shape& s, const double* F,
                      double* out, hipStream_t st) {
  if (s.R > 64 || s.L < 2 || s.L > 4) return 0;
  if (s.N / s.d[s.L - 1] > (int64_t)1 << 40) return 0;
  for (int l = 0; l < s.L; ++l)
    if (s.d[l] > 0x7fffffff) return 0;
  if (t_dtype == TFX_F32) {
    const float* Tp = (const float*)T;
    if (s.L == 2) return launch<float, 1>(ctx, Tp, s, F, out, st);
    if (s.L == 3) return launch<float, 2>(ctx, Tp, s, F, out, st);
    return launch<float, 3>(ctx, Tp, s, F, out, st);
  }
  if (t_dtype == TFX_F64) {
    const double* Tp = (const double*)T;
    if (s.L == 2) return launch<double, 1>(ctx, Tp, s, F, out, st);
    if (s.L == 3) return launch<double, 2>(ctx, Tp, s, F, out, st);
    return launch<double, 3>(ctx, Tp, s, F, out, st);
  }
  return 0;
}
