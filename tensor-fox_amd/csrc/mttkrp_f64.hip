// MTTKRP of the selected modes of an FP64 tensor on the FP64 matrix cores
// (v_mfma_f64_16x16x4f64), FP64 factors and outputs (GaussNewton.py:584-606 forms these products
// for the gradient; CPD MTTKRP N_l = T_(l) (KR of the other factors)).
//
//   N_l[i, r] = sum_c T_(l)[i, c] KR_l[c, r],   KR_l[c, r] = prod_{m != l} F_m[i_m(c), r]
//
// c enumerates the indices of the modes other than l in C order (last mode fastest).  A wavefront
// owns rows i0..i0+15 of mode l, the 64 ranks rc0..rc0+63 of the launch (four 16 x 16
// accumulators in registers) and a slice of c.  Per MFMA step (4 values of c) a lane loads one
// entry of T (the A operand: row i0 + lane % 16, c = c0 + lane / 16) and forms its B operand
// entries KR_l[c, r] for the four rank blocks from factor entries (cached: the factors are
// small next to T), so T is read once per mode and rank chunk.  The four wavefronts of a
// workgroup take four consecutive slices of the same rows and add their tiles in LDS (in wave
// order); the workgroup tiles are summed in a fixed order by k_m64_sum, so results do not depend
// on scheduling.  Slices are short (many wavefronts) so that load latency is covered by other
// wavefronts rather than by prefetch registers.
//
// Fragment layouts of v_mfma_f64_16x16x4f64 (lane l): A[l % 16][l / 16], B[l / 16][l % 16],
// result register j = D[4 j + l / 16][l % 16].
#include "common.h"

int tfx_mttkrp_f64mfma(tfx_ctx* ctx, const double* T, const Shape& s, const double* F,
                       unsigned mask, double* out, hipStream_t stream);
int tfx_mttkrp_f64_gemm3(tfx_ctx* ctx, const double* T, const Shape& s, const double* F,
                         unsigned mask, double* out, hipStream_t stream);

// Third-order tensors of at least this many entries take the two-pass GEMM form
// (mttkrp_f64_gemm.hip), whose workspace holds an (I0 I1) x 64 matrix.
constexpr int64_t F64_GEMM_MIN_N = int64_t(1) << 22;
constexpr int64_t F64_GEMM_MAX_W = int64_t(1) << 29;   // doubles of W (4 GiB)

namespace {

typedef double f64x4 __attribute__((ext_vector_type(4)));

constexpr int M64_NT = 256;      // 4 wavefronts (slices) per workgroup
constexpr int M64_RC = 64;       // ranks per launch
constexpr int M64_MINSLICE = 16; // c per wavefront at least (4 MFMA steps)
constexpr int M64_PART = 16 * M64_RC;

struct M64Plan {
  Shape s;
  int rc0;
  int nsel;                      // selected modes
  int modes[TFX_MAXL];           // selected modes in workgroup order
  int first[TFX_MAXL + 1];       // first workgroup of each selected mode
  int ngr[TFX_MAXL];             // slice groups (workgroups) per row block
  int64_t slice[TFX_MAXL];       // c per slice (multiple of 16)
  int64_t stride[TFX_MAXL];      // C-order strides of T
  const double* F;               // with P1 and grad: the sum also writes grad = N - F_l P1_l
  const double* P1;
  double* grad;                  // nullptr: N only
};

template <int L>
__global__ __launch_bounds__(M64_NT) void k_mttkrp_f64(const double* __restrict__ T, M64Plan P,
                                                       const double* __restrict__ F,
                                                       double* __restrict__ part) {
  __shared__ f64x4 red[3][4][64];
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63, li = lane & 15, lk = lane >> 4;
  const int wg = blockIdx.x;
  int ms = 0;
  while (ms + 1 < P.nsel && wg >= P.first[ms + 1]) ++ms;
  const int l = P.modes[ms];
  const int local = wg - P.first[ms];
  const int ib = local / P.ngr[ms], grp = local - (local / P.ngr[ms]) * P.ngr[ms];
  const int sp = 4 * grp + wave;
  const Shape& s = P.s;
  const int64_t Il = s.d[l];
  const int64_t Kl = s.N / Il;
  const int64_t cs = (int64_t)sp * P.slice[ms];
  const int64_t ce = min(Kl, cs + P.slice[ms]);
  // the other modes, slowest first (uniform)
  int64_t od[L - 1], ostr[L - 1];
  const double* fb[L - 1][4];
#pragma unroll
  for (int t = 0; t < L - 1; ++t) {
    const int q = t < l ? t : t + 1;
    od[t] = s.d[q];
    ostr[t] = P.stride[q];
#pragma unroll
    for (int b = 0; b < 4; ++b) {
      const int r = min(P.rc0 + 16 * b + li, s.R - 1);   // ranks past R: any valid entry
      fb[t][b] = F + s.off[q] + (int64_t)r * s.d[q];
    }
  }
  // rows past I_l read row I_l - 1 (their results are not stored)
  const int64_t abase = min((int64_t)ib * 16 + li, Il - 1) * P.stride[l];
  f64x4 acc[4];
#pragma unroll
  for (int b = 0; b < 4; ++b) acc[b] = f64x4{0.0, 0.0, 0.0, 0.0};
  for (int64_t c0 = cs; c0 < ce; c0 += 16) {
    double av[4], bv[4][4];
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const int64_t c = c0 + 4 * u + lk;
      const bool ok = c < ce;
      uint32_t cc = (uint32_t)(ok ? c : cs);
      int64_t aoff = abase;
      uint32_t ix[L - 1];
#pragma unroll
      for (int t = L - 2; t >= 1; --t) {
        const uint32_t d = (uint32_t)od[t];
        const uint32_t qd = cc / d;
        ix[t] = cc - qd * d;
        cc = qd;
      }
      ix[0] = cc;
#pragma unroll
      for (int t = 0; t < L - 1; ++t) aoff += (int64_t)ix[t] * ostr[t];
      double kr[4];
#pragma unroll
      for (int b = 0; b < 4; ++b) {
        kr[b] = fb[0][b][ix[0]];
#pragma unroll
        for (int t = 1; t < L - 1; ++t) kr[b] *= fb[t][b][ix[t]];
      }
      const double a = T[aoff];
      av[u] = ok ? a : 0.0;
#pragma unroll
      for (int b = 0; b < 4; ++b) bv[u][b] = kr[b];
    }
#pragma unroll
    for (int u = 0; u < 4; ++u)
#pragma unroll
      for (int b = 0; b < 4; ++b)
        acc[b] = __builtin_amdgcn_mfma_f64_16x16x4f64(av[u], bv[u][b], acc[b], 0, 0, 0);
  }
  // workgroup tile = wave 0 + wave 1 + wave 2 + wave 3, then part[wg][row][col] with
  // row = 4 j + lane / 16, col = 16 b + lane % 16
  if (wave > 0) {
#pragma unroll
    for (int b = 0; b < 4; ++b) red[wave - 1][b][lane] = acc[b];
  }
  __syncthreads();
  if (wave == 0) {
#pragma unroll
    for (int w = 0; w < 3; ++w)
#pragma unroll
      for (int b = 0; b < 4; ++b) acc[b] += red[w][b][lane];
    double* pj = part + (int64_t)wg * M64_PART;
#pragma unroll
    for (int b = 0; b < 4; ++b)
#pragma unroll
      for (int j = 0; j < 4; ++j) pj[(4 * j + lk) * M64_RC + 16 * b + li] = acc[b][j];
  }
}

// out[off_l + r I_l + i] = sum over the slices (in order) of the partial tiles; one thread per
// (selected mode, row, rank of the chunk), ranks fastest.
__global__ __launch_bounds__(256) void k_m64_sum(M64Plan P, const double* __restrict__ part,
                                                 double* __restrict__ out) {
  const Shape& s = P.s;
  int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x;
  int ms = 0;
  for (; ms < P.nsel; ++ms) {
    const int64_t cnt = s.d[P.modes[ms]] * M64_RC;
    if (e < cnt) break;
    e -= cnt;
  }
  if (ms >= P.nsel) return;
  const int l = P.modes[ms];
  const int64_t i = e / M64_RC;
  const int rr = (int)(e - i * M64_RC);
  const int r = P.rc0 + rr;
  if (r >= s.R) return;
  const int ib = (int)(i / 16), row = (int)(i - (int64_t)ib * 16);
  const int ng = P.ngr[ms];
  const double* src = part + ((int64_t)(P.first[ms] + ib * ng) * 16 + row) * M64_RC + rr;
  double acc = 0.0;
  int g = 0;
  for (; g + 8 <= ng; g += 8) {   // eight loads in flight, added in order
    double v[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) v[u] = src[(int64_t)(g + u) * M64_PART];
#pragma unroll
    for (int u = 0; u < 8; ++u) acc += v[u];
  }
  for (; g < ng; ++g) acc += src[(int64_t)g * M64_PART];
  const int64_t I = s.d[l];
  const int64_t e_out = s.off[l] + (int64_t)r * I + i;
  out[e_out] = acc;
  if (P.grad) {
    // grad = N - F_l P1_l with tfx_neg_grad's arithmetic: four partial sums combined in a fixed
    // order, so the fused and the separate forms give the same bits
    const double* Fl = P.F + s.off[l];
    const double* P1l = P.P1 + (int64_t)l * s.R * s.R;
    double a0 = 0.0, a1 = 0.0, a2 = 0.0, a3 = 0.0;
    int c = 0;
#pragma unroll 4   // four steps' loads in flight together; each accumulator's order unchanged
    for (; c + 4 <= s.R; c += 4) {
      a0 += Fl[(int64_t)c * I + i] * P1l[(int64_t)c * s.R + r];
      a1 += Fl[(int64_t)(c + 1) * I + i] * P1l[(int64_t)(c + 1) * s.R + r];
      a2 += Fl[(int64_t)(c + 2) * I + i] * P1l[(int64_t)(c + 2) * s.R + r];
      a3 += Fl[(int64_t)(c + 3) * I + i] * P1l[(int64_t)(c + 3) * s.R + r];
    }
    for (; c < s.R; ++c) a0 += Fl[(int64_t)c * I + i] * P1l[(int64_t)c * s.R + r];
    P.grad[e_out] = acc - ((a0 + a1) + (a2 + a3));
  }
}

template <int L>
void launch_m64(const double* T, const M64Plan& P, const double* F, double* part, int nwg,
                hipStream_t stream) {
  hipLaunchKernelGGL(k_mttkrp_f64<L>, dim3(nwg), dim3(M64_NT), 0, stream, T, P, F, part);
}

}  // namespace

// True when tfx_mttkrp_f64mfma uses the two-pass GEMM form for this shape (it then cannot fuse
// the gradient).
int tfx_mttkrp_f64_uses_gemm3(const Shape& s) {
  return s.L == 3 && s.N >= F64_GEMM_MIN_N && s.d[0] * s.d[1] * 64 <= F64_GEMM_MAX_W;
}

// grad != nullptr (with P1): every mode must be selected; the slice sum also writes
// grad = N - F_l P1_l.
int tfx_mttkrp_f64mfma_g(tfx_ctx* ctx, const double* T, const Shape& s, const double* F,
                         unsigned mask, double* out, const double* P1, double* grad,
                         hipStream_t stream) {
  const int L = s.L;
  if (L < 2 || L > TFX_MAXL) return tfx_set_error(ctx, TFX_E_UNSUPPORTED, "MTTKRP order %d", L);
  if (tfx_mttkrp_f64_uses_gemm3(s)) {
    if (grad) return tfx_set_error(ctx, TFX_E_ARG, "gradient fusion not available for this shape");
    return tfx_mttkrp_f64_gemm3(ctx, T, s, F, mask, out, stream);
  }
  M64Plan P;
  P.s = s;
  P.F = F;
  P.P1 = P1;
  P.grad = grad;
  P.stride[L - 1] = 1;
  for (int q = L - 2; q >= 0; --q) P.stride[q] = P.stride[q + 1] * s.d[q + 1];
  P.nsel = 0;
  int64_t rowblocks = 0;
  for (int l = 0; l < L; ++l) {
    if (!((mask >> l) & 1u)) continue;
    if (s.N / s.d[l] > 0x7fffffff)
      return tfx_set_error(ctx, TFX_E_UNSUPPORTED, "MTTKRP (FP64 MFMA) needs fewer than 2^31 columns per unfolding");
    P.modes[P.nsel++] = l;
    rowblocks += (s.d[l] + 15) / 16;
  }
  if (P.nsel == 0) return TFX_OK;
  // slices: about 16 wavefronts per CU over all modes, at least M64_MINSLICE values of c each,
  // in groups of four (one workgroup); the last slices of a row block may be empty
  const int64_t target = 16 * (int64_t)ctx->cu_count;
  int nwg = 0;
  for (int m = 0; m < P.nsel; ++m) {
    const int l = P.modes[m];
    const int64_t Kl = s.N / s.d[l];
    const int64_t nib = (s.d[l] + 15) / 16;
    int64_t nsp = std::max<int64_t>(1, target / rowblocks);
    nsp = std::min<int64_t>(nsp, (Kl + M64_MINSLICE - 1) / M64_MINSLICE);
    int64_t slice = (Kl + nsp - 1) / nsp;
    slice = (slice + 15) & ~int64_t(15);
    nsp = (Kl + slice - 1) / slice;
    const int64_t ng = (nsp + 3) / 4;
    P.first[m] = nwg;
    P.ngr[m] = (int)ng;
    P.slice[m] = slice;
    nwg += (int)(nib * ng);
  }
  P.first[P.nsel] = nwg;
  if (grad && P.nsel != L) return tfx_set_error(ctx, TFX_E_ARG, "gradient fusion needs every mode");
  void* wsp;
  int rc = tfx_workspace(ctx, sizeof(double) * (size_t)nwg * M64_PART + 256, stream, &wsp);
  if (rc) return rc;
  double* part = (double*)wsp;
  int64_t nout = 0;
  for (int m = 0; m < P.nsel; ++m) nout += s.d[P.modes[m]] * M64_RC;
  const unsigned nsum = (unsigned)((nout + 255) / 256);
  for (int rc0 = 0; rc0 < s.R; rc0 += M64_RC) {
    P.rc0 = rc0;
    switch (L) {
      case 2: launch_m64<2>(T, P, F, part, nwg, stream); break;
      case 3: launch_m64<3>(T, P, F, part, nwg, stream); break;
      case 4: launch_m64<4>(T, P, F, part, nwg, stream); break;
      case 5: launch_m64<5>(T, P, F, part, nwg, stream); break;
      case 6: launch_m64<6>(T, P, F, part, nwg, stream); break;
      case 7: launch_m64<7>(T, P, F, part, nwg, stream); break;
      default: launch_m64<8>(T, P, F, part, nwg, stream); break;
    }
    TFX_CHECK_LAUNCH(ctx);
    hipLaunchKernelGGL(k_m64_sum, dim3(nsum), dim3(256), 0, stream, P, part, out);
    TFX_CHECK_LAUNCH(ctx);
  }
  return TFX_OK;
}

int tfx_mttkrp_f64mfma(tfx_ctx* ctx, const double* T, const Shape& s, const double* F,
                       unsigned mask, double* out, hipStream_t stream) {
  return tfx_mttkrp_f64mfma_g(ctx, T, s, F, mask, out, nullptr, nullptr, stream);
}
