// MTTKRP of every mode in one pass over an FP32 tensor on the bf16 matrix cores (gfx950
// v_mfma_f32_16x16x32_bf16), with FP64 factors and outputs.
//
// Split precision: every FP32 operand x is written as hi + lo with hi = bf16(x) and
// lo = bf16(x - hi) (|x - hi - lo| <= 2^-18 |x|); a product is formed as
// hi*hi + hi*lo + lo*hi with FP32 accumulation (the dropped lo*lo term is ~2^-16 relative).
//
// T is read as rows of K = I_{L-1} contiguous values.  A row is (o, j): o the combined index of
// modes 0..L-3 ("outer"), j = i_{L-2}.  A workgroup owns one block of 64 consecutive j and a
// range of o; a tile is the 64 rows (o, j0..j0+63), streamed through LDS in chunks of 32 k.
// Per chunk and per wave w (ranks 16w..16w+15 of the current 64-rank chunk):
//   Y[rows, r]   += T[rows, k] F_{L-1}[k, r]           (GEMM 1, A = T row-major)
//   N_{L-1}[k, r] += T[rows, k]^T W[rows, r]           (GEMM 2, A = T transposed)
// with W[(o, j), r] = prod_{m < L-1} F_m[i_m, r].  After a tile, Y gives the other modes:
//   N_{L-2}[j, r] += Y Fo[r]                           (per-workgroup partial in LDS, FP64)
//   N_m[i_m, r]   += sum_j Y F_{L-2}[j, r] Fo_{!m}[r]  (m < L-2; FP64 atomics per tile)
// N_{L-1} partials (FP32 MFMA accumulators, one per workgroup) are summed in FP64 afterwards.
//
// Register fragment layouts of v_mfma_f32_16x16x32_bf16 (64 lanes, l = lane):
//   A (16 x 32): row l % 16, eight k values in the slot group l / 16
//   B (32 x 16): column l % 16, the same eight k values as A for the same slot group
//   D (16 x 16): column l % 16, rows 4 (l / 16) + 0..3
// Only the D layout matters for correctness: A and B share the k assignment of a slot group.
#include "common.h"

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

int tfx_mttkrp_native_fd(tfx_ctx* ctx, const float* T, const Shape& s, const double* F,
                         unsigned mask, double* out, hipStream_t stream);

namespace {

constexpr int TR = 64;        // rows per tile (block of j)
constexpr int KC = 32;        // k per chunk
constexpr int RC = 64;        // ranks per launch
constexpr int PD = 3;         // chunks prefetched ahead
constexpr int RM_P = 40;      // row-major tile pitch (bf16): 32 + 8
constexpr int CM_P = 72;      // transposed tile pitch (bf16): 64 + 8

struct MPlan {
  Shape s;
  int rc0;
  int64_t K, J, NO;
  int jblocks;
  int64_t ochunk;
  int nwg;
};

struct Lds {
  __bf16* rm;    // [2 buf][2 hi/lo][TR][RM_P]
  __bf16* cm;    // [2 buf][2 hi/lo][KC][CM_P]
  double* fj;    // [TR][RC]  F_{L-2} rows of the block
  double* pj;    // [TR][RC]  partial N_{L-2}
};

__device__ inline void split(float x, __bf16& h, __bf16& l) {
  h = (__bf16)x;
  l = (__bf16)(x - (float)h);
}

// global -> registers: 8 consecutive k of one row (zeros outside the tensor)
__device__ inline void load_chunk(const float* T, const MPlan& P, int64_t o, int j0, int c,
                                  float v[8]) {
  const int t = threadIdx.x;
  const int row = t >> 2, kq = t & 3;
  const int64_t j = j0 + row;
  const int64_t k = (int64_t)c * KC + 8 * kq;
  if (o < 0 || j >= P.J || k >= P.K) {
#pragma unroll
    for (int e = 0; e < 8; ++e) v[e] = 0.f;
    return;
  }
  const float* src = T + (o * P.J + j) * P.K + k;
  if (k + 8 <= P.K) {
    const float4 a = *(const float4*)src;
    const float4 b = *(const float4*)(src + 4);
    v[0] = a.x; v[1] = a.y; v[2] = a.z; v[3] = a.w;
    v[4] = b.x; v[5] = b.y; v[6] = b.z; v[7] = b.w;
  } else {
#pragma unroll
    for (int e = 0; e < 8; ++e) v[e] = (k + e < P.K) ? src[e] : 0.f;
  }
}

// registers -> LDS buffer `buf` as hi/lo bf16, row-major and transposed
__device__ inline void store_chunk(const Lds& S, int buf, const float v[8]) {
  const int t = threadIdx.x;
  const int row = t >> 2, kq = t & 3;
  bf16x8 h, l;
#pragma unroll
  for (int e = 0; e < 8; ++e) {
    __bf16 a, b;
    split(v[e], a, b);
    h[e] = a;
    l[e] = b;
  }
  __bf16* rmh = S.rm + ((buf * 2 + 0) * TR + row) * RM_P + 8 * kq;
  __bf16* rml = S.rm + ((buf * 2 + 1) * TR + row) * RM_P + 8 * kq;
  *(bf16x8*)rmh = h;
  *(bf16x8*)rml = l;
#pragma unroll
  for (int e = 0; e < 8; ++e) {
    S.cm[((buf * 2 + 0) * KC + 8 * kq + e) * CM_P + row] = h[e];
    S.cm[((buf * 2 + 1) * KC + 8 * kq + e) * CM_P + row] = l[e];
  }
}

__device__ inline f32x4 mfma3(bf16x8 ah, bf16x8 al, bf16x8 bh, bf16x8 bl, f32x4 c) {
  c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah, bh, c, 0, 0, 0);
  c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah, bl, c, 0, 0, 0);
  c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(al, bh, c, 0, 0, 0);
  return c;
}

// decompose the outer row index o over modes 0..L-3 (last fastest)
__device__ inline void outer_idx(const Shape& s, int64_t o, int64_t* idx) {
  for (int m = s.L - 3; m >= 0; --m) {
    idx[m] = o % s.d[m];
    o /= s.d[m];
  }
}

template <int KCH>
__global__ __launch_bounds__(256, 1) void k_mttkrp_mfma(const float* __restrict__ T, MPlan P,
                                                        const double* __restrict__ F,
                                                        float* __restrict__ npart,
                                                        double* __restrict__ out, unsigned mask) {
  extern __shared__ __align__(16) unsigned char smem[];
  Lds S;
  S.rm = (__bf16*)smem;
  S.cm = S.rm + 2 * 2 * TR * RM_P;
  S.fj = (double*)(S.cm + 2 * 2 * KC * CM_P);
  S.pj = S.fj + TR * RC;

  const Shape& s = P.s;
  const int L = s.L, R = s.R;
  const int t = threadIdx.x;
  const int w = t >> 6, lane = t & 63, ln = lane & 15, rg = lane >> 4;
  const int jb = blockIdx.x % P.jblocks;
  const int64_t ob = (int64_t)(blockIdx.x / P.jblocks) * P.ochunk;
  const int64_t oe = ob + P.ochunk < P.NO ? ob + P.ochunk : P.NO;
  const int j0 = jb * TR;
  const int rl = 16 * w + ln;          // rank within the chunk owned by this lane's column
  const int r = P.rc0 + rl;            // global rank
  const bool rok = r < R;
  const int Lk = L - 1, Lj = L - 2;
  const double* Fk = F + s.off[Lk];

  // B fragments of GEMM 1: F_{L-1}[32c + 8rg + e, r]
  bf16x8 fh[KCH], fl[KCH];
#pragma unroll
  for (int c = 0; c < KCH; ++c) {
#pragma unroll
    for (int e = 0; e < 8; ++e) {
      const int64_t k = (int64_t)c * KC + 8 * rg + e;
      const float v = (rok && k < P.K) ? (float)Fk[(int64_t)r * P.K + k] : 0.f;
      __bf16 a, b;
      split(v, a, b);
      fh[c][e] = a;
      fl[c][e] = b;
    }
  }
  f32x4 nacc[KCH][2];
#pragma unroll
  for (int c = 0; c < KCH; ++c) {
    nacc[c][0] = f32x4{0.f, 0.f, 0.f, 0.f};
    nacc[c][1] = f32x4{0.f, 0.f, 0.f, 0.f};
  }
  // F_{L-2} rows of this block and the N_{L-2} partial
  for (int e = t; e < TR * RC; e += 256) {
    const int jj = e / RC, rr = e % RC;
    const int64_t j = j0 + jj;
    double v = 0.0;
    if (L >= 2 && j < P.J && P.rc0 + rr < R) v = F[s.off[Lj] + (int64_t)(P.rc0 + rr) * P.J + j];
    S.fj[e] = v;
    S.pj[e] = 0.0;
  }

  // Chunk c of a tile always travels through register slot c % PD.  A slot is refilled right
  // after it is consumed: with chunk c + PD of the same tile, or, after the last chunk of its
  // residue class, with chunk c % PD of the next tile.
  float pre[PD][8];
#pragma unroll
  for (int q = 0; q < PD; ++q)
    if (q < KCH) load_chunk(T, P, ob < oe ? ob : -1, j0, q, pre[q]);
  __syncthreads();

  for (int64_t o = ob; o < oe; ++o) {
    // product of the outer-mode factor rows for this lane's rank
    int64_t idx[TFX_MAXL];
    outer_idx(s, o, idx);
    double Fo = 1.0;
    for (int m = 0; m < Lj; ++m) Fo *= rok ? F[s.off[m] + (int64_t)r * s.d[m] + idx[m]] : 0.0;
    // B fragments of GEMM 2: W[32 sv + 8rg + e, rl]
    bf16x8 wh[2], wl[2];
#pragma unroll
    for (int sv = 0; sv < 2; ++sv)
#pragma unroll
      for (int e = 0; e < 8; ++e) {
        const int jj = 32 * sv + 8 * rg + e;
        const float v = (float)(Fo * S.fj[jj * RC + rl]);
        __bf16 a, b;
        split(v, a, b);
        wh[sv][e] = a;
        wl[sv][e] = b;
      }
    f32x4 yacc[4];
#pragma unroll
    for (int rb = 0; rb < 4; ++rb) yacc[rb] = f32x4{0.f, 0.f, 0.f, 0.f};

#pragma unroll
    for (int c = 0; c < KCH; ++c) {
      const int buf = c & 1;
      store_chunk(S, buf, pre[c % PD]);
      if (c + PD < KCH)
        load_chunk(T, P, o, j0, c + PD, pre[c % PD]);
      else
        load_chunk(T, P, o + 1 < oe ? o + 1 : -1, j0, c % PD, pre[c % PD]);
      // one barrier per chunk: buffer c & 1 is written before it and read after it; the other
      // buffer was last read before this barrier
      __syncthreads();
      const __bf16* rmh = S.rm + (buf * 2 + 0) * TR * RM_P;
      const __bf16* rml = S.rm + (buf * 2 + 1) * TR * RM_P;
      const __bf16* cmh = S.cm + (buf * 2 + 0) * KC * CM_P;
      const __bf16* cml = S.cm + (buf * 2 + 1) * KC * CM_P;
#pragma unroll
      for (int rb = 0; rb < 4; ++rb) {
        const bf16x8 ah = *(const bf16x8*)(rmh + (16 * rb + ln) * RM_P + 8 * rg);
        const bf16x8 al = *(const bf16x8*)(rml + (16 * rb + ln) * RM_P + 8 * rg);
        yacc[rb] = mfma3(ah, al, fh[c], fl[c], yacc[rb]);
      }
#pragma unroll
      for (int kb = 0; kb < 2; ++kb)
#pragma unroll
        for (int sv = 0; sv < 2; ++sv) {
          const bf16x8 ah = *(const bf16x8*)(cmh + (16 * kb + ln) * CM_P + 32 * sv + 8 * rg);
          const bf16x8 al = *(const bf16x8*)(cml + (16 * kb + ln) * CM_P + 32 * sv + 8 * rg);
          nacc[c][kb] = mfma3(ah, al, wh[sv], wl[sv], nacc[c][kb]);
        }
    }

    // leading modes from Y: lane holds Y[16 rb + 4 rg + e, rl]
    double part[TFX_MAXL];
    for (int m = 0; m < Lj; ++m) part[m] = 0.0;
#pragma unroll
    for (int rb = 0; rb < 4; ++rb)
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        const int jj = 16 * rb + 4 * rg + e;
        if (j0 + jj < P.J && rok) {
          const double y = (double)yacc[rb][e];
          S.pj[jj * RC + rl] += y * Fo;
          const double yf = y * S.fj[jj * RC + rl];
          for (int m = 0; m < Lj; ++m) part[m] += yf;
        }
      }
    for (int m = 0; m < Lj; ++m) {
      if (!((mask >> m) & 1u)) continue;
      // multiply by the product of the outer rows other than m
      double fo = 1.0;
      for (int q = 0; q < Lj; ++q)
        if (q != m) fo *= rok ? F[s.off[q] + (int64_t)r * s.d[q] + idx[q]] : 0.0;
      double v = part[m] * fo;
      v += __shfl_xor(v, 16, 64);
      v += __shfl_xor(v, 32, 64);
      if (rg == 0 && rok) atomicAdd(&out[s.off[m] + (int64_t)r * s.d[m] + idx[m]], v);
    }
  }
  __syncthreads();
  // flush N_{L-2}
  if ((mask >> Lj) & 1u) {
    for (int e = t; e < TR * RC; e += 256) {
      const int jj = e / RC, rr = e % RC;
      const int64_t j = j0 + jj;
      if (j < P.J && P.rc0 + rr < R && S.pj[e] != 0.0)
        atomicAdd(&out[s.off[Lj] + (int64_t)(P.rc0 + rr) * P.J + j], S.pj[e]);
    }
  }
  // N_{L-1} partial of this workgroup: lane holds N[32c + 16kb + 4rg + e, rl]
  float* dst = npart + (int64_t)blockIdx.x * P.K * RC;
#pragma unroll
  for (int c = 0; c < KCH; ++c)
#pragma unroll
    for (int kb = 0; kb < 2; ++kb)
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        const int64_t k = (int64_t)c * KC + 16 * kb + 4 * rg + e;
        if (k < P.K) dst[k * RC + rl] = nacc[c][kb][e];
      }
}

__global__ void k_reduce_npart(const float* __restrict__ npart, int nwg, int64_t K, int rc0, int R,
                               int64_t off, double* __restrict__ out) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= K * RC) return;
  const int64_t k = e / RC;
  const int rr = (int)(e % RC);
  if (rc0 + rr >= R) return;
  double acc = 0.0;
  for (int b = 0; b < nwg; ++b) acc += (double)npart[(int64_t)b * K * RC + e];
  out[off + (int64_t)(rc0 + rr) * K + k] = acc;
}

template <int KCH>
int launch_kch(tfx_ctx* ctx, const float* T, const MPlan& P, const double* F, float* npart,
               double* out, unsigned mask, hipStream_t st) {
  const size_t lds = sizeof(__bf16) * (2 * 2 * TR * RM_P + 2 * 2 * KC * CM_P) + sizeof(double) * 2 * TR * RC;
  TFX_HIP(ctx, hipFuncSetAttribute((const void*)k_mttkrp_mfma<KCH>,
                                   hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
  hipLaunchKernelGGL(k_mttkrp_mfma<KCH>, dim3(P.nwg), dim3(256), lds, st, T, P, F, npart, out, mask);
  TFX_CHECK_LAUNCH(ctx);
  return TFX_OK;
}

typedef int (*launch_fn)(tfx_ctx*, const float*, const MPlan&, const double*, float*, double*,
                         unsigned, hipStream_t);

const launch_fn kLaunch[17] = {nullptr,
                               launch_kch<1>,  launch_kch<2>,  launch_kch<3>,  launch_kch<4>,
                               launch_kch<5>,  launch_kch<6>,  launch_kch<7>,  launch_kch<8>,
                               launch_kch<9>,  launch_kch<10>, launch_kch<11>, launch_kch<12>,
                               launch_kch<13>, launch_kch<14>, launch_kch<15>, launch_kch<16>};

}  // namespace

int tfx_mttkrp_split16(tfx_ctx* ctx, const float* T, const Shape& s, const double* F,
                       unsigned mask, double* out, hipStream_t stream) {
  const int L = s.L;
  const int64_t K = s.d[L - 1];
  // shapes outside the tiled kernel's envelope use the FMA path
  if (L < 2 || K % 4 != 0 || K > 16 * KC) return tfx_mttkrp_native_fd(ctx, T, s, F, mask, out, stream);
  MPlan P;
  P.s = s;
  P.K = K;
  P.J = s.d[L - 2];
  P.NO = s.N / (K * P.J);
  P.jblocks = (int)((P.J + TR - 1) / TR);
  const int64_t want = (int64_t)ctx->cu_count;
  int64_t per = want / P.jblocks;
  if (per < 1) per = 1;
  P.ochunk = (P.NO + per - 1) / per;
  P.nwg = (int)(P.jblocks * ((P.NO + P.ochunk - 1) / P.ochunk));
  const int kch = (int)((K + KC - 1) / KC);
  void* wsp;
  int rc = tfx_workspace(ctx, sizeof(float) * (size_t)P.nwg * K * RC + 256, stream, &wsp);
  if (rc) return rc;
  float* npart = (float*)wsp;
  for (int m = 0; m < L - 1; ++m)
    if ((mask >> m) & 1u)
      TFX_HIP(ctx, hipMemsetAsync(out + s.off[m], 0, sizeof(double) * s.d[m] * s.R, stream));
  for (int rc0 = 0; rc0 < s.R; rc0 += RC) {
    P.rc0 = rc0;
    rc = kLaunch[kch](ctx, T, P, F, npart, out, mask, stream);
    if (rc) return rc;
    if ((mask >> (L - 1)) & 1u) {
      const int64_t tot = K * RC;
      hipLaunchKernelGGL(k_reduce_npart, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, stream,
                         npart, P.nwg, K, rc0, s.R, s.off[L - 1], out);
      TFX_CHECK_LAUNCH(ctx);
    }
  }
  return TFX_OK;
}
