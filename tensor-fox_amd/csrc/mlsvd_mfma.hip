// MFMA forms of the MLSVD's two tensor passes (the reference computes both with FP64 BLAS in
// Compression.py:48-237 and MultilinearAlgebra.py:72-105):
//
//   mode Gram      G = X_(mode) X_(mode)^T               (I x I, FP64 result)
//   mode product   Y[p, s, q] = sum_a U[a, s] X[p, a, q]  (the TTM, FP64 result) and
//                  sig2[s] = sum_{p,q} Y[p, s, q]^2        (squared singular values of the step)
//
// The tensor is viewed around `mode` as (P, I, Q): element (p, a, q) at (p I + a) Q + q; the
// unfolding has rows a and columns n = p Q + q (K = P Q of them).  Nothing is unfolded in memory.
//
// Gram: one workgroup per (128 x 128 block pair with tb >= ta, slice of the K columns).  The rows
// of both blocks are staged per 32 columns through LDS (global loads for the next stage in flight
// during the MFMAs of the current one); each of the 4 waves owns a 64 x 64 sub-block.  FP32 data
// runs bf16x3 (hi*hi + hi*lo + lo*hi, v_mfma_f32_16x16x32_bf16, FP32 accumulation per slice),
// FP64 data runs v_mfma_f64_16x16x4f64.  Slices write partial blocks; a second kernel sums them
// in a fixed order and mirrors the triangle, so the result is deterministic.
//
// TTM: one wavefront per 64 consecutive columns n and up to 64 rows s of Y, looping over a in
// groups of 16.  Each lane owns 4 columns (n0 + 4 c + t, t = 0..3), so that for Q % 4 == 0 one
// 16-byte load brings the 4 tiles' values of one a.  FP32 data runs v_mfma_f32_16x16x4f32 with U
// rounded to FP32 (exact products, FP32 accumulation), FP64 data v_mfma_f64_16x16x4f64.  The
// wave also emits its columns' contribution to sig2, summed in a fixed order by a second kernel.
#include "common.h"

namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef double f64x4 __attribute__((ext_vector_type(4)));

struct PIQ {
  int64_t P, I, Q;
};

PIQ piq_of(const Shape& s, int mode) {
  PIQ v{1, s.d[mode], 1};
  for (int m = 0; m < mode; ++m) v.P *= s.d[m];
  for (int m = mode + 1; m < s.L; ++m) v.Q *= s.d[m];
  return v;
}

constexpr int NXCD = 8;

__device__ inline void split(float x, __bf16& h, __bf16& l) {
  h = (__bf16)x;
  l = (__bf16)(x - (float)h);
}

// ---------------------------------------------------------------------------------------------
// Mode Gram
// ---------------------------------------------------------------------------------------------

constexpr int GB = 128;    // rows of a block
constexpr int GKS = 32;    // columns per stage
constexpr int GNT = 256;   // 4 waves, 64 x 64 sub-block each
constexpr int BP = 40;     // bf16 LDS row pitch: 80 B rows keep 16-byte fragment reads conflict-free
constexpr int DP = 33;     // FP64 LDS row pitch (odd): 16 rows of one 8-byte read hit distinct banks
constexpr int GMAX_SPW = 256;   // stages per slice at most (FP32 partials over <= 8192 columns)

// Column layouts of the staging loads
constexpr int LAY_VEC = 0;    // Q % 8 == 0: 8 consecutive columns are contiguous (vector loads)
constexpr int LAY_ACONT = 1;  // Q == 1: rows contiguous, lanes walk a
constexpr int LAY_GEN = 2;    // anything else: scalar loads with per-element wrap of q

struct GPlan {
  int64_t I, Q, K;
  int64_t nstages;
  int nb, npairs, ksplit, spw, lay;
  int nwg, nlaunch;
};

template <typename S>
struct GramT;
template <>
struct GramT<float> {
  typedef f32x4 Acc;
  typedef float Part;
  static constexpr size_t lds = sizeof(__bf16) * 2 * 2 * GB * BP;   // [blk][hi/lo][row][BP]
};
template <>
struct GramT<double> {
  typedef f64x4 Acc;
  typedef double Part;
  static constexpr size_t lds = sizeof(double) * 2 * GB * DP;       // [blk][row][DP]
};

// Column tracker of one thread: the 8 columns n..n+7 it stages, with q = n % Q and base = the
// address of element (a = 0, n).
struct ColTrack {
  int64_t n, q, base;
};

__device__ inline ColTrack col_init(const GPlan& G, int64_t n) {
  ColTrack c;
  c.n = n;
  const int64_t p = n / G.Q;
  c.q = n - p * G.Q;
  c.base = p * G.I * G.Q + c.q;
  return c;
}

template <int LAY>
__device__ inline void col_advance(const GPlan& G, ColTrack& c) {
  c.n += GKS;
  if constexpr (LAY == LAY_ACONT) {
    c.base += (int64_t)GKS * G.I;
    return;
  }
  c.q += GKS;
  c.base += GKS;
  while (c.q >= G.Q) {
    c.q -= G.Q;
    c.base += (G.I - 1) * G.Q;
  }
}

// The 8 values (a, n..n+7) of the unfolding.  Branch-free: an element outside the tensor loads
// X[0] instead, and its bit in the returned mask is clear (the value is zeroed where it is used,
// so the loads stay in flight until then).
template <typename S, int LAY>
__device__ inline unsigned load8(const S* __restrict__ X, const GPlan& G, const ColTrack& c,
                                 int64_t a, S v[8]) {
  const bool aok = a < G.I;
  if constexpr (LAY == LAY_VEC) {
    const bool ok = aok && c.n < G.K;   // K % 8 == 0 here: the 8 columns are all in range
    const S* src = X + (ok ? c.base + a * G.Q : 0);
    if constexpr (sizeof(S) == 4) {
      const float4 x0 = *(const float4*)src, x1 = *(const float4*)(src + 4);
      v[0] = x0.x; v[1] = x0.y; v[2] = x0.z; v[3] = x0.w;
      v[4] = x1.x; v[5] = x1.y; v[6] = x1.z; v[7] = x1.w;
    } else {
#pragma unroll
      for (int e = 0; e < 8; e += 2) {
        const double2 x = *(const double2*)(src + e);
        v[e] = x.x;
        v[e + 1] = x.y;
      }
    }
    return ok ? 0xffu : 0u;
  } else if constexpr (LAY == LAY_ACONT) {
    unsigned m = 0;
#pragma unroll
    for (int e = 0; e < 8; ++e) {
      const bool ok = aok && c.n + e < G.K;
      v[e] = X[ok ? c.base + (int64_t)e * G.I + a : 0];
      m |= (unsigned)ok << e;
    }
    return m;
  } else {
    unsigned m = 0;
    int64_t qq = c.q, o = c.base + a * G.Q;
#pragma unroll
    for (int e = 0; e < 8; ++e) {
      const bool ok = aok && c.n + e < G.K;
      v[e] = X[ok ? o : 0];
      m |= (unsigned)ok << e;
      ++qq;
      ++o;
      if (qq == G.Q) {
        qq = 0;
        o += (G.I - 1) * G.Q;
      }
    }
    return m;
  }
}

__device__ inline void decode_pair(int pair, int nb, int& ta, int& tb) {
  ta = 0;
  while (pair >= nb - ta) {
    pair -= nb - ta;
    ++ta;
  }
  tb = ta + pair;
}

__device__ inline f32x4 mfma_bf16x3(bf16x8 ah, bf16x8 al, bf16x8 bh, bf16x8 bl, f32x4 c) {
  c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah, bh, c, 0, 0, 0);
  c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah, bl, c, 0, 0, 0);
  c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(al, bh, c, 0, 0, 0);
  return c;
}

// Staging items of thread t: 4 (block, row, column group) triples.  LAY_ACONT puts consecutive
// threads on consecutive rows (the contiguous direction); the others on consecutive columns.
template <int LAY>
__device__ inline void item_of(int t, int i, int& blk, int& row, int& cg, int& ng) {
  blk = i >> 1;
  const int u = i & 1;
  if constexpr (LAY == LAY_ACONT) {
    row = t & (GB - 1);
    ng = (t >> 7) + 2 * u;   // column groups of 8: 0..3
    cg = u;                  // two trackers (the two column groups)
  } else {
    ng = t & 3;
    row = (t >> 2) + 64 * u;
    cg = 0;
  }
}

template <typename S, int LAY>
__global__ __launch_bounds__(GNT) void k_gram_mfma(const S* __restrict__ X, GPlan G,
                                                   typename GramT<S>::Part* __restrict__ part) {
  typedef typename GramT<S>::Acc Acc;
  extern __shared__ __align__(16) unsigned char smem[];
  // consecutive slots (the block pairs of one column slice) run on one XCD and share its L2
  const int per_xcd = G.nlaunch / NXCD;
  const int g = (int)(blockIdx.x % NXCD) * per_xcd + (int)(blockIdx.x / NXCD);
  if (g >= G.nwg) return;   // uniform, before any barrier
  const int pair = g % G.npairs, ks = g / G.npairs;
  int ta, tb;
  decode_pair(pair, G.nb, ta, tb);
  const bool diag = ta == tb;
  const int64_t s0 = (int64_t)ks * G.spw;
  const int64_t s1 = s0 + G.spw < G.nstages ? s0 + G.spw : G.nstages;
  const int t = threadIdx.x, w = t >> 6, lane = t & 63, l16 = lane & 15, lg = lane >> 4;
  const int sr = w >> 1, sc = w & 1;
  // the lower-left sub-block of a diagonal pair mirrors the upper-right one: not computed
  const bool active = !(diag && sr == 1 && sc == 0) && ((int64_t)ta * GB + 64 * sr < G.I) &&
                      ((int64_t)tb * GB + 64 * sc < G.I);
  const int nitems = diag ? 2 : 4;

  ColTrack ct[2];
  int irow[4], ing[4], icg[4], iblk[4];
#pragma unroll
  for (int i = 0; i < 4; ++i) item_of<LAY>(t, i, iblk[i], irow[i], icg[i], ing[i]);
  ct[0] = col_init(G, s0 * GKS + 8 * ing[0]);
  ct[1] = col_init(G, s0 * GKS + 8 * ing[1]);
  int64_t arow[4];
#pragma unroll
  for (int i = 0; i < 4; ++i) arow[i] = (int64_t)(iblk[i] ? tb : ta) * GB + irow[i];

  S v[4][8];
  unsigned vm[4] = {0u, 0u, 0u, 0u};
  auto load_all = [&]() {
#pragma unroll
    for (int i = 0; i < 4; ++i)
      if (i < nitems) vm[i] = load8<S, LAY>(X, G, ct[icg[i]], arow[i], v[i]);
  };
  load_all();

  Acc acc[4][4];
#pragma unroll
  for (int r = 0; r < 4; ++r)
#pragma unroll
    for (int c = 0; c < 4; ++c)
#pragma unroll
      for (int j = 0; j < 4; ++j) acc[r][c][j] = 0;

  for (int64_t st = s0; st < s1; ++st) {
    __syncthreads();   // previous stage's fragment reads are done
    if constexpr (sizeof(S) == 4) {
      __bf16* L = (__bf16*)smem;
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        if (i < nitems) {
          bf16x8 h, l;
#pragma unroll
          for (int e = 0; e < 8; ++e) {
            __bf16 a, b;
            split(((vm[i] >> e) & 1u) ? v[i][e] : 0.f, a, b);
            h[e] = a;
            l[e] = b;
          }
          __bf16* rowp = L + ((size_t)(iblk[i] * 2) * GB + irow[i]) * BP + 8 * ing[i];
          *(bf16x8*)rowp = h;
          *(bf16x8*)(rowp + GB * BP) = l;
        }
      }
    } else {
      double* L = (double*)smem;
#pragma unroll
      for (int i = 0; i < 4; ++i)
        if (i < nitems) {
          double* rowp = L + ((size_t)iblk[i] * GB + irow[i]) * DP + 8 * ing[i];
#pragma unroll
          for (int e = 0; e < 8; ++e) rowp[e] = ((vm[i] >> e) & 1u) ? v[i][e] : 0.0;
        }
    }
    __syncthreads();
    if (st + 1 < s1) {   // next stage's loads overlap this stage's MFMAs
      col_advance<LAY>(G, ct[0]);
      if (LAY == LAY_ACONT) col_advance<LAY>(G, ct[1]);
      load_all();
    }
    if (active) {
      if constexpr (sizeof(S) == 4) {
        const __bf16* L = (const __bf16*)smem;
        const __bf16* HA = L;
        const __bf16* HB = diag ? L : L + 2 * GB * BP;
        bf16x8 ah[4], al[4], bh[4], bl[4];
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const __bf16* p = HA + (64 * sr + 16 * r + l16) * BP + 8 * lg;
          ah[r] = *(const bf16x8*)p;
          al[r] = *(const bf16x8*)(p + GB * BP);
        }
#pragma unroll
        for (int c = 0; c < 4; ++c) {
          const __bf16* p = HB + (64 * sc + 16 * c + l16) * BP + 8 * lg;
          bh[c] = *(const bf16x8*)p;
          bl[c] = *(const bf16x8*)(p + GB * BP);
        }
#pragma unroll
        for (int r = 0; r < 4; ++r)
#pragma unroll
          for (int c = 0; c < 4; ++c) acc[r][c] = mfma_bf16x3(ah[r], al[r], bh[c], bl[c], acc[r][c]);
      } else {
        const double* L = (const double*)smem;
        const double* DA = L;
        const double* DB = diag ? L : L + GB * DP;
#pragma unroll
        for (int m = 0; m < GKS / 4; ++m) {
          double a[4], b[4];
#pragma unroll
          for (int r = 0; r < 4; ++r) a[r] = DA[(64 * sr + 16 * r + l16) * DP + 4 * m + lg];
#pragma unroll
          for (int c = 0; c < 4; ++c) b[c] = DB[(64 * sc + 16 * c + l16) * DP + 4 * m + lg];
#pragma unroll
          for (int r = 0; r < 4; ++r)
#pragma unroll
            for (int c = 0; c < 4; ++c)
              acc[r][c] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[r], b[c], acc[r][c], 0, 0, 0);
        }
      }
    }
  }
  if (!active) return;
  typename GramT<S>::Part* P = part + (size_t)(pair * G.ksplit + ks) * GB * GB;
#pragma unroll
  for (int r = 0; r < 4; ++r)
#pragma unroll
    for (int c = 0; c < 4; ++c)
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        // f32 16x16 result: register j of lane l is row 4 (l / 16) + j; f64: row 4 j + l / 16
        const int rr = sizeof(S) == 4 ? 4 * lg + j : 4 * j + lg;
        P[(size_t)(64 * sr + 16 * r + rr) * GB + 64 * sc + 16 * c + l16] = acc[r][c][j];
      }
}

// G[i][j] = G[j][i] = sum over slices (fixed order) of the pair's partial entry, i <= j.
template <typename PT>
__global__ __launch_bounds__(256) void k_gram_reduce(const PT* __restrict__ part, GPlan G,
                                                     double* __restrict__ out) {
  const int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (e >= G.I * G.I) return;
  const int64_t i = e / G.I, j = e - i * G.I;
  if (i > j) return;
  const int ta = (int)(i / GB), tb = (int)(j / GB);
  const int pair = ta * G.nb - ta * (ta - 1) / 2 + (tb - ta);
  const int64_t r = i - (int64_t)ta * GB, c = j - (int64_t)tb * GB;
  const PT* p = part + (size_t)pair * G.ksplit * GB * GB + r * GB + c;
  double acc = 0.0;
  for (int ks = 0; ks < G.ksplit; ++ks) acc += (double)p[(size_t)ks * GB * GB];
  out[i * G.I + j] = acc;
  out[j * G.I + i] = acc;
}

template <typename S>
int gram_mfma(tfx_ctx* ctx, const S* X, const Shape& s, int mode, double* out, hipStream_t st) {
  const PIQ v = piq_of(s, mode);
  GPlan G{};
  G.I = v.I;
  G.Q = v.Q;
  G.K = v.P * v.Q;
  if (G.I > 8192) return tfx_set_error(ctx, TFX_E_UNSUPPORTED, "MFMA Gram limited to I <= 8192");
  G.nstages = (G.K + GKS - 1) / GKS;
  G.nb = (int)((G.I + GB - 1) / GB);
  G.npairs = G.nb * (G.nb + 1) / 2;
  const bool al = ((uintptr_t)X % 16) == 0;
  G.lay = (G.Q % 8 == 0 && al) ? LAY_VEC : (G.Q == 1 ? LAY_ACONT : LAY_GEN);
  // about two workgroups per CU, at least 4 and at most GMAX_SPW stages per slice
  const int64_t want = std::max<int64_t>(1, 2 * (int64_t)ctx->cu_count / G.npairs);
  int64_t spw = (G.nstages + want - 1) / want;
  spw = std::max<int64_t>(spw, 4);
  spw = std::min<int64_t>(spw, GMAX_SPW);
  const int64_t ksplit = (G.nstages + spw - 1) / spw;
  if ((int64_t)G.npairs * ksplit > 0x7fffffff / NXCD)
    return tfx_set_error(ctx, TFX_E_UNSUPPORTED, "MFMA Gram: too many slices");
  G.spw = (int)spw;
  G.ksplit = (int)ksplit;
  G.nwg = G.npairs * G.ksplit;
  G.nlaunch = ((G.nwg + NXCD - 1) / NXCD) * NXCD;
  typedef typename GramT<S>::Part PT;
  const size_t pbytes = sizeof(PT) * (size_t)G.nwg * GB * GB;
  void* wsp;
  int rc = tfx_workspace(ctx, pbytes, st, &wsp);
  if (rc) return rc;
  PT* part = (PT*)wsp;
  const size_t lds = GramT<S>::lds;
  const void* kern = G.lay == LAY_VEC     ? (const void*)k_gram_mfma<S, LAY_VEC>
                     : G.lay == LAY_ACONT ? (const void*)k_gram_mfma<S, LAY_ACONT>
                                          : (const void*)k_gram_mfma<S, LAY_GEN>;
  TFX_HIP(ctx, hipFuncSetAttribute(kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
  const dim3 grid((unsigned)G.nlaunch), block(GNT);
  if (G.lay == LAY_VEC)
    hipLaunchKernelGGL((k_gram_mfma<S, LAY_VEC>), grid, block, lds, st, X, G, part);
  else if (G.lay == LAY_ACONT)
    hipLaunchKernelGGL((k_gram_mfma<S, LAY_ACONT>), grid, block, lds, st, X, G, part);
  else
    hipLaunchKernelGGL((k_gram_mfma<S, LAY_GEN>), grid, block, lds, st, X, G, part);
  TFX_CHECK_LAUNCH(ctx);
  const int64_t tot = G.I * G.I;
  hipLaunchKernelGGL(k_gram_reduce<PT>, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, st, part,
                     G, out);
  TFX_CHECK_LAUNCH(ctx);
  return TFX_OK;
}

// ---------------------------------------------------------------------------------------------
// Mode product (TTM) with the squared row norms of the result
// ---------------------------------------------------------------------------------------------

constexpr int TNT = 256;   // 4 waves, one 64-column chunk each
constexpr int TLAY_QVEC = 0;    // Q % 4 == 0: the lane's 4 columns are contiguous
constexpr int TLAY_ACONT = 1;   // Q == 1, I % 4 == 0: 4 consecutive a of a column are contiguous
constexpr int TLAY_GEN = 2;

struct TPlan {
  int64_t I, Q, Nc;   // Nc = P Q columns
  int k, ngrp, nchunk, lay;
  int nsb, sbw;       // blocks of sbw = 16 x (s tiles per wave) rows of Y
};

// U fragments in MFMA operand order: UF[((sb 4 + st) ngrp + grp) 4 + m][lane] = U[s][a] with
// s = sbw sb + 16 st + lane % 16 and a = 16 grp + 4 (lane / 16) + m (zero outside the block,
// the rank or the mode).
template <typename S>
__global__ __launch_bounds__(256) void k_ttm_prep(const double* __restrict__ U, TPlan T,
                                                  S* __restrict__ UF) {
  const int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x;
  const int64_t tot = (int64_t)T.nsb * 4 * T.ngrp * 4 * 64;
  if (e >= tot) return;
  const int lane = (int)(e & 63), m = (int)((e >> 6) & 3);
  int64_t rest = e >> 8;
  const int grp = (int)(rest % T.ngrp);
  rest /= T.ngrp;
  const int stile = (int)(rest & 3), sb = (int)(rest >> 2);
  const int64_t s = (int64_t)T.sbw * sb + 16 * stile + (lane & 15);
  const int64_t a = 16 * (int64_t)grp + 4 * (lane >> 4) + m;
  UF[e] = (16 * stile < T.sbw && s < T.k && a < T.I) ? (S)U[s * T.I + a] : S(0);
}

template <typename S>
struct TtmT;
template <>
struct TtmT<float> {
  typedef f32x4 Acc;
  static __device__ inline Acc mfma(float a, float b, Acc c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
  }
  static __device__ inline int row(int lg, int j) { return 4 * lg + j; }
};
template <>
struct TtmT<double> {
  typedef f64x4 Acc;
  static __device__ inline Acc mfma(double a, double b, Acc c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
  }
  static __device__ inline int row(int lg, int j) { return 4 * j + lg; }
};

template <typename S>
__device__ inline void load4(const S* p, S o[4]) {
  if constexpr (sizeof(S) == 4) {
    const float4 x = *(const float4*)p;
    o[0] = x.x; o[1] = x.y; o[2] = x.z; o[3] = x.w;
  } else {
    const double2 x0 = *(const double2*)p, x1 = *(const double2*)(p + 2);
    o[0] = x0.x; o[1] = x0.y; o[2] = x1.x; o[3] = x1.y;
  }
}

// The lane's values of one group of 16 a (K-step m: a = a0 + m) for its 4 columns, [t][m], and
// the mask of the ones inside the tensor (bit 4 t + m); outside ones load X[0].
template <typename S, int LAY>
__device__ inline unsigned load_group(const S* __restrict__ X, const TPlan& T, const int64_t xb[4],
                                      const bool cv[4], int64_t a0, S v[4][4]) {
  unsigned msk = 0;
  if constexpr (LAY == TLAY_QVEC) {   // the 4 columns are contiguous: one vector per a
#pragma unroll
    for (int m = 0; m < 4; ++m) {
      const bool ok = cv[0] && a0 + m < T.I;
      S o[4];
      load4<S>(X + (ok ? xb[0] + (a0 + m) * T.Q : 0), o);
#pragma unroll
      for (int t = 0; t < 4; ++t) v[t][m] = o[t];
      msk |= ok ? (0x1111u << m) : 0u;
    }
  } else if constexpr (LAY == TLAY_ACONT) {   // I % 4 == 0: a0 < I means all 4 a are inside
#pragma unroll
    for (int t = 0; t < 4; ++t) {
      const bool ok = cv[t] && a0 < T.I;
      load4<S>(X + (ok ? xb[t] + a0 : 0), v[t]);
      msk |= ok ? (0xfu << (4 * t)) : 0u;
    }
  } else {
#pragma unroll
    for (int t = 0; t < 4; ++t)
#pragma unroll
      for (int m = 0; m < 4; ++m) {
        const bool ok = cv[t] && a0 + m < T.I;
        v[t][m] = X[ok ? xb[t] + (a0 + m) * T.Q : 0];
        msk |= (unsigned)ok << (4 * t + m);
      }
  }
  return msk;
}

template <typename S, int ST, int LAY>
__global__ __launch_bounds__(TNT) void k_ttm_mfma(const S* __restrict__ X, TPlan T,
                                                  const S* __restrict__ UF, double* __restrict__ Y,
                                                  double* __restrict__ part) {
  typedef typename TtmT<S>::Acc Acc;
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63, cc = lane & 15, lg = lane >> 4;
  const int chunk = blockIdx.x * 4 + w, sb = blockIdx.y;
  if (chunk >= T.nchunk) return;   // no barriers in this kernel
  const int64_t n0 = (int64_t)chunk * 64;
  int64_t xb[4], yb[4];
  bool cv[4];
#pragma unroll
  for (int t = 0; t < 4; ++t) {
    const int64_t n = n0 + 4 * cc + t;
    cv[t] = n < T.Nc;
    const int64_t p = n / T.Q, q = n - p * T.Q;
    xb[t] = p * T.I * T.Q + q;              // element (a = 0) of column n
    yb[t] = (p * T.k) * T.Q + q;            // element (s = 0) of column n in Y
  }
  Acc acc[ST][4];
#pragma unroll
  for (int i = 0; i < ST; ++i)
#pragma unroll
    for (int t = 0; t < 4; ++t)
#pragma unroll
      for (int j = 0; j < 4; ++j) acc[i][t][j] = 0;

  const S* uf = UF + ((size_t)(sb * 4) * T.ngrp * 4) * 64 + lane;
  auto load_u = [&](int grp, S u[ST][4]) {
#pragma unroll
    for (int i = 0; i < ST; ++i)
#pragma unroll
      for (int m = 0; m < 4; ++m) u[i][m] = uf[(((size_t)i * T.ngrp + grp) * 4 + m) * 64];
  };
  // One group at a time: loads (branch-free) then MFMAs; the other waves on the SIMD (3 at 4
  // tiles of FP32) cover the load latency.
  for (int grp = 0; grp < T.ngrp; ++grp) {
    S v[4][4];
    S u[ST][4];
    const unsigned msk = load_group<S, LAY>(X, T, xb, cv, 16 * (int64_t)grp + 4 * lg, v);
    load_u(grp, u);
#pragma unroll
    for (int m = 0; m < 4; ++m)
#pragma unroll
      for (int t = 0; t < 4; ++t) {
        const S b = ((msk >> (4 * t + m)) & 1u) ? v[t][m] : S(0);
#pragma unroll
        for (int i = 0; i < ST; ++i) acc[i][t] = TtmT<S>::mfma(u[i][m], b, acc[i][t]);
      }
  }

  // Y and the columns' share of sig2
  double* pp = part ? part + ((size_t)sb * T.nchunk + chunk) * 64 : nullptr;
#pragma unroll
  for (int i = 0; i < ST; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int sl = 16 * i + TtmT<S>::row(lg, j);
      const int64_t s = (int64_t)T.sbw * sb + sl;
      double sq = 0.0;
#pragma unroll
      for (int t = 0; t < 4; ++t) {
        const double y = (double)acc[i][t][j];
        sq += y * y;
      }
      if (s < T.k) {
        if constexpr (LAY == TLAY_QVEC) {
          if (cv[0]) {
            double* dst = Y + yb[0] + s * T.Q;
            *(double2*)dst = make_double2((double)acc[i][0][j], (double)acc[i][1][j]);
            *(double2*)(dst + 2) = make_double2((double)acc[i][2][j], (double)acc[i][3][j]);
          }
        } else {
#pragma unroll
          for (int t = 0; t < 4; ++t)
            if (cv[t]) Y[yb[t] + s * T.Q] = (double)acc[i][t][j];
        }
      }
      if (pp) {
        // sum over the 16 lanes of this row group (fixed butterfly order)
        sq += __shfl_xor(sq, 8, 64);
        sq += __shfl_xor(sq, 4, 64);
        sq += __shfl_xor(sq, 2, 64);
        sq += __shfl_xor(sq, 1, 64);
        if (cc == 0) pp[sl] = sq;
      }
    }
}

// sig2[s] = sum over chunks (fixed order) of the chunks' shares; one workgroup per s
__global__ __launch_bounds__(256) void k_ttm_sig2(const double* __restrict__ part, TPlan T,
                                                  int st_tiles, double* __restrict__ sig2) {
  __shared__ double red[4];
  const int s = blockIdx.x;
  const int sb = s / T.sbw, sl = s % T.sbw;
  double acc = 0.0;
  if (sl < 16 * st_tiles)
    for (int c = threadIdx.x; c < T.nchunk; c += 256) acc += part[((size_t)sb * T.nchunk + c) * 64 + sl];
  acc = block_sum(acc, red);
  if (threadIdx.x == 0) sig2[s] = acc;
}

template <typename S, int ST>
void launch_ttm(const S* X, const TPlan& T, const S* UF, double* Y, double* part, hipStream_t st) {
  const dim3 grid((unsigned)((T.nchunk + 3) / 4), (unsigned)T.nsb);
  if (T.lay == TLAY_QVEC) {
    hipLaunchKernelGGL((k_ttm_mfma<S, ST, TLAY_QVEC>), grid, dim3(TNT), 0, st, X, T, UF, Y, part);
  } else if (T.lay == TLAY_ACONT) {
    hipLaunchKernelGGL((k_ttm_mfma<S, ST, TLAY_ACONT>), grid, dim3(TNT), 0, st, X, T, UF, Y, part);
  } else {
    hipLaunchKernelGGL((k_ttm_mfma<S, ST, TLAY_GEN>), grid, dim3(TNT), 0, st, X, T, UF, Y, part);
  }
}

template <typename S>
int ttm_mfma(tfx_ctx* ctx, const S* X, const Shape& s, int mode, const double* U, int k, double* Y,
             double* sig2, hipStream_t st) {
  const PIQ v = piq_of(s, mode);
  TPlan T{};
  T.I = v.I;
  T.Q = v.Q;
  T.Nc = v.P * v.Q;
  T.k = k;
  T.ngrp = (int)((v.I + 15) / 16);
  const int64_t nchunk = (T.Nc + 63) / 64;
  if (nchunk > 0x7ffffff0 / 4 * 4 || v.I > (1 << 24))
    return tfx_set_error(ctx, TFX_E_UNSUPPORTED, "MFMA TTM: tensor too large");
  T.nchunk = (int)nchunk;
  const bool xal = ((uintptr_t)X % 16) == 0, yal = ((uintptr_t)Y % 16) == 0;
  if (v.Q % 4 == 0 && xal && yal)
    T.lay = TLAY_QVEC;
  else if (v.Q == 1 && v.I % 4 == 0 && xal)
    T.lay = TLAY_ACONT;
  else
    T.lay = TLAY_GEN;
  const int st_max = 4;   // s tiles per wave
  const int st_tiles = k > 16 * st_max ? st_max : (k + 15) / 16;
  T.sbw = 16 * st_tiles;
  T.nsb = (k + T.sbw - 1) / T.sbw;
  const size_t nuf = (size_t)T.nsb * 4 * T.ngrp * 4 * 64;
  const size_t uf_bytes = align256(sizeof(S) * nuf);
  const size_t part_bytes = sig2 ? sizeof(double) * (size_t)T.nsb * T.nchunk * 64 : 0;
  void* wsp;
  int rc = tfx_workspace(ctx, uf_bytes + part_bytes, st, &wsp);
  if (rc) return rc;
  S* UF = (S*)wsp;
  double* part = sig2 ? (double*)((char*)wsp + uf_bytes) : nullptr;
  hipLaunchKernelGGL(k_ttm_prep<S>, dim3((unsigned)((nuf + 255) / 256)), dim3(256), 0, st, U, T, UF);
  TFX_CHECK_LAUNCH(ctx);
  switch (st_tiles) {
    case 1: launch_ttm<S, 1>(X, T, UF, Y, part, st); break;
    case 2: launch_ttm<S, 2>(X, T, UF, Y, part, st); break;
    case 3: launch_ttm<S, 3>(X, T, UF, Y, part, st); break;
    default: launch_ttm<S, 4>(X, T, UF, Y, part, st); break;
  }
  TFX_CHECK_LAUNCH(ctx);
  if (sig2) {
    hipLaunchKernelGGL(k_ttm_sig2, dim3((unsigned)k), dim3(256), 0, st, part, T, st_tiles, sig2);
    TFX_CHECK_LAUNCH(ctx);
  }
  return TFX_OK;
}

}  // namespace

// FP32 data: bf16x3 MFMA Gram, FP32-MFMA mode product (entry points of the split precision).
int tfx_gram_split16(tfx_ctx* ctx, const float* X, const Shape& s, int mode, double* G,
                     hipStream_t st) {
  return gram_mfma<float>(ctx, X, s, mode, G, st);
}

int tfx_ttm_split16(tfx_ctx* ctx, const float* X, const Shape& s, int mode, const double* U, int k,
                    double* Y, double* sig2, hipStream_t st) {
  return ttm_mfma<float>(ctx, X, s, mode, U, k, Y, sig2, st);
}

// FP64 data: FP64 MFMA for both.
int tfx_gram_f64mfma(tfx_ctx* ctx, const double* X, const Shape& s, int mode, double* G,
                     hipStream_t st) {
  return gram_mfma<double>(ctx, X, s, mode, G, st);
}

int tfx_ttm_f64mfma(tfx_ctx* ctx, const double* X, const Shape& s, int mode, const double* U,
                    int k, double* Y, double* sig2, hipStream_t st) {
  return ttm_mfma<double>(ctx, X, s, mode, U, k, Y, sig2, st);
}
