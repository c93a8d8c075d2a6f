// MTTKRP of every mode in one pass over an FP32 tensor on the bf16 matrix cores (gfx950
// v_mfma_f32_16x16x32_bf16), with FP64 factors and outputs.
//
// Split precision: every FP32 operand x is written as hi + lo with hi = bf16(x) and
// lo = bf16(x - hi) (|x - hi - lo| <= 2^-18 |x|); a product is formed as
// hi*hi + hi*lo + lo*hi with FP32 accumulation (the dropped lo*lo term is ~2^-16 relative).
//
// Notation: T is viewed as T[o, j, k] with k = i_{L-1} (contiguous), j = i_{L-2} and o the
// combined index of modes 0..L-3 ("outer", o = 0 when L = 2).  Fo[o, r] = prod_{m<L-2} F_m[i_m, r].
//   N_{L-2}[j, r] = sum_o Fo[o, r] Y_o[j, r],   Y_o = T[o] F_{L-1}       (GEMM 1)
//   N_{L-1}[k, r] = sum_o Fo[o, r] Z_o[k, r],   Z_o = T[o]^T F_{L-2}     (GEMM 2)
//   P[o, r]       = sum_j F_{L-2}[j, r] Y_o[j, r]
//   N_m[i, r]     = sum_{o: i_m(o) = i} P[o, r] prod_{q<L-2, q!=m} F_q[i_q(o), r]   (m < L-2)
//
// Work decomposition: a wavefront owns a block of 32 j x 32 k and walks a range of o; there is
// no workgroup barrier in the main loop.  Per o the wave
//   - loads its 32 x 32 FP32 block straight from HBM into registers in the GEMM 1 A-operand
//     layout (lane row m = lane % 16 of row block rb, eight consecutive k: two 16-byte raw buffer
//     loads per row block) and splits it into hi / lo bf16;
//   - runs GEMM 1 on those registers (2 row blocks x 4 rank blocks x 3 products);
//   - stores the hi / lo block to its own 4 KB of LDS and reads it back transposed with the
//     gfx950 transposing read ds_read_b64_tr_b16 (LDS operations of one wave complete in order,
//     so no barrier is involved), then runs GEMM 2 (2 k blocks x 4 rank blocks x 3 products);
//   - scales both results by Fo[o, :] into its two block accumulators and reduces P over its 32
//     j in registers (lane sums, then three cross-row swaps).
// The next o's loads are issued right after the split, so they are in flight during the MFMAs.
// The rows of GEMM 1's A operand are permuted (lane row m of row block rb is j = 8 (m / 4) +
// 4 rb + m % 4) so that the rows a lane holds in GEMM 1's result are the eight j of its GEMM 2 B
// fragment: the FP32 factor values for P are those eight entries.
// A workgroup holds NB = 2 wave blocks (consecutive in k, sharing the 128-byte lines that cross
// block edges) x NP = 4 phases of o (wave w walks o = ob + w / NB, ob + w / NB + 4, ...), so the
// two waves that share a SIMD work on different o and their MFMA and VALU phases interleave.
// Each wave writes its P[o] to its block's LDS slot; at the end the slots are summed in a fixed
// order, the phases' accumulators are summed in LDS in a fixed order, and the workgroup writes
// FP32 partials that k_sum_parts adds in FP64 (results do not depend on scheduling).
// Consecutive groups of a range of o run on one XCD, so lines split between neighbouring k
// blocks are fetched once into that XCD's L2.
//
// Register fragment layouts of v_mfma_f32_16x16x32_bf16 (64 lanes, l = lane):
//   A (16 x 32): row l % 16, eight k values in the slot group l / 16
//   B (32 x 16): column l % 16, the same eight k values as A for the same slot group
//   D (16 x 16): column l % 16, rows 4 (l / 16) + 0..3
#include "common.h"

#include <mutex>

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef __bf16 bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef int i32x4 __attribute__((ext_vector_type(4)));

int tfx_mttkrp_native_fd(tfx_ctx* ctx, const float* T, const Shape& s, const double* F,
                         unsigned mask, double* out, hipStream_t stream);

namespace {

#ifndef TFX_MT_NW
#define TFX_MT_NW 8
#endif

constexpr int NW = TFX_MT_NW; // waves per workgroup (one workgroup per CU)
constexpr int NT = 64 * NW;   // threads per workgroup
constexpr int NB = 2;         // wave blocks per workgroup
constexpr int NP = NW / NB;   // o phases per wave block
constexpr int WB = 32;        // wave block: 32 j x 32 k
constexpr int RC = 64;        // ranks per launch
constexpr int MAXOC = 144;    // o per workgroup at most (LDS Fo and P tables)
constexpr int MAXO = TFX_MAXL - 2;

constexpr int NXCD = 8;       // XCDs; workgroup g runs on XCD g % NXCD

constexpr int XBYTES = 2 * WB * WB * 2;                  // a wave's hi + lo block (4 KB)
constexpr size_t LDS_X = (size_t)NW * XBYTES;            // transpose buffers
constexpr size_t LDS_P = (size_t)NB * MAXOC * RC * 4;    // P slots [block][o][rank]
constexpr size_t LDS_BJ = (size_t)NB * WB * RC * 4;      // F_{L-2} rows of each block, FP32
constexpr size_t LDS_TOTAL = LDS_X + LDS_P + LDS_BJ;
constexpr size_t LDS_COMB = (size_t)NW * WB * RC * 4;    // end: accumulators [wave][32][64]
static_assert(LDS_COMB <= LDS_X + LDS_P, "the combine area reuses the transpose buffers and P slots");
static_assert(LDS_TOTAL <= 160 * 1024, "LDS budget");

struct MPlan {
  Shape s;
  int rc0;
  int64_t K, J, NO;
  int kb, jb;        // 32-wide blocks along k and j
  int nblk, ngrp;    // wave blocks (kb x jb) and workgroup groups of NB of them
  int osplit;
  int64_t ochunk;
  int nwg;           // groups x o ranges
  int nlaunch;       // nwg rounded up to a multiple of NXCD
  int ngk, ngj;      // 8-row groups of the prepared F_{L-1} and F_{L-2} fragments (4 per block)
  // operands prepared by k_mttkrp_prep for rank chunk rc0:
  const float* fo;       // Fo[o][ln][cb] (FP32)
  const bf16x8* cf;      // F_{L-1} fragments [g8][cb][hi/lo][ln]
  const bf16x8* bf;      // F_{L-2} fragments [g8][cb][hi/lo][ln]
  const float* bjf;      // F_{L-2} in FP32 [g8][cb][ln][8]
};

__device__ inline void split(float x, __bf16& h, __bf16& l) {
  h = (__bf16)x;
  l = (__bf16)(x - (float)h);
}

typedef float f32x2 __attribute__((ext_vector_type(2)));
typedef __bf16 bf16x2 __attribute__((ext_vector_type(2)));
typedef unsigned u32x4 __attribute__((ext_vector_type(4)));

// split() of two values with packed conversions: hi pair by one v_cvt_pk_bf16_f32, its two
// halves back to FP32 by a shift and a mask, the remainders' pair by a second conversion
__device__ inline void split2(float x0, float x1, unsigned& h, unsigned& l) {
  h = __builtin_bit_cast(unsigned, __builtin_convertvector((f32x2){x0, x1}, bf16x2));
  const float h0 = __builtin_bit_cast(float, h << 16);
  const float h1 = __builtin_bit_cast(float, h & 0xffff0000u);
  l = __builtin_bit_cast(unsigned, __builtin_convertvector((f32x2){x0 - h0, x1 - h1}, bf16x2));
}

__device__ inline void split8(const float v[8], bf16x8& h, bf16x8& l) {
  u32x4 hh, ll;
#pragma unroll
  for (int e = 0; e < 4; ++e) {
    unsigned a, b;
    split2(v[2 * e], v[2 * e + 1], a, b);
    hh[e] = a;
    ll[e] = b;
  }
  h = __builtin_bit_cast(bf16x8, hh);
  l = __builtin_bit_cast(bf16x8, ll);
}

// Raw buffer loads: an offset at or past the descriptor's record count reads as zeros, without a
// branch or a select on the loaded data.
constexpr int RSRC_FLAGS = 0x00020000;     // gfx9 raw buffer descriptor word 3
constexpr int OFF_NONE = 0x7ffffff0;       // past any record count used here: reads as zeros

// j of lane row m in GEMM 1 row block rb (see the permutation above)
__device__ inline int jrow(int rb, int m) { return 8 * (m >> 2) + 4 * rb + (m & 3); }

// Byte offset of the 16-byte chunk `chunk` (k = 8 chunk..8 chunk+7) of row j in a wave's 32 x 32
// bf16 block (64-byte rows).  The chunk index is XORed with sx(j) = 2 ((j >> 3) & 1) +
// ((j >> 4) & 1): the 16 lanes of a store quarter write rows from all four row octets, which
// then fill 16 distinct 16-byte bank groups, and the 16-lane groups 0, 1 (and 2, 3) of a
// transposing read (rows 8 G + c) read disjoint 32-byte halves of their rows.
__device__ inline int xoff(int j, int chunk) {
  return j * 64 + ((chunk ^ (2 * ((j >> 3) & 1) + ((j >> 4) & 1))) << 4);
}

// Byte offsets of the two transposing reads that give a lane its GEMM 2 A fragment for k block
// kb2 (A = T^T: row k = 16 kb2 + lane % 16, eight j = 8 (lane / 16) + 0..7).
// ds_read_b64_tr_b16 works on groups of 16 lanes: lane i of a group receives element i % 4 of
// the 8-byte chunks addressed by lanes 4c + i / 4, c = 0..3, of its group.  Lane 4c + q of
// group G therefore addresses row 8 G + c (+ 4) at k = 16 kb2 + 4 q, and lane i gets
// T[8 G + c (+ 4), 16 kb2 + i]: the first read gives slots 0..3, the second slots 4..7.
__device__ inline void tr_offsets(int kb2, int lane, int off[2]) {
  const int G = lane >> 4, c = (lane >> 2) & 3, q = lane & 3;
  const int k = 16 * kb2 + 4 * q;
#pragma unroll
  for (int h = 0; h < 2; ++h) {
    const int j = 8 * G + c + 4 * h;
    off[h] = xoff(j, k >> 3) + (k & 7) * 2;
  }
}

// v_permlane32_swap_b32 a, b exchanges the upper 32 lanes of a with the lower 32 lanes of b;
// v_permlane16_swap_b32 a, b exchanges the odd 16-lane rows of a with the even rows of b.
// Written as inline asm (register permutes, no memory access; s_nop on both sides for the VALU
// hazards): this toolchain's builtins for these two-result instructions return the first result
// twice.
__device__ inline void swap32(float& a, float& b) {
  asm volatile("s_nop 1\n\tv_permlane32_swap_b32 %0, %1\n\ts_nop 1" : "+v"(a), "+v"(b));
}
__device__ inline void swap16(float& a, float& b) {
  asm volatile("s_nop 1\n\tv_permlane16_swap_b32 %0, %1\n\ts_nop 1" : "+v"(a), "+v"(b));
}

// p[cb] holds, in every lane, a partial sum for rank column 16 cb + lane % 16; the partials of
// the four 16-lane rows add up to the total.  Returns in lane (G, c) the total of column 16 G + c
// (rank 16 G + c = lane), with a fixed summation order.
__device__ inline float rows_total(const float p[4]) {
  float a = p[0], b = p[2];
  swap32(a, b);             // lower half: p0 own / p0 of row + 2; upper: p2 of row - 2 / p2 own
  const float s02 = a + b;  // rows 0, 1: rank block 0; rows 2, 3: rank block 2 (two rows each)
  float c = p[1], d = p[3];
  swap32(c, d);
  const float s13 = c + d;  // rows 0, 1: rank block 1; rows 2, 3: rank block 3
  float x = s02, y = s13;
  swap16(x, y);             // row 0: s02 of rows 0, 1; row 1: s13 of rows 0, 1; rows 2, 3 alike
  return x + y;
}

__device__ inline f32x4 mfma3(bf16x8 ah, bf16x8 al, bf16x8 bh, bf16x8 bl, f32x4 c) {
  c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah, bh, c, 0, 0, 0);
  c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah, bl, c, 0, 0, 0);
  c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(al, bh, c, 0, 0, 0);
  return c;
}

// decompose the outer row index o over modes 0..L-3 (last fastest); static indexing only
__device__ inline void outer_idx(const Shape& s, int64_t o, int64_t idx[MAXO]) {
#pragma unroll
  for (int m = MAXO - 1; m >= 0; --m) {
    if (m < s.L - 2) {
      idx[m] = o % s.d[m];
      o /= s.d[m];
    } else {
      idx[m] = 0;
    }
  }
}

// Loads of factor entries are unconditional (an entry outside the factor reads col[0], a valid
// address, and is zeroed afterwards) so that all of them are in flight together instead of each
// waiting behind a branch.
__device__ inline double load_or0(const double* col, int64_t x, bool ok) {
  const double v = col[ok ? x : 0];
  return ok ? v : 0.0;
}

// Operands of k_mttkrp_mfma for rank chunk rc0, written once per call (every workgroup of the
// main kernel reads them with a few contiguous 16-byte loads instead of gathering and splitting
// factor entries itself):
//   Fo[o][ln][cb] = prod_{m<L-2} F_m[i_m(o), r]  (FP32; 1 for order 2), r = rc0 + 16 cb + ln;
//   fragments of F_{L-1} and F_{L-2}: for the 8-row group g8 (rows 8 g8 + e, e = 0..7; rows past
//   the mode and ranks past R are zeros) the hi and lo bf16 of each rank column, laid out
//   [g8][cb][hi/lo][ln] (16 bytes each), and F_{L-2} in FP32 as [g8][cb][ln][8].
// Thread jobs: NO x 64 Fo entries, then ngk x 64 and ngj x 64 fragment groups.
__global__ void k_mttkrp_prep(const double* __restrict__ F, MPlan P, float* __restrict__ fo,
                              bf16x8* __restrict__ cf, bf16x8* __restrict__ bf,
                              float* __restrict__ bjf) {
  const Shape& s = P.s;
  const int L = s.L, R = s.R;
  int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e < P.NO * RC) {
    // consecutive threads take consecutive o of one rank: contiguous reads of the factor columns
    const int64_t o = e % P.NO;
    const int q = (int)(e / P.NO), ln = q & 15, cb = q >> 4;
    const int r = P.rc0 + q;
    float f = 0.f;
    if (r < R) {
      double fd = 1.0;
      int64_t oi = o;
      for (int m = L - 3; m >= 0; --m) {
        const int64_t dm = s.d[m];
        fd *= F[s.off[m] + (int64_t)r * dm + oi % dm];
        oi /= dm;
      }
      f = (float)fd;
    }
    fo[o * RC + ln * 4 + cb] = f;
    return;
  }
  e -= P.NO * RC;
  const bool isk = e < (int64_t)P.ngk * RC;
  if (!isk) e -= (int64_t)P.ngk * RC;
  if (!isk && e >= (int64_t)P.ngj * RC) return;
  const int g8 = (int)(e / RC), cb = (int)(e % RC) >> 4, ln = (int)e & 15;
  const int l = isk ? L - 1 : L - 2;
  const int64_t n = s.d[l];
  const int r = P.rc0 + 16 * cb + ln;
  float v[8];
#pragma unroll
  for (int u = 0; u < 8; ++u) {
    const int64_t i = 8 * (int64_t)g8 + u;
    v[u] = r < R && i < n ? (float)F[s.off[l] + (int64_t)r * n + i] : 0.f;
  }
  bf16x8 h, lo;
  split8(v, h, lo);
  bf16x8* dst = isk ? cf : bf;
  dst[((int64_t)(g8 * 4 + cb) * 2 + 0) * 16 + ln] = h;
  dst[((int64_t)(g8 * 4 + cb) * 2 + 1) * 16 + ln] = lo;
  if (!isk) {
    f32x4* b4 = (f32x4*)(bjf + ((int64_t)(g8 * 4 + cb) * 16 + ln) * 8);
    b4[0] = f32x4{v[0], v[1], v[2], v[3]};
    b4[1] = f32x4{v[4], v[5], v[6], v[7]};
  }
}

__global__ __launch_bounds__(NT, 1) __attribute__((amdgpu_waves_per_eu(1, NW / 4)))
void k_mttkrp_mfma(const float* __restrict__ T, MPlan P,
                                                       const double* __restrict__ F,
                                                       float* __restrict__ Pk,
                                                       float* __restrict__ Pj,
                                                       float* __restrict__ Po) {
  extern __shared__ __align__(16) unsigned char smem[];
  const Shape& s = P.s;
  const int L = s.L, R = s.R;
  const int t = threadIdx.x;
  // the wave index through readfirstlane: everything derived from it (block, phase, o) is then
  // known to be wave-uniform and lives in scalar registers (buffer descriptors included)
  const int w = __builtin_amdgcn_readfirstlane(t >> 6);
  const int lane = t & 63, ln = lane & 15, G = lane >> 4;
  // consecutive indices g (the groups of one o range, k fastest) go to one XCD
  const int per_xcd = P.nlaunch / NXCD;
  const int g = (int)(blockIdx.x % NXCD) * per_xcd + (int)(blockIdx.x / NXCD);
  if (g >= P.nwg) return;                   // padding workgroup (uniform, before any barrier)
  const int grp = g % P.ngrp, oc = g / P.ngrp;
  const int64_t ob = (int64_t)oc * P.ochunk;
  const int64_t oe = ob + P.ochunk < P.NO ? ob + P.ochunk : P.NO;
  const int nO = (int)(oe - ob);
  const int bl = w % NB, phase = w / NB;
  const int blk = grp * NB + bl;
  const bool bvalid = blk < P.nblk;         // wave-uniform
  const int jbi = bvalid ? blk / P.kb : 0, kbi = bvalid ? blk % P.kb : 0;
  const int64_t j0 = (int64_t)jbi * WB, k0 = (int64_t)kbi * WB;

  unsigned char* xb = smem + w * XBYTES;                 // this wave's block: hi, then lo
  float* po_t = (float*)(smem + LDS_X);                  // [NB][MAXOC][RC]
  // [NB][G][u][ln][4]: F_{L-2}[j0 + 8 G + 4 (u & 1) + c, rc0 + 16 (u >> 1) + ln] at c, so that
  // a lane reads its 32 values with eight conflict-free 16-byte reads
  float* bj_t = (float*)(smem + LDS_X + LDS_P) + (bl * 4 + G) * 8 * 64 + ln * 4;

  for (int e = t; e < NB * MAXOC * RC; e += NT) po_t[e] = 0.f;

  // B fragments, rank r = rc0 + 16 cb + ln (prepared by k_mttkrp_prep):
  //   GEMM 1: F_{L-1}[k0 + 8 G + e, r],  GEMM 2: F_{L-2}[j0 + 8 G + e, r] (also in FP32 for P:
  //   they are the rows 8 G + 4 rb + i this lane holds in GEMM 1's result)
  bf16x8 ch[4], cl[4], bh[4], blo[4];
  {
    const int gk = bvalid ? 4 * kbi + G : 0, gj = bvalid ? 4 * jbi + G : 0;
#pragma unroll
    for (int cb = 0; cb < 4; ++cb) {
      ch[cb] = P.cf[((gk * 4 + cb) * 2 + 0) * 16 + ln];
      cl[cb] = P.cf[((gk * 4 + cb) * 2 + 1) * 16 + ln];
      bh[cb] = P.bf[((gj * 4 + cb) * 2 + 0) * 16 + ln];
      blo[cb] = P.bf[((gj * 4 + cb) * 2 + 1) * 16 + ln];
    }
    if (phase == 0) {
      f32x4 bv[4][2];
#pragma unroll
      for (int cb = 0; cb < 4; ++cb)
#pragma unroll
        for (int u = 0; u < 2; ++u)
          bv[cb][u] = *(const f32x4*)(P.bjf + ((int64_t)(gj * 4 + cb) * 16 + ln) * 8 + 4 * u);
#pragma unroll
      for (int cb = 0; cb < 4; ++cb)
#pragma unroll
        for (int u = 0; u < 2; ++u) *(f32x4*)(bj_t + (2 * cb + u) * 64) = bv[cb][u];
    }
  }

  // this lane's T loads: rows j0 + jrow(rb, ln), k = k0 + 8 G + 0..7 (two 16-byte halves); rows
  // past J lie past the record count J K 4 of T[o]'s descriptor and read as zeros, k past K (K %
  // 4 == 0: whole halves) gets OFF_NONE
  const int64_t kq = k0 + 8 * G;
  int toff[2][2];
#pragma unroll
  for (int rb = 0; rb < 2; ++rb) {
    const int off = (int)(((j0 + jrow(rb, ln)) * P.K + kq) * 4);
    toff[rb][0] = kq < P.K ? off : OFF_NONE;
    toff[rb][1] = kq + 4 < P.K ? off + 16 : OFF_NONE;
  }
  const int trec = (int)(P.J * P.K * 4);
  auto load = [&](int64_t o, f32x4 v[4]) __attribute__((always_inline)) {
    const bool ok = o < oe;
    const __amdgpu_buffer_rsrc_t rs = __builtin_amdgcn_make_buffer_rsrc(
        (void*)(T + (ok ? o : 0) * P.J * P.K), (short)0, ok ? trec : 0, RSRC_FLAGS);
    // (convert whole vectors: an element-wise bit_cast of the load result has been miscompiled
    // into a single-dword load by this toolchain)
#pragma unroll
    for (int rb = 0; rb < 2; ++rb)
#pragma unroll
      for (int h = 0; h < 2; ++h)
        v[2 * rb + h] = __builtin_bit_cast(
            f32x4, __builtin_amdgcn_raw_buffer_load_b128(rs, toff[rb][h], 0, 0));
  };
  int woff[2], roff[2][2];
#pragma unroll
  for (int rb = 0; rb < 2; ++rb) woff[rb] = xoff(jrow(rb, ln), G);
#pragma unroll
  for (int kb2 = 0; kb2 < 2; ++kb2) tr_offsets(kb2, lane, roff[kb2]);
  typedef __attribute__((address_space(3))) bf16x4 lds_bf16x4;
  __syncthreads();

  // block accumulators: acc1 = N_{L-2} rows j0 + 8 G + 4 rb + i, acc2 = N_{L-1} rows
  // k0 + 16 kb2 + 4 G + i, rank column 16 cb + ln
  f32x4 acc1[2][4], acc2[2][4];
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int cb = 0; cb < 4; ++cb) {
      acc1[a][cb] = f32x4{0.f, 0.f, 0.f, 0.f};
      acc2[a][cb] = f32x4{0.f, 0.f, 0.f, 0.f};
    }
  if (bvalid) {
    f32x4 v[4];
    load(ob + phase, v);
    const f32x4* fop = (const f32x4*)P.fo + ln;
    f32x4 fo_next = fop[(ob + phase < oe ? ob + phase : 0) * 16];
    for (int64_t o = ob + phase; o < oe; o += NP) {
      const int oo = (int)(o - ob);
      const f32x4 fo = fo_next;
      fo_next = fop[(o + NP < oe ? o + NP : o) * 16];
      bf16x8 ah[2], al[2];
#pragma unroll
      for (int rb = 0; rb < 2; ++rb) {
        float x[8];
#pragma unroll
        for (int e = 0; e < 4; ++e) {
          x[e] = v[2 * rb][e];
          x[4 + e] = v[2 * rb + 1][e];
        }
        split8(x, ah[rb], al[rb]);
      }
      load(o + NP, v);        // next tile in flight during this one's MFMAs
      // GEMM 1
      f32x4 y[2][4];
#pragma unroll
      for (int rb = 0; rb < 2; ++rb)
#pragma unroll
        for (int cb = 0; cb < 4; ++cb)
          y[rb][cb] = mfma3(ah[rb], al[rb], ch[cb], cl[cb], f32x4{0.f, 0.f, 0.f, 0.f});
      // stage the block in this wave's LDS and read it back transposed; LDS operations of a
      // wave complete in order, so the reads see the stores and the previous o's reads are done
      // before these stores land
      asm volatile("" ::: "memory");
#pragma unroll
      for (int rb = 0; rb < 2; ++rb) {
        *(bf16x8*)(xb + woff[rb]) = ah[rb];
        *(bf16x8*)(xb + 2048 + woff[rb]) = al[rb];
      }
      asm volatile("" ::: "memory");
      bf16x8 a2h[2], a2l[2];
#pragma unroll
      for (int kb2 = 0; kb2 < 2; ++kb2) {
        a2h[kb2] = __builtin_shufflevector(
            __builtin_amdgcn_ds_read_tr16_b64_v4bf16((lds_bf16x4*)(xb + roff[kb2][0])),
            __builtin_amdgcn_ds_read_tr16_b64_v4bf16((lds_bf16x4*)(xb + roff[kb2][1])),
            0, 1, 2, 3, 4, 5, 6, 7);
        a2l[kb2] = __builtin_shufflevector(
            __builtin_amdgcn_ds_read_tr16_b64_v4bf16((lds_bf16x4*)(xb + 2048 + roff[kb2][0])),
            __builtin_amdgcn_ds_read_tr16_b64_v4bf16((lds_bf16x4*)(xb + 2048 + roff[kb2][1])),
            0, 1, 2, 3, 4, 5, 6, 7);
      }
      f32x4 bjv[2][4];
#pragma unroll
      for (int cb = 0; cb < 4; ++cb)
#pragma unroll
        for (int rb = 0; rb < 2; ++rb) bjv[rb][cb] = *(const f32x4*)(bj_t + (2 * cb + rb) * 64);
      // GEMM 2, and GEMM 1's epilogue (Fo-scaled block accumulator and this lane's share of P),
      // independent of each other
      f32x4 z[2][4];
#pragma unroll
      for (int kb2 = 0; kb2 < 2; ++kb2)
#pragma unroll
        for (int cb = 0; cb < 4; ++cb)
          z[kb2][cb] = mfma3(a2h[kb2], a2l[kb2], bh[cb], blo[cb], f32x4{0.f, 0.f, 0.f, 0.f});
      float p[4];
#pragma unroll
      for (int cb = 0; cb < 4; ++cb) {
        float pc = 0.f;
#pragma unroll
        for (int rb = 0; rb < 2; ++rb)
#pragma unroll
          for (int i = 0; i < 4; ++i) {
            acc1[rb][cb][i] += fo[cb] * y[rb][cb][i];
            pc += bjv[rb][cb][i] * y[rb][cb][i];
          }
        p[cb] = pc;
      }
      if (L > 2) po_t[bl * MAXOC * RC + oo * RC + lane] = rows_total(p);
#pragma unroll
      for (int kb2 = 0; kb2 < 2; ++kb2)
#pragma unroll
        for (int cb = 0; cb < 4; ++cb)
#pragma unroll
          for (int i = 0; i < 4; ++i) acc2[kb2][cb][i] += fo[cb] * z[kb2][cb][i];
    }
  }
  __syncthreads();

  // P partial of this group: the NB block slots summed in a fixed order
  if (L > 2)
    for (int e = t; e < nO * RC; e += NT) {
      float a = 0.f;
#pragma unroll
      for (int b = 0; b < NB; ++b) a += po_t[b * MAXOC * RC + e];
      Po[((int64_t)grp * P.NO + ob) * RC + e] = a;
    }
  __syncthreads();          // the combine area below overlaps the P slots
  // block accumulators: the NP phases of each block summed in LDS in a fixed order
  float* cm = (float*)smem;                 // [NW][32][RC]
#pragma unroll
  for (int pass = 0; pass < 2; ++pass) {
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
      for (int cb = 0; cb < 4; ++cb)
#pragma unroll
        for (int i = 0; i < 4; ++i) {
          const int x = pass == 0 ? 8 * G + 4 * a + i : 16 * a + 4 * G + i;
          cm[(w * WB + x) * RC + 16 * cb + ln] = pass == 0 ? acc1[a][cb][i] : acc2[a][cb][i];
        }
    __syncthreads();
    for (int e = t; e < NB * WB * RC; e += NT) {
      const int b = e / (WB * RC), x = (e / RC) % WB, q = e % RC;
      const int bk = grp * NB + b;
      if (bk >= P.nblk) continue;
      const int64_t jb = bk / P.kb, kb = bk % P.kb;
      const int64_t row = (pass == 0 ? jb : kb) * WB + x;
      if (row >= (pass == 0 ? P.J : P.K)) continue;
      float a = 0.f;
#pragma unroll
      for (int ph = 0; ph < NP; ++ph) a += cm[((b + NB * ph) * WB + x) * RC + q];
      if (pass == 0)
        Pj[((kb * P.osplit + oc) * P.J + row) * RC + q] = a;
      else
        Pk[((jb * P.osplit + oc) * P.K + row) * RC + q] = a;
    }
    __syncthreads();
  }
}

// Sums of the brick partials, all jobs in one launch.  Job b covers rows x < rows_b of a
// partial array part_b[np_b][rows_b][RC]; its FP64 sum goes to out[off_b + r * rows_b + x]
// (r = rc0 + q < R), or, for to_ps, to Ps[x * RC + q].  A workgroup (4 waves) owns 64
// consecutive entries of one job (job sizes are multiples of RC = 64); wave w adds the partials
// [w np / 4, (w + 1) np / 4) in order, eight loads in flight per lane, and wave 0 adds the four
// sums in wave order (a fixed order: results do not depend on scheduling).
struct SumJob {
  const float* part;
  int np;
  int to_ps;
  int64_t rows, off;
};
struct SumJobs {
  SumJob job[3];
  int n;
};

constexpr int SUM_WAVES = 4;

constexpr int SUM_ROWS = 4;   // rows per workgroup: 16 lanes x 4 ranks per row, 16-byte loads

// Output (x, q) of a job = the sum of its np partials part[p][x][q], p ascending: wave w sums
// the partials of quarter w, and wave 0 adds the quarters in wave order.  A workgroup takes rows
// x0 .. x0 + 3 of one job, lane (x0 + lane / 16, ranks 4 (lane % 16) .. + 3).
__global__ __launch_bounds__(64 * SUM_WAVES) void k_sum_parts(SumJobs J, int rc0, int R,
                                                             double* __restrict__ out,
                                                             double* __restrict__ Ps) {
  __shared__ double red[SUM_WAVES - 1][4][64];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  // job b and its first tile, with constant indices into the job table
  const int64_t blk = blockIdx.x;
  int b = 0;
  int64_t base = 0;
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    const int64_t t = (J.job[k].rows + SUM_ROWS - 1) / SUM_ROWS;
    if (k < J.n && b == k && blk >= base + t) {
      base += t;
      b = k + 1;
    }
  }
  if (b >= J.n) return;     // uniform over the workgroup
  const float* part = b == 0 ? J.job[0].part : (b == 1 ? J.job[1].part : J.job[2].part);
  const int np = b == 0 ? J.job[0].np : (b == 1 ? J.job[1].np : J.job[2].np);
  const int to_ps = b == 0 ? J.job[0].to_ps : (b == 1 ? J.job[1].to_ps : J.job[2].to_ps);
  const int64_t rows = b == 0 ? J.job[0].rows : (b == 1 ? J.job[1].rows : J.job[2].rows);
  const int64_t off = b == 0 ? J.job[0].off : (b == 1 ? J.job[1].off : J.job[2].off);
  const int64_t x = (blk - base) * SUM_ROWS + (lane >> 4);
  const int qd = lane & 15;
  const bool xok = x < rows;
  const int64_t stride = rows * RC;
  const int p0 = (int)((int64_t)np * w / SUM_WAVES);
  const int p1 = (int)((int64_t)np * (w + 1) / SUM_WAVES);
  const float* src = part + (xok ? x : 0) * RC + 4 * qd;
  typedef float f32x4 __attribute__((ext_vector_type(4)));
  double acc[4] = {0.0, 0.0, 0.0, 0.0};
  int p = p0;
  for (; p + 8 <= p1; p += 8) {
    f32x4 v[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) v[u] = *(const f32x4*)(src + (int64_t)(p + u) * stride);
#pragma unroll
    for (int u = 0; u < 8; ++u)
#pragma unroll
      for (int h = 0; h < 4; ++h) acc[h] += (double)v[u][h];
  }
  for (; p < p1; ++p) {
    const f32x4 v = *(const f32x4*)(src + (int64_t)p * stride);
#pragma unroll
    for (int h = 0; h < 4; ++h) acc[h] += (double)v[h];
  }
  if (w > 0)
#pragma unroll
    for (int h = 0; h < 4; ++h) red[w - 1][h][lane] = acc[h];
  __syncthreads();
  if (w != 0 || !xok) return;
#pragma unroll
  for (int h = 0; h < 4; ++h) {
#pragma unroll
    for (int u = 0; u < SUM_WAVES - 1; ++u) acc[h] += red[u][h][lane];
    const int q = 4 * qd + h;
    if (to_ps)
      Ps[x * RC + q] = acc[h];
    else if (rc0 + q < R)
      out[off + (int64_t)(rc0 + q) * rows + x] = acc[h];
  }
}

// Leading modes of an order >= 4 tensor from the summed P[o, r] (FP64, [NO][RC]):
// out_m[i, r] = sum_{o: i_m(o) = i} P[o, r] prod_{q<L-2, q!=m} F_q[i_q(o), r]
__global__ void k_outer_modes(const double* __restrict__ Ps, Shape s, int64_t NO, int rc0,
                              const double* __restrict__ F, unsigned mask,
                              double* __restrict__ out) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  // e enumerates (m, i, q) over the outer modes
  int64_t rem = e;
  int m = -1;
  for (int mm = 0; mm < s.L - 2; ++mm) {
    const int64_t n = s.d[mm] * RC;
    if (rem < n) { m = mm; break; }
    rem -= n;
  }
  if (m < 0 || !((mask >> m) & 1u)) return;
  const int64_t i = rem / RC;
  const int q = (int)(rem % RC);
  const int r = rc0 + q;
  if (r >= s.R) return;
  // strides of the outer index
  int64_t stride[MAXO];
  int64_t st = 1;
  for (int mm = s.L - 3; mm >= 0; --mm) {
    stride[mm] = st;
    st *= s.d[mm];
  }
  const int64_t nrest = NO / s.d[m];
  double acc = 0.0;
  for (int64_t u = 0; u < nrest; ++u) {
    // u enumerates the other outer indices (last fastest); insert i at position m
    int64_t v = u, o = i * stride[m];
    double f = 1.0;
    for (int mm = s.L - 3; mm >= 0; --mm) {
      if (mm == m) continue;
      const int64_t im = v % s.d[mm];
      v /= s.d[mm];
      o += im * stride[mm];
      f *= F[s.off[mm] + (int64_t)r * s.d[mm] + im];
    }
    acc += Ps[o * RC + q] * f;
  }
  out[s.off[m] + (int64_t)r * s.d[m] + i] = acc;
}

inline unsigned nblk(int64_t n) { return (unsigned)((n + 255) / 256); }

}  // namespace

namespace {

// Byte offsets of the brick partials in the kernel's scratch buffer.
struct SplitScratch {
  size_t k, j, o, s, fo, cf, bf, bjf, total;
};

// Plans the kernel for shape s (no device call): fills P and the scratch layout.
void split16_plan(const tfx_ctx* ctx, const Shape& s, MPlan& P, SplitScratch& W) {
  const int L = s.L;
  const int64_t K = s.d[L - 1];
  P.s = s;
  P.rc0 = 0;
  P.K = K;
  P.J = s.d[L - 2];
  P.NO = s.N / (K * P.J);
  P.kb = (int)((K + WB - 1) / WB);
  P.jb = (int)((P.J + WB - 1) / WB);
  P.nblk = P.kb * P.jb;
  P.ngrp = (P.nblk + NB - 1) / NB;
  // one workgroup per CU when the groups allow it; at most MAXOC values of o per workgroup
  int64_t osplit = std::max<int64_t>(1, ctx->cu_count / P.ngrp);
  osplit = std::min<int64_t>(osplit, P.NO);
  osplit = std::max<int64_t>(osplit, (P.NO + MAXOC - 1) / MAXOC);
  P.ochunk = (P.NO + osplit - 1) / osplit;
  P.osplit = (int)((P.NO + P.ochunk - 1) / P.ochunk);
  P.nwg = P.ngrp * P.osplit;
  P.nlaunch = (P.nwg + NXCD - 1) / NXCD * NXCD;
  P.ngk = 4 * P.kb;
  P.ngj = 4 * P.jb;
  P.fo = nullptr;
  P.cf = P.bf = nullptr;
  P.bjf = nullptr;
  const size_t nk = (size_t)P.jb * P.osplit * K * RC;
  const size_t nj = (size_t)P.kb * P.osplit * P.J * RC;
  const size_t no = L > 2 ? (size_t)P.ngrp * P.NO * RC : 0;
  const size_t ns = L > 3 ? (size_t)P.NO * RC : 0;
  W.k = 0;
  W.j = W.k + align256(4 * nk);
  W.o = W.j + align256(4 * nj);
  W.s = W.o + align256(4 * no);
  W.fo = W.s + align256(8 * ns);
  W.cf = W.fo + align256(4 * (size_t)P.NO * RC);
  W.bf = W.cf + align256(sizeof(bf16x8) * 8 * 16 * (size_t)P.ngk);
  W.bjf = W.bf + align256(sizeof(bf16x8) * 8 * 16 * (size_t)P.ngj);
  W.total = W.bjf + align256(4 * 8 * 4 * 16 * (size_t)P.ngj);
}

// Launches the kernel and the partial sums with the scratch buffer at wsp (every partial is
// written before it is read, so the buffer needs no clearing).
int split16_run(tfx_ctx* ctx, const float* T, MPlan P, const SplitScratch& W, const double* F,
                unsigned mask, double* out, void* wsp, hipStream_t stream) {
  const Shape& s = P.s;
  const int L = s.L;
  const int64_t K = P.K;
  float* Pk = (float*)((char*)wsp + W.k);
  float* Pj = (float*)((char*)wsp + W.j);
  float* Po = (float*)((char*)wsp + W.o);
  double* Ps = (double*)((char*)wsp + W.s);
  float* fo = (float*)((char*)wsp + W.fo);
  bf16x8* cf = (bf16x8*)((char*)wsp + W.cf);
  bf16x8* bf = (bf16x8*)((char*)wsp + W.bf);
  float* bjf = (float*)((char*)wsp + W.bjf);
  P.fo = fo;
  P.cf = cf;
  P.bf = bf;
  P.bjf = bjf;
  TFX_HIP(ctx, hipFuncSetAttribute((const void*)k_mttkrp_mfma,
                                   hipFuncAttributeMaxDynamicSharedMemorySize, (int)LDS_TOTAL));
  for (int rc0 = 0; rc0 < s.R; rc0 += RC) {
    P.rc0 = rc0;
    const int64_t nprep = (P.NO + P.ngk + P.ngj) * RC;
    hipLaunchKernelGGL(k_mttkrp_prep, dim3(nblk(nprep)), dim3(256), 0, stream, F, P, fo, cf, bf,
                       bjf);
    TFX_CHECK_LAUNCH(ctx);
    hipLaunchKernelGGL(k_mttkrp_mfma, dim3(P.nlaunch), dim3(NT), LDS_TOTAL, stream, T, P, F, Pk,
                       Pj, Po);
    TFX_CHECK_LAUNCH(ctx);
    SumJobs J;
    J.n = 0;
    if ((mask >> (L - 1)) & 1u) J.job[J.n++] = SumJob{Pk, P.jb * P.osplit, 0, K, s.off[L - 1]};
    if ((mask >> (L - 2)) & 1u) J.job[J.n++] = SumJob{Pj, P.kb * P.osplit, 0, P.J, s.off[L - 2]};
    const bool outer = L > 2 && (mask & ((1u << (L - 2)) - 1u));
    if (outer) J.job[J.n++] = SumJob{Po, P.ngrp, L > 3 ? 1 : 0, P.NO, s.off[0]};
    int64_t tot = 0;
    for (int b = 0; b < J.n; ++b) tot += (J.job[b].rows + SUM_ROWS - 1) / SUM_ROWS;
    if (tot > 0) {
      hipLaunchKernelGGL(k_sum_parts, dim3((unsigned)tot), dim3(64 * SUM_WAVES), 0, stream,
                         J, rc0, s.R, out, Ps);
      TFX_CHECK_LAUNCH(ctx);
    }
    if (L > 3 && outer) {
      int64_t tot = 0;
      for (int m = 0; m < L - 2; ++m) tot += s.d[m] * RC;
      hipLaunchKernelGGL(k_outer_modes, dim3(nblk(tot)), dim3(256), 0, stream, Ps, s, P.NO, rc0, F,
                         mask, out);
      TFX_CHECK_LAUNCH(ctx);
    }
  }
  return TFX_OK;
}

// Scratch of tfx_mttkrp_noalloc: one copy per GPU, held until the process ends.  Host code uses
// the device address from hipGetSymbolAddress, never the array's name.
constexpr size_t NOALLOC_BYTES = (size_t)TFX_NOALLOC_SCRATCH_MB << 20;
static __device__ __align__(256) unsigned char scratch[NOALLOC_BYTES];

constexpr int MAX_DEVICES = 64;
void* g_scratch_ptr[MAX_DEVICES];
std::mutex g_scratch_mu;

int scratch_address(tfx_ctx* ctx, void** out) {
  if (ctx->device < 0 || ctx->device >= MAX_DEVICES)
    return tfx_set_error(ctx, TFX_E_ARG, "device %d outside [0, %d)", ctx->device, MAX_DEVICES);
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch_ptr[ctx->device]) {
    int cur = 0;
    TFX_HIP(ctx, hipGetDevice(&cur));
    TFX_HIP(ctx, hipSetDevice(ctx->device));
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    TFX_HIP(ctx, hipSetDevice(cur));
    if (e != hipSuccess || !p)
      return tfx_set_error(ctx, TFX_E_HIP, "hipGetSymbolAddress(scratch) failed: %s",
                           hipGetErrorString(e));
    g_scratch_ptr[ctx->device] = p;
  }
  *out = g_scratch_ptr[ctx->device];
  return TFX_OK;
}

}  // namespace

int tfx_mttkrp_split16_fits(const Shape& s) {
  const int L = s.L;
  if (L < 2) return 0;
  const int64_t K = s.d[L - 1];
  // T[o] slabs are addressed by 32-bit buffer offsets (j * K + k) * 4 with j < J + 32 (rows past
  // J read as zeros), and the slab's record count J K 4 must stay below OFF_NONE
  const int64_t J = s.d[L - 2];
  return K % 4 == 0 && (J + WB) * K * 4 <= (int64_t)OFF_NONE && s.N / (J * K) <= 0x7fffffff;
}

int tfx_mttkrp_split16(tfx_ctx* ctx, const float* T, const Shape& s, const double* F,
                       unsigned mask, double* out, hipStream_t stream) {
  if (!tfx_mttkrp_split16_fits(s))
    return tfx_set_error(ctx, TFX_E_ARG, "shape outside the tiled split16 kernel");
  MPlan P;
  SplitScratch W;
  split16_plan(ctx, s, P, W);
  void* wsp;
  int rc = tfx_workspace(ctx, W.total + 256, stream, &wsp);
  if (rc) return rc;
  return split16_run(ctx, T, P, W, F, mask, out, wsp, stream);
}

extern "C" int tfx_mttkrp_noalloc(tfx_ctx* ctx, const void* T, int L, const int64_t* dims,
                                  const void* F_flat, int R, unsigned mode_mask, void* out_flat,
                                  hipStream_t stream) {
  if (!ctx || !T || !F_flat || !out_flat) return tfx_set_error(ctx, TFX_E_ARG, "null argument");
  Shape s;
  int rc = tfx_make_shape(ctx, L, dims, R, &s);
  if (rc) return rc;
  if (mode_mask == 0) return TFX_OK;
  if (!tfx_mttkrp_split16_fits(s))
    return tfx_set_error(ctx, TFX_E_UNSUPPORTED,
                         "tfx_mttkrp_noalloc: the last dimension must be a multiple of 4 and "
                         "(d[L-2] + 32) d[L-1] 4 below 2^31 (got %lld, %lld)",
                         (long long)s.d[L - 2], (long long)s.d[L - 1]);
  MPlan P;
  SplitScratch W;
  split16_plan(ctx, s, P, W);
  if (W.total > NOALLOC_BYTES)
    return tfx_set_error(ctx, TFX_E_UNSUPPORTED,
                         "tfx_mttkrp_noalloc: this shape needs %zu bytes of scratch, more than the "
                         "static %zu", W.total, NOALLOC_BYTES);
  void* wsp;
  rc = scratch_address(ctx, &wsp);
  if (rc) return rc;
  return split16_run(ctx, (const float*)T, P, W, (const double*)F_flat, mode_mask,
                     (double*)out_flat, wsp, stream);
}
