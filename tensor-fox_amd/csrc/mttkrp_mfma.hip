// MTTKRP of every mode in one pass over an FP32 tensor on the bf16 matrix cores (gfx950
// v_mfma_f32_16x16x32_bf16), with FP64 factors and outputs.
//
// Split precision: every FP32 operand x is written as hi + lo with hi = bf16(x) and
// lo = bf16(x - hi) (|x - hi - lo| <= 2^-18 |x|); a product is formed as
// hi*hi + hi*lo + lo*hi with FP32 accumulation (the dropped lo*lo term is ~2^-16 relative).
//
// Notation: T is viewed as T[o, j, k] with k = i_{L-1} (contiguous), j = i_{L-2} and o the
// combined index of modes 0..L-3 ("outer", o = 0 when L = 2).  Fo[o, r] = prod_{m<L-2} F_m[i_m, r].
//   N_{L-1}[k, r] = sum_o Fo[o, r] Z_o[k, r],   Z_o = T[o]^T F_{L-2}     (GEMM 2)
//   N_{L-2}[j, r] = sum_o Fo[o, r] Y_o[j, r],   Y_o = T[o] F_{L-1}       (GEMM 1)
//   P[o, r]       = sum_j F_{L-2}[j, r] Y_o[j, r]
//   N_m[i, r]     = sum_{o: i_m(o) = i} P[o, r] prod_{q<L-2, q!=m} F_q[i_q(o), r]   (m < L-2)
// A workgroup owns a brick of 64 k x 64 j x a range of o and streams the tiles T[o, j-block,
// k-block] (64 x 64 FP32 = 16 KB) through LDS, stored once, row-major, as hi/lo bf16.  GEMM 1
// reads its A operand (rows of T) with plain 16-byte LDS reads; GEMM 2 needs T transposed and
// reads it with the gfx950 transposing LDS read (ds_read_b64_tr_b16, see tr_offsets).  Both GEMMs
// have a B operand that is fixed for the workgroup (the F_{L-1} and F_{L-2} rows of the brick,
// held in registers as hi/lo fragments), and Fo enters as a per-tile scale of the MFMA results.
// Bricks that share a (j-block, o range) run on one XCD, so that the 128-byte lines split
// between neighbouring k-blocks are fetched once into that XCD's L2.  The 8 waves split a tile as
//   waves 0..3: GEMM 1 on rows 16 w..16 w+15, all 64 ranks of the chunk (4 column blocks)
//   waves 4..7: GEMM 2 on k    16 (w-4)..16 (w-4)+15, all 64 ranks
// so no two waves read the same A operand, each wave runs 4 independent MFMA chains, every SIMD
// holds one wave of each kind, and every wave keeps its brick accumulator (N_{L-2} rows or
// N_{L-1} rows) in registers for the whole range of o.  P is summed over the 4 GEMM 1 waves with
// LDS atomics.  The brick partials (FP32) are written once per workgroup and summed in FP64 by
// k_sum_parts.
//
// Register fragment layouts of v_mfma_f32_16x16x32_bf16 (64 lanes, l = lane):
//   A (16 x 32): row l % 16, eight k values in the slot group l / 16
//   B (32 x 16): column l % 16, the same eight k values as A for the same slot group
//   D (16 x 16): column l % 16, rows 4 (l / 16) + 0..3
// Only the D layout matters for correctness: A and B share the k assignment of a slot group.
#include "common.h"

#include <mutex>

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef __bf16 bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef int i32x4 __attribute__((ext_vector_type(4)));

int tfx_mttkrp_native_fd(tfx_ctx* ctx, const float* T, const Shape& s, const double* F,
                         unsigned mask, double* out, hipStream_t stream);

namespace {

constexpr int NT = 512;       // threads per workgroup (8 waves)
constexpr int TB = 64;        // tile: 64 j x 64 k
constexpr int RC = 64;        // ranks per launch
constexpr int PD = 2;         // tiles prefetched ahead (even: the LDS buffer is (o - ob) & 1)
constexpr int MAXOC = 128;    // o per workgroup at most (LDS Fo and P tables)
constexpr int MAXO = TFX_MAXL - 2;
constexpr int TP = 64;        // tile pitch in bf16 (rows are swizzled, see swz)

constexpr int NXCD = 8;       // XCDs; workgroup g runs on XCD g % NXCD

constexpr size_t LDS_TILE = sizeof(__bf16) * 2 * 2 * TB * TP;   // [2 buf][2 hi/lo][64][TP]
constexpr size_t LDS_TAB = sizeof(float) * MAXOC * RC;
constexpr size_t LDS_TOTAL = LDS_TILE + 2 * LDS_TAB;
static_assert(LDS_TOTAL <= 160 * 1024, "LDS budget");

struct MPlan {
  Shape s;
  int rc0;
  int64_t K, J, NO;
  int kblocks, jblocks, osplit;
  int64_t ochunk;
  int nwg;       // bricks x o ranges
  int nlaunch;   // nwg rounded up to a multiple of NXCD
};

__device__ inline void split(float x, __bf16& h, __bf16& l) {
  h = (__bf16)x;
  l = (__bf16)(x - (float)h);
}

__device__ inline void split8(const float v[8], bf16x8& h, bf16x8& l) {
#pragma unroll
  for (int e = 0; e < 8; ++e) {
    __bf16 a, b;
    split(v[e], a, b);
    h[e] = a;
    l[e] = b;
  }
}

// global -> registers: thread t reads 8 consecutive k (8 (t & 7)..) of tile row t >> 3 with two
// 16-byte raw buffer loads.  An out-of-range half gets an offset past the buffer's record count,
// for which the hardware returns zeros: no branch and no select on the loaded data, so the
// compiler leaves the loads in flight until the tile is stored.  Needs K % 4 == 0 (a row's tail
// is then a whole half).
constexpr int RSRC_FLAGS = 0x00020000;     // gfx9 raw buffer descriptor word 3
constexpr int RSRC_RECORDS = 0x7ffffff0;
constexpr int OFF_NONE = 0x7ffffff0;       // >= RSRC_RECORDS: reads as zeros

__device__ inline void load_tile(const float* T, const MPlan& P, int64_t o, int64_t j0,
                                 int64_t k0, float v[8]) {
  const int t = threadIdx.x;
  const int row = t >> 3;
  const int64_t j = j0 + row;
  const int64_t k = k0 + 8 * (t & 7);
  const bool ok0 = o >= 0 && j < P.J && k < P.K;
  const bool ok1 = ok0 && k + 8 <= P.K;
  const float* base = T + (o >= 0 ? (o * P.J + j0) * P.K + k0 : 0);
  const __amdgpu_buffer_rsrc_t rs =
      __builtin_amdgcn_make_buffer_rsrc((void*)base, (short)0, RSRC_RECORDS, RSRC_FLAGS);
  const int off = (int)((int64_t)row * P.K + 8 * (t & 7)) * 4;
  // (convert whole vectors: an element-wise bit_cast of the load result has been miscompiled
  // into a single-dword load by this toolchain)
  const f32x4 a = __builtin_bit_cast(
      f32x4, __builtin_amdgcn_raw_buffer_load_b128(rs, ok0 ? off : OFF_NONE, 0, 0));
  const f32x4 b = __builtin_bit_cast(
      f32x4, __builtin_amdgcn_raw_buffer_load_b128(rs, ok1 ? off + 16 : OFF_NONE, 0, 0));
#pragma unroll
  for (int e = 0; e < 4; ++e) {
    v[e] = a[e];
    v[4 + e] = b[e];
  }
}

// Offset of the 16-byte chunk `chunk` (k = 8 chunk..8 chunk+7) of tile row `row`.  Chunks are
// permuted within a row by an XOR with g(row) = 2 (row & 3) + ((row >> 2) & 1), which keeps
// the three access patterns free of LDS bank conflicts: 16-byte stores and 16-byte row reads
// (eight consecutive rows, one chunk each: g takes 8 distinct values on 8 consecutive rows) and
// the transposing reads (rows j, j+1, j+2, j+3 with j % 8 in {0, 4}, two adjacent chunks).
__device__ inline int swz(int row, int chunk) {
  return row * TP + ((chunk ^ (2 * (row & 3) + ((row >> 2) & 1))) << 3);
}

// registers -> LDS buffer `buf` as hi/lo bf16, rm[j][k]
__device__ inline void store_tile(__bf16* rm, int buf, const float v[8]) {
  const int t = threadIdx.x;
  const int row = t >> 3, kq = t & 7;
  bf16x8 h, l;
  split8(v, h, l);
  *(bf16x8*)(rm + (buf * 2 + 0) * TB * TP + swz(row, kq)) = h;
  *(bf16x8*)(rm + (buf * 2 + 1) * TB * TP + swz(row, kq)) = l;
}

// Element offsets of the two transposing reads that give a lane its GEMM 2 A fragment
// (A = T^T: row k0 + lane % 16, eight j = j0 + 8 (lane / 16) + 0..7) from the row-major tile
// [j][k].  ds_read_b64_tr_b16 works on groups of 16 lanes: lane i of a group receives element
// i % 4 of the 8-byte chunks addressed by lanes 4c + i / 4, c = 0..3, of its group.  Lane
// 4c + q of group G therefore addresses [j0 + 8G + c (+ 4)][k0 + 4q], and lane i gets
// T[j0 + 8G + c (+ 4), k0 + i]: the first read gives slots 0..3, the second slots 4..7.
__device__ inline void tr_offsets(int j0, int k0, int lane, int off[2]) {
  const int G = lane >> 4, c = (lane >> 2) & 3, q = lane & 3;
  const int row = j0 + 8 * G + c, k = k0 + 4 * q;     // k0 % 8 == 0
  off[0] = swz(row, k >> 3) + (k & 7);
  off[1] = swz(row + 4, k >> 3) + (k & 7);
}

// x[l] + x[l ^ 16] + x[l ^ 32] + x[l ^ 48] in every lane, in VALU: v_permlane16_swap swaps the
// odd 16-lane rows of its first operand with the even rows of its second, v_permlane32_swap the
// upper half of the first with the lower half of the second, so with both operands = x the two
// results add up to x plus its partner row.  Written as inline asm (register permutes, no memory
// access; s_nop on both sides for the VALU hazards): this toolchain's builtins for these
// two-result instructions return the first result twice.
__device__ inline float swap_add16(float x) {
  float a = x, b = x;
  asm volatile("s_nop 1\n\tv_permlane16_swap_b32 %0, %1\n\ts_nop 1" : "+v"(a), "+v"(b));
  return a + b;
}

__device__ inline float swap_add32(float x) {
  float a = x, b = x;
  asm volatile("s_nop 1\n\tv_permlane32_swap_b32 %0, %1\n\ts_nop 1" : "+v"(a), "+v"(b));
  return a + b;
}

__device__ inline float sum_rows4(float x) { return swap_add32(swap_add16(x)); }

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

// B fragment: 8 consecutive rows x0 + e of a factor column (zeros outside), split hi/lo
__device__ inline void factor_frag(const double* col, int64_t n, int64_t x0, bool ok, bf16x8& h,
                                   bf16x8& l) {
  float v[8];
#pragma unroll
  for (int e = 0; e < 8; ++e) v[e] = (float)load_or0(col, x0 + e, ok && x0 + e < n);
  split8(v, h, l);
}

__global__ __launch_bounds__(NT, 1) void k_mttkrp_mfma(const float* __restrict__ T, MPlan P,
                                                       const double* __restrict__ F,
                                                       float* __restrict__ Pk,
                                                       float* __restrict__ Pj,
                                                       float* __restrict__ Po) {
  extern __shared__ __align__(16) unsigned char smem[];
  __bf16* rm = (__bf16*)smem;
  float* fo_t = (float*)(smem + LDS_TILE);             // [MAXOC][RC] Fo of the range
  float* po_t = (float*)(smem + LDS_TILE + LDS_TAB);   // [MAXOC][RC] P of the range

  const Shape& s = P.s;
  const int L = s.L, R = s.R;
  const int t = threadIdx.x;
  const int w = t >> 6, lane = t & 63, ln = lane & 15, rg = lane >> 4;
  const bool g1 = w < 4;          // GEMM 1 wave (else GEMM 2)
  const int sl = w & 3;           // its 16-row (GEMM 1) or 16-k (GEMM 2) slice of the tile
  // brick index: consecutive indices (the k-blocks of one (j-block, o range), k-block fastest)
  // go to one XCD
  const int per_xcd = P.nlaunch / NXCD;
  const int g = (int)(blockIdx.x % NXCD) * per_xcd + (int)(blockIdx.x / NXCD);
  if (g >= P.nwg) return;                   // padding workgroup (uniform, before any barrier)
  const int kb = g % P.kblocks;
  const int jb = (g / P.kblocks) % P.jblocks;
  const int oc = g / (P.kblocks * P.jblocks);
  const int64_t j0 = (int64_t)jb * TB, k0 = (int64_t)kb * TB;
  const int64_t ob = (int64_t)oc * P.ochunk;
  const int64_t oe = ob + P.ochunk < P.NO ? ob + P.ochunk : P.NO;
  const int nO = (int)(oe - ob);

  // B fragments (all 64 ranks of the chunk, column block cb = ranks 16 cb..16 cb+15):
  //   GEMM 1: F_{L-1}[k0 + 32 h + 8 rg + e, r],  GEMM 2: F_{L-2}[j0 + 32 h + 8 rg + e, r]
  const int Lb = g1 ? L - 1 : L - 2;
  const int64_t nb = g1 ? P.K : P.J, xb = g1 ? k0 : j0;
  bf16x8 bh[2][4], bl[2][4];
  float fjo[4][4];      // GEMM 1 waves: F_{L-2}[j0 + 16 sl + 4 rg + e, r] at their result rows
#pragma unroll
  for (int cb = 0; cb < 4; ++cb) {
    const int r = P.rc0 + 16 * cb + ln;
    const bool rok = r < R;
    const double* col = F + s.off[Lb] + (int64_t)(rok ? r : 0) * nb;
#pragma unroll
    for (int h = 0; h < 2; ++h) factor_frag(col, nb, xb + 32 * h + 8 * rg, rok, bh[h][cb], bl[h][cb]);
    const double* colj = F + s.off[L - 2] + (int64_t)(rok ? r : 0) * P.J;
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      const int64_t j = j0 + 16 * sl + 4 * rg + e;
      fjo[cb][e] = (float)load_or0(colj, j, g1 && rok && j < P.J);
    }
  }
  // Fo over the o range (order 3: one factor entry, order 2: 1), P accumulators
  constexpr int FOU = MAXOC * RC / NT;
  if (L <= 3) {
    float fv[FOU];
#pragma unroll
    for (int u = 0; u < FOU; ++u) {
      const int e = t + NT * u;
      const int oo = e / RC, r = P.rc0 + e % RC;
      const bool ok = oo < nO && r < R;
      fv[u] = L == 3 ? (float)load_or0(F + s.off[0], (int64_t)r * s.d[0] + ob + oo, ok)
                     : (ok ? 1.f : 0.f);
    }
#pragma unroll
    for (int u = 0; u < FOU; ++u) {
      const int e = t + NT * u;
      if (e < nO * RC) {
        fo_t[e] = fv[u];
        po_t[e] = 0.f;
      }
    }
  } else {
    for (int e = t; e < nO * RC; e += NT) {
      const int oo = e / RC, q = e % RC;
      const int r = P.rc0 + q;
      double f = r < R ? 1.0 : 0.0;
      if (r < R) {
        int64_t idx[MAXO];
        outer_idx(s, ob + oo, idx);
#pragma unroll
        for (int m = 0; m < MAXO; ++m)
          if (m < L - 2) f *= F[s.off[m] + (int64_t)r * s.d[m] + idx[m]];
      }
      fo_t[e] = (float)f;
      po_t[e] = 0.f;
    }
  }
  float pre[PD][8];
#pragma unroll
  for (int q = 0; q < PD; ++q) {
    load_tile(T, P, ob + q < oe ? ob + q : -1, j0, k0, pre[q]);
    __builtin_amdgcn_sched_barrier(0);     // keep the slots' loads in slot order
  }
  __syncthreads();

  // brick accumulator: N_{L-2} rows (GEMM 1 waves) or N_{L-1} rows (GEMM 2 waves),
  // lane holds rows 16 sl + 4 rg + e of rank column 16 cb + ln
  f32x4 bacc[4];
#pragma unroll
  for (int cb = 0; cb < 4; ++cb) bacc[cb] = f32x4{0.f, 0.f, 0.f, 0.f};

  // LDS offsets (elements) of this lane's A operand reads in buffer 0: GEMM 1 rows
  // 16 sl + ln (16-byte reads), GEMM 2 columns k = 16 sl + lane % 16 (transposing reads)
  int a_off[2][2];
#pragma unroll
  for (int h = 0; h < 2; ++h) {
    if (g1) {
      a_off[h][0] = swz(16 * sl + ln, 4 * h + rg);
      a_off[h][1] = 0;
    } else {
      tr_offsets(32 * h, 16 * sl, lane, a_off[h]);
    }
  }
  typedef __attribute__((address_space(3))) bf16x4 lds_bf16x4;

  // Tile o travels through register slot q = (o - ob) % PD, refilled with tile o + PD right
  // after it is stored.  Full groups of PD tiles run in a loop without early exits (so that the
  // compiler's count of loads in flight carries across its back edge); the rest follows.
  auto tile = [&](const int q, const int64_t o) __attribute__((always_inline)) {
      const int buf = q & 1;
      store_tile(rm, buf, pre[q]);
      load_tile(T, P, o + PD < oe ? o + PD : -1, j0, k0, pre[q]);
      // one barrier per tile: buffer q & 1 is written before it and read after it; the other
      // buffer was last read before this barrier
      __syncthreads();
      const int oo = (int)(o - ob);
      float fo[4];
#pragma unroll
      for (int cb = 0; cb < 4; ++cb) fo[cb] = fo_t[oo * RC + 16 * cb + ln];
      // A operands (rows and columns outside the tensor are zeros in LDS)
      const __bf16* th = rm + (buf * 2 + 0) * TB * TP;
      const __bf16* tl = rm + (buf * 2 + 1) * TB * TP;
      bf16x8 ah[2], al[2];
#pragma unroll
      for (int h = 0; h < 2; ++h) {
        if (g1) {
          ah[h] = *(const bf16x8*)(th + a_off[h][0]);
          al[h] = *(const bf16x8*)(tl + a_off[h][0]);
        } else {
          ah[h] = __builtin_shufflevector(
              __builtin_amdgcn_ds_read_tr16_b64_v4bf16((lds_bf16x4*)(th + a_off[h][0])),
              __builtin_amdgcn_ds_read_tr16_b64_v4bf16((lds_bf16x4*)(th + a_off[h][1])),
              0, 1, 2, 3, 4, 5, 6, 7);
          al[h] = __builtin_shufflevector(
              __builtin_amdgcn_ds_read_tr16_b64_v4bf16((lds_bf16x4*)(tl + a_off[h][0])),
              __builtin_amdgcn_ds_read_tr16_b64_v4bf16((lds_bf16x4*)(tl + a_off[h][1])),
              0, 1, 2, 3, 4, 5, 6, 7);
        }
      }
      // 4 independent accumulator chains (one per column block), interleaved
      f32x4 acc[4];
#pragma unroll
      for (int cb = 0; cb < 4; ++cb) acc[cb] = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
      for (int h = 0; h < 2; ++h)
#pragma unroll
        for (int term = 0; term < 3; ++term)
#pragma unroll
          for (int cb = 0; cb < 4; ++cb)
            acc[cb] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(term == 2 ? al[h] : ah[h],
                                                              term == 1 ? bl[h][cb] : bh[h][cb],
                                                              acc[cb], 0, 0, 0);
      // Fo scale into the brick accumulator; GEMM 1 waves also add P over their 16 rows (lane
      // rows summed in VALU) to the LDS table
#pragma unroll
      for (int cb = 0; cb < 4; ++cb) {
#pragma unroll
        for (int e = 0; e < 4; ++e) bacc[cb][e] += fo[cb] * acc[cb][e];
        if (g1) {
          float p = 0.f;
#pragma unroll
          for (int e = 0; e < 4; ++e) p += fjo[cb][e] * acc[cb][e];
          p = sum_rows4(p);
          if (rg == 0) atomicAdd(&po_t[oo * RC + 16 * cb + ln], p);
        }
      }
  };
  const int64_t ofull = ob + (int64_t)(nO / PD) * PD;
  for (int64_t o0 = ob; o0 < ofull; o0 += PD) {
#pragma unroll
    for (int q = 0; q < PD; ++q) tile(q, o0 + q);
  }
#pragma unroll
  for (int q = 0; q < PD; ++q)
    if (ofull + q < oe) tile(q, ofull + q);
  __syncthreads();

  // brick partials: lane holds rows 16 sl + 4 rg + e of rank columns 16 cb + ln
#pragma unroll
  for (int cb = 0; cb < 4; ++cb)
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      const int64_t x = 16 * sl + 4 * rg + e;
      const int rr = 16 * cb + ln;
      if (!g1 && k0 + x < P.K)
        Pk[(((int64_t)jb * P.osplit + oc) * P.K + k0 + x) * RC + rr] = bacc[cb][e];
      if (g1 && j0 + x < P.J)
        Pj[(((int64_t)kb * P.osplit + oc) * P.J + j0 + x) * RC + rr] = bacc[cb][e];
    }
  if (L > 2)
    for (int e = t; e < nO * RC; e += NT)
      Po[(((int64_t)kb * P.jblocks + jb) * P.NO + ob) * RC + e] = po_t[e];
}

// Sums of the brick partials, all jobs in one launch.  Job b covers rows x < rows_b of a
// partial array part_b[np_b][rows_b][RC]; its FP64 sum goes to out[off_b + r * rows_b + x]
// (r = rc0 + q < R), or, for to_ps, to Ps[x * RC + q].
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

__global__ void k_sum_parts(SumJobs J, int rc0, int R, double* __restrict__ out,
                            double* __restrict__ Ps) {
  int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  int b = 0;
  while (b < J.n && e >= J.job[b].rows * RC) e -= J.job[b++].rows * RC;
  if (b >= J.n) return;
  const SumJob& jb = J.job[b];
  const int64_t x = e / RC;
  const int q = (int)(e % RC);
  if (!jb.to_ps && rc0 + q >= R) return;
  const int64_t stride = jb.rows * RC;
  const float* p0 = jb.part + e;
  double acc = 0.0;
  int p = 0;
  for (; p + 8 <= jb.np; p += 8) {          // eight loads in flight per thread
    float v[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) v[u] = p0[(int64_t)(p + u) * stride];
#pragma unroll
    for (int u = 0; u < 8; ++u) acc += (double)v[u];
  }
  for (; p < jb.np; ++p) acc += (double)p0[(int64_t)p * stride];
  if (jb.to_ps)
    Ps[e] = acc;
  else
    out[jb.off + (int64_t)(rc0 + q) * jb.rows + x] = acc;
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
  size_t k, j, o, s, total;
};

// Plans the tiled kernel for shape s (no device call): fills P and the scratch layout.
void split16_plan(const tfx_ctx* ctx, const Shape& s, MPlan& P, SplitScratch& W) {
  const int L = s.L;
  const int64_t K = s.d[L - 1];
  P.s = s;
  P.rc0 = 0;
  P.K = K;
  P.J = s.d[L - 2];
  P.NO = s.N / (K * P.J);
  P.kblocks = (int)((K + TB - 1) / TB);
  P.jblocks = (int)((P.J + TB - 1) / TB);
  const int64_t bricks = (int64_t)P.kblocks * P.jblocks;
  // one workgroup per CU when the bricks allow it; at most MAXOC values of o per workgroup
  int64_t osplit = std::max<int64_t>(1, ctx->cu_count / bricks);
  osplit = std::min<int64_t>(osplit, P.NO);
  osplit = std::max<int64_t>(osplit, (P.NO + MAXOC - 1) / MAXOC);
  P.ochunk = (P.NO + osplit - 1) / osplit;
  P.osplit = (int)((P.NO + P.ochunk - 1) / P.ochunk);
  P.nwg = (int)(bricks * P.osplit);
  P.nlaunch = (P.nwg + NXCD - 1) / NXCD * NXCD;
  const size_t nk = (size_t)P.jblocks * P.osplit * K * RC;
  const size_t nj = (size_t)P.kblocks * P.osplit * P.J * RC;
  const size_t no = L > 2 ? (size_t)bricks * P.NO * RC : 0;
  const size_t ns = L > 3 ? (size_t)P.NO * RC : 0;
  W.k = 0;
  W.j = W.k + align256(4 * nk);
  W.o = W.j + align256(4 * nj);
  W.s = W.o + align256(4 * no);
  W.total = W.s + align256(8 * ns);
}

// Launches the tiled kernel and the partial sums with the scratch buffer at wsp (every partial
// is written before it is read, so the buffer needs no clearing).
int split16_run(tfx_ctx* ctx, const float* T, MPlan P, const SplitScratch& W, const double* F,
                unsigned mask, double* out, void* wsp, hipStream_t stream) {
  const Shape& s = P.s;
  const int L = s.L;
  const int64_t K = P.K;
  const int64_t bricks = (int64_t)P.kblocks * P.jblocks;
  float* Pk = (float*)((char*)wsp + W.k);
  float* Pj = (float*)((char*)wsp + W.j);
  float* Po = (float*)((char*)wsp + W.o);
  double* Ps = (double*)((char*)wsp + W.s);
  TFX_HIP(ctx, hipFuncSetAttribute((const void*)k_mttkrp_mfma,
                                   hipFuncAttributeMaxDynamicSharedMemorySize, (int)LDS_TOTAL));
  for (int rc0 = 0; rc0 < s.R; rc0 += RC) {
    P.rc0 = rc0;
    hipLaunchKernelGGL(k_mttkrp_mfma, dim3(P.nlaunch), dim3(NT), LDS_TOTAL, stream, T, P, F, Pk,
                       Pj, Po);
    TFX_CHECK_LAUNCH(ctx);
    SumJobs J;
    J.n = 0;
    if ((mask >> (L - 1)) & 1u) J.job[J.n++] = SumJob{Pk, P.jblocks * P.osplit, 0, K, s.off[L - 1]};
    if ((mask >> (L - 2)) & 1u)
      J.job[J.n++] = SumJob{Pj, P.kblocks * P.osplit, 0, P.J, s.off[L - 2]};
    const bool outer = L > 2 && (mask & ((1u << (L - 2)) - 1u));
    if (outer) J.job[J.n++] = SumJob{Po, (int)bricks, L > 3 ? 1 : 0, P.NO, s.off[0]};
    int64_t tot = 0;
    for (int b = 0; b < J.n; ++b) tot += J.job[b].rows * RC;
    if (tot > 0) {
      hipLaunchKernelGGL(k_sum_parts, dim3(nblk(tot)), dim3(256), 0, stream, J, rc0, s.R, out, Ps);
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
  // rows of a 64-row tile are addressed by 32-bit buffer offsets (row * K + k) * 4 bytes from the
  // tile's first row: 64 rows of K floats must stay below 2^31 bytes
  return K % 4 == 0 && (int64_t)TB * K * 4 < ((int64_t)1 << 31);
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
                         "tfx_mttkrp_noalloc: the last dimension must be a multiple of 4 below "
                         "2^23 (got %lld)", (long long)s.d[L - 1]);
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
