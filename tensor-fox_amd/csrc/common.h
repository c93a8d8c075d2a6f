// Shared definitions of libtfxhip: context, error reporting, workspace, tensor shape descriptor,
// block reductions.  All kernels target gfx950 (64-wide wavefronts).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>
#include <stdio.h>
#include <string.h>

#include <algorithm>

#include "tfxhip.h"

struct tfx_ctx {
  int device;
  int cu_count;
  int arch_major, arch_minor;
  char err[512];
  void* ws;           // device workspace, grown on demand
  size_t ws_size;
  double* dscal;      // device scratch for cross-workgroup sums (TFX_DSCAL_BYTES, zeroed at
                      // creation; calls that use it on one context are stream-ordered)
};

// Context scratch: 8 doubles of counters, then the per-rank partials of tfx_dgn_update (4 doubles
// per rank).  Counters (each the first 4 bytes of its double slot, reset by the kernel that uses
// it): slot 0 the rank-update kernel's arrivals, slot TFX_DSCAL_RES_COUNTER the residual's,
// slot TFX_DSCAL_GRAM_COUNTER the Gram products'.  The space after the counters holds one
// kernel's transient partials at a time (stream order).
constexpr size_t TFX_DSCAL_BYTES = (size_t)64 << 10;
constexpr int TFX_DSCAL_RES_COUNTER = 1;
constexpr int TFX_DSCAL_GRAM_COUNTER = 2;
constexpr int TFX_UPD_MAX_RANKS = (int)((TFX_DSCAL_BYTES / 8 - 8) / 4);

int tfx_set_error(tfx_ctx* ctx, int code, const char* fmt, ...);

#define TFX_HIP(ctx, call)                                                                 \
  do {                                                                                     \
    hipError_t e_ = (call);                                                                \
    if (e_ != hipSuccess)                                                                  \
      return tfx_set_error(ctx, TFX_E_HIP, "%s failed: %s", #call, hipGetErrorString(e_)); \
  } while (0)

#define TFX_CHECK_LAUNCH(ctx) TFX_HIP(ctx, hipGetLastError())

// Workspace of at least `bytes` bytes (256-byte aligned carving is the caller's job).
int tfx_workspace(tfx_ctx* ctx, size_t bytes, hipStream_t stream, void** out);

// Shape of a dense C-order tensor together with the flat factor layout offsets.
struct Shape {
  int L;
  int R;
  int64_t d[TFX_MAXL];    // dims
  int64_t off[TFX_MAXL];  // offset of factor l in the flat layout: R * sum_{m<l} d[m]
  int64_t N;              // prod d
  int64_t n;              // R * sum d
};

int tfx_make_shape(tfx_ctx* ctx, int L, const int64_t* dims, int R, Shape* s);

// Up to four copies src[k] -> dst[k] of n[k] doubles in one launch on `stream` (library-internal;
// a destination may be pinned host memory).
int tfx_copies(tfx_ctx* ctx, int np, const double* const* src, double* const* dst, const int64_t* n,
               hipStream_t stream);

static inline size_t align256(size_t x) { return (x + 255) & ~size_t(255); }

// Sum over the 64 lanes of a wavefront, returned in every lane; all lanes must be active.  DPP
// row rotations (8, 4, 2, 1) give every lane its 16-lane row's sum, then the four row sums are
// read from lanes 0, 16, 32, 48 into scalar registers: no LDS traffic and a fixed summation order.
template <int CTRL>
__device__ inline unsigned dpp_u32(unsigned v) {
  return (unsigned)__builtin_amdgcn_update_dpp(0, (int)v, CTRL, 0xf, 0xf, false);
}
template <int CTRL>
__device__ inline double dpp_f64(double x) {
  const unsigned long long u = __builtin_bit_cast(unsigned long long, x);
  const unsigned lo = dpp_u32<CTRL>((unsigned)u), hi = dpp_u32<CTRL>((unsigned)(u >> 32));
  return __builtin_bit_cast(double, ((unsigned long long)hi << 32) | lo);
}
__device__ inline double readlane_f64(double x, int l) {
  const unsigned long long u = __builtin_bit_cast(unsigned long long, x);
  const unsigned lo = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)u, l);
  const unsigned hi = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)(u >> 32), l);
  return __builtin_bit_cast(double, ((unsigned long long)hi << 32) | lo);
}
__device__ inline double wave_sum(double x) {
  x += dpp_f64<0x128>(x);   // row_ror:8
  x += dpp_f64<0x124>(x);   // row_ror:4
  x += dpp_f64<0x122>(x);   // row_ror:2
  x += dpp_f64<0x121>(x);   // row_ror:1
  return (readlane_f64(x, 0) + readlane_f64(x, 16)) + (readlane_f64(x, 32) + readlane_f64(x, 48));
}
__device__ inline float wave_sum(float x) {
  x += __builtin_bit_cast(float, dpp_u32<0x128>(__builtin_bit_cast(unsigned, x)));
  x += __builtin_bit_cast(float, dpp_u32<0x124>(__builtin_bit_cast(unsigned, x)));
  x += __builtin_bit_cast(float, dpp_u32<0x122>(__builtin_bit_cast(unsigned, x)));
  x += __builtin_bit_cast(float, dpp_u32<0x121>(__builtin_bit_cast(unsigned, x)));
  const unsigned u = __builtin_bit_cast(unsigned, x);
  const float r0 = __builtin_bit_cast(float, (unsigned)__builtin_amdgcn_readlane((int)u, 0));
  const float r1 = __builtin_bit_cast(float, (unsigned)__builtin_amdgcn_readlane((int)u, 16));
  const float r2 = __builtin_bit_cast(float, (unsigned)__builtin_amdgcn_readlane((int)u, 32));
  const float r3 = __builtin_bit_cast(float, (unsigned)__builtin_amdgcn_readlane((int)u, 48));
  return (r0 + r1) + (r2 + r3);
}

// Sum over a workgroup of blockDim.x threads (multiple of 64).  `red` must hold blockDim.x/64
// values.  Returns the total in every thread.
template <typename V>
__device__ inline V block_sum(V v, V* red) {
  v = wave_sum(v);
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63, nw = blockDim.x >> 6;
  __syncthreads();
  if (lane == 0) red[w] = v;
  __syncthreads();
  V t = 0;
  for (int i = 0; i < nw; ++i) t += red[i];
  return t;
}

template <typename V>
__device__ inline V block_max(V v, V* red) {
  for (int o = 32; o > 0; o >>= 1) {
    V u = __shfl_xor(v, o, 64);
    v = u > v ? u : v;
  }
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63, nw = blockDim.x >> 6;
  __syncthreads();
  if (lane == 0) red[w] = v;
  __syncthreads();
  V t = red[0];
  for (int i = 1; i < nw; ++i) t = red[i] > t ? red[i] : t;
  return t;
}
