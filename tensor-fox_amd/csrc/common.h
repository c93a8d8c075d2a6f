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
  double* dscal;      // small device scratch for internal reductions
};

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

static inline size_t align256(size_t x) { return (x + 255) & ~size_t(255); }

// Sum over a workgroup of blockDim.x threads (multiple of 64).  `red` must hold blockDim.x/64
// values.  Returns the total in every thread.
template <typename V>
__device__ inline V block_sum(V v, V* red) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
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
