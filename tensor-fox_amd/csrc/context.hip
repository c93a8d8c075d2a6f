// Context, error reporting and workspace management of libtfxhip.
#include <stdarg.h>
#include <stdlib.h>

#include "common.h"

int tfx_set_error(tfx_ctx* ctx, int code, const char* fmt, ...) {
  if (ctx) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(ctx->err, sizeof(ctx->err), fmt, ap);
    va_end(ap);
  }
  return code;
}

int tfx_workspace(tfx_ctx* ctx, size_t bytes, hipStream_t stream, void** out) {
  if (bytes <= ctx->ws_size) {
    *out = ctx->ws;
    return TFX_OK;
  }
  // Work already queued may still read the old buffer: wait before releasing it.
  if (ctx->ws) {
    TFX_HIP(ctx, hipDeviceSynchronize());
    TFX_HIP(ctx, hipFree(ctx->ws));
    ctx->ws = nullptr;
    ctx->ws_size = 0;
  }
  size_t want = align256(bytes + bytes / 4);
  if (hipMalloc(&ctx->ws, want) != hipSuccess) {
    (void)hipGetLastError();
    ctx->ws = nullptr;
    return tfx_set_error(ctx, TFX_E_NOMEM, "workspace allocation of %zu bytes failed", want);
  }
  ctx->ws_size = want;
  *out = ctx->ws;
  return TFX_OK;
}

int tfx_make_shape(tfx_ctx* ctx, int L, const int64_t* dims, int R, Shape* s) {
  if (L < 2 || L > TFX_MAXL) return tfx_set_error(ctx, TFX_E_ARG, "order %d outside [2, %d]", L, TFX_MAXL);
  if (R < 1) return tfx_set_error(ctx, TFX_E_ARG, "rank %d < 1", R);
  s->L = L;
  s->R = R;
  s->N = 1;
  int64_t acc = 0;
  for (int l = 0; l < TFX_MAXL; ++l) {
    s->d[l] = l < L ? dims[l] : 1;
    s->off[l] = (int64_t)R * acc;
    if (l < L) {
      if (dims[l] < 1) return tfx_set_error(ctx, TFX_E_ARG, "dims[%d] = %lld < 1", l, (long long)dims[l]);
      acc += dims[l];
      s->N *= dims[l];
    }
  }
  s->n = (int64_t)R * acc;
  return TFX_OK;
}

extern "C" int tfx_ctx_create(int device, tfx_ctx** out) {
  if (!out) return TFX_E_ARG;
  *out = nullptr;
  tfx_ctx* ctx = (tfx_ctx*)calloc(1, sizeof(tfx_ctx));
  if (!ctx) return TFX_E_NOMEM;
  ctx->device = device;
  hipError_t e = hipSetDevice(device);
  if (e != hipSuccess) {
    free(ctx);
    return TFX_E_HIP;
  }
  hipDeviceProp_t prop;
  e = hipGetDeviceProperties(&prop, device);
  if (e != hipSuccess) {
    free(ctx);
    return TFX_E_HIP;
  }
  ctx->cu_count = prop.multiProcessorCount;
  ctx->arch_major = prop.major;
  ctx->arch_minor = prop.minor;
  if (hipMalloc(&ctx->dscal, TFX_DSCAL_BYTES) != hipSuccess) {
    free(ctx);
    return TFX_E_NOMEM;
  }
  if (hipMemset(ctx->dscal, 0, TFX_DSCAL_BYTES) != hipSuccess ||
      hipDeviceSynchronize() != hipSuccess) {
    (void)hipFree(ctx->dscal);
    free(ctx);
    return TFX_E_HIP;
  }
  *out = ctx;
  return TFX_OK;
}

extern "C" int tfx_ctx_destroy(tfx_ctx* ctx) {
  if (!ctx) return TFX_OK;
  (void)hipDeviceSynchronize();
  if (ctx->ws) (void)hipFree(ctx->ws);
  if (ctx->dscal) (void)hipFree(ctx->dscal);
  free(ctx);
  return TFX_OK;
}

extern "C" const char* tfx_last_error(const tfx_ctx* ctx) { return ctx ? ctx->err : "null context"; }

extern "C" int tfx_device_info(tfx_ctx* ctx, int* cu_count, int* arch_major, int* arch_minor) {
  if (!ctx) return TFX_E_ARG;
  if (cu_count) *cu_count = ctx->cu_count;
  if (arch_major) *arch_major = ctx->arch_major;
  if (arch_minor) *arch_minor = ctx->arch_minor;
  return TFX_OK;
}
