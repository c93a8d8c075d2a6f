// Split-precision MFMA paths of the MLSVD (FP32 data, 16-bit hi+lo operands, FP32 accumulation).
// These entry points report TFX_E_UNSUPPORTED until their MFMA kernels land, so that the split16
// precision fails loudly instead of silently using another path.
#include "common.h"

int tfx_gram_split16(tfx_ctx* ctx, const float*, const Shape&, int, double*, hipStream_t) {
  return tfx_set_error(ctx, TFX_E_UNSUPPORTED, "split16 Gram not built");
}

int tfx_ttm_split16(tfx_ctx* ctx, const float*, const Shape&, int, const double*, int, double*,
                    double*, hipStream_t) {
  return tfx_set_error(ctx, TFX_E_UNSUPPORTED, "split16 TTM not built");
}
