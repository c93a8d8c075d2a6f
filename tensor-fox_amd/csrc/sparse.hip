// Sparse COO input T = [data, idxs, dims] (TensorFox.py:256-280): the passes of the sparse
// pipeline over the stored entries.
//
//   tfx_sparse_pack      user coordinates (nnz x L, uint64, row-major, original mode order) ->
//                        per-mode int32 columns in the sorted mode order, with a bounds check
//                        (Auxiliar.py:99-106, Conversion.py:293-297)
//   tfx_sparse_gram      G = T_(m) T_(m)^T, dense I_m x I_m FP64 (Compression.py:123-130,
//                        211-215): entries sorted by their column key (the other coordinates),
//                        then every pair of entries that share a column adds d_a d_b to G
//   tfx_sparse_ttm0      Y = T x_1 U_0^T, dense (k, I_1, ..., I_{L-1}) FP64: the first step of
//                        the sparse multilinear product (MultilinearAlgebra.py:145-175); the
//                        other modes are dense mode products (tfx_ttm)
//   tfx_sparse_residual  sum_e (data_e - sum_r prod_l F_l[c_le, r])^2 over the stored entries
//                        (Critical.py:25-53)
//
// Layout: coordinates are L columns of nnz int32 (column l = sorted mode l), data is nnz FP64 in
// the same entry order.  The first sorted mode is the longest, so contracting it first gives the
// smallest dense intermediate.  Gram and first-mode product sum in a fixed order (entries sorted
// by column with stable radix sorts, one thread per output row or per column run), so their
// results do not depend on scheduling; the truncation that follows compares roundoff-sized
// singular values, so this keeps whole runs reproducible.

#include <rocprim/device/device_radix_sort.hpp>

#include "common.h"

namespace {

constexpr int kThreads = 256;

int grid_for(const tfx_ctx* ctx, int64_t n, int per_cu = 16) {
  int64_t nb = (n + kThreads - 1) / kThreads;
  nb = std::min<int64_t>(nb, (int64_t)ctx->cu_count * per_cu);
  return (int)std::max<int64_t>(nb, 1);
}

__global__ __launch_bounds__(kThreads) void k_fill_zero(double* __restrict__ x, int64_t n) {
  for (int64_t i = blockIdx.x * (int64_t)kThreads + threadIdx.x; i < n; i += (int64_t)gridDim.x * kThreads)
    x[i] = 0.0;
}

struct PackArgs {
  int L;
  int32_t src[TFX_MAXL];   // column of the user's coordinate row read for sorted mode l
  int64_t dim[TFX_MAXL];   // sorted dims
};

__global__ __launch_bounds__(kThreads) void k_sparse_pack(const uint64_t* __restrict__ in, int64_t nnz,
                                                          PackArgs a, int32_t* __restrict__ out,
                                                          int* __restrict__ bad) {
  int local_bad = 0;
  for (int64_t e = blockIdx.x * (int64_t)kThreads + threadIdx.x; e < nnz; e += (int64_t)gridDim.x * kThreads) {
    const uint64_t* row = in + e * a.L;
    for (int l = 0; l < a.L; ++l) {
      const uint64_t c = row[a.src[l]];
      const bool ok = c < (uint64_t)a.dim[l];
      local_bad |= !ok;
      out[(int64_t)l * nnz + e] = ok ? (int32_t)c : 0;
    }
  }
  if (local_bad) atomicOr(bad, 1);
}

__global__ __launch_bounds__(kThreads) void k_zero_int(int* p) {
  if (threadIdx.x == 0 && blockIdx.x == 0) *p = 0;
}

struct KeyArgs {
  int L, mode;
  int64_t stride[TFX_MAXL];   // key stride of each mode other than `mode` (0 for `mode`)
};

__global__ __launch_bounds__(kThreads) void k_col_keys(const int32_t* __restrict__ idx, int64_t nnz, KeyArgs a,
                                                       uint64_t* __restrict__ key, int32_t* __restrict__ ent) {
  for (int64_t e = blockIdx.x * (int64_t)kThreads + threadIdx.x; e < nnz; e += (int64_t)gridDim.x * kThreads) {
    uint64_t k = 0;
    for (int l = 0; l < a.L; ++l) k += (uint64_t)idx[(int64_t)l * nnz + e] * (uint64_t)a.stride[l];
    key[e] = k;
    ent[e] = (int32_t)e;
  }
}

// Row keys of the column-sorted entries: rkey[s] = row of the entry at column-sorted position s,
// rpos[s] = s (sorted next by row, stably, so each row lists its entries in column order).
__global__ __launch_bounds__(kThreads) void k_row_keys(const int32_t* __restrict__ ent,
                                                       const int32_t* __restrict__ rowc, int64_t nnz,
                                                       uint32_t* __restrict__ rkey, int32_t* __restrict__ rpos) {
  for (int64_t s = blockIdx.x * (int64_t)kThreads + threadIdx.x; s < nnz; s += (int64_t)gridDim.x * kThreads) {
    rkey[s] = (uint32_t)rowc[ent[s]];
    rpos[s] = (int32_t)s;
  }
}

// One thread per row i of G: for each entry e of row i (column order) and each entry f in the
// run of e's column (column-sorted order), G[i][row f] += d_e d_f where row f >= i.  Each
// unordered pair of rows is summed by the smaller row's thread alone, in a fixed order (pairs
// inside row i come in both orders, so duplicates add up as (sum d)^2); k_gram_mirror copies the
// upper triangle down.
__global__ __launch_bounds__(kThreads) void k_gram_rows(const uint64_t* __restrict__ key,
                                                        const int32_t* __restrict__ ent,
                                                        const uint32_t* __restrict__ rkey,
                                                        const int32_t* __restrict__ rpos,
                                                        const int32_t* __restrict__ rowc,
                                                        const double* __restrict__ data, int64_t nnz,
                                                        int64_t I, double* __restrict__ G) {
  for (int64_t i = blockIdx.x * (int64_t)kThreads + threadIdx.x; i < I; i += (int64_t)gridDim.x * kThreads) {
    // the range of row i in the row-sorted keys (lower and upper bound)
    int64_t lo = 0, hi = nnz;
    while (lo < hi) {
      const int64_t mid = (lo + hi) / 2;
      if ((int64_t)rkey[mid] < i) lo = mid + 1; else hi = mid;
    }
    int64_t end = lo;
    hi = nnz;
    while (end < hi) {
      const int64_t mid = (end + hi) / 2;
      if ((int64_t)rkey[mid] <= i) end = mid + 1; else hi = mid;
    }
    double* Gi = G + i * I;
    for (int64_t q = lo; q < end; ++q) {
      const int64_t s = rpos[q];
      const uint64_t k = key[s];
      const double d = data[ent[s]];
      int64_t t = s;
      while (t > 0 && key[t - 1] == k) --t;
      for (; t < nnz && key[t] == k; ++t) {
        const int32_t f = ent[t];
        const int64_t j = rowc[f];
        if (j >= i) Gi[j] += d * data[f];
      }
    }
  }
}

__global__ __launch_bounds__(kThreads) void k_gram_mirror(int64_t I, double* __restrict__ G) {
  for (int64_t g = blockIdx.x * (int64_t)kThreads + threadIdx.x; g < I * I; g += (int64_t)gridDim.x * kThreads) {
    const int64_t r = g / I, c = g - r * I;
    if (r > c) G[g] = G[c * I + r];
  }
}

struct Ttm0Args {
  int L, k;
  int64_t I0;                 // rows of U (column-major I0 x k)
  int64_t C;                  // prod_{l>=1} I_l
  int64_t stride[TFX_MAXL];   // C-order strides of modes 1..L-1 inside C (stride[0] unused)
};

// rest(e): C-order linear index of the coordinates of modes 1..L-1 (the output column).
__global__ __launch_bounds__(kThreads) void k_rest_keys(const int32_t* __restrict__ idx, int64_t nnz,
                                                        Ttm0Args a, uint64_t* __restrict__ key,
                                                        int32_t* __restrict__ ent) {
  for (int64_t e = blockIdx.x * (int64_t)kThreads + threadIdx.x; e < nnz; e += (int64_t)gridDim.x * kThreads) {
    uint64_t rest = 0;
    for (int l = 1; l < a.L; ++l) rest += (uint64_t)idx[(int64_t)l * nnz + e] * (uint64_t)a.stride[l];
    key[e] = rest;
    ent[e] = (int32_t)e;
  }
}

// Flat index g = p * k + s over (rest-sorted position, output row): the position that starts a
// run of equal rest sums d_e U[c_0e, s] over the run in sorted order into Y[s, rest].
__global__ __launch_bounds__(kThreads) void k_sparse_ttm0(const uint64_t* __restrict__ key,
                                                          const int32_t* __restrict__ ent,
                                                          const int32_t* __restrict__ idx0,
                                                          const double* __restrict__ data, int64_t nnz,
                                                          const double* __restrict__ U, Ttm0Args a,
                                                          double* __restrict__ Y) {
  const int64_t total = nnz * a.k;
  for (int64_t g = blockIdx.x * (int64_t)kThreads + threadIdx.x; g < total; g += (int64_t)gridDim.x * kThreads) {
    const int64_t p = g / a.k;
    const int s = (int)(g - p * a.k);
    const uint64_t k = key[p];
    if (p > 0 && key[p - 1] == k) continue;
    double acc = 0.0;
    for (int64_t t = p; t < nnz && key[t] == k; ++t) {
      const int32_t e = ent[t];
      acc += data[e] * U[(int64_t)s * a.I0 + idx0[e]];
    }
    Y[(int64_t)s * a.C + (int64_t)k] = acc;
  }
}

struct ResArgs {
  int L, R;
  int64_t dim[TFX_MAXL];
  int64_t off[TFX_MAXL];      // offset of factor l in the flat layout (column-major I_l x R)
};

__global__ __launch_bounds__(kThreads) void k_sparse_residual(const int32_t* __restrict__ idx,
                                                              const double* __restrict__ data, int64_t nnz,
                                                              const double* __restrict__ F, ResArgs a,
                                                              double* __restrict__ part) {
  __shared__ double red[kThreads / 64];
  double acc = 0.0;
  for (int64_t e = blockIdx.x * (int64_t)kThreads + threadIdx.x; e < nnz; e += (int64_t)gridDim.x * kThreads) {
    const double* base[TFX_MAXL];
    for (int l = 0; l < a.L; ++l) base[l] = F + a.off[l] + idx[(int64_t)l * nnz + e];
    double model = 0.0;
    for (int r = 0; r < a.R; ++r) {
      double p = base[0][(int64_t)r * a.dim[0]];
      for (int l = 1; l < a.L; ++l) p *= base[l][(int64_t)r * a.dim[l]];
      model += p;
    }
    const double d = data[e] - model;
    acc += d * d;
  }
  acc = block_sum(acc, red);
  if (threadIdx.x == 0) part[blockIdx.x] = acc;
}

__global__ __launch_bounds__(kThreads) void k_sum_part(const double* __restrict__ part, int n,
                                                       double* __restrict__ out) {
  __shared__ double red[kThreads / 64];
  double acc = 0.0;
  for (int i = threadIdx.x; i < n; i += kThreads) acc += part[i];
  acc = block_sum(acc, red);
  if (threadIdx.x == 0) out[0] = acc;
}

int check_dims(tfx_ctx* ctx, int L, const int64_t* dims) {
  if (L < 2 || L > TFX_MAXL) return tfx_set_error(ctx, TFX_E_ARG, "order %d outside [2, %d]", L, TFX_MAXL);
  for (int l = 0; l < L; ++l)
    if (dims[l] < 1 || dims[l] > INT32_MAX)
      return tfx_set_error(ctx, TFX_E_ARG, "dimension %lld of mode %d outside [1, 2^31)", (long long)dims[l], l);
  return TFX_OK;
}

}  // namespace

extern "C" int tfx_sparse_pack(tfx_ctx* ctx, int64_t nnz, int L, const uint64_t* coords, const int32_t* ordering,
                               const int64_t* sorted_dims, int32_t* idx, int* bad, hipStream_t stream) {
  if (!ctx || !ordering || !sorted_dims || !bad || (nnz > 0 && (!coords || !idx)))
    return tfx_set_error(ctx, TFX_E_ARG, "null argument");
  int rc = check_dims(ctx, L, sorted_dims);
  if (rc) return rc;
  if (nnz < 0 || nnz > INT32_MAX) return tfx_set_error(ctx, TFX_E_ARG, "nnz %lld outside [0, 2^31)", (long long)nnz);
  PackArgs a{};
  a.L = L;
  for (int l = 0; l < L; ++l) {
    if (ordering[l] < 0 || ordering[l] >= L) return tfx_set_error(ctx, TFX_E_ARG, "bad mode ordering");
    a.src[l] = ordering[l];
    a.dim[l] = sorted_dims[l];
  }
  hipLaunchKernelGGL(k_zero_int, dim3(1), dim3(64), 0, stream, bad);
  TFX_CHECK_LAUNCH(ctx);
  if (nnz == 0) return TFX_OK;
  hipLaunchKernelGGL(k_sparse_pack, dim3(grid_for(ctx, nnz)), dim3(kThreads), 0, stream, coords, nnz, a, idx, bad);
  TFX_CHECK_LAUNCH(ctx);
  return TFX_OK;
}

extern "C" int tfx_sparse_gram(tfx_ctx* ctx, int64_t nnz, int L, const int64_t* dims, const int32_t* idx,
                               const double* data, int mode, double* G, hipStream_t stream) {
  if (!ctx || !dims || !G || (nnz > 0 && (!idx || !data))) return tfx_set_error(ctx, TFX_E_ARG, "null argument");
  int rc = check_dims(ctx, L, dims);
  if (rc) return rc;
  if (mode < 0 || mode >= L) return tfx_set_error(ctx, TFX_E_ARG, "mode %d outside [0, %d)", mode, L);
  if (nnz < 0 || nnz > INT32_MAX) return tfx_set_error(ctx, TFX_E_ARG, "nnz %lld outside [0, 2^31)", (long long)nnz);
  const int64_t I = dims[mode];
  if (I > 65536) return tfx_set_error(ctx, TFX_E_UNSUPPORTED, "Gram of a mode of size %lld", (long long)I);
  hipLaunchKernelGGL(k_fill_zero, dim3(grid_for(ctx, I * I)), dim3(kThreads), 0, stream, G, I * I);
  TFX_CHECK_LAUNCH(ctx);
  if (nnz == 0) return TFX_OK;
  KeyArgs ka{};
  ka.L = L;
  ka.mode = mode;
  // key = linear index of the other coordinates (the unfolding's column up to a permutation)
  double span = 1.0;
  uint64_t s = 1;
  for (int l = 0; l < L; ++l) {
    if (l == mode) continue;
    ka.stride[l] = (int64_t)s;
    span *= (double)dims[l];
    s *= (uint64_t)dims[l];
  }
  if (span > 1.8e19) return tfx_set_error(ctx, TFX_E_UNSUPPORTED, "column space of mode %d exceeds 64-bit keys", mode);
  unsigned end_bit = 1;
  while (end_bit < 64 && (uint64_t(1) << end_bit) < s) ++end_bit;

  unsigned row_bits = 1;
  while (row_bits < 32 && (int64_t(1) << row_bits) < I) ++row_bits;
  size_t temp = 0, temp2 = 0;
  TFX_HIP(ctx, rocprim::radix_sort_pairs(nullptr, temp, (uint64_t*)nullptr, (uint64_t*)nullptr,
                                         (int32_t*)nullptr, (int32_t*)nullptr, (size_t)nnz, 0u, end_bit, stream));
  TFX_HIP(ctx, rocprim::radix_sort_pairs(nullptr, temp2, (uint32_t*)nullptr, (uint32_t*)nullptr,
                                         (int32_t*)nullptr, (int32_t*)nullptr, (size_t)nnz, 0u, row_bits, stream));
  if (temp2 > temp) temp = temp2;
  const size_t kb = align256(sizeof(uint64_t) * nnz), vb = align256(sizeof(int32_t) * nnz);
  void* wsp;
  rc = tfx_workspace(ctx, 2 * kb + 6 * vb + align256(temp), stream, &wsp);
  if (rc) return rc;
  char* w = (char*)wsp;
  uint64_t* k_in = (uint64_t*)w;
  uint64_t* k_out = (uint64_t*)(w + kb);
  int32_t* v_in = (int32_t*)(w + 2 * kb);
  int32_t* v_out = (int32_t*)(w + 2 * kb + vb);
  uint32_t* rk_in = (uint32_t*)(w + 2 * kb + 2 * vb);
  uint32_t* rk_out = (uint32_t*)(w + 2 * kb + 3 * vb);
  int32_t* rp_in = (int32_t*)(w + 2 * kb + 4 * vb);
  int32_t* rp_out = (int32_t*)(w + 2 * kb + 5 * vb);
  void* tmp = w + 2 * kb + 6 * vb;
  const int32_t* rowc = idx + (int64_t)mode * nnz;
  const int nb = grid_for(ctx, nnz);
  hipLaunchKernelGGL(k_col_keys, dim3(nb), dim3(kThreads), 0, stream, idx, nnz, ka, k_in, v_in);
  TFX_CHECK_LAUNCH(ctx);
  TFX_HIP(ctx, rocprim::radix_sort_pairs(tmp, temp, k_in, k_out, v_in, v_out, (size_t)nnz, 0u, end_bit, stream));
  hipLaunchKernelGGL(k_row_keys, dim3(nb), dim3(kThreads), 0, stream, v_out, rowc, nnz, rk_in, rp_in);
  TFX_CHECK_LAUNCH(ctx);
  TFX_HIP(ctx, rocprim::radix_sort_pairs(tmp, temp, rk_in, rk_out, rp_in, rp_out, (size_t)nnz, 0u, row_bits, stream));
  hipLaunchKernelGGL(k_gram_rows, dim3(grid_for(ctx, I)), dim3(kThreads), 0, stream, k_out, v_out, rk_out,
                     rp_out, rowc, data, nnz, I, G);
  TFX_CHECK_LAUNCH(ctx);
  hipLaunchKernelGGL(k_gram_mirror, dim3(grid_for(ctx, I * I)), dim3(kThreads), 0, stream, I, G);
  TFX_CHECK_LAUNCH(ctx);
  return TFX_OK;
}

extern "C" int tfx_sparse_ttm0(tfx_ctx* ctx, int64_t nnz, int L, const int64_t* dims, const int32_t* idx,
                               const double* data, const double* U, int k, double* Y, hipStream_t stream) {
  if (!ctx || !dims || !U || !Y || (nnz > 0 && (!idx || !data))) return tfx_set_error(ctx, TFX_E_ARG, "null argument");
  int rc = check_dims(ctx, L, dims);
  if (rc) return rc;
  if (k < 1 || k > dims[0]) return tfx_set_error(ctx, TFX_E_ARG, "k=%d outside [1, %lld]", k, (long long)dims[0]);
  if (nnz < 0 || nnz > INT32_MAX) return tfx_set_error(ctx, TFX_E_ARG, "nnz %lld outside [0, 2^31)", (long long)nnz);
  Ttm0Args a{};
  a.L = L;
  a.k = k;
  a.I0 = dims[0];
  int64_t C = 1;
  for (int l = L - 1; l >= 1; --l) {
    a.stride[l] = C;
    C *= dims[l];
  }
  a.C = C;
  if ((double)C * (double)k > 8.0e9) return tfx_set_error(ctx, TFX_E_UNSUPPORTED, "first-mode product of %g entries", (double)C * k);
  hipLaunchKernelGGL(k_fill_zero, dim3(grid_for(ctx, C * k)), dim3(kThreads), 0, stream, Y, C * k);
  TFX_CHECK_LAUNCH(ctx);
  if (nnz == 0) return TFX_OK;
  // entries sorted (stably) by output column; one sum per (column run, output row)
  unsigned end_bit = 1;
  while (end_bit < 64 && (uint64_t(1) << end_bit) < (uint64_t)C) ++end_bit;
  size_t temp = 0;
  TFX_HIP(ctx, rocprim::radix_sort_pairs(nullptr, temp, (uint64_t*)nullptr, (uint64_t*)nullptr,
                                         (int32_t*)nullptr, (int32_t*)nullptr, (size_t)nnz, 0u, end_bit, stream));
  const size_t kb = align256(sizeof(uint64_t) * nnz), vb = align256(sizeof(int32_t) * nnz);
  void* wsp;
  rc = tfx_workspace(ctx, 2 * kb + 2 * vb + align256(temp), stream, &wsp);
  if (rc) return rc;
  char* w = (char*)wsp;
  uint64_t* k_in = (uint64_t*)w;
  uint64_t* k_out = (uint64_t*)(w + kb);
  int32_t* v_in = (int32_t*)(w + 2 * kb);
  int32_t* v_out = (int32_t*)(w + 2 * kb + vb);
  void* tmp = w + 2 * kb + 2 * vb;
  const int nb = grid_for(ctx, nnz);
  hipLaunchKernelGGL(k_rest_keys, dim3(nb), dim3(kThreads), 0, stream, idx, nnz, a, k_in, v_in);
  TFX_CHECK_LAUNCH(ctx);
  TFX_HIP(ctx, rocprim::radix_sort_pairs(tmp, temp, k_in, k_out, v_in, v_out, (size_t)nnz, 0u, end_bit, stream));
  hipLaunchKernelGGL(k_sparse_ttm0, dim3(grid_for(ctx, nnz * k)), dim3(kThreads), 0, stream, k_out, v_out, idx,
                     data, nnz, U, a, Y);
  TFX_CHECK_LAUNCH(ctx);
  return TFX_OK;
}

extern "C" int tfx_sparse_residual(tfx_ctx* ctx, int64_t nnz, int L, const int64_t* dims, const int32_t* idx,
                                   const double* data, const double* F, int R, double* out, hipStream_t stream) {
  if (!ctx || !dims || !F || !out || (nnz > 0 && (!idx || !data))) return tfx_set_error(ctx, TFX_E_ARG, "null argument");
  int rc = check_dims(ctx, L, dims);
  if (rc) return rc;
  if (R < 1) return tfx_set_error(ctx, TFX_E_ARG, "rank %d < 1", R);
  if (nnz < 0 || nnz > INT32_MAX) return tfx_set_error(ctx, TFX_E_ARG, "nnz %lld outside [0, 2^31)", (long long)nnz);
  ResArgs a{};
  a.L = L;
  a.R = R;
  int64_t off = 0;
  for (int l = 0; l < L; ++l) {
    a.dim[l] = dims[l];
    a.off[l] = off;
    off += (int64_t)R * dims[l];
  }
  const int nb = grid_for(ctx, nnz, 4);
  void* wsp;
  rc = tfx_workspace(ctx, sizeof(double) * nb + 256, stream, &wsp);
  if (rc) return rc;
  double* part = (double*)wsp;
  hipLaunchKernelGGL(k_sparse_residual, dim3(nb), dim3(kThreads), 0, stream, idx, data, nnz, F, a, part);
  TFX_CHECK_LAUNCH(ctx);
  hipLaunchKernelGGL(k_sum_part, dim3(1), dim3(kThreads), 0, stream, part, nb, out);
  TFX_CHECK_LAUNCH(ctx);
  return TFX_OK;
}
