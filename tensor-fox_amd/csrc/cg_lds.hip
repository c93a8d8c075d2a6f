// Single-problem CG with every operand in one workgroup's LDS and the matvec on FP64 MFMA
// (GaussNewton.py:315-398, 551-581).  Used by tfx_cg for FP64 problems that fit: the compressed
// cores of the dGN main stage (e.g. 24^4 at rank 24: 137.5 KB of LDS).
#include "lds_cg_body.h"

namespace {

constexpr int NW = 8;       // waves
constexpr int JZMAX = 2;    // z jobs (16 x 16 output tiles) per wave
constexpr int JWMAX = 1;    // (c, r) tiles of tmp per wave

// SHORT: 8 when every MFMA chain has at most 8 steps (rows per mode and R at most 32): the z
// phase then runs its chains as two unrolled chunks (the tmp phase keeps its loop: unrolling its
// L chains at once exceeds the register budget); 0 for the general chunked chains
template <int SHORT>
__global__ __launch_bounds__(NW * 64) void k_cg_lds(ldscg::Geom g, const double* __restrict__ F,
                                                    const double* __restrict__ P1,
                                                    const double* __restrict__ P2,
                                                    const double* __restrict__ gamma, double damp,
                                                    const double* __restrict__ grad, int maxiter,
                                                    double tol, double* __restrict__ y,
                                                    double* __restrict__ jtjg,
                                                    double* __restrict__ scal) {
  extern __shared__ __align__(16) double lds_cg_mem[];
  __shared__ double red[2][NW];
  const ldscg::Lds s = ldscg::carve(g, lds_cg_mem);
  ldscg::zero_lds<NW * 64>(g, s);
  __syncthreads();
  ldscg::stage<NW * 64>(g, s, F, grad, P1, P2);
  __syncthreads();
  int itn;
  double rn;
  ldscg::cg_body<NW, JZMAX, JWMAX, 0, SHORT>(g, s, gamma, damp, grad, maxiter, tol, y, jtjg,
                                             red, &itn, &rn);
  if (threadIdx.x == 0) {
    scal[0] = (double)itn;
    scal[1] = rn;
  }
}

}  // namespace

// Launches the LDS CG when the problem fits (returns 1), returns 0 when it does not, or an
// error code.
int cg_lds_try(tfx_ctx* ctx, const Shape& s, const double* F, const double* P1, const double* P2,
               const double* gamma, double damp, const double* grad, int maxiter, double tol,
               double* y, double* jtjg, double* scal, hipStream_t st) {
  if (s.L < 2 || s.L > ldscg::MAXL || s.R < 1) return 0;
  int I[ldscg::MAXL];
  for (int l = 0; l < s.L; ++l) {
    if (s.d[l] > 4096) return 0;
    I[l] = (int)s.d[l];
  }
  ldscg::Geom g;
  ldscg::make_geom(s.L, s.R, I, &g);
  if (g.nz > NW * JZMAX || g.nrt * g.nrt > NW * JWMAX) return 0;
  const size_t lds = sizeof(double) * (size_t)ldscg::lds_doubles(g);
  if (lds > 150 * 1024) return 0;
  int maxrows = 0;
  for (int l = 0; l < s.L; ++l) maxrows = std::max(maxrows, g.ip4[l]);
  if (maxrows <= 32 && g.R4 <= 32) {
    TFX_HIP(ctx, hipFuncSetAttribute((const void*)k_cg_lds<8>,
                                     hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    hipLaunchKernelGGL(k_cg_lds<8>, dim3(1), dim3(NW * 64), lds, st, g, F, P1, P2, gamma, damp,
                       grad, maxiter, tol, y, jtjg, scal);
  } else {
    TFX_HIP(ctx, hipFuncSetAttribute((const void*)k_cg_lds<0>,
                                     hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    hipLaunchKernelGGL(k_cg_lds<0>, dim3(1), dim3(NW * 64), lds, st, g, F, P1, P2, gamma, damp,
                       grad, maxiter, tol, y, jtjg, scal);
  }
  TFX_CHECK_LAUNCH(ctx);
  return 1;
}
