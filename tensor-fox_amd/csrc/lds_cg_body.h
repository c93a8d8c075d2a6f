// Staging and CG iteration of the LDS-resident FP64 MFMA CG (layouts in lds_cg.h).
//
// cg_body follows GaussNewton.py:315-398: J^T J grad first (for the dogleg), then the split
// preconditioned CG on (J^T J + damp Gamma) with M = 1/sqrt(gamma^2 + damp^2 gamma^2), starting
// from y = 0, r = M grad, p = r; the first direction update is p_1 = r_1 + beta_1 r_1 (the
// reference's aliasing); stops after maxiter iterations or when ||r||^2 <= tol; returns M y.
// Problems with n <= 100 unknowns run with tol = 0, larger ones with maxiter capped at n
// (GaussNewton.py:117-121 zeroes the tolerances of small problems; cg() never runs more than n).
//
// The CG vectors stay in registers: wave w owns z jobs w, w + NW, ..., and each lane holds the
// four elements of a job that its MFMA accumulator holds.
#pragma once

#include "lds_cg.h"

namespace ldscg {

// Zero Fs, qs, Ts, P1s; the caller synchronizes before staging.
template <int NT>
__device__ inline void zero_lds(const Geom& g, const Lds& s) {
  const int64_t nz = 2 * (int64_t)g.nrows * g.PR + 2 * (int64_t)g.L * g.R4 * g.PR;
  for (int64_t e = threadIdx.x; e < nz; e += NT) s.Fs[e] = 0.0;
}

// Fs <- F, qs <- q0 (flat layout), P1s <- P1 (L x R x R), P2s <- P2[a][b] for a < b
// (L x L x R x R).  Pointers may address global memory or LDS.
template <int NT>
__device__ inline void stage(const Geom& g, const Lds& s, const double* F, const double* q0,
                             const double* P1, const double* P2) {
  for (int64_t e = threadIdx.x; e < g.n; e += NT) {
    int l = 0;
    while (l + 1 < g.L && e >= g.off[l + 1]) ++l;
    const int64_t w = e - g.off[l];
    const int r = (int)(w / g.I[l]);
    const int i = (int)(w - (int64_t)r * g.I[l]);
    const int64_t a = (int64_t)(g.row[l] + i) * g.PR + r;
    s.Fs[a] = F[e];
    s.qs[a] = q0[e];
  }
  const int RR = g.R * g.R;
  for (int e = threadIdx.x; e < g.L * RR; e += NT) {
    const int l = e / RR, cr = e - l * RR, c = cr / g.R, r = cr - c * g.R;
    s.P1s[(int64_t)(l * g.R4 + c) * g.PR + r] = P1[e];
  }
  for (int e = threadIdx.x; e < g.npairs * RR; e += NT) {
    const int p = e / RR, cr = e - p * RR;
    int a = 0, b = 1;
    for (int q = 0; q < p; ++q) {   // p-th pair in (0,1), (0,2), .., (1,2), ..
      if (++b == g.L) {
        ++a;
        b = a + 1;
      }
    }
    s.P2s[e] = P2[(int64_t)(a * g.L + b) * RR + cr];
  }
}

__device__ inline double inv_m64(double gm, double d) {
  return 1.0 / sqrt(gm * gm + d * d * (gm * gm));
}

// Needs: Fs, qs (= grad), P1s, P2s staged; Ts zero.  Writes jtjg (if not null) and y (flat).
// Every thread of the workgroup calls it.  Returns the iteration count as the reference reports
// it (itn + 1) and the final ||r||^2.  SHORT_T / SHORT_Z: the chain bound of the tmp and z
// phases (see chain(); 0 = any length).
template <int NW, int JZMAX, int JWMAX, int SHORT_T = 0, int SHORT_Z = 0>
__device__ inline void cg_body(const Geom& g, const Lds& s, const double* gamma, double damp,
                               const double* grad, int maxiter, double tol, double* y,
                               double* jtjg, double (*red)[NW], int* itn_out, double* rn_out) {
  const int t = threadIdx.x, wave = t >> 6, lane = t & 63;
  const int li = lane & 15, lg = lane >> 4;
  int buf = 0;
  // job geometry and diagonal scalings
  int jl[JZMAX], ji0[JZMAX], jr0[JZMAX];
  bool jon[JZMAX];
  double Mv[JZMAX], Gv[JZMAX];
#pragma unroll
  for (int k = 0; k < JZMAX; ++k) {
    const int j = wave + NW * k;
    jon[k] = j < g.nz;
    jl[k] = 0;
    ji0[k] = 0;
    jr0[k] = 0;
    if (jon[k]) zjob(g, j, &jl[k], &ji0[k], &jr0[k]);
    const int r = jr0[k] + li;
    const bool okr = jon[k] && r < g.R;
    Gv[k] = okr ? gamma[jl[k] * g.R + r] : 0.0;
    Mv[k] = okr ? inv_m64(Gv[k], damp) : 0.0;
  }
  auto valid = [&](int k, int j) {
    return jon[k] && jr0[k] + li < g.R && ji0[k] + 4 * j + lg < g.I[jl[k]];
  };
  auto flat = [&](int k, int j) {
    return g.off[jl[k]] + (int64_t)(jr0[k] + li) * g.I[jl[k]] + ji0[k] + 4 * j + lg;
  };
  auto qaddr = [&](int k, int j) {
    return (int64_t)(g.row[jl[k]] + ji0[k] + 4 * j + lg) * g.PR + jr0[k] + li;
  };

  // J^T J grad
  tmp_phase<NW, JWMAX, SHORT_T>(g, s, wave, lane);
  __syncthreads();
  dbl4 z[JZMAX];
#pragma unroll
  for (int k = 0; k < JZMAX; ++k) {
    if (jon[k]) {
      z[k] = z_job<SHORT_Z>(g, s, jl[k], ji0[k], jr0[k], lane);
      if (jtjg) {
#pragma unroll
        for (int j = 0; j < 4; ++j)
          if (valid(k, j)) jtjg[flat(k, j)] = z[k][j];
      }
    }
  }
  if (g.n > 100) {
    if ((int64_t)maxiter > g.n) maxiter = (int)g.n;
  } else {
    tol = 0.0;
  }
  double rv[JZMAX][4], pv[JZMAX][4], yv[JZMAX][4];
  double acc = 0.0;
#pragma unroll
  for (int k = 0; k < JZMAX; ++k)
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const double gr = valid(k, j) ? grad[flat(k, j)] : 0.0;
      rv[k][j] = Mv[k] * gr;
      pv[k][j] = rv[k][j];
      yv[k][j] = 0.0;
      acc += rv[k][j] * rv[k][j];
    }
  double rn = wg_sum<NW>(acc, red, buf);   // its barrier also ends the z-phase reads of qs
  if (rn == 0.0) rn = 1e-6;
  int itn;
  for (itn = 0; itn < maxiter; ++itn) {
#pragma unroll
    for (int k = 0; k < JZMAX; ++k)
#pragma unroll
      for (int j = 0; j < 4; ++j)
        if (valid(k, j)) s.qs[qaddr(k, j)] = Mv[k] * pv[k][j];
    __syncthreads();
    tmp_phase<NW, JWMAX, SHORT_T>(g, s, wave, lane);
    __syncthreads();
    acc = 0.0;
#pragma unroll
    for (int k = 0; k < JZMAX; ++k) {
      if (jon[k]) {
        const dbl4 zp = z_job<SHORT_Z>(g, s, jl[k], ji0[k], jr0[k], lane);
#pragma unroll
        for (int j = 0; j < 4; ++j) {
          const double qe = Mv[k] * pv[k][j];
          z[k][j] = valid(k, j) ? Mv[k] * (zp[j] + damp * Gv[k] * qe) : 0.0;
          acc += pv[k][j] * z[k][j];
        }
      }
    }
    double den = wg_sum<NW>(acc, red, buf);
    if (den == 0.0) den = 1e-6;
    const double alpha = rn / den;
    acc = 0.0;
#pragma unroll
    for (int k = 0; k < JZMAX; ++k) {
      if (jon[k]) {
#pragma unroll
        for (int j = 0; j < 4; ++j) {
          yv[k][j] += alpha * pv[k][j];
          rv[k][j] -= alpha * z[k][j];
          acc += rv[k][j] * rv[k][j];
        }
      }
    }
    const double rn_new = wg_sum<NW>(acc, red, buf);
    const double beta = rn_new / rn;
    rn = rn_new;
#pragma unroll
    for (int k = 0; k < JZMAX; ++k)
#pragma unroll
      for (int j = 0; j < 4; ++j) pv[k][j] = rv[k][j] + beta * (itn == 0 ? rv[k][j] : pv[k][j]);
    if (rn <= tol) break;
  }
  if (itn == maxiter) itn = maxiter - 1;
#pragma unroll
  for (int k = 0; k < JZMAX; ++k)
#pragma unroll
    for (int j = 0; j < 4; ++j)
      if (valid(k, j)) y[flat(k, j)] = Mv[k] * yv[k][j];
  *itn_out = itn + 1;
  *rn_out = rn;
}

}  // namespace ldscg
