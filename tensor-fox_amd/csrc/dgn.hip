// Device side of one damped Gauss-Newton iteration (GaussNewton.py:254-627): Gram matrices and
// their Hadamard products, regularization, minus the gradient, the split-preconditioned CG with
// the Gram-Hadamard approximate-Hessian matvec, the point update with equalization, and the
// small vector kernels of the dogleg fallback.
//
// Flat layout (x, grad, y, factors): element (i, r) of mode l at off_l + r*I_l + i.
#include "common.h"

#include <type_traits>

// multi-kernel CG (cg_multi.hip) for problems whose operands do not fit one workgroup's LDS
template <typename S>
int cg_multi_t(tfx_ctx* ctx, const Shape& s, const S* F, const S* P1, const S* P2, const S* gamma,
               double damp, const S* grad, int maxiter, double tol, S* y, S* jtjg, double* scal,
               hipStream_t st);

// single-workgroup FP64 MFMA CG (cg_lds.hip): 1 if launched, 0 if the problem does not fit
int cg_lds_try(tfx_ctx* ctx, const Shape& s, const double* F, const double* P1, const double* P2,
               const double* gamma, double damp, const double* grad, int maxiter, double tol,
               double* y, double* jtjg, double* scal, hipStream_t st);

namespace {

// Mode of flat element e, with its offset within the mode and (optionally) the mode's size and
// offset.  Every comparison uses a constant index into the shape, so the shape is read with
// scalar loads at kernel entry; a per-element index would read the kernel arguments with
// dependent vector loads.
__device__ inline int mode_of(const Shape& s, int64_t e, int64_t* within, int64_t* dim = nullptr,
                              int64_t* offset = nullptr) {
  int l = 0;
  int64_t o = s.off[0], d = s.d[0];
#pragma unroll
  for (int q = 1; q < TFX_MAXL; ++q) {
    const bool in = q < s.L && e >= s.off[q];
    l = in ? q : l;
    o = in ? s.off[q] : o;
    d = in ? s.d[q] : d;
  }
  *within = e - o;
  if (dim) *dim = d;
  if (offset) *offset = o;
  return l;
}

// Gr[l][r][c] = sum_i F_l[i, r] F_l[i, c].  One workgroup per (l, r).
template <typename S>
__global__ __launch_bounds__(256) void k_gram_factors(Shape s, const S* __restrict__ F,
                                                       S* __restrict__ Gr) {
  __shared__ S red[4][64];
  const int l = blockIdx.x / s.R, r = blockIdx.x % s.R;
  const int64_t I = s.d[l];
  const S* Fl = F + s.off[l];
  const int lane = threadIdx.x & 63, grp = threadIdx.x >> 6;
  for (int c0 = 0; c0 < s.R; c0 += 64) {
    const int c = c0 + lane;
    S acc = S(0);
    if (c < s.R) {
      const S* fr = Fl + (int64_t)r * I;
      const S* fc = Fl + (int64_t)c * I;
      int64_t i = grp;
      for (; i + 12 < I; i += 16) {   // four rows per step, loads issued together
        const S a0 = fr[i], a1 = fr[i + 4], a2 = fr[i + 8], a3 = fr[i + 12];
        const S b0 = fc[i], b1 = fc[i + 4], b2 = fc[i + 8], b3 = fc[i + 12];
        acc += a0 * b0;
        acc += a1 * b1;
        acc += a2 * b2;
        acc += a3 * b3;
      }
      for (; i < I; i += 4) acc += fr[i] * fc[i];
    }
    __syncthreads();
    red[grp][lane] = acc;
    __syncthreads();
    if (grp == 0 && c < s.R)
      Gr[((int64_t)l * s.R + r) * s.R + c] = (red[0][lane] + red[1][lane]) + (red[2][lane] + red[3][lane]);
  }
}

// P1[l] = o_{k != l} Gr_k, P2[l][m] = o_{k != l, m} Gr_k (ones on l == m), and
// gamma[l, r] = max|P1_{L-1}| * sqrt|P1_l[r, r]| (GaussNewton.py:584-606 scales with the LAST
// mode's max).  One workgroup.
template <typename S>
__global__ __launch_bounds__(256) void k_gram_hadamard(Shape s, const S* __restrict__ Gr,
                                                        S* __restrict__ P1, S* __restrict__ P2,
                                                        S* __restrict__ gamma, S* __restrict__ pmax,
                                                        unsigned* __restrict__ counter) {
  // one thread per (l, entry): the L mode Grams of the entry are loaded first, then
  // P1[l] = prod_{k != l} Gr_k and P2[l][m] = prod_{k != l, m} Gr_k (mode order, as the reference
  // forms them); workgroup 0 also forms gamma from max |P1_{L-1}| (GaussNewton.py:584-606),
  // recomputing the products it needs from Gr
  __shared__ S red[4];
  const int L = s.L, R = s.R;
  const int64_t RR = (int64_t)R * R;
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  S lmx = S(0);   // max |P1_{L-1}| over this workgroup's entries
  if (e < (int64_t)L * RR) {
    const int l = (int)(e / RR);
    const int64_t rc = e - (int64_t)l * RR;
    S g[TFX_MAXL];
#pragma unroll
    for (int k = 0; k < TFX_MAXL; ++k) g[k] = Gr[(int64_t)(k < L ? k : 0) * RR + rc];
    S p = S(1);
#pragma unroll
    for (int k = 0; k < TFX_MAXL; ++k)
      if (k < L && k != l) p *= g[k];
    P1[e] = p;
    if (l == L - 1) lmx = fabs(p);
    for (int m = 0; m < L; ++m) {
      S q = S(1);
      if (m != l)
#pragma unroll
        for (int k = 0; k < TFX_MAXL; ++k)
          if (k < L && k != l && k != m) q *= g[k];
      P2[((int64_t)l * L + m) * RR + rc] = q;
    }
  }
  S mx = S(0);
  if (counter) {
    // every workgroup stores the max of the P1_{L-1} entries it formed; the last one to arrive
    // takes the max of those (a max is exact, so the value is the one max |P1_{L-1}| has)
    lmx = block_max(lmx, red);
    __shared__ int is_last;
    if (threadIdx.x == 0) {
      pmax[blockIdx.x] = lmx;
      __threadfence();
      is_last = atomicAdd(counter, 1u) == gridDim.x - 1;
    }
    __syncthreads();
    if (!is_last) return;
    __threadfence();
    for (int b = threadIdx.x; b < (int)gridDim.x; b += blockDim.x) {
      const S v = __builtin_nontemporal_load(pmax + b);
      mx = v > mx ? v : mx;
    }
    mx = block_max(mx, red);
  } else {
    if (blockIdx.x != 0) return;   // uniform: the rest is workgroup 0's
    for (int64_t rc = threadIdx.x; rc < RR; rc += blockDim.x) {
      S g[TFX_MAXL];
#pragma unroll
      for (int k = 0; k < TFX_MAXL; ++k) g[k] = Gr[(int64_t)(k < L - 1 ? k : 0) * RR + rc];
      S p = S(1);
#pragma unroll
      for (int k = 0; k < TFX_MAXL; ++k)
        if (k < L - 1) p *= g[k];
      const S v = fabs(p);
      mx = v > mx ? v : mx;
    }
    mx = block_max(mx, red);
  }
  for (int e2 = threadIdx.x; e2 < L * R; e2 += blockDim.x) {
    const int l = e2 / R, r = e2 % R;
    S p = S(1);
    for (int k = 0; k < L; ++k)
      if (k != l) p *= Gr[k * RR + (int64_t)r * R + r];
    gamma[e2] = mx * sqrt(fabs(p));
  }
  if (counter && threadIdx.x == 0) *counter = 0u;
}

// grad = N - F_l P1_l (minus the gradient), flat layout.
template <typename S>
__global__ void k_neg_grad(Shape s, const S* __restrict__ F, const S* __restrict__ P1,
                           const S* __restrict__ N, S* __restrict__ g) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= s.n) return;
  int64_t w;
  int64_t I, offl;
  const int l = mode_of(s, e, &w, &I, &offl);
  const int r = (int)(w / I);
  const int64_t i = w % I;
  const S* Fl = F + offl;
  const S* P = P1 + (int64_t)l * s.R * s.R;
  // four partial sums (independent loads and FMA chains), combined in a fixed order
  S a0 = S(0), a1 = S(0), a2 = S(0), a3 = S(0);
  int c = 0;
#pragma unroll 4   // four steps' loads in flight together; each accumulator's order unchanged
  for (; c + 4 <= s.R; c += 4) {
    a0 += Fl[(int64_t)c * I + i] * P[(int64_t)c * s.R + r];
    a1 += Fl[(int64_t)(c + 1) * I + i] * P[(int64_t)(c + 1) * s.R + r];
    a2 += Fl[(int64_t)(c + 2) * I + i] * P[(int64_t)(c + 2) * s.R + r];
    a3 += Fl[(int64_t)(c + 3) * I + i] * P[(int64_t)(c + 3) * s.R + r];
  }
  for (; c < s.R; ++c) a0 += Fl[(int64_t)c * I + i] * P[(int64_t)c * s.R + r];
  g[e] = N[e] - ((a0 + a1) + (a2 + a3));
}

// ---------------------------------------------------------------------------------------------
// CG for small problems (GaussNewton.py:315-398, 551-581): one workgroup, every vector held in
// registers as 4 x 4 tiles (rows i, ranks r) of one mode, F / q / tmp / P1 staged row-major in
// LDS, the matvec GEMMs register-tiled.  Larger problems use the multi-kernel CG (cg_multi.hip).
// ---------------------------------------------------------------------------------------------
constexpr int SMALL_NT = 512;

struct TileMap {
  int ntile;                 // total number of 4x4 tiles over all modes
  int first[TFX_MAXL + 1];   // first tile of each mode
  int rb;                    // tiles along R
};

__device__ inline void tile_of(const Shape& s, const TileMap& tm, int t, int* l, int* i0, int* r0) {
  int m = 0;
  while (m + 1 < s.L && t >= tm.first[m + 1]) ++m;
  const int w = t - tm.first[m];
  *l = m;
  *i0 = 4 * (w / tm.rb);
  *r0 = 4 * (w % tm.rb);
}

template <typename S>
__device__ inline S inv_m(S g, S d) {
  return S(1) / sqrt(g * g + d * d * (g * g));
}

// z-tile = F_l tmp_l + q_l P1_l for the tile (l, i0, r0); arrays row-major [i][r] per mode.
template <typename S>
__device__ inline void z_tile(const Shape& s, int l, int i0, int r0, const S* F, const S* q,
                              const S* tmp, const S* P1, S z[16]) {
  const int R = s.R;
  const int64_t I = s.d[l];
  const int64_t off = s.off[l];
  const int64_t RR = (int64_t)R * R;
  const S* Fl = F + off;
  const S* ql = q + off;
  const S* tl = tmp + l * RR;
  const S* pl = P1 + l * RR;
#pragma unroll
  for (int e = 0; e < 16; ++e) z[e] = S(0);
  for (int c = 0; c < R; ++c) {
    S f[4], qq[4], tv[4], pv[4];
#pragma unroll
    for (int a = 0; a < 4; ++a) {
      const bool ok = i0 + a < I;
      f[a] = ok ? Fl[(int64_t)(i0 + a) * R + c] : S(0);
      qq[a] = ok ? ql[(int64_t)(i0 + a) * R + c] : S(0);
      const bool okr = r0 + a < R;
      tv[a] = okr ? tl[(int64_t)c * R + r0 + a] : S(0);
      pv[a] = okr ? pl[(int64_t)c * R + r0 + a] : S(0);
    }
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
      for (int b = 0; b < 4; ++b) z[a * 4 + b] += f[a] * tv[b] + qq[a] * pv[b];
  }
}

// tmp_l = sum_{m != l} P2[l][m] o (q_m^T F_m), computed per 4 x 4 (c, r) tile: each mode's
// product is formed once and added to the tiles of the other modes straight away (the tile is
// owned by one thread, so the read-modify-write of tmp has no race).
template <typename S>
__device__ inline void tmp_phase(const Shape& s, const S* F, const S* q, const S* __restrict__ P2,
                                 S* tmp) {
  const int R = s.R, L = s.L;
  const int rb = (R + 3) / 4;
  const int64_t RR = (int64_t)R * R;
  for (int t = threadIdx.x; t < rb * rb; t += blockDim.x) {
    const int c0 = 4 * (t / rb), r0 = 4 * (t % rb);
    for (int m = 0; m < L; ++m) {
      const int64_t I = s.d[m];
      const S* Fm = F + s.off[m];
      const S* qm = q + s.off[m];
      S acc[16];
#pragma unroll
      for (int e = 0; e < 16; ++e) acc[e] = S(0);
      for (int64_t i = 0; i < I; ++i) {
        S qa[4], fb[4];
#pragma unroll
        for (int a = 0; a < 4; ++a) {
          qa[a] = c0 + a < R ? qm[i * R + c0 + a] : S(0);
          fb[a] = r0 + a < R ? Fm[i * R + r0 + a] : S(0);
        }
#pragma unroll
        for (int a = 0; a < 4; ++a)
#pragma unroll
          for (int b = 0; b < 4; ++b) acc[a * 4 + b] += qa[a] * fb[b];
      }
      for (int l = 0; l < L; ++l) {
#pragma unroll
        for (int a = 0; a < 4; ++a)
#pragma unroll
          for (int b = 0; b < 4; ++b) {
            const int c = c0 + a, r = r0 + b;
            if (c < R && r < R) {
              const int64_t te = l * RR + (int64_t)c * R + r;
              const S contrib = m == l ? S(0) : P2[((int64_t)l * L + m) * RR + (int64_t)c * R + r] * acc[a * 4 + b];
              tmp[te] = (m == 0 ? S(0) : tmp[te]) + contrib;
            }
          }
      }
    }
  }
}

template <typename S>
__global__ __launch_bounds__(SMALL_NT) void k_cg_small(
    Shape s, TileMap tm, const S* __restrict__ Fg, const S* __restrict__ P1g,
    const S* __restrict__ P2, const S* __restrict__ gamma, double damp,
    const S* __restrict__ grad, int maxiter, double tol, S* __restrict__ y, S* __restrict__ jtjg,
    double* __restrict__ scal, S* gws) {
  extern __shared__ __align__(16) unsigned char smem[];
  __shared__ double red[SMALL_NT / 64];
  const int L = s.L, R = s.R;
  const int64_t n = s.n;
  const int64_t LRR = (int64_t)L * R * R;
  // q, F, tmp, P1 in LDS (carved from the dynamic allocation, so that their accesses compile to
  // LDS instructions)
  S* q = (S*)smem;
  S* F = q + n;
  S* tmp = F + n;
  S* P1 = tmp + LRR;
  S* ycg = gws;           // CG iterate, flat layout, in global memory (fewer registers)
  const int t = threadIdx.x;
  // stage F (row-major per mode) and P1
  for (int64_t e = t; e < n; e += blockDim.x) {
    int64_t w;
    int64_t I, offl;
    const int l = mode_of(s, e, &w, &I, &offl);
    const int r = (int)(w / I);
    const int64_t i = w % I;
    F[offl + i * R + r] = Fg[e];
    q[offl + i * R + r] = grad[e];   // J^T J grad first (no preconditioner)
  }
  for (int64_t e = t; e < LRR; e += blockDim.x) P1[e] = P1g[e];
  __syncthreads();

  const bool own = t < tm.ntile;
  int tl = 0, ti0 = 0, tr0 = 0;
  if (own) tile_of(s, tm, t, &tl, &ti0, &tr0);
  const int64_t I = s.d[tl];
  const S dampS = (S)damp;
  S Mv[4], Gv[4];
#pragma unroll
  for (int b = 0; b < 4; ++b) {
    const bool ok = own && tr0 + b < R;
    Gv[b] = ok ? gamma[tl * R + tr0 + b] : S(0);
    Mv[b] = ok ? inv_m(Gv[b], dampS) : S(0);
  }
  auto valid = [&](int a, int b) { return own && ti0 + a < I && tr0 + b < R; };
  auto flat = [&](int a, int b) { return s.off[tl] + (int64_t)(tr0 + b) * I + ti0 + a; };

  // J^T J grad (dogleg) -> jtjg
  tmp_phase<S>(s, F, q, P2, tmp);
  __syncthreads();
  S z[16];
  if (own) {
    z_tile<S>(s, tl, ti0, tr0, F, q, tmp, P1, z);
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
      for (int b = 0; b < 4; ++b)
        if (valid(a, b)) jtjg[flat(a, b)] = z[a * 4 + b];
  }
  if (n > 100) {
    if ((int64_t)maxiter > n) maxiter = (int)n;
  } else {
    tol = 0.0;
  }
  S rv[16], pv[16];
  double acc = 0.0;
#pragma unroll
  for (int a = 0; a < 4; ++a)
#pragma unroll
    for (int b = 0; b < 4; ++b) {
      const int e = a * 4 + b;
      const S g = valid(a, b) ? grad[flat(a, b)] : S(0);
      rv[e] = Mv[b] * g;
      pv[e] = rv[e];
      if (valid(a, b)) ycg[flat(a, b)] = S(0);
      acc += (double)rv[e] * (double)rv[e];
    }
  double rn = block_sum(acc, red);
  if (rn == 0.0) rn = 1e-6;
  int itn;
  for (itn = 0; itn < maxiter; ++itn) {
    __syncthreads();  // previous readers of q / tmp are done
    if (own) {
#pragma unroll
      for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int b = 0; b < 4; ++b)
          if (valid(a, b)) q[s.off[tl] + (int64_t)(ti0 + a) * R + tr0 + b] = Mv[b] * pv[a * 4 + b];
    }
    __syncthreads();
    tmp_phase<S>(s, F, q, P2, tmp);
    __syncthreads();
    acc = 0.0;
    if (own) {
      z_tile<S>(s, tl, ti0, tr0, F, q, tmp, P1, z);
#pragma unroll
      for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int b = 0; b < 4; ++b) {
          const int e = a * 4 + b;
          const S qe = Mv[b] * pv[e];
          z[e] = valid(a, b) ? Mv[b] * (z[e] + dampS * Gv[b] * qe) : S(0);
          acc += (double)pv[e] * (double)z[e];
        }
    }
    double den = block_sum(acc, red);
    if (den == 0.0) den = 1e-6;
    const double alpha = rn / den;
    const S aS = (S)alpha;
    acc = 0.0;
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
      for (int b = 0; b < 4; ++b) {
        const int e = a * 4 + b;
        if (valid(a, b)) ycg[flat(a, b)] += aS * pv[e];
        rv[e] -= aS * z[e];
        acc += (double)rv[e] * (double)rv[e];
      }
    const double rn_new = block_sum(acc, red);
    const double beta = rn_new / rn;
    rn = rn_new;
    const S bS = (S)beta;
#pragma unroll
    for (int e = 0; e < 16; ++e) pv[e] = rv[e] + bS * (itn == 0 ? rv[e] : pv[e]);
    if (rn <= tol) break;
  }
  if (itn == maxiter) itn = maxiter - 1;
  if (own) {
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
      for (int b = 0; b < 4; ++b)
        if (valid(a, b)) y[flat(a, b)] = Mv[b] * ycg[flat(a, b)];
  }
  if (t == 0) {
    scal[0] = (double)(itn + 1);
    scal[1] = rn;
  }
}

// x_new = x + alpha y; F = x2cpd(x_new) (+ equalize) then transform; gain-ratio scalars.
template <typename S>
__global__ __launch_bounds__(1024) void k_dgn_update(Shape s, const S* __restrict__ x,
                                                      const S* __restrict__ y, double alpha,
                                                      const S* __restrict__ g, double damp, int eq,
                                                      int symm, double factors_norm,
                                                      S* __restrict__ xn, S* __restrict__ F,
                                                      double* __restrict__ scal) {
  extern __shared__ __align__(16) unsigned char smem[];
  double* cn = (double*)smem;   // L x R column norms, then L x R equalization scales
  __shared__ double red[16];
  __shared__ double fnorm[TFX_MAXL];
  const int64_t n = s.n;
  const int L = s.L, R = s.R;
  const int t = threadIdx.x, nt = blockDim.x;
  const int lane = t & 63, wave = t >> 6, nwave = nt >> 6;
  double gain = 0.0, dy2 = 0.0, x2 = 0.0, gmax = 0.0;
  // one wave per UB consecutive factor columns (l, r) at a time: the update, the step statistics
  // and the columns' norms; the loads of the UB columns are in flight together
  constexpr int UB = 4;
  for (int cb = wave * UB; cb < L * R; cb += nwave * UB) {
    int64_t base[UB], Ic[UB];
    int64_t Imax = 0;
#pragma unroll
    for (int k = 0; k < UB; ++k) {
      const int c = cb + k < L * R ? cb + k : L * R - 1;
      const int l = c / R, r = c - (c / R) * R;
      Ic[k] = cb + k < L * R ? s.d[l] : 0;
      base[k] = s.off[l] + (int64_t)r * s.d[l];
      Imax = Ic[k] > Imax ? Ic[k] : Imax;
    }
    double ss[UB];
#pragma unroll
    for (int k = 0; k < UB; ++k) ss[k] = 0.0;
    for (int64_t i0 = 0; i0 < Imax; i0 += 64) {
      S xv[UB], yv[UB], gv[UB];
#pragma unroll
      for (int k = 0; k < UB; ++k) {
        const int64_t e = i0 + lane < Ic[k] ? base[k] + i0 + lane : 0;
        xv[k] = x[e];
        yv[k] = y[e];
        gv[k] = g[e];
      }
#pragma unroll
      for (int k = 0; k < UB; ++k) {
        if (i0 + lane < Ic[k]) {
          const int64_t e = base[k] + i0 + lane;
          const double xo = (double)xv[k];
          const S v = xv[k] + (S)alpha * yv[k];
          xn[e] = v;
          F[e] = v;
          const double dy = (double)(v - xv[k]);   // x_new - x_old as the reference forms it
          gain += dy * ((double)gv[k] + damp * dy);
          dy2 += dy * dy;
          x2 += xo * xo;
          const double ag = fabs((double)gv[k]);
          gmax = ag > gmax ? ag : gmax;
          ss[k] += (double)v * (double)v;
        }
      }
    }
    if (eq) {
#pragma unroll
      for (int k = 0; k < UB; ++k) {
        const double a = wave_sum(ss[k]);
        if (lane == 0 && cb + k < L * R) cn[cb + k] = sqrt(a);
      }
    }
  }
  gain = block_sum(gain, red);
  dy2 = block_sum(dy2, red);
  x2 = block_sum(x2, red);
  gmax = block_max(gmax, red);
  if (t == 0) {
    scal[0] = gain;
    scal[1] = dy2;
    scal[2] = x2;
    scal[3] = gmax;
  }
  __syncthreads();
  if (eq && symm != 2) {
    // column r of mode l scaled by (prod_k ||f_k^r||)^(1/L) / ||f_l^r|| (Conversion.py:361-387);
    // a column whose norm product is zero is left as it is (scale 0 marks it)
    double* sc = cn + L * R;
    for (int c = t; c < L * R; c += nt) {
      const int l = c / R, r = c - (c / R) * R;
      double prod = 1.0;
      for (int k = 0; k < L; ++k) prod *= cn[k * R + r];
      sc[c] = prod != 0.0 ? pow(prod, 1.0 / L) / cn[l * R + r] : 0.0;
    }
    __syncthreads();
    for (int cb = wave * UB; cb < L * R; cb += nwave * UB) {
      int64_t base[UB], Ic[UB];
      double f[UB];
      int64_t Imax = 0;
#pragma unroll
      for (int k = 0; k < UB; ++k) {
        const int c = cb + k < L * R ? cb + k : L * R - 1;
        const int l = c / R, r = c - (c / R) * R;
        f[k] = sc[c];
        Ic[k] = (cb + k < L * R && f[k] != 0.0) ? s.d[l] : 0;
        base[k] = s.off[l] + (int64_t)r * s.d[l];
        Imax = Ic[k] > Imax ? Ic[k] : Imax;
      }
      for (int64_t i0 = 0; i0 < Imax; i0 += 64) {
        S fv[UB];
#pragma unroll
        for (int k = 0; k < UB; ++k) fv[k] = F[i0 + lane < Ic[k] ? base[k] + i0 + lane : 0];
#pragma unroll
        for (int k = 0; k < UB; ++k)
          if (i0 + lane < Ic[k]) F[base[k] + i0 + lane] = (S)(f[k] * (double)fv[k]);
      }
    }
    __syncthreads();
  }
  if (symm == 1) {
    // transform on separate factor arrays (Conversion.py:432-437): s = F_0 accumulates the other
    // modes left to right, F_0 = s / L, every mode takes F_0
    const int64_t I = s.d[0];
    for (int64_t e = t; e < (int64_t)R * I; e += nt) {
      S acc = F[e];
      for (int l = 1; l < L; ++l) acc += F[s.off[l] + e];
      acc = acc / (S)L;
      for (int l = 0; l < L; ++l) F[s.off[l] + e] = acc;
    }
    __syncthreads();
  } else if (symm == 2) {
    // the reference's factor list after one symmetric transform holds L references to a single
    // array (Conversion.py:437-438).  Then x2cpd writes every mode's block into that array in
    // turn, so only the last mode's block remains (Conversion.py:39-44); equalize sees L equal
    // column norms n and scales the column L times by (n^L)^(1/L) / n (Conversion.py:378-385);
    // transform's in-place `s += factors[l]` doubles the array L - 1 times before the division
    // by L (Conversion.py:433-436).  The result is copied to every mode.
    const int64_t I = s.d[0];
    const int64_t last = s.off[L - 1];
    double* sc = cn + L * R;
    if (eq) {
      for (int r = t; r < R; r += nt) {
        const double nr = cn[(L - 1) * R + r];
        double prod = 1.0;
        for (int k = 0; k < L; ++k) prod *= nr;
        sc[r] = prod != 0.0 ? pow(prod, 1.0 / L) / nr : 0.0;
      }
      __syncthreads();
    }
    const double dbl = (double)(1 << (L - 1));
    for (int64_t e = t; e < (int64_t)R * I; e += nt) {
      S v = F[last + e];
      if (eq) {
        const double f = sc[e / I];
        if (f != 0.0)
          for (int k = 0; k < L; ++k) v = (S)(f * (double)v);
      }
      v = (S)((double)v * dbl) / (S)L;
      for (int l = 0; l < L; ++l) F[s.off[l] + e] = v;
    }
    __syncthreads();
  }
  if (factors_norm > 0) {
    for (int l = 0; l < L; ++l) {
      const int64_t len = (int64_t)R * s.d[l];
      double a = 0.0;
      for (int64_t e = t; e < len; e += nt) a += (double)F[s.off[l] + e] * (double)F[s.off[l] + e];
      a = block_sum(a, red);
      if (t == 0) fnorm[l] = sqrt(a);
      __syncthreads();
    }
    for (int64_t e = t; e < n; e += nt) {
      int64_t w;
      const int l = mode_of(s, e, &w);
      F[e] = (S)(factors_norm * (1.0 / fnorm[l]) * (double)F[e]);
    }
  }
}


// k_dgn_update with one workgroup per rank r (factors_norm == 0): every step of the update --
// x_new = x + alpha y, F = x2cpd(x_new), the column norms, equalize (Conversion.py:361-387) and
// the symmetric transforms (Conversion.py:412-444) -- mixes only the L columns (l, r) of one
// rank, so a workgroup does all of it for its rank without exchanging data with the others.
// The four step scalars are summed over the ranks by the last workgroup to finish (an atomic
// counter in the context's scratch, reset by that workgroup; no workgroup waits for another),
// in rank order, so the result does not depend on scheduling.  Columns of up to 4 x 256 rows
// keep their new values in registers for the scaling; longer ones re-read F.
typedef double f64x2 __attribute__((ext_vector_type(2)));
constexpr int UPD_NT = 256;
constexpr int UPD_VC = 4;

template <typename S>
__global__ __launch_bounds__(UPD_NT) void k_dgn_update_rk(
    Shape s, const S* __restrict__ x, const S* __restrict__ y, double alpha,
    const S* __restrict__ g, double damp, int eq, int symm, S* __restrict__ xn,
    S* __restrict__ F, double* __restrict__ scal, unsigned* __restrict__ counter,
    double* __restrict__ part) {
  const int r = blockIdx.x;
  const int L = s.L, R = s.R;
  const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
  __shared__ double red[UPD_NT / 64][TFX_MAXL + 4];
  __shared__ double cn[TFX_MAXL], sc[TFX_MAXL];
  __shared__ int is_last;
  int64_t maxd = 0;
#pragma unroll
  for (int l = 0; l < TFX_MAXL; ++l)   // constant indices: the shape is read at kernel entry
    if (l < L) maxd = s.d[l] > maxd ? s.d[l] : maxd;
  const bool cached = maxd <= (int64_t)UPD_NT * UPD_VC;
  double gain = 0.0, dy2 = 0.0, x2 = 0.0, gmax = 0.0;
  double ss[TFX_MAXL];
  S vc[TFX_MAXL][UPD_VC];
#pragma unroll
  for (int l = 0; l < TFX_MAXL; ++l) {
    ss[l] = 0.0;
#pragma unroll
    for (int u = 0; u < UPD_VC; ++u) vc[l][u] = S(0);
  }
#pragma unroll
  for (int l = 0; l < TFX_MAXL; ++l) {
    if (l < L) {
      const int64_t I = s.d[l], base = s.off[l] + (int64_t)r * I;
      for (int64_t i0 = 0; i0 < I; i0 += (int64_t)UPD_NT * UPD_VC) {
        S xv[UPD_VC], yv[UPD_VC], gv[UPD_VC];
#pragma unroll
        for (int u = 0; u < UPD_VC; ++u) {
          const int64_t i = i0 + t + UPD_NT * u;
          const int64_t e = i < I ? base + i : base;
          xv[u] = x[e];
          yv[u] = y[e];
          gv[u] = g[e];
        }
#pragma unroll
        for (int u = 0; u < UPD_VC; ++u) {
          const int64_t i = i0 + t + UPD_NT * u;
          if (i < I) {
            const int64_t e = base + i;
            const double xo = (double)xv[u];
            const S v = xv[u] + (S)alpha * yv[u];
            xn[e] = v;
            F[e] = v;
            vc[l][u] = v;
            const double dy = (double)(v - xv[u]);   // x_new - x_old as the reference forms it
            gain += dy * ((double)gv[u] + damp * dy);
            dy2 += dy * dy;
            x2 += xo * xo;
            const double ag = fabs((double)gv[u]);
            gmax = ag > gmax ? ag : gmax;
            ss[l] += (double)v * (double)v;
          }
        }
      }
    }
  }
  // column norms and this rank's scalars: wave sums, then the waves in order
#pragma unroll
  for (int l = 0; l < TFX_MAXL; ++l)
    if (l < L) ss[l] = wave_sum(ss[l]);
  gain = wave_sum(gain);
  dy2 = wave_sum(dy2);
  x2 = wave_sum(x2);
  for (int o = 32; o > 0; o >>= 1) {
    const double u = __shfl_xor(gmax, o, 64);
    gmax = u > gmax ? u : gmax;
  }
  if (lane == 0) {
#pragma unroll
    for (int l = 0; l < TFX_MAXL; ++l)
      if (l < L) red[wave][l] = ss[l];
    red[wave][TFX_MAXL + 0] = gain;
    red[wave][TFX_MAXL + 1] = dy2;
    red[wave][TFX_MAXL + 2] = x2;
    red[wave][TFX_MAXL + 3] = gmax;
  }
  __syncthreads();
  if (t == 0) {
    double q[4] = {0.0, 0.0, 0.0, red[0][TFX_MAXL + 3]};
    for (int w = 0; w < UPD_NT / 64; ++w) {
      q[0] += red[w][TFX_MAXL + 0];
      q[1] += red[w][TFX_MAXL + 1];
      q[2] += red[w][TFX_MAXL + 2];
      q[3] = red[w][TFX_MAXL + 3] > q[3] ? red[w][TFX_MAXL + 3] : q[3];
    }
    for (int l = 0; l < L; ++l) {
      double a = 0.0;
      for (int w = 0; w < UPD_NT / 64; ++w) a += red[w][l];
      cn[l] = sqrt(a);
    }
#pragma unroll
    for (int k = 0; k < 4; ++k) part[(int64_t)r * 4 + k] = q[k];
    // scales: equalize (symm != 2) or the shared-array rule on the last mode (symm == 2)
    for (int l = 0; l < L; ++l) {
      double prod = 1.0;
      if (symm == 2) {
        const double nr = cn[L - 1];
        for (int k = 0; k < L; ++k) prod *= nr;
        sc[l] = prod != 0.0 ? pow(prod, 1.0 / L) / nr : 0.0;
      } else {
        for (int k = 0; k < L; ++k) prod *= cn[k];
        sc[l] = prod != 0.0 ? pow(prod, 1.0 / L) / cn[l] : 0.0;
      }
    }
  }
  __syncthreads();
  if (eq && symm != 2) {
#pragma unroll
    for (int l = 0; l < TFX_MAXL; ++l) {
      if (l < L) {
        const double f = sc[l];
        const int64_t I = s.d[l], base = s.off[l] + (int64_t)r * I;
        if (f != 0.0) {
          if (cached) {
#pragma unroll
            for (int u = 0; u < UPD_VC; ++u) {
              const int64_t i = t + UPD_NT * u;
              if (i < I) F[base + i] = (S)(f * (double)vc[l][u]);
            }
          } else {
            for (int64_t i = t; i < I; i += UPD_NT) F[base + i] = (S)(f * (double)F[base + i]);
          }
        }
      }
    }
    if (symm == 1) __syncthreads();
  }
  if (symm == 1) {
    // transform on separate factor arrays (Conversion.py:432-437), this rank's rows
    const int64_t I = s.d[0];
    for (int64_t i = t; i < I; i += UPD_NT) {
      const int64_t e = (int64_t)r * I + i;
      S acc = F[e];
      for (int l = 1; l < L; ++l) acc += F[s.off[l] + e];
      acc = acc / (S)L;
      for (int l = 0; l < L; ++l) F[s.off[l] + e] = acc;
    }
  } else if (symm == 2) {
    // the shared-array rule of k_dgn_update (Conversion.py:39-44, 378-385, 433-438)
    const int64_t I = s.d[0];
    const int64_t last = s.off[L - 1];
    const double f = sc[0];
    const double dbl = (double)(1 << (L - 1));
    for (int64_t i = t; i < I; i += UPD_NT) {
      const int64_t e = (int64_t)r * I + i;
      S v = F[last + e];
      if (eq && f != 0.0)
        for (int k = 0; k < L; ++k) v = (S)(f * (double)v);
      v = (S)((double)v * dbl) / (S)L;
      for (int l = 0; l < L; ++l) F[s.off[l] + e] = v;
    }
  }
  // the last workgroup sums the ranks' scalars in rank order
  if (t == 0) {
    __threadfence();
    const unsigned prev = atomicAdd(counter, 1u);
    is_last = prev == (unsigned)(gridDim.x - 1);
  }
  __syncthreads();
  if (is_last) {
    // all threads load a chunk of ranks' partials at once, then thread 0 adds them in order
    __shared__ double pbuf[UPD_NT][4];
    __threadfence();
    double q[4] = {0.0, 0.0, 0.0, 0.0};
    for (int r0 = 0; r0 < R; r0 += UPD_NT) {
      if (r0 + t < R) {
        const f64x2* pr = (const f64x2*)(part + (int64_t)(r0 + t) * 4);
        const f64x2 a = __builtin_nontemporal_load(pr), b = __builtin_nontemporal_load(pr + 1);
        pbuf[t][0] = a[0];
        pbuf[t][1] = a[1];
        pbuf[t][2] = b[0];
        pbuf[t][3] = b[1];
      }
      __syncthreads();
      if (t == 0) {
        const int n = R - r0 < UPD_NT ? R - r0 : UPD_NT;
        for (int u = 0; u < n; ++u) {
          q[0] += pbuf[u][0];
          q[1] += pbuf[u][1];
          q[2] += pbuf[u][2];
          q[3] = (r0 == 0 && u == 0) || pbuf[u][3] > q[3] ? pbuf[u][3] : q[3];
        }
      }
      __syncthreads();
    }
    if (t == 0) {
#pragma unroll
      for (int k = 0; k < 4; ++k) scal[k] = q[k];
      *counter = 0u;
    }
  }
}

template <typename S>
__global__ __launch_bounds__(1024) void k_dogleg_dots(int64_t n, const S* __restrict__ y,
                                                       const S* __restrict__ g,
                                                       const S* __restrict__ jg,
                                                       double* __restrict__ dots) {
  __shared__ double red[16];
  double a = 0, b = 0, c = 0, d = 0;
  for (int64_t e = threadIdx.x; e < n; e += blockDim.x) {
    const double ye = y[e], ge = g[e], je = jg[e];
    a += ye * ye;
    b += ye * ge;
    c += ge * ge;
    d += ge * je;
  }
  a = block_sum(a, red);
  b = block_sum(b, red);
  c = block_sum(c, red);
  d = block_sum(d, red);
  if (threadIdx.x == 0) {
    dots[0] = a;
    dots[1] = b;
    dots[2] = c;
    dots[3] = d;
  }
}

template <typename S>
__global__ void k_axpby(int64_t n, double a, const S* __restrict__ y, double b,
                        const S* __restrict__ g, S* __restrict__ out) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e < n) out[e] = (S)(a * (double)y[e] + b * (double)g[e]);
}

// Up to four copies of n doubles in one launch (pairs past `np` are unused); a destination may be
// pinned host memory (the per-iteration scalars of the dGN loop).
struct Copies {
  const double* src[4];
  double* dst[4];
  int64_t n[4];
  int np;
};

__global__ void k_copies(Copies C) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
#pragma unroll
  for (int k = 0; k < 4; ++k)
    if (k < C.np && e < C.n[k]) C.dst[k][e] = C.src[k][e];
}

template <typename A, typename B>
__global__ void k_convert(int64_t n, const A* __restrict__ src, B* __restrict__ dst) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e < n) dst[e] = (B)src[e];
}

template <typename S>
int gramians_t(tfx_ctx* ctx, const Shape& s, const S* F, S* Gr, S* P1, S* P2, S* gamma,
               hipStream_t st) {
  hipLaunchKernelGGL(k_gram_factors<S>, dim3(s.L * s.R), dim3(256), 0, st, s, F, Gr);
  TFX_CHECK_LAUNCH(ctx);
  // the per-workgroup maxima go to the context scratch after its counters (transient, like the
  // update's partials there); very large products fall back to workgroup 0 forming the max
  const int64_t nwg = ((int64_t)s.L * s.R * s.R + 255) / 256;
  const bool part = nwg <= (int64_t)(TFX_DSCAL_BYTES / 8 - 8);
  S* pmax = part ? (S*)(ctx->dscal + 8) : nullptr;
  unsigned* counter = part ? (unsigned*)(ctx->dscal + TFX_DSCAL_GRAM_COUNTER) : nullptr;
  hipLaunchKernelGGL(k_gram_hadamard<S>, dim3((unsigned)nwg), dim3(256), 0, st, s, Gr, P1, P2, gamma,
                     pmax, counter);
  TFX_CHECK_LAUNCH(ctx);
  return TFX_OK;
}

template <typename S>
int cg_t(tfx_ctx* ctx, const Shape& s, const S* F, const S* P1, const S* P2, const S* gamma,
         double damp, const S* grad, int maxiter, double tol, S* y, S* jtjg, double* scal,
         hipStream_t st) {
  const int64_t LRR = (int64_t)s.L * s.R * s.R;
  if constexpr (std::is_same<S, double>::value) {
    const int rc = cg_lds_try(ctx, s, F, P1, P2, gamma, damp, grad, maxiter, tol, y, jtjg, scal, st);
    if (rc != 0) return rc < 0 ? rc : TFX_OK;
  }
  // small problems: register-tiled single-workgroup kernel
  TileMap tm;
  tm.rb = (s.R + 3) / 4;
  int nt = 0;
  for (int l = 0; l < s.L; ++l) {
    tm.first[l] = nt;
    nt += (int)((s.d[l] + 3) / 4) * tm.rb;
  }
  tm.first[s.L] = nt;
  tm.ntile = nt;
  if (nt <= SMALL_NT && s.L <= 4) {
    // q, F, tmp and P1 must all sit in LDS: from global memory the register-tiled loops wait on
    // every load
    const size_t lds = sizeof(S) * (size_t)(2 * s.n + 2 * LRR);
    if (lds <= 150 * 1024) {
      void* wsp = nullptr;
      int rc = tfx_workspace(ctx, sizeof(S) * (size_t)s.n + 256, st, &wsp);
      if (rc) return rc;
      TFX_HIP(ctx, hipFuncSetAttribute((const void*)k_cg_small<S>,
                                       hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
      hipLaunchKernelGGL(k_cg_small<S>, dim3(1), dim3(SMALL_NT), lds, st, s, tm, F, P1, P2, gamma,
                         damp, grad, maxiter, tol, y, jtjg, scal, (S*)wsp);
      TFX_CHECK_LAUNCH(ctx);
      return TFX_OK;
    }
  }
  // larger problems: stream-ordered kernels over many workgroups
  if (s.n > 100) {
    if ((int64_t)maxiter > s.n) maxiter = (int)s.n;
  } else {
    tol = 0.0;
  }
  return cg_multi_t<S>(ctx, s, F, P1, P2, gamma, damp, grad, maxiter, tol, y, jtjg, scal, st);
}

}  // namespace

#define TFX_SHAPE_OR_RETURN(s)                       \
  Shape s;                                           \
  {                                                  \
    int rc_ = tfx_make_shape(ctx, L, dims, R, &s);   \
    if (rc_) return rc_;                             \
  }

extern "C" int tfx_gramians(tfx_ctx* ctx, int dtype, int L, const int64_t* dims, int R,
                            const void* F_flat, void* Gr, void* P1, void* P2, void* gamma,
                            hipStream_t stream) {
  if (!ctx || !F_flat || !Gr || !P1 || !P2 || !gamma) return tfx_set_error(ctx, TFX_E_ARG, "null argument");
  TFX_SHAPE_OR_RETURN(s);
  if (dtype == TFX_F32)
    return gramians_t<float>(ctx, s, (const float*)F_flat, (float*)Gr, (float*)P1, (float*)P2, (float*)gamma, stream);
  if (dtype == TFX_F64)
    return gramians_t<double>(ctx, s, (const double*)F_flat, (double*)Gr, (double*)P1, (double*)P2, (double*)gamma, stream);
  return tfx_set_error(ctx, TFX_E_ARG, "bad dtype %d", dtype);
}

extern "C" int tfx_neg_grad(tfx_ctx* ctx, int dtype, int L, const int64_t* dims, int R,
                            const void* F_flat, const void* P1, const void* N_flat,
                            void* grad_flat, hipStream_t stream) {
  if (!ctx || !F_flat || !P1 || !N_flat || !grad_flat) return tfx_set_error(ctx, TFX_E_ARG, "null argument");
  TFX_SHAPE_OR_RETURN(s);
  const unsigned nb = (unsigned)((s.n + 255) / 256);
  if (dtype == TFX_F32)
    hipLaunchKernelGGL(k_neg_grad<float>, dim3(nb), dim3(256), 0, stream, s, (const float*)F_flat,
                       (const float*)P1, (const float*)N_flat, (float*)grad_flat);
  else if (dtype == TFX_F64)
    hipLaunchKernelGGL(k_neg_grad<double>, dim3(nb), dim3(256), 0, stream, s, (const double*)F_flat,
                       (const double*)P1, (const double*)N_flat, (double*)grad_flat);
  else
    return tfx_set_error(ctx, TFX_E_ARG, "bad dtype %d", dtype);
  TFX_CHECK_LAUNCH(ctx);
  return TFX_OK;
}

extern "C" int tfx_cg(tfx_ctx* ctx, int dtype, int L, const int64_t* dims, int R,
                      const void* F_flat, const void* P1, const void* P2, const void* gamma,
                      double damp, const void* grad_flat, int maxiter, double tol, void* y_flat,
                      void* jtjg_flat, double* scal, hipStream_t stream) {
  if (!ctx || !F_flat || !P1 || !P2 || !gamma || !grad_flat || !y_flat || !jtjg_flat || !scal)
    return tfx_set_error(ctx, TFX_E_ARG, "null argument");
  if (maxiter < 1) return tfx_set_error(ctx, TFX_E_ARG, "maxiter %d < 1", maxiter);
  TFX_SHAPE_OR_RETURN(s);
  if (dtype == TFX_F32)
    return cg_t<float>(ctx, s, (const float*)F_flat, (const float*)P1, (const float*)P2,
                       (const float*)gamma, damp, (const float*)grad_flat, maxiter, tol,
                       (float*)y_flat, (float*)jtjg_flat, scal, stream);
  if (dtype == TFX_F64)
    return cg_t<double>(ctx, s, (const double*)F_flat, (const double*)P1, (const double*)P2,
                        (const double*)gamma, damp, (const double*)grad_flat, maxiter, tol,
                        (double*)y_flat, (double*)jtjg_flat, scal, stream);
  return tfx_set_error(ctx, TFX_E_ARG, "bad dtype %d", dtype);
}

extern "C" int tfx_dgn_update(tfx_ctx* ctx, int dtype, int L, const int64_t* dims, int R,
                              const void* x_flat, const void* y_flat, double alpha,
                              const void* grad_flat, double damp, int eq, int symm,
                              double factors_norm, void* x_new_flat, void* F_flat, double* scal,
                              hipStream_t stream) {
  if (!ctx || !x_flat || !y_flat || !grad_flat || !x_new_flat || !F_flat || !scal)
    return tfx_set_error(ctx, TFX_E_ARG, "null argument");
  TFX_SHAPE_OR_RETURN(s);
  if (symm)
    for (int l = 1; l < L; ++l)
      if (dims[l] != dims[0]) return tfx_set_error(ctx, TFX_E_ARG, "symm needs equal dims");
  if (factors_norm <= 0.0 && R <= TFX_UPD_MAX_RANKS) {
    // per-rank form: scalars summed through the context scratch (counter + R x 4 partials)
    unsigned* counter = (unsigned*)ctx->dscal;
    double* part = ctx->dscal + 8;
    if (dtype == TFX_F32)
      hipLaunchKernelGGL(k_dgn_update_rk<float>, dim3(R), dim3(UPD_NT), 0, stream, s,
                         (const float*)x_flat, (const float*)y_flat, alpha, (const float*)grad_flat,
                         damp, eq, symm, (float*)x_new_flat, (float*)F_flat, scal, counter, part);
    else if (dtype == TFX_F64)
      hipLaunchKernelGGL(k_dgn_update_rk<double>, dim3(R), dim3(UPD_NT), 0, stream, s,
                         (const double*)x_flat, (const double*)y_flat, alpha,
                         (const double*)grad_flat, damp, eq, symm, (double*)x_new_flat,
                         (double*)F_flat, scal, counter, part);
    else
      return tfx_set_error(ctx, TFX_E_ARG, "bad dtype %d", dtype);
    TFX_CHECK_LAUNCH(ctx);
    return TFX_OK;
  }
  const size_t lds = sizeof(double) * 2 * (size_t)L * R;
  if (lds > 120 * 1024) return tfx_set_error(ctx, TFX_E_UNSUPPORTED, "L*R too large");
  if (lds > 64 * 1024) {
    TFX_HIP(ctx, hipFuncSetAttribute((const void*)k_dgn_update<float>,
                                     hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    TFX_HIP(ctx, hipFuncSetAttribute((const void*)k_dgn_update<double>,
                                     hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
  }
  if (dtype == TFX_F32)
    hipLaunchKernelGGL(k_dgn_update<float>, dim3(1), dim3(1024), lds, stream, s,
                       (const float*)x_flat, (const float*)y_flat, alpha, (const float*)grad_flat,
                       damp, eq, symm, factors_norm, (float*)x_new_flat, (float*)F_flat, scal);
  else if (dtype == TFX_F64)
    hipLaunchKernelGGL(k_dgn_update<double>, dim3(1), dim3(1024), lds, stream, s,
                       (const double*)x_flat, (const double*)y_flat, alpha,
                       (const double*)grad_flat, damp, eq, symm, factors_norm,
                       (double*)x_new_flat, (double*)F_flat, scal);
  else
    return tfx_set_error(ctx, TFX_E_ARG, "bad dtype %d", dtype);
  TFX_CHECK_LAUNCH(ctx);
  return TFX_OK;
}

extern "C" int tfx_dogleg_dots(tfx_ctx* ctx, int dtype, int64_t n, const void* y, const void* g,
                               const void* jg, double* dots, hipStream_t stream) {
  if (!ctx || !y || !g || !jg || !dots) return tfx_set_error(ctx, TFX_E_ARG, "null argument");
  if (dtype == TFX_F32)
    hipLaunchKernelGGL(k_dogleg_dots<float>, dim3(1), dim3(1024), 0, stream, n, (const float*)y,
                       (const float*)g, (const float*)jg, dots);
  else if (dtype == TFX_F64)
    hipLaunchKernelGGL(k_dogleg_dots<double>, dim3(1), dim3(1024), 0, stream, n, (const double*)y,
                       (const double*)g, (const double*)jg, dots);
  else
    return tfx_set_error(ctx, TFX_E_ARG, "bad dtype %d", dtype);
  TFX_CHECK_LAUNCH(ctx);
  return TFX_OK;
}

extern "C" int tfx_axpby(tfx_ctx* ctx, int dtype, int64_t n, double a, const void* y, double b,
                         const void* g, void* out, hipStream_t stream) {
  if (!ctx || !y || !g || !out) return tfx_set_error(ctx, TFX_E_ARG, "null argument");
  const unsigned nb = (unsigned)((n + 255) / 256);
  if (n == 0) return TFX_OK;
  if (dtype == TFX_F32)
    hipLaunchKernelGGL(k_axpby<float>, dim3(nb), dim3(256), 0, stream, n, a, (const float*)y, b,
                       (const float*)g, (float*)out);
  else if (dtype == TFX_F64)
    hipLaunchKernelGGL(k_axpby<double>, dim3(nb), dim3(256), 0, stream, n, a, (const double*)y, b,
                       (const double*)g, (double*)out);
  else
    return tfx_set_error(ctx, TFX_E_ARG, "bad dtype %d", dtype);
  TFX_CHECK_LAUNCH(ctx);
  return TFX_OK;
}

int tfx_copies(tfx_ctx* ctx, int np, const double* const* src, double* const* dst, const int64_t* n,
               hipStream_t stream) {
  if (np < 1 || np > 4) return tfx_set_error(ctx, TFX_E_ARG, "copies: 1 to 4 pairs");
  Copies C{};
  int64_t nmax = 0;
  for (int k = 0; k < np; ++k) {
    C.src[k] = src[k];
    C.dst[k] = dst[k];
    C.n[k] = n[k];
    nmax = n[k] > nmax ? n[k] : nmax;
  }
  C.np = np;
  if (nmax == 0) return TFX_OK;
  hipLaunchKernelGGL(k_copies, dim3((unsigned)((nmax + 255) / 256)), dim3(256), 0, stream, C);
  TFX_CHECK_LAUNCH(ctx);
  return TFX_OK;
}

extern "C" int tfx_convert(tfx_ctx* ctx, int64_t n, const void* src, int src_dtype, void* dst,
                           int dst_dtype, hipStream_t stream) {
  if (!ctx || !src || !dst) return tfx_set_error(ctx, TFX_E_ARG, "null argument");
  if (n == 0) return TFX_OK;
  const unsigned nb = (unsigned)((n + 255) / 256);
  if (src_dtype == TFX_F64 && dst_dtype == TFX_F32)
    hipLaunchKernelGGL((k_convert<double, float>), dim3(nb), dim3(256), 0, stream, n, (const double*)src, (float*)dst);
  else if (src_dtype == TFX_F32 && dst_dtype == TFX_F64)
    hipLaunchKernelGGL((k_convert<float, double>), dim3(nb), dim3(256), 0, stream, n, (const float*)src, (double*)dst);
  else if (src_dtype == dst_dtype) {
    TFX_HIP(ctx, hipMemcpyAsync(dst, src, n * (src_dtype == TFX_F64 ? 8 : 4), hipMemcpyDeviceToDevice, stream));
    return TFX_OK;
  } else
    return tfx_set_error(ctx, TFX_E_ARG, "bad dtypes %d -> %d", src_dtype, dst_dtype);
  TFX_CHECK_LAUNCH(ctx);
  return TFX_OK;
}
