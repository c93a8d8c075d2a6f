// Split-preconditioned CG (GaussNewton.py:315-398) for problems whose operands do not fit the
// LDS of one workgroup: the refinement problem on the full tensor and large compressed cores.
// Every CG iteration j is two stream-ordered kernels over many workgroups:
//   K1  the update of iteration j - 1 (alpha = rn_{j-1} / (p_{j-1} . z_{j-1}): y += alpha p,
//       r_j = r_{j-1} - alpha z), then p_j = r_j + beta_j (j == 1 ? r_j : p_{j-1}) and
//       A_m = q_m^T F_m with q = M p_j (split over i, FP64 partials per split)
//   K2  z_l = M (F_l tmp_l + q_l P1_l + damp Gamma q_l),  tmp_l = sum_{m != l} P2[l][m] o A_m,
//       and the per-workgroup partials of p_j . z_j, r_j . z_j, z_j . z_j and r_j . r_j
// beta_j = rn_j / rn_{j-1} takes rn_j = rn_{j-1} - 2 alpha (r . z) + alpha^2 (z . z) (the exact
// r_j . r_j is only known after K2(j)); alpha and the stop test use the exact dot products.
// A dot product is left as one partial per workgroup; every workgroup of a later kernel that
// needs it sums the partials itself in a fixed order (the kernel boundary has published them),
// so no atomics, fences or waits are involved and all workgroups get the same value.  Partials
// and r are kept per iteration parity, so a kernel reads the previous iteration's while writing
// its own.  K1(j) stops the CG when rn_{j-1} <= tol (the update of iteration j - 2 was the last)
// by setting a flag; kernels launched after it return at once, so the host launches maxiter
// iterations without reading anything back.  The end kernel applies the last iteration's update
// (unless its residual already met the tolerance).  Dots and norms are FP64; vectors are in the
// working type S.  J^T J grad (for the dogleg) is the same K1/K2 pair without M and damping.
#include "common.h"

namespace {

typedef double f64x4 __attribute__((ext_vector_type(4)));

// EL consecutive elements of type S loaded or stored as one vector (EL = 2: 16-byte accesses,
// which need an even element offset)
template <typename S, int EL>
struct Lanes {
  typedef S T __attribute__((ext_vector_type(EL)));
};
template <typename S>
struct Lanes<S, 1> {
  typedef S T;
};
// a[i] for a run-time i, from an array read with constant indices.  For the kernel-argument
// arrays this reads every entry in one burst of scalar loads at kernel entry, where a[i] would be
// one more dependent scalar-load round trip (each costs a miss on the argument block).
template <typename T, int N>
__device__ inline T pick(const T* a, int i) {
  T v = a[0];
#pragma unroll
  for (int k = 1; k < N; ++k) v = i == k ? a[k] : v;
  return v;
}

template <typename S, int EL>
__device__ inline S lane_at(const typename Lanes<S, EL>::T& v, int h) {
  if constexpr (EL == 1) {
    return v;
  } else {
    return v[h];
  }
}

constexpr int CT = 16;        // c / r tile of the R x R products
constexpr int ICH = 128;      // i per workgroup in K1 (split of the reduction over I_m)
constexpr int KI = 64;        // i staged per step in K1 (a 50-row mode in one step)
constexpr int ZI = 16;        // rows i per workgroup in K2 (one MFMA row tile)
constexpr int ZC = 64;        // c per LDS chunk in K2
constexpr int NTH = 256;

// device scalars: stop flag and the last iteration whose update was applied
enum { SC_DONE = 0, SC_ITN, SC_N };
// K2 dot products, each nk2 partials per iteration parity
enum { D_PZ = 0, D_RZ, D_ZZ, D_RR, D_N };

struct CgPlan {
  Shape s;
  int nrb;                       // r (and c) blocks of 16
  int nsplit[TFX_MAXL];          // K1 splits of I_m
  int k1first[TFX_MAXL + 1];     // first K1 block of each mode
  int k2first[TFX_MAXL + 1];     // first K2 block of each mode
  int64_t aoff[TFX_MAXL];        // offset of mode m's partials in A_part: [split][c][r]
  int nbv, nk2;                  // workgroups of the vector kernels (K0, end) and of K2
  double damp, tol;
};

template <typename S>
struct CgBufs {
  const S* F;
  const S* P1;
  const S* P2;
  const S* gamma;
  const S* grad;
  S* r[2];          // r_j in r[j & 1]
  S* p[2];          // p_j in p[j & 1]
  S* z;
  S* yc;
  double* A;        // A_part
  double* Mv;       // preconditioner M[l][c] = 1 / sqrt(gamma^2 + damp^2 gamma^2), L x R
  double* dots[2];  // [parity][D_N][nk2] partials of K2
  double* rq0;      // K0 partials of r_0 . r_0 (read only when maxiter is 0)
  double* sc;
};

// Mode of flat element e, its offset within the mode and the mode's size, compared with
// constant indices into the shape (scalar loads at entry rather than per-element vector loads
// of the kernel arguments).
__device__ inline int mode_of_e(const Shape& s, int64_t e, int64_t* within, int64_t* dim) {
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
  *dim = d;
  return l;
}

template <typename S>
__device__ inline double inv_m(S g, double damp) {
  const double gd = (double)g;
  return 1.0 / sqrt(gd * gd + damp * damp * (gd * gd));
}

__device__ inline double wg_sum(double v, double* red) {
  v = wave_sum(v);
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
  __syncthreads();
  if (lane == 0) red[w] = v;
  __syncthreads();
  double t = 0.0;
  for (int i = 0; i < (int)(blockDim.x >> 6); ++i) t += red[i];
  return t;
}

// Sum of n per-workgroup partials written by an earlier kernel: wave 0 sums them in a fixed
// lane order and shares the result through `slot` (LDS).
__device__ inline double sum_parts(const double* part, int n, double* slot) {
  if (threadIdx.x < 64) {
    double t = 0.0;
    for (int i = threadIdx.x; i < n; i += 64) t += part[i];
    t = wave_sum(t);
    if (threadIdx.x == 0) *slot = t;
  }
  __syncthreads();
  return *slot;
}

// The scalars of iteration jj from K2(jj)'s partials: rn = r_jj . r_jj (0 becomes 1e-6 at
// jj = 0), alpha = rn / (p . z) (a zero p . z becomes 1e-6), and the recurrence for the next
// r . r.  `slots` holds D_N LDS doubles.
struct IterScalars {
  double rn, alpha, rn_next;
};

template <typename S>
__device__ inline IterScalars iter_scalars(const CgPlan& P, const CgBufs<S>& B, int jj,
                                           double* slots) {
  const double* d = B.dots[jj & 1];
  if (threadIdx.x < 64) {
    double t[D_N] = {0.0, 0.0, 0.0, 0.0};
    for (int i = threadIdx.x; i < P.nk2; i += 64)
#pragma unroll
      for (int q = 0; q < D_N; ++q) t[q] += d[q * P.nk2 + i];
#pragma unroll
    for (int q = 0; q < D_N; ++q) t[q] = wave_sum(t[q]);
    if (threadIdx.x == 0)
#pragma unroll
      for (int q = 0; q < D_N; ++q) slots[q] = t[q];
  }
  __syncthreads();
  IterScalars v;
  v.rn = slots[D_RR];
  if (jj == 0 && v.rn == 0.0) v.rn = 1e-6;
  double pz = slots[D_PZ];
  if (pz == 0.0) pz = 1e-6;
  v.alpha = v.rn / pz;
  v.rn_next = v.rn - 2.0 * v.alpha * slots[D_RZ] + v.alpha * v.alpha * slots[D_ZZ];
  return v;
}

// K0 (launched only when maxiter is 0; otherwise K1(0) forms the start): r_0 = M grad, p_0 = r_0,
// y = 0, partials of r_0 . r_0.
template <typename S>
__global__ __launch_bounds__(NTH) void k_cgm_start(CgPlan P, CgBufs<S> B) {
  __shared__ double red[NTH / 64];
  const Shape& s = P.s;
  double acc = 0.0;
  for (int64_t e = (int64_t)blockIdx.x * NTH + threadIdx.x; e < s.n; e += (int64_t)gridDim.x * NTH) {
    int64_t w, I;
    const int l = mode_of_e(s, e, &w, &I);
    const int rk = (int)(w / I);
    const double M = B.Mv[l * s.R + rk];
    const S rr = (S)(M * (double)B.grad[e]);
    B.r[0][e] = rr;
    B.p[0][e] = rr;
    B.yc[e] = S(0);
    acc += (double)rr * (double)rr;
  }
  const double v = wg_sum(acc, red);
  if (threadIdx.x == 0) B.rq0[blockIdx.x] = v;
}

// K1: A_part[m][split][c][r] = sum_{i in split} q_m[i][c] F_m[i][r].
//   raw: q = grad (J^T J grad); else q = M p_j.  At j = 0, p_0 = r_0 = M grad is formed from grad
//   and the r-block-0 workgroups store r_0, p_0 and y = 0.  For j >= 1 every wave forms the
//   scalars of iteration j - 1 itself (stopping the CG when rn_{j-1} <= tol, j >= 2) while its
//   staging loads are in flight, then updates its elements: the r-block-0 workgroups store y,
//   r_j and p_j.
//   Block 0 of the raw pass clears the stop flag and fills Mv.
template <typename S, int EL>
__global__ __launch_bounds__(NTH) void k_cgm_A(CgPlan P, CgBufs<S> B, int raw, int j) {
  __shared__ double qs[KI][CT + 1];
  __shared__ double fs[KI][CT + 1];
  const Shape& s = P.s;
  if (raw && blockIdx.x == 0 && threadIdx.x == 0) {
    B.sc[SC_DONE] = 0.0;
    B.sc[SC_ITN] = 0.0;
  }
  if (raw && blockIdx.x == 0)
    for (int e = threadIdx.x; e < s.L * s.R; e += NTH) B.Mv[e] = inv_m(B.gamma[e], P.damp);
  const bool scal = !raw && j >= 1;
  const int lane = threadIdx.x & 63;
  // block -> mode, and every argument used below, with constant indices (see pick)
  int m = 0;
#pragma unroll
  for (int q = 1; q < TFX_MAXL; ++q) m += (q < s.L && (int)blockIdx.x >= P.k1first[q]) ? 1 : 0;
  const int nrb = P.nrb, nk2 = P.nk2;
  const int64_t I = pick<int64_t, TFX_MAXL>(s.d, m);
  const int64_t off = pick<int64_t, TFX_MAXL>(s.off, m);
  const int64_t aoffm = pick<int64_t, TFX_MAXL>(P.aoff, m);
  const int local = (int)blockIdx.x - pick<int, TFX_MAXL>(P.k1first, m);
  const int odd = j & 1;
  const S* pold = odd ? B.p[0] : B.p[1];
  S* pnew = odd ? B.p[1] : B.p[0];
  S* rnew = odd ? B.r[1] : B.r[0];
  const S* src1 = (raw || j == 0) ? B.grad : (odd ? B.r[0] : B.r[1]);
  const double* dprev = odd ? B.dots[0] : B.dots[1];
  // K2(j - 1)'s dot-product partials (lanes < nk2; the rest after the staging loads)
  double dq[D_N] = {0.0, 0.0, 0.0, 0.0};
  if (scal && lane < nk2) {
#pragma unroll
    for (int q = 0; q < D_N; ++q) dq[q] = dprev[q * nk2 + lane];
  }
  double alpha = 0.0, beta = 0.0;
  S aS = S(0);
  const int split = local / (nrb * nrb);
  const int cb = (local / nrb) % nrb, rbk = local % nrb;
  const int c0 = cb * CT, r0 = rbk * CT;
  const int64_t ib = (int64_t)split * ICH;
  const int64_t ie = ib + ICH < I ? ib + ICH : I;
  const int R = s.R;
  const S* Fm = B.F + off;
  // the r-block-0 workgroups store their elements: at j = 0 the start r_0 = p_0 = M grad and
  // y = 0 (formed here from grad), at j >= 1 the update of iteration j - 1
  const bool writer = !raw && rbk == 0;
  const int t = threadIdx.x;
  // the 16 x 16 block on FP64 MFMA, wave w taking steps 4 w .. 4 w + 3 of each 64-row chunk:
  // lane (lg, li) supplies q[4 k + lg][c0 + li] and F[4 k + lg][r0 + li] at step k; register jj
  // holds c = c0 + 4 jj + lg, r = r0 + li; wave 0 adds the waves' blocks in wave order
  const int li = t & 15, lg = (t >> 4) & 3;
  const int wv = __builtin_amdgcn_readfirstlane(t >> 6);
  f64x4 acc = {0.0, 0.0, 0.0, 0.0};
  // EL consecutive rows per load (EL = 2 when every mode size is even, so that element offsets
  // are even and a row range [i, i + 2) is either inside [ib, ie) or outside it)
  typedef typename Lanes<S, EL>::T SV;
  for (int64_t i0 = ib; i0 < ie; i0 += KI) {
    // stage q[i0..i0+63][c0..c0+15] and F[i0..i0+63][r0..r0+15] (coalesced along i); loads
    // first (clamped indices), values after
    constexpr int NK1 = KI * CT / NTH / EL;
    SV v1[NK1], vo[NK1], vz[NK1], vy[NK1], vf[NK1];
    double vm[NK1];
#pragma unroll
    for (int k = 0; k < NK1; ++k) {
      const int u = t + NTH * k;
      const int cc = u / (KI / EL), ii = u % (KI / EL) * EL;
      const int64_t i = i0 + ii;
      const int c = c0 + cc, rr = r0 + cc;
      const int64_t e = (i < ie && c < R) ? off + (int64_t)c * I + i : 0;
      v1[k] = *(const SV*)(src1 + e);
      vo[k] = *(const SV*)(pold + e);
      vz[k] = *(const SV*)(B.z + e);
      vy[k] = *(const SV*)(B.yc + e);
      vf[k] = *(const SV*)(Fm + ((i < ie && rr < R) ? (int64_t)rr * I + i : 0));
      vm[k] = raw ? 0.0 : B.Mv[m * R + (c < R ? c : 0)];
    }
    if (i0 == ib && !raw) {
      // the stop flag and the scalars, every wave alike: workgroups of this kernel may set the
      // flag, but only when every wave of every workgroup also finds rn_{j-1} <= tol (sums of the
      // same partials in iter_scalars' order)
      if (B.sc[SC_DONE] != 0.0) return;
      if (scal) {
        double tq[D_N];
#pragma unroll
        for (int q = 0; q < D_N; ++q) {
          double v = 0.0 + dq[q];
          for (int i = lane + 64; i < nk2; i += 64) v += dprev[q * nk2 + i];
          tq[q] = wave_sum(v);
        }
        double rn = tq[D_RR];
        if (j - 1 == 0 && rn == 0.0) rn = 1e-6;
        double pz = tq[D_PZ];
        if (pz == 0.0) pz = 1e-6;
        alpha = rn / pz;
        const double rn_next = rn - 2.0 * alpha * tq[D_RZ] + alpha * alpha * tq[D_ZZ];
        if (j >= 2 && rn <= P.tol) {
          if (threadIdx.x == 0) {
            B.sc[SC_DONE] = 1.0;
            B.sc[SC_ITN] = (double)(j - 2);
          }
          return;
        }
        beta = rn_next / rn;
        aS = (S)alpha;
      }
    }
#pragma unroll
    for (int k = 0; k < NK1; ++k) {
      const int u = t + NTH * k;
      const int cc = u / (KI / EL), ii = u % (KI / EL) * EL;
      const int64_t i = i0 + ii;
      const int c = c0 + cc, rr = r0 + cc;
      const bool okq = i < ie && c < R;   // the same for all EL rows of the slot
      SV ry, rr_, rp;
#pragma unroll
      for (int h = 0; h < EL; ++h) {
        double qv = 0.0;
        if (okq) {
          const S a1 = lane_at<S, EL>(v1[k], h), ao = lane_at<S, EL>(vo[k], h);
          if (raw) {
            qv = (double)a1;
          } else {
            S pv;
            if (j == 0) {
              pv = (S)(vm[k] * (double)a1);   // r_0 = p_0 = M grad
              if constexpr (EL == 1) {
                ry = S(0), rr_ = pv, rp = pv;
              } else {
                ry[h] = S(0), rr_[h] = pv, rp[h] = pv;
              }
            } else {
              const S rv = a1 - aS * lane_at<S, EL>(vz[k], h);
              pv = (S)((double)rv + beta * (double)(j == 1 ? rv : ao));
              const S yv = lane_at<S, EL>(vy[k], h) + aS * ao;
              if constexpr (EL == 1) {
                ry = yv, rr_ = rv, rp = pv;
              } else {
                ry[h] = yv, rr_[h] = rv, rp[h] = pv;
              }
            }
            qv = vm[k] * (double)pv;
          }
        }
        qs[ii + h][cc] = qv;
        fs[ii + h][cc] = (i < ie && rr < R) ? (double)lane_at<S, EL>(vf[k], h) : 0.0;
      }
      if (okq && writer) {
        const int64_t e = off + (int64_t)c * I + i;
        *(SV*)(B.yc + e) = ry;
        *(SV*)(rnew + e) = rr_;
        *(SV*)(pnew + e) = rp;
      }
    }
    __syncthreads();
#pragma unroll
    for (int k = 4 * wv; k < 4 * wv + 4; ++k)
      acc = __builtin_amdgcn_mfma_f64_16x16x4f64(qs[4 * k + lg][li], fs[4 * k + lg][li], acc, 0, 0, 0);
    __syncthreads();
  }
  __shared__ double Ap[NTH / 64][4][64];
#pragma unroll
  for (int jj = 0; jj < 4; ++jj) Ap[wv][jj][lane] = acc[jj];
  __syncthreads();
  if (wv == 0) {
    const int rr = r0 + li;
#pragma unroll
    for (int jj = 0; jj < 4; ++jj) {
      const int c = c0 + 4 * jj + lg;
      const double v = ((Ap[0][jj][lane] + Ap[1][jj][lane]) + Ap[2][jj][lane]) + Ap[3][jj][lane];
      if (c < R && rr < R) B.A[aoffm + ((int64_t)split * R + c) * R + rr] = v;
    }
  }
}

// K2: z_l[i][r] for rows i0..i0+15 and ranks r0..r0+15 of mode l; c in chunks of ZC.  The four
// waves take c slices 16 w .. 16 w + 15 of each chunk (8 MFMAs each), and wave 0 adds their
// tiles in wave order and finishes the block.
template <typename S, int EL>
__global__ __launch_bounds__(NTH) void k_cgm_Z(CgPlan P, CgBufs<S> B, int raw, int j, S* out) {
  __shared__ double Fs[ZI][ZC + 1];
  __shared__ double Qs[ZI][ZC + 1];
  __shared__ double Ts[ZC][CT];
  __shared__ double P1s[ZC][CT];
  const Shape& s = P.s;
  // block -> mode, and the arguments used below, with constant indices (see pick)
  int l = 0;
#pragma unroll
  for (int q = 1; q < TFX_MAXL; ++q) l += (q < s.L && (int)blockIdx.x >= P.k2first[q]) ? 1 : 0;
  const int nrb = P.nrb;
  const int local = (int)blockIdx.x - pick<int, TFX_MAXL>(P.k2first, l);
  const int ibk = local / nrb, rbk = local % nrb;
  const int64_t I = pick<int64_t, TFX_MAXL>(s.d, l);
  const int64_t i0 = (int64_t)ibk * ZI;
  const int r0 = rbk * CT;
  const int R = s.R, L = s.L;
  const int64_t RR = (int64_t)R * R;
  const int64_t off = pick<int64_t, TFX_MAXL>(s.off, l);
  const S* pj = j & 1 ? B.p[1] : B.p[0];
  double* dout = j & 1 ? B.dots[1] : B.dots[0];
  const int t = threadIdx.x;
  // the 16 x 16 output block as one FP64 MFMA tile: register a of lane (li, lg) holds row
  // i0 + 4 a + lg, column r0 + li
  const int wv = __builtin_amdgcn_readfirstlane(t >> 6), li = t & 15, lg = (t >> 4) & 3;
  // wave 0's epilogue operands and the first chunk's staging operands are loaded before the stop
  // flag is read (K1 of this iteration sets it, so every wave reads the same value)
  double eg = 0.0, eM = 0.0, ep[4] = {0.0, 0.0, 0.0, 0.0}, er[4] = {0.0, 0.0, 0.0, 0.0};
  if (wv == 0) {
    const int rq = r0 + li < R ? r0 + li : 0;
    eg = (double)B.gamma[l * R + rq];
    eM = B.Mv[l * R + rq];
    const S* rjp = j & 1 ? B.r[1] : B.r[0];
#pragma unroll
    for (int a = 0; a < 4; ++a) {
      const int64_t i = i0 + 4 * a + lg;
      const int64_t e = (i < I && r0 + li < R) ? off + (int64_t)(r0 + li) * I + i : 0;
      ep[a] = (double)pj[e];
      er[a] = (double)rjp[e];
    }
  }
  f64x4 acc = {0.0, 0.0, 0.0, 0.0};
  const S* qsrc = raw ? B.grad : pj;
  for (int cc0 = 0; cc0 < R; cc0 += ZC) {
    // Loads first, values after: every loop below has a compile-time trip count and
    // clamped indices, so all of a thread's loads are in flight together.
    // F_l and q_l rows i0..i0+63, columns cc0..cc0+63 (coalesced along i)
    // (EL consecutive rows per load, as in K1)
    typedef typename Lanes<S, EL>::T SV;
    constexpr int NFQ = ZI * ZC / NTH / EL;
    SV fv[NFQ], qv[NFQ];
    double mq[NFQ];   // M of column c
#pragma unroll
    for (int k = 0; k < NFQ; ++k) {
      const int u = t + NTH * k;
      const int cc = u / (ZI / EL), ii = u % (ZI / EL) * EL;
      const int64_t i = i0 + ii;
      const int c = cc0 + cc;
      const int64_t e = (i < I && c < R) ? off + (int64_t)c * I + i : 0;
      fv[k] = *(const SV*)(B.F + e);
      qv[k] = *(const SV*)(qsrc + e);
      mq[k] = raw ? 0.0 : B.Mv[l * R + (c < R ? c : 0)];
    }
    // tmp_l[c][r] = sum_{m != l} P2[l][m][c][r] sum_split A_m[split][c][r], and P1_l.
    // Modes 0..3: split 0 of A_m and P2[l][m] of the modes m != l, loaded with F and q;
    // further splits and modes >= 4 follow in the same summation order
    constexpr int NT = ZC * CT / NTH;
    // (EL consecutive ranks r per load: R is even when EL = 2)
    typedef typename Lanes<double, EL>::T DV;
    constexpr int NTS = NT / EL;
    DV a0[NTS][4];
    SV p2[NTS][4], pv[NTS];
#pragma unroll
    for (int k = 0; k < NTS; ++k) {
      const int u = t + NTH * k;
      const int cc = u / (CT / EL), rq = u % (CT / EL) * EL;
      const int c = cc0 + cc, rr = r0 + rq;
      const bool ok = c < R && rr < R;   // the same for the EL ranks of the slot
      const int64_t cr = ok ? (int64_t)c * R + rr : 0;
#pragma unroll
      for (int m = 0; m < 4; ++m) {
        if (m < L && m != l) {   // uniform in the workgroup
          a0[k][m] = *(const DV*)(B.A + P.aoff[m] + cr);
          p2[k][m] = *(const SV*)(B.P2 + ((int64_t)l * L + m) * RR + cr);
        } else {
          a0[k][m] = DV(0.0);
          p2[k][m] = SV(0);
        }
      }
      pv[k] = *(const SV*)(B.P1 + (int64_t)l * RR + cr);
    }
    if (cc0 == 0 && !raw && B.sc[SC_DONE] != 0.0) return;
#pragma unroll
    for (int k = 0; k < NTS; ++k) {
      const int u = t + NTH * k;
      const int cc = u / (CT / EL), rq = u % (CT / EL) * EL;
      const int c = cc0 + cc, rr = r0 + rq;
      const bool ok = c < R && rr < R;
#pragma unroll
      for (int h = 0; h < EL; ++h) {
        const int64_t cr = ok ? (int64_t)c * R + rr + h : 0;
        double acc2 = 0.0;
#pragma unroll
        for (int m = 0; m < 4; ++m) {
          if (m >= L || m == l) continue;   // uniform in the workgroup
          double a = lane_at<double, EL>(a0[k][m], h);
          for (int sp = 1; sp < P.nsplit[m]; ++sp) a += B.A[P.aoff[m] + (int64_t)sp * RR + cr];
          acc2 += (double)lane_at<S, EL>(p2[k][m], h) * a;
        }
        for (int m = 4; m < L; ++m) {
          if (m == l) continue;
          double a = 0.0;
          for (int sp = 0; sp < P.nsplit[m]; ++sp) a += B.A[P.aoff[m] + (int64_t)sp * RR + cr];
          acc2 += (double)B.P2[((int64_t)l * L + m) * RR + cr] * a;
        }
        Ts[cc][rq + h] = ok ? acc2 : 0.0;
        P1s[cc][rq + h] = ok ? (double)lane_at<S, EL>(pv[k], h) : 0.0;
      }
    }
#pragma unroll
    for (int k = 0; k < NFQ; ++k) {
      const int u = t + NTH * k;
      const int cc = u / (ZI / EL), ii = u % (ZI / EL) * EL;
      const int64_t i = i0 + ii;
      const int c = cc0 + cc;
      const bool ok = i < I && c < R;
#pragma unroll
      for (int h = 0; h < EL; ++h) {
        const double fh = (double)lane_at<S, EL>(fv[k], h), qh = (double)lane_at<S, EL>(qv[k], h);
        Fs[ii + h][cc] = ok ? fh : 0.0;
        Qs[ii + h][cc] = ok ? (raw ? qh : mq[k] * qh) : 0.0;
      }
    }
    __syncthreads();
    // z[i][r] += sum_c F[i][c] tmp[c][r] + q[i][c] P1[c][r] over this wave's c slice (columns
    // past R are zeros)
#pragma unroll
    for (int k4 = 0; k4 < ZC / 16; ++k4) {
      const int cc = 16 * wv + 4 * k4 + lg;
      acc = __builtin_amdgcn_mfma_f64_16x16x4f64(Fs[li][cc], Ts[cc][li], acc, 0, 0, 0);
      acc = __builtin_amdgcn_mfma_f64_16x16x4f64(Qs[li][cc], P1s[cc][li], acc, 0, 0, 0);
    }
    __syncthreads();
  }
  // the waves' tiles, added by wave 0 in wave order
  __shared__ double Zp[NTH / 64][4][64];
  const int lane = t & 63;
#pragma unroll
  for (int a = 0; a < 4; ++a) Zp[wv][a][lane] = acc[a];
  __syncthreads();
  if (wv != 0) return;
#pragma unroll
  for (int a = 0; a < 4; ++a) acc[a] = ((Zp[0][a][lane] + Zp[1][a][lane]) + Zp[2][a][lane]) + Zp[3][a][lane];
  double dpz = 0.0, drz = 0.0, dzz = 0.0, drr = 0.0;
  const int rr = r0 + li;
#pragma unroll
  for (int a = 0; a < 4; ++a) {
    const int64_t i = i0 + 4 * a + lg;
    if (i < I && rr < R) {
      const int64_t e = off + (int64_t)rr * I + i;
      if (raw) {
        out[e] = (S)acc[a];
      } else {
        const double g = eg;
        const double M = eM;
        const double q = M * ep[a];
        const S zv = (S)(M * (acc[a] + P.damp * g * q));
        out[e] = zv;
        const double zd = (double)zv, rd = er[a];
        dpz += ep[a] * zd;
        drz += rd * zd;
        dzz += zd * zd;
        drr += rd * rd;
      }
    }
  }
  if (raw) return;
  // the block's four dot products (wave 0 alone)
  dpz = wave_sum(dpz);
  drz = wave_sum(drz);
  dzz = wave_sum(dzz);
  drr = wave_sum(drr);
  if (lane == 0) {
    dout[D_PZ * P.nk2 + blockIdx.x] = dpz;
    dout[D_RZ * P.nk2 + blockIdx.x] = drz;
    dout[D_ZZ * P.nk2 + blockIdx.x] = dzz;
    dout[D_RR * P.nk2 + blockIdx.x] = drr;
  }
}

// y_out = M (y + the last update unless the CG stopped); scal = [iterations done, final
// residual norm^2].
template <typename S>
__global__ __launch_bounds__(NTH) void k_cgm_end(CgPlan P, CgBufs<S> B, int maxiter, S* y,
                                                 double* scal) {
  __shared__ double slots[D_N];
  const Shape& s = P.s;
  int itn;
  double rn;
  bool last_upd = false;
  S aS = S(0);
  if (B.sc[SC_DONE] != 0.0) {
    // stopped in K1(itn + 2): rn_{itn+1} is the exact residual after the last update
    itn = (int)B.sc[SC_ITN];
    rn = iter_scalars(P, B, itn + 1, slots).rn;
  } else if (maxiter <= 0) {
    itn = -1;
    rn = sum_parts(B.rq0, P.nbv, &slots[0]);
    if (rn == 0.0) rn = 1e-6;
  } else {
    const IterScalars v = iter_scalars(P, B, maxiter - 1, slots);
    if (maxiter >= 2 && v.rn <= P.tol) {
      itn = maxiter - 2;    // the update of iteration maxiter - 2 already met the tolerance
      rn = v.rn;
    } else {
      itn = maxiter - 1;
      rn = v.rn_next;
      last_upd = true;
      aS = (S)v.alpha;
    }
  }
  const S* plast = B.p[(maxiter - 1) & 1];
  for (int64_t e = (int64_t)blockIdx.x * NTH + threadIdx.x; e < s.n; e += (int64_t)gridDim.x * NTH) {
    int64_t w, I;
    const int l = mode_of_e(s, e, &w, &I);
    const int rk = (int)(w / I);
    const S yv = last_upd ? B.yc[e] + aS * plast[e] : B.yc[e];
    y[e] = (S)(B.Mv[l * s.R + rk] * (double)yv);
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    scal[0] = (double)(itn + 1);
    scal[1] = rn;
  }
}

inline unsigned nblocks_vec(int64_t n) {
  int64_t b = (n + NTH - 1) / NTH;
  if (b < 1) b = 1;
  if (b > 1024) b = 1024;
  return (unsigned)b;
}

}  // namespace

// Multi-kernel CG; same arguments and results as the single-workgroup kernels of tfx_cg.
template <typename S>
int cg_multi_t(tfx_ctx* ctx, const Shape& s, const S* F, const S* P1, const S* P2, const S* gamma,
               double damp, const S* grad, int maxiter, double tol, S* y, S* jtjg, double* scal,
               hipStream_t st) {
  CgPlan P;
  P.s = s;
  P.damp = damp;
  P.tol = tol;
  P.nrb = (s.R + CT - 1) / CT;
  int k1 = 0, k2 = 0;
  int64_t aoff = 0;
  for (int m = 0; m < s.L; ++m) {
    P.nsplit[m] = (int)((s.d[m] + ICH - 1) / ICH);
    P.k1first[m] = k1;
    P.k2first[m] = k2;
    P.aoff[m] = aoff;
    k1 += P.nsplit[m] * P.nrb * P.nrb;
    k2 += (int)((s.d[m] + ZI - 1) / ZI) * P.nrb;
    aoff += (int64_t)P.nsplit[m] * s.R * s.R;
  }
  P.k1first[s.L] = k1;
  P.k2first[s.L] = k2;
  const unsigned nbv = nblocks_vec(s.n);
  P.nbv = (int)nbv;
  P.nk2 = k2;
  const size_t vec = align256(sizeof(S) * (size_t)s.n);
  const size_t bytes = 6 * vec + align256(sizeof(double) * (size_t)aoff) +
                       align256(sizeof(double) * (size_t)s.L * s.R) +
                       2 * align256(sizeof(double) * D_N * (size_t)k2) +
                       align256(sizeof(double) * (size_t)nbv) + align256(sizeof(double) * SC_N);
  void* wsp;
  int rc = tfx_workspace(ctx, bytes + 256, st, &wsp);
  if (rc) return rc;
  char* w = (char*)wsp;
  CgBufs<S> B;
  B.F = F;
  B.P1 = P1;
  B.P2 = P2;
  B.gamma = gamma;
  B.grad = grad;
  B.r[0] = (S*)w; w += vec;
  B.r[1] = (S*)w; w += vec;
  B.p[0] = (S*)w; w += vec;
  B.p[1] = (S*)w; w += vec;
  B.z = (S*)w; w += vec;
  B.yc = (S*)w; w += vec;
  B.A = (double*)w; w += align256(sizeof(double) * (size_t)aoff);
  B.Mv = (double*)w; w += align256(sizeof(double) * (size_t)s.L * s.R);
  for (int q = 0; q < 2; ++q) {
    B.dots[q] = (double*)w;
    w += align256(sizeof(double) * D_N * (size_t)k2);
  }
  B.rq0 = (double*)w; w += align256(sizeof(double) * (size_t)nbv);
  B.sc = (double*)w;

  // paired loads when every mode size and R are even (FP64)
  bool even = sizeof(S) == 8 && s.R % 2 == 0;
  for (int m = 0; m < s.L; ++m) even = even && s.d[m] % 2 == 0;
  const auto kA = even ? k_cgm_A<S, 2> : k_cgm_A<S, 1>;
  const auto kZ = even ? k_cgm_Z<S, 2> : k_cgm_Z<S, 1>;
  // J^T J grad for the dogleg (also clears the stop flag and fills Mv)
  hipLaunchKernelGGL(kA, dim3(k1), dim3(NTH), 0, st, P, B, 1, 0);
  TFX_CHECK_LAUNCH(ctx);
  hipLaunchKernelGGL(kZ, dim3(k2), dim3(NTH), 0, st, P, B, 1, 0, jtjg);
  TFX_CHECK_LAUNCH(ctx);
  if (maxiter <= 0) {   // otherwise K1(0) forms r_0, p_0 and y = 0 itself
    hipLaunchKernelGGL(k_cgm_start<S>, dim3(nbv), dim3(NTH), 0, st, P, B);
    TFX_CHECK_LAUNCH(ctx);
  }
  for (int jj = 0; jj < maxiter; ++jj) {
    hipLaunchKernelGGL(kA, dim3(k1), dim3(NTH), 0, st, P, B, 0, jj);
    TFX_CHECK_LAUNCH(ctx);
    hipLaunchKernelGGL(kZ, dim3(k2), dim3(NTH), 0, st, P, B, 0, jj, B.z);
    TFX_CHECK_LAUNCH(ctx);
  }
  hipLaunchKernelGGL(k_cgm_end<S>, dim3(nbv), dim3(NTH), 0, st, P, B, maxiter, y, scal);
  TFX_CHECK_LAUNCH(ctx);
  return TFX_OK;
}

template int cg_multi_t<float>(tfx_ctx*, const Shape&, const float*, const float*, const float*,
                               const float*, double, const float*, int, double, float*, float*,
                               double*, hipStream_t);
template int cg_multi_t<double>(tfx_ctx*, const Shape&, const double*, const double*,
                                const double*, const double*, double, const double*, int, double,
                                double*, double*, double*, hipStream_t);
