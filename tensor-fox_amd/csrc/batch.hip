// Batched dense CPDs: many small tensors of uneven shape and rank, each taken through the
// reference pipeline with default options (TensorFox.py:424-655: sorted modes, MLSVD 'seq',
// random start on the core, dGN with CG, uncompression, error on the full tensor; no
// refinement).  Three launches cover a whole batch:
//   k_bmlsvd  one workgroup per member: per mode, the Gram matrix of the unfolding, its full FP64
//             Jacobi eigendecomposition, the top components (largest-|.| entry made positive),
//             the mode product with the squared row norms and the truncation rule
//             (Compression.py:141-156, 192-301).
//   k_bdgn    one wavefront per member: the dGN loop of GaussNewton.py:29-312 on the core with
//             Gram/Hadamard products, MTTKRP, gradient, CG (lds_cg_body.h), update, equalization
//             and error all on FP64 MFMA; gain ratio, damping, best point, stopping codes and
//             the dogleg fallback on the device.  A launch runs iterations [it_begin, it_end)
//             and keeps the member's state in global memory, so one launch can cover the whole
//             loop or one iteration can be a graph node.
//   k_bfinish one workgroup per member: uncompression U_l W_l and ||T - [[F]]||^2, ||T||^2 on
//             the full member tensor (FP64).
// Members are described by int64 records (fields BM_*) built by the host; `order` lists the
// members in launch order (largest first).
#include "common.h"
#include "jacobi.h"
#include "lds_cg_body.h"

namespace {

enum {
  BM_L = 0, BM_R = 1, BM_D = 2,             // order, rank, sorted dims (BM_D .. BM_D + 3)
  BM_T = 6,                                  // tensor offset (elements)
  BM_U = 7,                                  // U blocks: mode l at + sum_{q<l} I_q min(R, I_q)
  BM_S = 8,                                  // core offset
  BM_X = 9, BM_NX = 10,                      // work vectors: NV vectors of BM_NX doubles
  BM_F = 11,                                 // full factors (R * sum I)
  BM_H = 12,                                 // histories: 4 x maxiter
  BM_W = 13, BM_WS = 14,                     // MLSVD scratch: two buffers of BM_WS doubles
  BM_STRIDE = 16
};

// member result/state record (doubles)
enum {
  RS_SSIZE = 0, RS_SMEAN = 1, RS_DAMP = 2, RS_ERR = 3, RS_BEST = 4, RS_IT = 5, RS_STOP = 6,
  RS_DONE = 7, RS_XI = 8, RS_STEPS = 9, RS_DOGLEG = 10, RS_E2 = 11, RS_T2 = 12,
  RS_STRIDE = 16
};

// work vectors of a member (each BM_NX doubles)
enum { V_X0 = 0, V_X1, V_F, V_BEST, V_Y, V_G, V_JG, V_T1, V_T2, V_T3, V_T4, V_T5, V_T6, V_T7, NV };

struct BOpts {
  int maxiter;
  int symm;
  double tol, tol_step, tol_improv, tol_grad, tol_jump, init_damp, cg_tol, factors_norm;
};

typedef ldscg::dbl4 dbl4;
using ldscg::mfma64;

// ---------------------------------------------------------------------------------------------
// MLSVD
// ---------------------------------------------------------------------------------------------
constexpr int MT = 256;
constexpr int MMAX = 64;     // largest mode size
constexpr int KMAX = 16;     // largest number of components (R <= 16)
constexpr int XCP = 65;      // pitch of a staged unfolding chunk

template <typename TX>
__device__ inline void gram_unfold(const TX* X, int m, int64_t post, int64_t n_rest, double* G,
                                   double* Xc) {
  const int t = threadIdx.x;
  const int a0 = 4 * (t >> 4), b0 = 4 * (t & 15);
  double acc[4][4];
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j) acc[i][j] = 0.0;
  for (int64_t u0 = 0; u0 < n_rest; u0 += 64) {
    for (int e = t; e < m * 64; e += MT) {
      const int a = e >> 6, uu = e & 63;
      const int64_t u = u0 + uu;
      double v = 0.0;
      if (u < n_rest) {
        const int64_t p = u / post, c = u - p * post;
        v = (double)X[(p * m + a) * post + c];
      }
      Xc[a * XCP + uu] = v;
    }
    __syncthreads();
    if (a0 < m && b0 < m) {
      for (int uu = 0; uu < 64; ++uu) {
        double xa[4], xb[4];
#pragma unroll
        for (int i = 0; i < 4; ++i) {
          xa[i] = a0 + i < m ? Xc[(a0 + i) * XCP + uu] : 0.0;
          xb[i] = b0 + i < m ? Xc[(b0 + i) * XCP + uu] : 0.0;
        }
#pragma unroll
        for (int i = 0; i < 4; ++i)
#pragma unroll
          for (int j = 0; j < 4; ++j) acc[i][j] += xa[i] * xb[j];
      }
    }
    __syncthreads();
  }
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j)
      if (a0 + i < m && b0 + j < m) G[(a0 + i) * m + b0 + j] = acc[i][j];
  __syncthreads();
}

// Y[(p * k + j) * post + c] = sum_a Us[j * m + a] X[(p * m + a) * post + c]; sig2[j] = sum Y^2
template <typename TX>
__device__ inline void ttm_unfold(const TX* X, int m, int64_t pre, int64_t post, const double* Us,
                                  int k, double* Y, double* sg, double* red) {
  const int t = threadIdx.x;
  double s2[KMAX];
#pragma unroll
  for (int j = 0; j < KMAX; ++j) s2[j] = 0.0;
  const int64_t npc = pre * post;
  for (int64_t u = t; u < npc; u += MT) {
    const int64_t p = u / post, c = u - p * post;
    double acc[KMAX];
#pragma unroll
    for (int j = 0; j < KMAX; ++j) acc[j] = 0.0;
    for (int a = 0; a < m; ++a) {
      const double x = (double)X[(p * m + a) * post + c];
#pragma unroll
      for (int j = 0; j < KMAX; ++j)
        if (j < k) acc[j] += Us[j * m + a] * x;
    }
#pragma unroll
    for (int j = 0; j < KMAX; ++j)
      if (j < k) {
        Y[(p * k + j) * post + c] = acc[j];
        s2[j] += acc[j] * acc[j];
      }
  }
#pragma unroll
  for (int j = 0; j < KMAX; ++j) {
    if (j < k) {
      const double v = block_sum(s2[j], red);
      if (t == 0) sg[j] = v;
    }
  }
  __syncthreads();
}

template <typename TS>
__global__ __launch_bounds__(MT) void k_bmlsvd(const TS* __restrict__ T,
                                               const int64_t* __restrict__ meta,
                                               const int32_t* __restrict__ order, double tol,
                                               int32_t* __restrict__ kdims, double* __restrict__ U,
                                               double* __restrict__ Sc, double* __restrict__ W) {
  __shared__ double G[MMAX * MMAX];
  __shared__ double V[MMAX * XCP];
  __shared__ double Us[MMAX * KMAX];
  __shared__ __align__(16) double cs[4 * MMAX];
  __shared__ double red[MT / 64];
  __shared__ double sg[KMAX];
  __shared__ int ordr[MMAX];
  const int mb = order[blockIdx.x];
  const int64_t* md = meta + (int64_t)mb * BM_STRIDE;
  const int L = (int)md[BM_L], R = (int)md[BM_R];
  int64_t Sd[4], I[4];
  for (int l = 0; l < 4; ++l) I[l] = Sd[l] = l < L ? md[BM_D + l] : 1;
  const TS* X0 = T + md[BM_T];
  double* Wa = W + md[BM_W];
  double* Wb = Wa + md[BM_WS];
  double* Ub = U + md[BM_U];
  const double* Xd = nullptr;     // current FP64 tensor (null: the input)
  int cur = 0;                    // 0: input, 1: Wa, 2: Wb
  int64_t uoff = 0;
  const int t = threadIdx.x;
  for (int l = 0; l < L; ++l) {
    const int m = (int)Sd[l];
    int64_t pre = 1, post = 1;
    for (int q = 0; q < l; ++q) pre *= Sd[q];
    for (int q = l + 1; q < L; ++q) post *= Sd[q];
    const int64_t n_rest = pre * post;
    int nc = (int)min((int64_t)min(R, (int)I[l]), min((int64_t)m, n_rest));
    if (cur == 0) gram_unfold<TS>(X0, m, post, n_rest, G, V);
    else gram_unfold<double>(Xd, m, post, n_rest, G, V);
    tfxjac::wg_jacobi(G, m, m, V, cs, red);
    for (int i = t; i < m; i += MT) {
      const double di = G[i * m + i];
      int rk = 0;
      for (int j = 0; j < m; ++j) {
        const double dj = G[j * m + j];
        if (dj > di || (dj == di && j < i)) ++rk;
      }
      ordr[rk] = i;
    }
    __syncthreads();
    for (int e = t; e < m * nc; e += MT) {
      const int j = e / m, a = e - j * m;
      Us[j * m + a] = V[ordr[j] * m + a];
    }
    __syncthreads();
    // the largest-|.| entry of each column made positive (first index on ties)
    if (t < nc) {
      double best = -1.0;
      int idx = 0;
      for (int a = 0; a < m; ++a) {
        const double v = fabs(Us[t * m + a]);
        if (v > best) {
          best = v;
          idx = a;
        }
      }
      if (Us[t * m + idx] < 0.0)
        for (int a = 0; a < m; ++a) Us[t * m + a] = -Us[t * m + a];
    }
    __syncthreads();
    for (int e = t; e < m * nc; e += MT) Ub[uoff + e] = Us[e];
    double* Y = cur == 1 ? Wb : Wa;
    if (cur == 0) ttm_unfold<TS>(X0, m, pre, post, Us, nc, Y, sg, red);
    else ttm_unfold<double>(Xd, m, pre, post, Us, nc, Y, sg, red);
    // truncation: smallest i with sum(sigma[i:]^2) / sum(sigma^2) < tol / L
    double sig[KMAX];
    double tot = 0.0;
#pragma unroll
    for (int j = 0; j < KMAX; ++j) {
      sig[j] = j < nc ? sqrt(fmax(sg[j], 0.0)) : 0.0;
      tot += sig[j] * sig[j];
    }
    int dim = nc;
    const double eps = tol / L;
    for (int i = 1; i < nc; ++i) {
      double tail = 0.0;
      for (int j = i; j < nc; ++j) tail += sig[j] * sig[j];
      if (tail / tot < eps) {
        dim = i;
        break;
      }
    }
    double* Yout = Y;
    if (dim < nc) {
      double* Z = Y == Wa ? Wb : Wa;
      const int64_t tot_out = pre * dim * post;
      for (int64_t e = t; e < tot_out; e += MT) {
        const int64_t p = e / (dim * post), rem = e - p * dim * post;
        const int64_t j = rem / post, c = rem - j * post;
        Z[e] = Y[(p * nc + j) * post + c];
      }
      __syncthreads();
      Yout = Z;
    }
    Xd = Yout;
    cur = Yout == Wa ? 1 : 2;
    Sd[l] = dim;
    if (t == 0) kdims[mb * 4 + l] = dim;
    uoff += I[l] * min(R, (int)I[l]);
    __syncthreads();
  }
  int64_t ns = 1;
  for (int l = 0; l < L; ++l) ns *= Sd[l];
  double* So = Sc + md[BM_S];
  for (int64_t e = t; e < ns; e += MT) So[e] = Xd[e];
  for (int l = L + t; l < 4; l += MT) kdims[mb * 4 + l] = 1;
}

// ---------------------------------------------------------------------------------------------
// dGN, one wavefront per member
// ---------------------------------------------------------------------------------------------
__device__ inline double wave_max(double x) {
  for (int o = 32; o > 0; o >>= 1) {
    const double u = __shfl_xor(x, o, 64);
    x = u > x ? u : x;
  }
  return x;
}

// The member's geometry, shared by the wave in LDS (indexed by mode at run time)
struct Member {
  ldscg::Geom g;
  int64_t st[4];          // C-order strides of the core
  int ns;                 // core elements
  double Ssize;           // ||S||
};

// The member's buffers, held in registers: each pointer is computed from the kernel's LDS base or
// from its global arguments, so its accesses compile to LDS or global instructions (a pointer
// loaded from memory would be generic)
struct MemP {
  ldscg::Lds s;
  const double* S;        // core, C order
  double* vec[NV];
  double* gam;            // L x R (LDS)
  double* cn;             // L x R column norms (LDS)
  double* herr;           // errors history (LDS)
  double* himp;           // improv history (LDS)
};

// Gram matrices of the factors (from Fs), P1, P2 pairs (LDS) and gamma (GaussNewton.py:520-548,
// 584-606).  Lane layout of an R x R tile: register j <-> row 4 j + lane / 16, column lane % 16.
template <int L>
__device__ inline void b_gramians(const Member& M, const MemP& P) {
  const ldscg::Geom& g = M.g;
  const int lane = threadIdx.x & 63, li = lane & 15, lg = lane >> 4;
  const int R = g.R;
  dbl4 gr[L];
#pragma unroll
  for (int l = 0; l < L; ++l) {
    const double* fa = P.s.Fs + (int64_t)(g.row[l] + lg) * g.PR + li;
    gr[l] = ldscg::mfma_chain_short<4>(fa, 4 * g.PR, fa, 4 * g.PR, g.ip4[l] >> 2,
                                       dbl4{0.0, 0.0, 0.0, 0.0});
  }
  double mx = 0.0;
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    const int r = 4 * j + lg, c = li;
    const bool ok = r < R && c < R;
#pragma unroll
    for (int l = 0; l < L; ++l) {
      double p = 1.0;
#pragma unroll
      for (int k = 0; k < L; ++k)
        if (k != l) p *= gr[k][j];
      if (ok) P.s.P1s[(int64_t)(l * g.R4 + r) * g.PR + c] = p;
      if (l == L - 1 && ok) mx = fmax(mx, fabs(p));
      if (ok && r == c) P.cn[l * R + r] = p;      // diagonal of P1_l (temporary)
#pragma unroll
      for (int m2 = l + 1; m2 < L; ++m2) {
        double q = 1.0;
#pragma unroll
        for (int k = 0; k < L; ++k)
          if (k != l && k != m2) q *= gr[k][j];
        if (ok) P.s.P2s[(int64_t)ldscg::pair_index(l, m2, L) * R * R + r * R + c] = q;
      }
    }
  }
  mx = wave_max(mx);
  __syncthreads();
  if (lane < L * R) P.gam[lane] = mx * sqrt(fabs(P.cn[lane]));
  __syncthreads();
}

// N_l = S_(l) KR_l for one mode, in the R x R-tile layout (row i = 4 j + lane / 16, rank lane % 16)
template <int L>
__device__ inline dbl4 b_mttkrp_mode(const Member& M, const MemP& P, int l) {
  const ldscg::Geom& g = M.g;
  const int lane = threadIdx.x & 63, li = lane & 15, lg = lane >> 4;
  int oq[L - 1], od[L - 1];
  int64_t ost[L - 1];
  int n_rest = 1;
  {
    int c = 0;
#pragma unroll
    for (int q = 0; q < L; ++q)
      if (q != l) {
        oq[c] = q;
        od[c] = g.I[q];
        ost[c] = M.st[q];
        n_rest *= g.I[q];
        ++c;
      }
  }
  const int kl = g.I[l];
  const int64_t sl = M.st[l];
  const double* Fs = P.s.Fs;
  dbl4 acc{0.0, 0.0, 0.0, 0.0};
  for (int e0 = 0; e0 < n_rest; e0 += 16) {
    double av[4], bv[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const int e = e0 + 4 * u + lg;
      const bool ok = e < n_rest;
      int rem = ok ? e : 0;
      int64_t off = 0;
      double b = 1.0;
#pragma unroll
      for (int c = L - 2; c >= 0; --c) {
        const int q = rem / od[c];
        const int ix = rem - q * od[c];
        rem = q;
        off += ix * ost[c];
        b *= Fs[(int64_t)(g.row[oq[c]] + ix) * g.PR + li];
      }
      const int ic = li < kl ? li : kl - 1;
      const double a = P.S[ic * sl + off];
      av[u] = (ok && li < kl) ? a : 0.0;
      bv[u] = ok ? b : 0.0;
    }
#pragma unroll
    for (int u = 0; u < 4; ++u) acc = mfma64(av[u], bv[u], acc);
  }
  return acc;
}

// grad (minus the gradient, N - F P1) into vec[V_G] (flat) and qs; returns max |grad| over the
// wave's entries (reduced by the caller when needed)
template <int L>
__device__ inline void b_grad(const Member& M, const MemP& P) {
  const ldscg::Geom& g = M.g;
  const int lane = threadIdx.x & 63, li = lane & 15, lg = lane >> 4;
#pragma unroll
  for (int l = 0; l < L; ++l) {
    const dbl4 N = b_mttkrp_mode<L>(M, P, l);
    const double* fa = P.s.Fs + (int64_t)(g.row[l] + li) * g.PR + lg;
    const double* pb = P.s.P1s + (int64_t)(l * g.R4 + lg) * g.PR + li;
    const dbl4 fp = ldscg::mfma_chain_short<4>(fa, 4, pb, 4 * g.PR, g.R4 >> 2,
                                               dbl4{0.0, 0.0, 0.0, 0.0});
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int i = 4 * j + lg, r = li;
      if (i < g.I[l] && r < g.R) {
        const double v = N[j] - fp[j];
        P.vec[V_G][g.off[l] + (int64_t)r * g.I[l] + i] = v;
        P.s.qs[(int64_t)(g.row[l] + i) * g.PR + r] = v;
      }
    }
  }
  __syncthreads();
}

// ||S - [[F]]||^2 with F from Fs: rows over modes 0..L-2 (16 at a time), columns = mode L-1
template <int L>
__device__ inline double b_err2(const Member& M, const MemP& P) {
  const ldscg::Geom& g = M.g;
  const int lane = threadIdx.x & 63, li = lane & 15, lg = lane >> 4;
  const double* Fs = P.s.Fs;
  const int K = g.I[L - 1];
  int NR = 1;
#pragma unroll
  for (int q = 0; q < L - 1; ++q) NR *= g.I[q];
  const int nq = g.R4 >> 2;
  double bq[4];
#pragma unroll
  for (int q = 0; q < 4; ++q)
    bq[q] = q < nq ? Fs[(int64_t)(g.row[L - 1] + li) * g.PR + 4 * q + lg] : 0.0;
  double e2 = 0.0;
  for (int rt = 0; rt < NR; rt += 16) {
    const int rho = rt + li;
    const bool ok = rho < NR;
    int ix[L - 1];
    {
      int rem = ok ? rho : 0;
#pragma unroll
      for (int q = L - 2; q >= 0; --q) {
        const int d = rem / g.I[q];
        ix[q] = rem - d * g.I[q];
        rem = d;
      }
    }
    // four k steps whatever nq (bq is zero past it; Fs columns up to 15 are in the padded row),
    // so the chain has no branches
    dbl4 acc{0.0, 0.0, 0.0, 0.0};
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      double a = 1.0;
#pragma unroll
      for (int p = 0; p < L - 1; ++p) a *= Fs[(int64_t)(g.row[p] + ix[p]) * g.PR + 4 * q + lg];
      acc = mfma64(ok ? a : 0.0, bq[q], acc);
    }
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int row = rt + 4 * j + lg;
      const bool v = row < NR && li < K;
      const int rc = row < NR ? row : NR - 1, cc = li < K ? li : K - 1;
      const double d = P.S[(int64_t)rc * K + cc] - acc[j];
      e2 += v ? d * d : 0.0;
    }
  }
  return wave_sum(e2);
}

// Fs <- flat factors
__device__ inline void b_stage_fs(const Member& M, const MemP& P, const double* F) {
  const ldscg::Geom& g = M.g;
  for (int64_t e = threadIdx.x; e < g.n; e += 64) {
    int l = 0;
    while (l + 1 < g.L && e >= g.off[l + 1]) ++l;
    const int64_t w = e - g.off[l];
    const int r = (int)(w / g.I[l]);
    const int i = (int)(w - (int64_t)r * g.I[l]);
    P.s.Fs[(int64_t)(g.row[l] + i) * g.PR + r] = F[e];
  }
  __syncthreads();
}

struct UpdOut {
  double gain, dx2, x2, gmax;
};

// xn = x + y, F = transform(equalize(xn)) (Conversion.py:361-444), the gain-ratio sums
// (k_dgn_update in dgn.hip); Fs gets F.
__device__ inline UpdOut b_update(const Member& M, const MemP& P, const double* x, const double* y,
                                  const double* gv, double damp, bool eq, int symm,
                                  double fnorm, double* xn, double* F) {
  const ldscg::Geom& g = M.g;
  const int t = threadIdx.x;
  const int L = g.L, R = g.R;
  double gain = 0.0, dy2 = 0.0, x2 = 0.0, gmax = 0.0;
  for (int64_t e = t; e < g.n; e += 64) {
    const double xo = x[e];
    const double v = xo + y[e];
    xn[e] = v;
    F[e] = v;
    const double dy = v - xo;
    gain += dy * (gv[e] + damp * dy);
    dy2 += dy * dy;
    x2 += xo * xo;
    gmax = fmax(gmax, fabs(gv[e]));
  }
  UpdOut o;
  o.gain = wave_sum(gain);
  o.dx2 = wave_sum(dy2);
  o.x2 = wave_sum(x2);
  o.gmax = wave_max(gmax);
  __syncthreads();
  if (eq) {
    if (t < L * R) {
      const int l = t / R, r = t - (t / R) * R;
      const double* col = F + g.off[l] + (int64_t)r * g.I[l];
      double a = 0.0;
      for (int i = 0; i < g.I[l]; ++i) a += col[i] * col[i];
      P.cn[t] = sqrt(a);
    }
    __syncthreads();
    for (int64_t e = t; e < g.n; e += 64) {
      int l = 0;
      while (l + 1 < L && e >= g.off[l + 1]) ++l;
      const int r = (int)((e - g.off[l]) / g.I[l]);
      double prod = 1.0;
      for (int k = 0; k < L; ++k) prod *= P.cn[k * R + r];
      if (prod != 0.0) F[e] = pow(prod, 1.0 / L) / P.cn[l * R + r] * F[e];
    }
    __syncthreads();
  }
  if (symm) {
    const int64_t I0 = g.I[0];
    for (int64_t e = t; e < (int64_t)R * I0; e += 64) {
      double acc = F[e];
      for (int l = 1; l < L; ++l) acc += F[g.off[l] + e];
      acc = acc / (double)L;
      for (int l = 0; l < L; ++l) F[g.off[l] + e] = acc;
    }
    __syncthreads();
  }
  if (fnorm > 0) {
    double fn[ldscg::MAXL];
    for (int l = 0; l < L; ++l) {
      const int64_t len = (int64_t)R * g.I[l];
      double a = 0.0;
      for (int64_t e = t; e < len; e += 64) a += F[g.off[l] + e] * F[g.off[l] + e];
      fn[l] = sqrt(wave_sum(a));
    }
    for (int64_t e = t; e < g.n; e += 64) {
      int l = 0;
      while (l + 1 < L && e >= g.off[l + 1]) ++l;
      F[e] = fnorm * (1.0 / fn[l]) * F[e];
    }
    __syncthreads();
  }
  b_stage_fs(M, P, F);
  return o;
}

__device__ inline double gain_ratio(double old_error, double error, double Tsize, double den) {
  const double num = Tsize * Tsize * (old_error * old_error - error * error);
  if (den != 0.0) return num / den;
  if (num != 0.0) return num > 0.0 ? INFINITY : -INFINITY;
  return -INFINITY;
}

__device__ inline void vcopy(double* dst, const double* src, int64_t n) {
  for (int64_t e = threadIdx.x; e < n; e += 64) dst[e] = src[e];
  __syncthreads();
}

// compute_dogleg_steps (GaussNewton.py:713-763) with dogleg (:766-795) and update_delta
// (:506-517): from x_back (in vec[V_T1]) and the CG step y, up to 11 trial points on the dogleg
// path.  Leaves the best trial's factors in vec[V_F] / Fs, its x in xout and its step in
// vec[V_Y]; returns its error.
template <int L>
__device__ inline double b_dogleg(const Member& M, const MemP& P, double error, double damp, const BOpts& o,
                                  double* xout) {
  const ldscg::Geom& g = M.g;
  const int64_t n = g.n;
  double* cx = P.vec[V_T1];
  double* cy = P.vec[V_T2];
  double* nx = P.vec[V_T3];
  double* ny = P.vec[V_T4];
  double* bx = P.vec[V_T5];
  double* by = P.vec[V_T6];
  double* bF = P.vec[V_T7];
  vcopy(cy, P.vec[V_Y], n);
  vcopy(bx, cx, n);
  vcopy(by, cy, n);
  vcopy(bF, P.vec[V_F], n);
  double best_error = error;
  double gr = 1.0, delta = 1.0;
  int count = 0;
  const double* gv = P.vec[V_G];
  const double* jg = P.vec[V_JG];
  while (gr > 0.0) {
    const double old_error = error;
    double yy = 0.0, yg = 0.0, gg = 0.0, gjg = 0.0;
    for (int64_t e = threadIdx.x; e < n; e += 64) {
      const double ye = cy[e], ge = gv[e], je = jg[e];
      yy += ye * ye;
      yg += ye * ge;
      gg += ge * ge;
      gjg += ge * je;
    }
    yy = wave_sum(yy);
    yg = wave_sum(yg);
    gg = wave_sum(gg);
    gjg = wave_sum(gjg);
    double ca, cb;
    {
      const double alpha = gg / gjg;
      const double c1 = fabs(alpha) * sqrt(gg);
      const double c2 = alpha * yg - alpha * alpha * gg;
      const double c3 = yy - 2 * alpha * yg + alpha * alpha * gg;
      const double c4 = delta * delta - c1 * c1;
      const double mm = c2 * c2 + c3 * c4;
      if (sqrt(yy) < delta) {
        ca = 1.0;
        cb = 0.0;
      } else if (c1 >= delta) {
        ca = 0.0;
        cb = (delta / c1) * alpha;
      } else if (mm < 0) {
        ca = 1.0;
        cb = 0.0;
      } else {
        const double c5 = sqrt(mm);
        const double beta = c2 <= 0 ? (-c2 + c5) / c3 : c4 / (c2 + c5);
        ca = beta;
        cb = alpha * (1 - beta);
      }
    }
    for (int64_t e = threadIdx.x; e < n; e += 64) ny[e] = ca * cy[e] + cb * gv[e];
    __syncthreads();
    const UpdOut u = b_update(M, P, cx, ny, gv, damp, false, o.symm, o.factors_norm, nx, P.vec[V_F]);
    error = sqrt(b_err2<L>(M, P)) / M.Ssize;
    gr = gain_ratio(old_error, error, M.Ssize, u.gain);
    if (error < old_error) {
      vcopy(bx, nx, n);
      vcopy(by, ny, n);
      vcopy(bF, P.vec[V_F], n);
      best_error = error;
    }
    const double step = sqrt(u.dx2);
    if (gr > 0.75) {
      delta = fmax(delta, 2 * step);
    } else {
      const double sg = 0.25 / (1 + exp(-14 * (gr - 0.75))) + 0.75;
      delta = fmin(sg * delta, delta);
    }
    double* tx = cx;
    cx = nx;
    nx = tx;
    double* ty = cy;
    cy = ny;
    ny = ty;
    if (++count > 10) break;
  }
  vcopy(P.vec[V_F], bF, n);
  vcopy(xout, bx, n);
  vcopy(P.vec[V_Y], by, n);
  b_stage_fs(M, P, P.vec[V_F]);
  return best_error;
}

// numpy's mean of a[0..n) (pairwise summation, 8 accumulators for 8 <= n <= 128)
__device__ inline double np_sum_block(const double* a, int n) {
  if (n < 8) {
    double s = 0.0;
    for (int i = 0; i < n; ++i) s += a[i];
    return s;
  }
  double r[8];
  for (int j = 0; j < 8; ++j) r[j] = a[j];
  int i = 8;
  for (; i < n - (n % 8); i += 8)
    for (int j = 0; j < 8; ++j) r[j] += a[i + j];
  double s = ((r[0] + r[1]) + (r[2] + r[3])) + ((r[4] + r[5]) + (r[6] + r[7]));
  for (; i < n; ++i) s += a[i];
  return s;
}
__device__ inline double np_mean(const double* a, int n) {
  double s;
  if (n <= 128) {
    s = np_sum_block(a, n);
  } else {
    int n2 = n / 2;
    n2 -= n2 % 8;
    s = np_sum_block(a, n2) + np_sum_block(a + n2, n - n2);
  }
  return s / n;
}

template <int L>
__global__ __launch_bounds__(64) void k_bdgn(const int64_t* __restrict__ meta,
                                             const int32_t* __restrict__ order,
                                             const int32_t* __restrict__ kdims,
                                             const double* __restrict__ Sc, double* work,
                                             double* __restrict__ hist, double* __restrict__ res,
                                             const int32_t* __restrict__ caps, BOpts o,
                                             int it_begin, int it_end) {
  extern __shared__ __align__(16) double bsm[];
  __shared__ double red[2][1];
  const int mb = order[blockIdx.x];
  const int64_t* md = meta + (int64_t)mb * BM_STRIDE;
  double* rs = res + (int64_t)mb * RS_STRIDE;
  if (it_begin > 0 && rs[RS_DONE] != 0.0) return;
  const int lane = threadIdx.x;
  // the member's geometry and pointers live in LDS: they are indexed by mode at run time
  __shared__ Member Ms;
  if (lane == 0) {
    const int R = (int)md[BM_R];
    int k[ldscg::MAXL];
    for (int l = 0; l < ldscg::MAXL; ++l) k[l] = l < L ? kdims[mb * 4 + l] : 1;
    ldscg::make_geom(L, R, k, &Ms.g);
    Ms.st[L - 1] = 1;
    for (int q = L - 2; q >= 0; --q) Ms.st[q] = Ms.st[q + 1] * k[q + 1];
    Ms.ns = (int)(Ms.st[0] * k[0]);
    Ms.Ssize = it_begin > 0 ? rs[RS_SSIZE] : 0.0;
  }
  __syncthreads();
  Member& M = Ms;
  MemP P;
  P.s = ldscg::carve(M.g, bsm);
  {
    double* tail = bsm + ldscg::lds_doubles(M.g);
    P.gam = tail;
    P.cn = tail + 64;
    P.herr = tail + 128;
    P.himp = P.herr + o.maxiter;
  }
  P.S = Sc + md[BM_S];
  {
    const int64_t x0 = md[BM_X], nx = md[BM_NX];
#pragma unroll
    for (int v = 0; v < NV; ++v) P.vec[v] = work + x0 + (int64_t)v * nx;
  }
  double* H = hist + md[BM_H];
  const int maxiter = o.maxiter;
  double damp, error, best_error;
  int it, stop, xi, dogleg;
  if (it_begin == 0) {
    double s2 = 0.0, sa = 0.0;
    for (int e = lane; e < M.ns; e += 64) {
      const double v = P.S[e];
      s2 += v * v;
      sa += fabs(v);
    }
    s2 = wave_sum(s2);
    sa = wave_sum(sa);
    __syncthreads();
    if (lane == 0) M.Ssize = sqrt(s2);
    __syncthreads();
    damp = o.init_damp * (sa / M.ns);
    error = 1.0;
    best_error = INFINITY;
    it = 0;
    stop = 5;
    xi = 0;
    dogleg = 0;
    vcopy(P.vec[V_F], P.vec[V_X0], M.g.n);
    vcopy(P.vec[V_BEST], P.vec[V_X0], M.g.n);
    if (lane == 0) {
      rs[RS_SSIZE] = M.Ssize;
      rs[RS_SMEAN] = sa / M.ns;
    }
  } else {
    damp = rs[RS_DAMP];
    error = rs[RS_ERR];
    best_error = rs[RS_BEST];
    it = (int)rs[RS_IT];
    stop = (int)rs[RS_STOP];
    xi = (int)rs[RS_XI];
    dogleg = (int)rs[RS_DOGLEG];
  }
  for (int i = lane; i < it; i += 64) {
    P.herr[i] = H[i];
    P.himp[i] = H[2 * maxiter + i];
  }
  ldscg::zero_lds<64>(M.g, P.s);
  __syncthreads();
  b_stage_fs(M, P, P.vec[V_F]);
  double tol = o.tol, tol_step = o.tol_step, tol_improv = o.tol_improv, tol_grad = o.tol_grad;
  if (M.g.n <= 100) tol = tol_step = tol_improv = tol_grad = 0.0;
  const int cst = 1 + maxiter / 10;
  bool done = false;
  int steps = 0;
  for (; it < it_end && it < maxiter; ++it) {
    const double old_error = error;
    const int cap = caps[(int64_t)mb * maxiter + it];
    b_gramians<L>(M, P);
    b_grad<L>(M, P);
    int itn;
    double rn;
    ldscg::cg_body<1, 4, 1, 4, 4>(M.g, P.s, P.gam, damp, P.vec[V_G], cap, o.cg_tol, P.vec[V_Y],
                            P.vec[V_JG], red, &itn, &rn);
    __syncthreads();
    double* x = xi ? P.vec[V_X0 + 1] : P.vec[V_X0];
    double* xn = xi ? P.vec[V_X0] : P.vec[V_X0 + 1];
    const UpdOut u = b_update(M, P, x, P.vec[V_Y], P.vec[V_G], damp, true, o.symm, o.factors_norm,
                              xn, P.vec[V_F]);
    error = sqrt(b_err2<L>(M, P)) / M.Ssize;
    xi = 1 - xi;
    double gain_den = u.gain, dx2 = u.dx2;
    if (it > 3 && error > o.tol_jump * old_error) {
      ++dogleg;
      double* xc = xi ? P.vec[V_X0 + 1] : P.vec[V_X0];
      const double* ox = xi ? P.vec[V_X0] : P.vec[V_X0 + 1];
      for (int64_t e = lane; e < M.g.n; e += 64) P.vec[V_T1][e] = xc[e] - P.vec[V_Y][e];
      __syncthreads();
      error = b_dogleg<L>(M, P, error, damp, o, xc);
      double gd = 0.0, d2 = 0.0;
      for (int64_t e = lane; e < M.g.n; e += 64) {
        const double d = xc[e] - ox[e];
        const double v = ox[e] + d;
        const double dy = v - ox[e];
        gd += dy * (P.vec[V_G][e] + damp * dy);
        d2 += dy * dy;
      }
      gain_den = wave_sum(gd);
      dx2 = wave_sum(d2);
    }
    const double gr = gain_ratio(old_error, error, M.Ssize, gain_den);
    if (gr < 0.25) damp = 3 * damp;
    else if (gr > 0.75) damp = damp / 3;
    if (error < best_error) {
      best_error = error;
      vcopy(P.vec[V_BEST], P.vec[V_F], M.g.n);
    }
    const double step = sqrt(dx2) / sqrt(u.x2);
    const double imp = it == 0 ? error : fabs(error - P.herr[it - 1]);
    if (lane == 0) {
      P.herr[it] = error;
      P.himp[it] = imp;
      H[it] = error;
      H[maxiter + it] = step;
      H[2 * maxiter + it] = imp;
      H[3 * maxiter + it] = u.gmax;
    }
    __syncthreads();
    if (it > 1) {
      int st = -1;
      if (error < tol) st = 0;
      else if (step < tol_step) st = 1;
      else if (imp < tol_improv) st = 2;
      else if (u.gmax < tol_grad) st = 3;
      else if (it > 2 * cst && it % cst == 0) {
        const double mean1 = np_mean(P.herr + it - 2 * cst, cst);
        const double mean2 = np_mean(P.herr + it - cst, cst);
        if (mean1 - mean2 <= tol_improv) {
          st = 4;
        } else {
          const double mean3 = np_mean(P.himp + it - cst, cst);
          if (mean3 < 1e-3 * mean2) st = 7;
        }
      }
      if (st < 0 && error > fmax(1.0, M.Ssize * M.Ssize) / (1e-16 + tol)) st = 6;
      if (st >= 0) {
        stop = st;
        done = true;
        steps = it + 1;
        ++it;
        break;
      }
    }
  }
  if (!done && it >= maxiter) {
    done = true;
    steps = maxiter;
  }
  if (lane == 0) {
    rs[RS_DAMP] = damp;
    rs[RS_ERR] = error;
    rs[RS_BEST] = best_error;
    rs[RS_IT] = it;
    rs[RS_STOP] = stop;
    rs[RS_XI] = xi;
    rs[RS_DOGLEG] = dogleg;
    rs[RS_DONE] = done ? 1.0 : 0.0;
    rs[RS_STEPS] = steps;
  }
}

// Uncompression F_l = U_l W_l and the FP64 error of the factors on the full member tensor.
template <typename TS, int L>
__global__ __launch_bounds__(MT) void k_bfinish(const TS* __restrict__ T,
                                                const int64_t* __restrict__ meta,
                                                const int32_t* __restrict__ order,
                                                const int32_t* __restrict__ kdims,
                                                const double* __restrict__ U,
                                                const double* __restrict__ work,
                                                double* __restrict__ Ffull,
                                                double* __restrict__ res) {
  __shared__ double Fl[4 * MMAX * KMAX];
  __shared__ double red[MT / 64];
  const int mb = order[blockIdx.x];
  const int64_t* md = meta + (int64_t)mb * BM_STRIDE;
  const int R = (int)md[BM_R];
  int I[L], k[L];
  int64_t uo[L], co[L], fo[L];
  int rowF[L];
  {
    int64_t u = 0, c = 0, f = 0;
    int rf = 0;
    for (int l = 0; l < L; ++l) {
      I[l] = (int)md[BM_D + l];
      k[l] = kdims[mb * 4 + l];
      uo[l] = u;
      co[l] = c;
      fo[l] = f;
      rowF[l] = rf;
      u += (int64_t)I[l] * min(R, I[l]);
      c += (int64_t)R * k[l];
      f += (int64_t)R * I[l];
      rf += I[l];
    }
  }
  const double* Ub = U + md[BM_U];
  const double* Wc = work + md[BM_X] + (int64_t)V_BEST * md[BM_NX];
  double* Fo = Ffull + md[BM_F];
  const int t = threadIdx.x;
  int64_t nfull = 0;
  for (int l = 0; l < L; ++l) nfull += (int64_t)R * I[l];
  for (int64_t e = t; e < nfull; e += MT) {
    int l = 0;
    while (l + 1 < L && e >= fo[l + 1]) ++l;
    const int64_t w = e - fo[l];
    const int r = (int)(w / I[l]), i = (int)(w - (int64_t)r * I[l]);
    double v = 0.0;
    for (int j = 0; j < k[l]; ++j) v += Ub[uo[l] + (int64_t)j * I[l] + i] * Wc[co[l] + (int64_t)r * k[l] + j];
    Fo[e] = v;
    Fl[(rowF[l] + i) * R + r] = v;
  }
  __syncthreads();
  const TS* Tm = T + md[BM_T];
  const int K = I[L - 1];
  int NR = 1;
  for (int q = 0; q < L - 1; ++q) NR *= I[q];
  const int wave = t >> 6, lane = t & 63;
  double e2 = 0.0, t2 = 0.0;
  for (int rho = wave; rho < NR; rho += MT / 64) {
    int ix[L];
    int rem = rho;
    for (int q = L - 2; q >= 0; --q) {
      const int d = rem / I[q];
      ix[q] = rem - d * I[q];
      rem = d;
    }
    double wr[KMAX];
#pragma unroll
    for (int r = 0; r < KMAX; ++r) {
      double w = 0.0;
      if (r < R) {
        w = 1.0;
        for (int q = 0; q < L - 1; ++q) w *= Fl[(rowF[q] + ix[q]) * R + r];
      }
      wr[r] = w;
    }
    for (int kk = lane; kk < K; kk += 64) {
      double mdl = 0.0;
#pragma unroll
      for (int r = 0; r < KMAX; ++r)
        if (r < R) mdl += wr[r] * Fl[(rowF[L - 1] + kk) * R + r];
      const double tv = (double)Tm[(int64_t)rho * K + kk];
      const double d = tv - mdl;
      e2 += d * d;
      t2 += tv * tv;
    }
  }
  e2 = block_sum(e2, red);
  t2 = block_sum(t2, red);
  if (t == 0) {
    res[(int64_t)mb * RS_STRIDE + RS_E2] = e2;
    res[(int64_t)mb * RS_STRIDE + RS_T2] = t2;
  }
}

template <int L>
int launch_bdgn(tfx_ctx* ctx, int nbatch, const int64_t* meta, const int32_t* order,
                const int32_t* kdims, const double* S, double* work, double* hist, double* res,
                const int32_t* caps, const BOpts& o, int it_begin, int it_end, int lds,
                hipStream_t st) {
  TFX_HIP(ctx, hipFuncSetAttribute((const void*)k_bdgn<L>,
                                   hipFuncAttributeMaxDynamicSharedMemorySize, lds));
  hipLaunchKernelGGL(k_bdgn<L>, dim3((unsigned)nbatch), dim3(64), (unsigned)lds, st, meta, order,
                     kdims, S, work, hist, res, caps, o, it_begin, it_end);
  TFX_CHECK_LAUNCH(ctx);
  return TFX_OK;
}

template <typename TS, int L>
int launch_bfinish(tfx_ctx* ctx, const void* T, int nbatch, const int64_t* meta,
                   const int32_t* order, const int32_t* kdims, const double* U,
                   const double* work, double* Ffull, double* res, hipStream_t st) {
  hipLaunchKernelGGL((k_bfinish<TS, L>), dim3((unsigned)nbatch), dim3(MT), 0, st, (const TS*)T,
                     meta, order, kdims, U, work, Ffull, res);
  TFX_CHECK_LAUNCH(ctx);
  return TFX_OK;
}

}  // namespace

extern "C" int tfx_batch_mlsvd(tfx_ctx* ctx, const void* T, int dtype, int nbatch,
                               const int64_t* meta, const int32_t* order, double tol_mlsvd,
                               int32_t* kdims, double* U, double* S, double* W,
                               hipStream_t stream) {
  if (!ctx || !T || !meta || !order || !kdims || !U || !S || !W)
    return tfx_set_error(ctx, TFX_E_ARG, "null argument");
  if (nbatch < 1) return tfx_set_error(ctx, TFX_E_ARG, "nbatch %d < 1", nbatch);
  if (dtype == TFX_F32)
    hipLaunchKernelGGL(k_bmlsvd<float>, dim3((unsigned)nbatch), dim3(MT), 0, stream,
                       (const float*)T, meta, order, tol_mlsvd, kdims, U, S, W);
  else if (dtype == TFX_F64)
    hipLaunchKernelGGL(k_bmlsvd<double>, dim3((unsigned)nbatch), dim3(MT), 0, stream,
                       (const double*)T, meta, order, tol_mlsvd, kdims, U, S, W);
  else
    return tfx_set_error(ctx, TFX_E_ARG, "bad dtype %d", dtype);
  TFX_CHECK_LAUNCH(ctx);
  return TFX_OK;
}

extern "C" int tfx_batch_dgn(tfx_ctx* ctx, int L, int nbatch, const int64_t* meta,
                             const int32_t* order, const int32_t* kdims, const double* S,
                             double* work, double* hist, double* res, const int32_t* caps,
                             const tfx_batch_opts* opts, int it_begin, int it_end, int lds_bytes,
                             hipStream_t stream) {
  if (!ctx || !meta || !order || !kdims || !S || !work || !hist || !res || !caps || !opts)
    return tfx_set_error(ctx, TFX_E_ARG, "null argument");
  if (nbatch < 1 || opts->maxiter < 1 || it_begin < 0 || it_end <= it_begin)
    return tfx_set_error(ctx, TFX_E_ARG, "bad batch/iteration range");
  if (lds_bytes < 1 || lds_bytes > 160 * 1024)
    return tfx_set_error(ctx, TFX_E_UNSUPPORTED, "member LDS %d bytes out of range", lds_bytes);
  BOpts o;
  o.maxiter = opts->maxiter;
  o.symm = opts->symm;
  o.tol = opts->tol;
  o.tol_step = opts->tol_step;
  o.tol_improv = opts->tol_improv;
  o.tol_grad = opts->tol_grad;
  o.tol_jump = opts->tol_jump;
  o.init_damp = opts->init_damp;
  o.cg_tol = opts->cg_tol;
  o.factors_norm = opts->factors_norm;
  if (L == 2) return launch_bdgn<2>(ctx, nbatch, meta, order, kdims, S, work, hist, res, caps, o, it_begin, it_end, lds_bytes, stream);
  if (L == 3) return launch_bdgn<3>(ctx, nbatch, meta, order, kdims, S, work, hist, res, caps, o, it_begin, it_end, lds_bytes, stream);
  if (L == 4) return launch_bdgn<4>(ctx, nbatch, meta, order, kdims, S, work, hist, res, caps, o, it_begin, it_end, lds_bytes, stream);
  return tfx_set_error(ctx, TFX_E_UNSUPPORTED, "batched dGN supports order 2-4, got %d", L);
}

extern "C" int tfx_batch_finish(tfx_ctx* ctx, const void* T, int dtype, int L, int nbatch,
                                const int64_t* meta, const int32_t* order, const int32_t* kdims,
                                const double* U, const double* work, double* Ffull, double* res,
                                hipStream_t stream) {
  if (!ctx || !T || !meta || !order || !kdims || !U || !work || !Ffull || !res)
    return tfx_set_error(ctx, TFX_E_ARG, "null argument");
  if (nbatch < 1) return tfx_set_error(ctx, TFX_E_ARG, "nbatch %d < 1", nbatch);
  if (dtype == TFX_F32) {
    if (L == 2) return launch_bfinish<float, 2>(ctx, T, nbatch, meta, order, kdims, U, work, Ffull, res, stream);
    if (L == 3) return launch_bfinish<float, 3>(ctx, T, nbatch, meta, order, kdims, U, work, Ffull, res, stream);
    if (L == 4) return launch_bfinish<float, 4>(ctx, T, nbatch, meta, order, kdims, U, work, Ffull, res, stream);
  } else if (dtype == TFX_F64) {
    if (L == 2) return launch_bfinish<double, 2>(ctx, T, nbatch, meta, order, kdims, U, work, Ffull, res, stream);
    if (L == 3) return launch_bfinish<double, 3>(ctx, T, nbatch, meta, order, kdims, U, work, Ffull, res, stream);
    if (L == 4) return launch_bfinish<double, 4>(ctx, T, nbatch, meta, order, kdims, U, work, Ffull, res, stream);
  } else {
    return tfx_set_error(ctx, TFX_E_ARG, "bad dtype %d", dtype);
  }
  return tfx_set_error(ctx, TFX_E_UNSUPPORTED, "batched finish supports order 2-4, got %d", L);
}
