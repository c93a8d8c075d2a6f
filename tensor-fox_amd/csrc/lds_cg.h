// Workgroup-level pieces of the split-preconditioned CG of GaussNewton.py:315-398 with the
// Gram-Hadamard matvec of GaussNewton.py:551-581, for problems whose factors fit in LDS.  The
// matvec GEMMs run on FP64 MFMA (v_mfma_f64_16x16x4_f64).  Operand layout of that instruction
// (gfx950): lane l supplies A[l % 16][l / 16] and B[l / 16][l % 16]; accumulator register j of
// lane l holds D[4 j + l / 16][l % 16].
//
// LDS layout, all FP64, row pitch PR = Rp + 1 (odd: 16 lanes reading a column hit distinct banks):
//   Fs, qs : rows row[l] .. row[l] + ip4[l] - 1 hold F_l[i][c] / q_l[i][c] (i < I_l, c < R), zero
//            elsewhere; 16 spare zero rows follow the last mode, so a 16-row output tile may run
//            past its mode (those output rows are discarded).
//   Ts, P1s: rows l * R4 + c hold tmp_l[c][r] / P1_l[c][r] (c, r < R), zero elsewhere.
//   P2s    : one R x R block per mode pair a < b, P2[a][b] (= P2[b][a]), pitch R.
// Used by the single-problem CG (cg_lds.hip) and by the batched dGN, one member per workgroup.
#pragma once

#include "common.h"

namespace ldscg {

constexpr int MAXL = 4;

typedef double dbl4 __attribute__((ext_vector_type(4)));

struct Geom {
  int L, R, Rp, R4, PR;
  int nrt;                  // Rp / 16 rank tiles
  int I[MAXL];
  int ip4[MAXL];            // I rounded up to 4 (MFMA k steps over rows)
  int row[MAXL + 1];        // first LDS row of each mode in Fs / qs
  int nrows;                // row[L] + 16
  int zfirst[MAXL + 1];     // first z job (16-row x 16-rank output tile) of each mode
  int nz;                   // z jobs
  int npairs;               // L (L - 1) / 2
  int64_t off[MAXL];        // flat offsets: element (i, r) of mode l at off[l] + r * I[l] + i
  int64_t n;                // R * sum I
};

__host__ __device__ inline void make_geom(int L, int R, const int* I, Geom* g) {
  g->L = L;
  g->R = R;
  g->Rp = (R + 15) & ~15;
  g->R4 = (R + 3) & ~3;
  g->PR = g->Rp + 1;
  g->nrt = g->Rp / 16;
  int row = 0, z = 0;
  int64_t off = 0;
  for (int l = 0; l < L; ++l) {
    g->I[l] = I[l];
    g->ip4[l] = (I[l] + 3) & ~3;
    g->row[l] = row;
    g->zfirst[l] = z;
    g->off[l] = off;
    row += g->ip4[l];
    z += ((I[l] + 15) / 16) * g->nrt;
    off += (int64_t)R * I[l];
  }
  g->row[L] = row;
  g->zfirst[L] = z;
  g->nrows = row + 16;
  g->nz = z;
  g->npairs = L * (L - 1) / 2;
  g->n = off;
}

// doubles of LDS: Fs, qs, Ts, P1s, P2s
__host__ __device__ inline int64_t lds_doubles(const Geom& g) {
  return 2 * (int64_t)g.nrows * g.PR + 2 * (int64_t)g.L * g.R4 * g.PR + (int64_t)g.npairs * g.R * g.R;
}

__host__ __device__ inline int pair_index(int a, int b, int L) {   // a < b
  return a * L - a * (a + 1) / 2 + (b - a - 1);
}

__device__ inline dbl4 mfma64(double a, double b, dbl4 c) {
  return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}

// acc += sum_q A_q B_q over `steps` MFMA steps, lane operands a[q * sa], b[q * sb].  Chunks of
// U: the chunk's loads are issued before its MFMAs; past the end the address is clamped and the
// A operand zeroed (MFMA is convergent, so the compiler cannot unroll with a remainder loop).
template <int U>
__device__ inline dbl4 mfma_chain(const double* a, int sa, const double* b, int sb, int steps,
                                  dbl4 acc) {
  for (int q0 = 0; q0 < steps; q0 += U) {
    double av[U], bv[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const bool ok = q0 + u < steps;
      const int q = ok ? q0 + u : q0;
      av[u] = a[q * sa];
      bv[u] = b[q * sb];
      av[u] = ok ? av[u] : 0.0;
    }
#pragma unroll
    for (int u = 0; u < U; ++u) acc = mfma64(av[u], bv[u], acc);
  }
  return acc;
}

// acc += sum_q A_q B_q for steps <= U: one chunk, no loop (address clamped, A zeroed past the
// end), for callers whose chains are known to be that short.
template <int U>
__device__ inline dbl4 mfma_chain_short(const double* a, int sa, const double* b, int sb,
                                        int steps, dbl4 acc) {
  double av[U], bv[U];
#pragma unroll
  for (int u = 0; u < U; ++u) {
    const bool ok = u < steps;
    const int q = ok ? u : 0;
    av[u] = a[q * sa];
    bv[u] = b[q * sb];
    av[u] = ok ? av[u] : 0.0;
  }
#pragma unroll
  for (int u = 0; u < U; ++u) acc = mfma64(av[u], bv[u], acc);
  return acc;
}

// acc += sum_q A_q B_q for steps <= 4 C: C unrolled chunks of 4 steps with no loop or branch
// (addresses past the end clamped to step 0, A zeroed there).
template <int C>
__device__ inline dbl4 mfma_chain_chunks(const double* a, int sa, const double* b, int sb,
                                         int steps, dbl4 acc) {
#pragma unroll
  for (int c = 0; c < C; ++c) {
    double av[4], bv[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const bool ok = 4 * c + u < steps;
      const int q = ok ? 4 * c + u : 0;
      av[u] = a[q * sa];
      bv[u] = b[q * sb];
      av[u] = ok ? av[u] : 0.0;
    }
#pragma unroll
    for (int u = 0; u < 4; ++u) acc = mfma64(av[u], bv[u], acc);
  }
  return acc;
}

// SHORT > 0: every chain has at most SHORT steps (cores of at most 4 SHORT rows per mode and
// R <= 4 SHORT), run as SHORT / 4 unrolled chunks; 0: any length, in a loop of chunks of 4
template <int SHORT>
__device__ inline dbl4 chain(const double* a, int sa, const double* b, int sb, int steps,
                             dbl4 acc) {
  if constexpr (SHORT == 4)
    return mfma_chain_short<4>(a, sa, b, sb, steps, acc);
  else if constexpr (SHORT > 0)
    return mfma_chain_chunks<SHORT / 4>(a, sa, b, sb, steps, acc);
  else
    return mfma_chain<4>(a, sa, b, sb, steps, acc);
}

struct Lds {
  double *Fs, *qs, *Ts, *P1s, *P2s;
};

__device__ inline Lds carve(const Geom& g, double* base) {
  Lds s;
  s.Fs = base;
  s.qs = s.Fs + (int64_t)g.nrows * g.PR;
  s.Ts = s.qs + (int64_t)g.nrows * g.PR;
  s.P1s = s.Ts + (int64_t)g.L * g.R4 * g.PR;
  s.P2s = s.P1s + (int64_t)g.L * g.R4 * g.PR;
  return s;
}

// (mode, first row, first rank) of z job j
__device__ inline void zjob(const Geom& g, int j, int* l, int* i0, int* r0) {
  int m = 0;
  while (m + 1 < g.L && j >= g.zfirst[m + 1]) ++m;
  const int w = j - g.zfirst[m];
  *l = m;
  *i0 = 16 * (w / g.nrt);
  *r0 = 16 * (w % g.nrt);
}

// tmp_l = sum_{m != l} P2[l][m] o (q_m^T F_m) into Ts.  Job (c-tile, r-tile) per wave, every
// mode's product formed as an MFMA chain over the rows and combined in registers.
template <int NW, int JWMAX, int SHORT = 0>
__device__ inline void tmp_phase(const Geom& g, const Lds& s, int wave, int lane) {
  const int li = lane & 15, lg = lane >> 4;
  const int nwj = g.nrt * g.nrt;
  const int RR = g.R * g.R;
#pragma unroll
  for (int k = 0; k < JWMAX; ++k) {
    const int jw = wave + NW * k;
    if (jw < nwj) {
      const int c0 = 16 * (jw / g.nrt), r0 = 16 * (jw % g.nrt);
      dbl4 acc[MAXL];
#pragma unroll
      for (int m = 0; m < MAXL; ++m) {
        acc[m] = dbl4{0.0, 0.0, 0.0, 0.0};
        if (m < g.L) {
          const double* qa = s.qs + (int64_t)(g.row[m] + lg) * g.PR + c0 + li;
          const double* fb = s.Fs + (int64_t)(g.row[m] + lg) * g.PR + r0 + li;
          acc[m] = chain<SHORT>(qa, 4 * g.PR, fb, 4 * g.PR, g.ip4[m] >> 2, acc[m]);
        }
      }
      const int r = r0 + li;
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const int c = c0 + 4 * j + lg;
        if (c < g.R && r < g.R) {
#pragma unroll
          for (int l = 0; l < MAXL; ++l) {
            if (l < g.L) {
              double t = 0.0;
#pragma unroll
              for (int m = 0; m < MAXL; ++m) {
                if (m < g.L && m != l) {
                  const int p = l < m ? pair_index(l, m, g.L) : pair_index(m, l, g.L);
                  t += s.P2s[(int64_t)p * RR + c * g.R + r] * acc[m][j];
                }
              }
              s.Ts[(int64_t)(l * g.R4 + c) * g.PR + r] = t;
            }
          }
        }
      }
    }
  }
}

// F_l tmp_l + q_l P1_l for the z job (l, i0, r0): register j holds row i0 + 4 j + lane / 16,
// rank r0 + lane % 16.
template <int SHORT = 0>
__device__ inline dbl4 z_job(const Geom& g, const Lds& s, int l, int i0, int r0, int lane) {
  const int li = lane & 15, lg = lane >> 4;
  const double* fa = s.Fs + (int64_t)(g.row[l] + i0 + li) * g.PR + lg;
  const double* qa = s.qs + (int64_t)(g.row[l] + i0 + li) * g.PR + lg;
  const double* tb = s.Ts + (int64_t)(l * g.R4 + lg) * g.PR + r0 + li;
  const double* pb = s.P1s + (int64_t)(l * g.R4 + lg) * g.PR + r0 + li;
  const dbl4 zero{0.0, 0.0, 0.0, 0.0};
  const int steps = g.R4 >> 2;
  const dbl4 a1 = chain<SHORT>(fa, 4, tb, 4 * g.PR, steps, zero);
  const dbl4 a2 = chain<SHORT>(qa, 4, pb, 4 * g.PR, steps, zero);
  return a1 + a2;
}

// Sum over the NW waves of the workgroup, the same value in every thread, fixed order.  `red`
// alternates between two buffers so consecutive sums need one barrier each.
template <int NW>
__device__ inline double wg_sum(double v, double (*red)[NW], int& buf) {
  v = wave_sum(v);
  if ((threadIdx.x & 63) == 0) red[buf][threadIdx.x >> 6] = v;
  __syncthreads();
  double t = 0.0;
#pragma unroll
  for (int w = 0; w < NW; ++w) t += red[buf][w];
  buf ^= 1;
  return t;
}

}  // namespace ldscg
