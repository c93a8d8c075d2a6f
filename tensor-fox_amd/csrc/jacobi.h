// Parallel cyclic Jacobi eigensolver for one workgroup (blockDim.x a multiple of 64), shared by
// the MLSVD eigensolve (mlsvd.hip) and the batched MLSVD (batch.hip).
#pragma once

#include "common.h"

namespace tfxjac {

// Parallel cyclic Jacobi (round-robin pairing) on the symmetric m x m matrix H (row-major, in
// place); V (m x m, col-major) receives the eigenvectors.  Eigenvalues end on the diagonal.
__device__ inline void wg_jacobi(double* H, int m, double* V, double* cs, double* red) {
  // Cyclic parallel Jacobi (round-robin ordering: every round rotates mm/2 disjoint pairs).
  // A round computes its rotations from the current 2 x 2 diagonal blocks, then applies both
  // sides at once: the 2 x 2 block of rows (a1, b1) and columns (a2, b2) becomes R1^T B R2, each
  // block read and written by one thread; V's columns follow.  Index mm - 1 is a dummy player
  // when m is odd.  cs holds 2 * (mm / 2) doubles and 2 * (mm / 2) ints (pair indices).
  const int mm = (m + 1) & ~1;
  const int hp = mm / 2;
  double2* csk = (double2*)cs;       // (cos, sin) per pair
  int2* pab = (int2*)(cs + 2 * hp);  // (first, second) index per pair (second >= m: dummy)
  const int t = threadIdx.x, nt = blockDim.x;
  const int lane = t & 63, wave = t >> 6, nwave = nt >> 6;
  for (int e = t; e < m * m; e += nt) {
    const int i = e / m, j = e - (e / m) * m;
    V[e] = i == j ? 1.0 : 0.0;
  }
  __syncthreads();
  for (int sweep = 0; sweep < 30; ++sweep) {
    double off = 0.0, dia = 0.0;
    for (int i = wave; i < m; i += nwave)
      for (int j = lane; j < m; j += 64) {
        const double h = H[i * m + j];
        if (i == j) dia += h * h;
        else off += h * h;
      }
    off = block_sum(off, red);
    dia = block_sum(dia, red);
    // converged when the off-diagonal norm is below 1e-13 of the diagonal's
    if (off <= 1e-26 * dia || off == 0.0) break;
    for (int round = 0; round < mm - 1; ++round) {
      for (int k = t; k < hp; k += nt) {
        int a, b;
        if (k == 0) {
          a = round;
          b = mm - 1;
        } else {
          a = round + k;
          if (a >= mm - 1) a -= mm - 1;
          b = round - k;
          if (b < 0) b += mm - 1;
        }
        if (a > b) {
          const int x = a;
          a = b;
          b = x;
        }
        double c = 1.0, sn = 0.0;
        if (b < m) {
          const double hab = H[a * m + b];
          if (hab != 0.0) {
            // t = tan(theta) of the symmetric Schur rotation, as sign(tau) / (|tau| +
            // sqrt(1 + tau^2)) with tau = d / (2 hab), d = hbb - haa, written with one square
            // root and one division: t = 2 hab sgn(d) / (|d| + sqrt(d^2 + 4 hab^2)), sgn(0) = 1
            const double d = H[b * m + b] - H[a * m + a];
            const double num = d >= 0.0 ? 2.0 * hab : -2.0 * hab;
            const double tt = d == 0.0 ? 1.0 : num / (fabs(d) + sqrt(d * d + 4.0 * hab * hab));
            // c = 1 / sqrt(1 + t^2) from the hardware reciprocal square root plus one Newton step
            const double x = 1.0 + tt * tt;
            double r = __builtin_amdgcn_rsq(x);
            r = r * (1.5 - 0.5 * x * r * r);
            c = r;
            sn = tt * c;
          }
        }
        csk[k] = make_double2(c, sn);
        pab[k] = make_int2(a, b);
      }
      __syncthreads();
      if (hp <= 64) {
        // H <- J^T H J by 2 x 2 blocks and V <- V J with shift-only index maps: item q of the H
        // pass is the block (k1, k2) = (q >> sh, q & (2^sh - 1)) with 2^sh >= hp, item q of
        // the V pass the pair k = q >> 6 and rows (q & 63) + 64 u; the pair indices follow
        // from the round directly
        auto pair_of = [&](int k, int& a, int& b) {
          if (k == 0) {
            a = round;
            b = mm - 1;
          } else {
            a = round + k;
            if (a >= mm - 1) a -= mm - 1;
            b = round - k;
            if (b < 0) b += mm - 1;
          }
          if (a > b) {
            const int x = a;
            a = b;
            b = x;
          }
        };
        const int sh = hp <= 32 ? 5 : 6;
        const int msk = (1 << sh) - 1;
        for (int q = t; q < (hp << sh); q += nt) {
          const int k1 = q >> sh, k2 = q & msk;
          if (k2 >= hp) continue;
          int a1, b1, a2, b2;
          pair_of(k1, a1, b1);
          pair_of(k2, a2, b2);
          const double2 r1 = csk[k1], r2 = csk[k2];
          const double c1 = r1.x, s1 = r1.y, c2 = r2.x, s2 = r2.y;
          const bool v1 = b1 < m, v2 = b2 < m;
          const double h00 = H[a1 * m + a2];
          const double h01 = v2 ? H[a1 * m + b2] : 0.0;
          const double h10 = v1 ? H[b1 * m + a2] : 0.0;
          const double h11 = (v1 && v2) ? H[b1 * m + b2] : 0.0;
          const double x00 = c1 * h00 - s1 * h10, x01 = c1 * h01 - s1 * h11;
          const double x10 = s1 * h00 + c1 * h10, x11 = s1 * h01 + c1 * h11;
          H[a1 * m + a2] = c2 * x00 - s2 * x01;
          if (v2) H[a1 * m + b2] = s2 * x00 + c2 * x01;
          if (v1) H[b1 * m + a2] = c2 * x10 - s2 * x11;
          if (v1 && v2) H[b1 * m + b2] = s2 * x10 + c2 * x11;
        }
        for (int q = t; q < (hp << 6); q += nt) {
          const int k = q >> 6;
          int a, b;
          pair_of(k, a, b);
          if (b >= m) continue;   // dummy pair
          const double2 rk = csk[k];
          const double c = rk.x, sn = rk.y;
          for (int row = q & 63; row < m; row += 64) {
            const double va = V[a * m + row], vb = V[b * m + row];
            V[a * m + row] = c * va - sn * vb;
            V[b * m + row] = sn * va + c * vb;
          }
        }
      } else {
      // H <- J^T H J by 2 x 2 blocks (rows of pair k1, columns of pair k2), block q = k1 hp + k2
      for (int q = t; q < hp * hp; q += nt) {
        const int k1 = q / hp, k2 = q - (q / hp) * hp;
        {
          const int2 p1 = pab[k1], p2 = pab[k2];
          const double2 r1 = csk[k1], r2 = csk[k2];
          const int a1 = p1.x, b1 = p1.y, a2 = p2.x, b2 = p2.y;
          const double c1 = r1.x, s1 = r1.y, c2 = r2.x, s2 = r2.y;
          const bool v1 = b1 < m;
          const bool v2 = b2 < m;
          const double h00 = H[a1 * m + a2];
          const double h01 = v2 ? H[a1 * m + b2] : 0.0;
          const double h10 = v1 ? H[b1 * m + a2] : 0.0;
          const double h11 = (v1 && v2) ? H[b1 * m + b2] : 0.0;
          // rows: x_a = c1 h_a - s1 h_b, x_b = s1 h_a + c1 h_b
          const double x00 = c1 * h00 - s1 * h10, x01 = c1 * h01 - s1 * h11;
          const double x10 = s1 * h00 + c1 * h10, x11 = s1 * h01 + c1 * h11;
          // columns: y_a = c2 x_a - s2 x_b, y_b = s2 x_a + c2 x_b
          H[a1 * m + a2] = c2 * x00 - s2 * x01;
          if (v2) H[a1 * m + b2] = s2 * x00 + c2 * x01;
          if (v1) H[b1 * m + a2] = c2 * x10 - s2 * x11;
          if (v1 && v2) H[b1 * m + b2] = s2 * x10 + c2 * x11;
        }
      }
      // V <- V J (V[col][row] holds column col), item q = k m + row
      for (int q = t; q < hp * m; q += nt) {
        const int k = q / m, row = q - (q / m) * m;
        const int2 pk = pab[k];
        const int a = pk.x, b = pk.y;
        if (b >= m) continue;
        const double2 rk = csk[k];
        const double c = rk.x, sn = rk.y;
        const double va = V[a * m + row], vb = V[b * m + row];
        V[a * m + row] = c * va - sn * vb;
        V[b * m + row] = sn * va + c * vb;
      }
      }
      __syncthreads();
    }
  }
}

}  // namespace tfxjac
