// Parallel cyclic Jacobi eigensolver for one workgroup (blockDim.x a multiple of 64), shared by
// the MLSVD eigensolve (mlsvd.hip) and the batched MLSVD (batch.hip).
#pragma once

#include "common.h"

namespace tfxjac {

// Parallel cyclic Jacobi (round-robin pairing) on the symmetric m x m matrix H (row-major with
// row stride ld >= m, in place); V (m x m, col-major) receives the eigenvectors.  Eigenvalues
// end on the diagonal.  An odd ld keeps the column accesses of H (the mirrored writes) free of
// LDS bank conflicts.
__device__ inline void wg_jacobi(double* H, int m, int ld, double* V, double* cs, double* red) {
  // Cyclic parallel Jacobi (round-robin ordering: every round rotates mm/2 disjoint pairs).
  // A round computes its rotations from the current 2 x 2 diagonal blocks, then applies both
  // sides at once: the 2 x 2 block of rows (a1, b1) and columns (a2, b2) becomes R1^T B R2, each
  // block read and written by one thread; V's columns follow.  Index mm - 1 is a dummy player
  // when m is odd.  cs holds at least 2 mm doubles: for mm <= 128, two (cos, sin) buffers of
  // mm / 2 pairs; above, one such buffer and the pair indices as ints.
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
        const double h = H[i * ld + j];
        if (i == j) dia += h * h;
        else off += h * h;
      }
    off = block_sum(off, red);
    dia = block_sum(dia, red);
    // converged when the off-diagonal norm is below 1e-13 of the diagonal's
    if (off <= 1e-26 * dia || off == 0.0) break;
    // rotation of pair k in round `round` from the current 2 x 2 diagonal block, into out[k]
    auto rotation = [&](int round, int k, double2* out, int2* idx) {
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
        const double hab = H[a * ld + b];
        if (hab != 0.0) {
          // t = tan(theta) of the symmetric Schur rotation, as sign(tau) / (|tau| +
          // sqrt(1 + tau^2)) with tau = d / (2 hab), d = hbb - haa, written with one square
          // root and one division: t = 2 hab sgn(d) / (|d| + sqrt(d^2 + 4 hab^2)), sgn(0) = 1
          const double d = H[b * ld + b] - H[a * ld + a];
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
      out[k] = make_double2(c, sn);
      if (idx) idx[k] = make_int2(a, b);
    };
    if (hp <= 64) {
      // Rounds are pipelined: after the H update of round r, wave 0 computes the rotations of
      // round r + 1 (into the other (cos, sin) buffer) while the other waves apply round r to V.
      // H <- J^T H J on the blocks (k1, k2) with k1 <= k2 only: H stays exactly symmetric, so
      // each block also writes its mirror (k2, k1), whose entries no other item reads in this
      // round.  Items are dense in the upper triangle: row pair u holds block row u (blocks
      // (u, u..hp-1), hp - u items) followed by block row hp-1-u (u + 1 items), hp + 1 items
      // in 2^sh slots.  V <- V J: item q is the pair k = q >> 5 and rows (q & 31) + 32 v.
      // The pair indices follow from the round directly (shift-only index maps).
      auto pair_of = [&](int rnd, int k, int& a, int& b) {
        if (k == 0) {
          a = rnd;
          b = mm - 1;
        } else {
          a = rnd + k;
          if (a >= mm - 1) a -= mm - 1;
          b = rnd - k;
          if (b < 0) b += mm - 1;
        }
        if (a > b) {
          const int x = a;
          a = b;
          b = x;
        }
      };
      const int sh = hp < 32 ? 5 : (hp < 64 ? 6 : 7);
      const int msk = (1 << sh) - 1;
      const int npr = (hp + 1) >> 1;
      // V items run on waves 1.. (all of a one-wave workgroup)
      const int v0 = nwave > 1 ? t - 64 : t, vstep = nwave > 1 ? nt - 64 : nt;
      for (int k = t; k < hp; k += nt) rotation(0, k, csk, nullptr);
      __syncthreads();
      for (int round = 0; round < mm - 1; ++round) {
        const double2* cur = (round & 1) ? csk + hp : csk;
        double2* nxt = (round & 1) ? csk : csk + hp;
        for (int q = t; q < (npr << sh); q += nt) {
          const int u = q >> sh, x = q & msk;
          int k1, k2;
          if (x < hp - u) {
            k1 = u;
            k2 = u + x;
          } else {
            k1 = hp - 1 - u;
            k2 = k1 + (x - (hp - u));
            if (k1 == u || x > hp) continue;   // middle row of an odd hp; padding slots
          }
          int a1, b1, a2, b2;
          pair_of(round, k1, a1, b1);
          pair_of(round, k2, a2, b2);
          const double2 r1 = cur[k1], r2 = cur[k2];
          const double c1 = r1.x, s1 = r1.y, c2 = r2.x, s2 = r2.y;
          const bool v1 = b1 < m, v2 = b2 < m;
          const double h00 = H[a1 * ld + a2];
          const double h01 = v2 ? H[a1 * ld + b2] : 0.0;
          const double h10 = v1 ? H[b1 * ld + a2] : 0.0;
          const double h11 = (v1 && v2) ? H[b1 * ld + b2] : 0.0;
          const double x00 = c1 * h00 - s1 * h10, x01 = c1 * h01 - s1 * h11;
          const double x10 = s1 * h00 + c1 * h10, x11 = s1 * h01 + c1 * h11;
          const double y00 = c2 * x00 - s2 * x01, y01 = s2 * x00 + c2 * x01;
          const double y10 = c2 * x10 - s2 * x11, y11 = s2 * x10 + c2 * x11;
          if (k1 == k2) {
            // diagonal block: its two off-diagonal entries are one value
            H[a1 * ld + a1] = y00;
            if (v1) {
              H[a1 * ld + b1] = y01;
              H[b1 * ld + a1] = y01;
              H[b1 * ld + b1] = y11;
            }
          } else {
            H[a1 * ld + a2] = y00;
            H[a2 * ld + a1] = y00;
            if (v2) {
              H[a1 * ld + b2] = y01;
              H[b2 * ld + a1] = y01;
            }
            if (v1) {
              H[b1 * ld + a2] = y10;
              H[a2 * ld + b1] = y10;
            }
            if (v1 && v2) {
              H[b1 * ld + b2] = y11;
              H[b2 * ld + b1] = y11;
            }
          }
        }
        __syncthreads();
        if (wave == 0 && round + 1 < mm - 1)
          for (int k = lane; k < hp; k += 64) rotation(round + 1, k, nxt, nullptr);
        if (nwave == 1 || wave > 0) {
          for (int q = v0; q < (hp << 5); q += vstep) {
            const int k = q >> 5;
            int a, b;
            pair_of(round, k, a, b);
            if (b >= m) continue;   // dummy pair
            const double2 rk = cur[k];
            const double c = rk.x, sn = rk.y;
            for (int row = q & 31; row < m; row += 32) {
              const double va = V[a * m + row], vb = V[b * m + row];
              V[a * m + row] = c * va - sn * vb;
              V[b * m + row] = sn * va + c * vb;
            }
          }
        }
        __syncthreads();
      }
    } else {
      for (int round = 0; round < mm - 1; ++round) {
        for (int k = t; k < hp; k += nt) rotation(round, k, csk, pab);
        __syncthreads();
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
            const double h00 = H[a1 * ld + a2];
            const double h01 = v2 ? H[a1 * ld + b2] : 0.0;
            const double h10 = v1 ? H[b1 * ld + a2] : 0.0;
            const double h11 = (v1 && v2) ? H[b1 * ld + b2] : 0.0;
            // rows: x_a = c1 h_a - s1 h_b, x_b = s1 h_a + c1 h_b
            const double x00 = c1 * h00 - s1 * h10, x01 = c1 * h01 - s1 * h11;
            const double x10 = s1 * h00 + c1 * h10, x11 = s1 * h01 + c1 * h11;
            // columns: y_a = c2 x_a - s2 x_b, y_b = s2 x_a + c2 x_b
            H[a1 * ld + a2] = c2 * x00 - s2 * x01;
            if (v2) H[a1 * ld + b2] = s2 * x00 + c2 * x01;
            if (v1) H[b1 * ld + a2] = c2 * x10 - s2 * x11;
            if (v1 && v2) H[b1 * ld + b2] = s2 * x10 + c2 * x11;
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
        __syncthreads();
      }
    }
  }
}

}  // namespace tfxjac
