// Host-side random draws of a batch of CPDs, one member per seed, reproducing the stream of
// NumPy's legacy global generator (RandomState: MT19937 seeded with init_genrand, Gaussians by
// the polar method with the second value cached, bounded integers by masked rejection on 32-bit
// outputs) in the order the reference consumes it for one `cpd` call:
//   MLSVD sketches, one Gaussian block per mode (Compression.py:337-409), discarded here;
//   starting factors randn(k_l, R) per mode (Initialization.py:20-94);
//   the CG cap of every dGN iteration, randint(1 + it^0.4, 2 + it^0.9) (GaussNewton.py:263-266).
// The starting factors are then equalized (Conversion.py:361-387) and scaled to factors_norm
// when it is positive (Conversion.py:412-444), and written in the flat factor layout.
// Members run in parallel on host threads.
#include <math.h>
#include <sched.h>
#include <stdint.h>
#include <stdlib.h>

#include <algorithm>
#include <thread>
#include <vector>

#include "tfxhip.h"

namespace {

struct MT {
  uint32_t key[624];
  int pos;
  bool has_gauss;
  double gauss;

  void seed(uint32_t s) {
    key[0] = s;
    for (int i = 1; i < 624; ++i) key[i] = 1812433253u * (key[i - 1] ^ (key[i - 1] >> 30)) + (uint32_t)i;
    pos = 624;
    has_gauss = false;
    gauss = 0.0;
  }

  void regen() {
    const uint32_t up = 0x80000000u, lo = 0x7fffffffu, mag = 0x9908b0dfu;
    int i = 0;
    for (; i < 624 - 397; ++i) {
      const uint32_t y = (key[i] & up) | (key[i + 1] & lo);
      key[i] = key[i + 397] ^ (y >> 1) ^ ((y & 1u) ? mag : 0u);
    }
    for (; i < 623; ++i) {
      const uint32_t y = (key[i] & up) | (key[i + 1] & lo);
      key[i] = key[i + 397 - 624] ^ (y >> 1) ^ ((y & 1u) ? mag : 0u);
    }
    const uint32_t y = (key[623] & up) | (key[0] & lo);
    key[623] = key[396] ^ (y >> 1) ^ ((y & 1u) ? mag : 0u);
    pos = 0;
  }

  uint32_t next32() {
    if (pos == 624) regen();
    uint32_t y = key[pos++];
    y ^= (y >> 11);
    y ^= (y << 7) & 0x9d2c5680u;
    y ^= (y << 15) & 0xefc60000u;
    y ^= (y >> 18);
    return y;
  }

  double next_double() {   // 53-bit uniform in [0, 1)
    const int32_t a = (int32_t)(next32() >> 5), b = (int32_t)(next32() >> 6);
    return (a * 67108864.0 + b) / 9007199254740992.0;
  }

  double gauss_polar() {
    if (has_gauss) {
      has_gauss = false;
      const double t = gauss;
      gauss = 0.0;
      return t;
    }
    double x1, x2, r2;
    do {
      x1 = 2.0 * next_double() - 1.0;
      x2 = 2.0 * next_double() - 1.0;
      r2 = x1 * x1 + x2 * x2;
    } while (r2 >= 1.0 || r2 == 0.0);
    const double f = sqrt(-2.0 * log(r2) / r2);
    gauss = f * x1;
    has_gauss = true;
    return f * x2;
  }

  // Advances the stream exactly as cnt calls of gauss_polar would, forming only the value left
  // in the cache (the discarded values need no logarithm or square root).
  void skip_gauss(int64_t cnt) {
    if (cnt <= 0) return;
    if (has_gauss) {
      has_gauss = false;
      gauss = 0.0;
      --cnt;
    }
    for (int64_t p = 0; p < cnt / 2; ++p) {   // pairs whose two values are both discarded
      double x1, x2, r2;
      do {
        x1 = 2.0 * next_double() - 1.0;
        x2 = 2.0 * next_double() - 1.0;
        r2 = x1 * x1 + x2 * x2;
      } while (r2 >= 1.0 || r2 == 0.0);
    }
    if (cnt & 1) gauss_polar();   // its first value discarded, the second cached
  }

  // integer in [0, rng] (rng < 2^32 - 1), masked rejection
  uint32_t bounded(uint32_t rng) {
    if (rng == 0) return 0;
    uint32_t mask = rng;
    mask |= mask >> 1;
    mask |= mask >> 2;
    mask |= mask >> 4;
    mask |= mask >> 8;
    mask |= mask >> 16;
    uint32_t v;
    while ((v = (next32() & mask)) > rng) {
    }
    return v;
  }
};

}  // namespace

extern "C" int tfx_batch_draws(int nbatch, int L, const uint32_t* seeds, const int32_t* dims,
                               const int32_t* kdims, const int32_t* ranks, int maxiter,
                               const int64_t* cap_low, const int64_t* cap_rng, double cg_factor,
                               int cg_static, int cg_maxiter, double factors_norm,
                               const int64_t* x_off, double* x0, int32_t* caps,
                               int32_t* has_zero, int nthreads) {
  if (nbatch < 0 || L < 2 || L > TFX_MAXL || maxiter < 1 || !seeds || !dims || !kdims || !ranks ||
      !x_off || !x0 || !caps || !has_zero || (!cg_static && (!cap_low || !cap_rng)))
    return TFX_E_ARG;
  auto work = [&](int m0, int m1) {
    MT mt;
    std::vector<double> f;
    for (int m = m0; m < m1; ++m) {
      mt.seed(seeds[m]);
      const int R = ranks[m];
      const int32_t* d = dims + (int64_t)m * L;
      const int32_t* k = kdims + (int64_t)m * 4;
      // sketch shapes follow the sequential compression: mode l sees modes < l compressed
      int64_t sd[TFX_MAXL];
      for (int l = 0; l < L; ++l) sd[l] = d[l];
      for (int l = 0; l < L; ++l) {
        int64_t prod = 1;
        for (int q = 0; q < L; ++q) prod *= sd[q];
        const int64_t n_rest = prod / sd[l];
        const int64_t rows = sd[l] < n_rest ? sd[l] : n_rest;
        const int64_t cnt = rows * (std::min<int64_t>(R, d[l]) + 10);
        mt.skip_gauss(cnt);
        sd[l] = k[l];
      }
      // starting factors, drawn row-major (k_l x R) per mode
      int64_t nf = 0;
      for (int l = 0; l < L; ++l) nf += (int64_t)k[l] * R;
      f.assign(nf, 0.0);
      int zero = 0;
      for (int64_t i = 0; i < nf; ++i) {
        f[i] = mt.gauss_polar();
        if (f[i] == 0.0) zero = 1;
      }
      has_zero[m] = zero;
      // equalize: per rank-one term, every mode's column gets the geometric mean of the norms
      for (int r = 0; r < R; ++r) {
        double nr[TFX_MAXL];
        double prod = 1.0;
        int64_t o = 0;
        for (int l = 0; l < L; ++l) {
          double s = 0.0;
          for (int i = 0; i < k[l]; ++i) s += f[o + (int64_t)i * R + r] * f[o + (int64_t)i * R + r];
          nr[l] = sqrt(s);
          prod *= nr[l];
          o += (int64_t)k[l] * R;
        }
        if (prod != 0.0) {
          const double num = pow(prod, 1.0 / L);
          o = 0;
          for (int l = 0; l < L; ++l) {
            const double sc = num / nr[l];
            for (int i = 0; i < k[l]; ++i) f[o + (int64_t)i * R + r] *= sc;
            o += (int64_t)k[l] * R;
          }
        }
      }
      if (factors_norm > 0) {
        int64_t o = 0;
        for (int l = 0; l < L; ++l) {
          double s = 0.0;
          const int64_t len = (int64_t)k[l] * R;
          for (int64_t i = 0; i < len; ++i) s += f[o + i] * f[o + i];
          const double sc = factors_norm * (1.0 / sqrt(s));
          for (int64_t i = 0; i < len; ++i) f[o + i] *= sc;
          o += len;
        }
      }
      // flat layout: mode l at R * sum_{q<l} k_q, element (i, r) at + r * k_l + i
      double* x = x0 + x_off[m];
      int64_t o = 0;
      for (int l = 0; l < L; ++l) {
        for (int i = 0; i < k[l]; ++i)
          for (int r = 0; r < R; ++r) x[o + (int64_t)r * k[l] + i] = f[o + (int64_t)i * R + r];
        o += (int64_t)k[l] * R;
      }
      int32_t* c = caps + (int64_t)m * maxiter;
      for (int it = 0; it < maxiter; ++it) {
        if (cg_static) {
          c[it] = cg_maxiter;
        } else {
          const int64_t draw = cap_low[it] + (int64_t)mt.bounded((uint32_t)cap_rng[it]);
          c[it] = 1 + (L - 2) * (int32_t)trunc(cg_factor * (double)draw);
        }
      }
    }
  };
  int nt = nthreads;
  if (nt <= 0) {
    // the CPUs this process may run on, and at most OMP_NUM_THREADS when that is set
    nt = (int)std::thread::hardware_concurrency();
    cpu_set_t set;
    if (sched_getaffinity(0, sizeof(set), &set) == 0) nt = std::min(nt, (int)CPU_COUNT(&set));
    if (const char* e = getenv("OMP_NUM_THREADS")) {
      const int v = atoi(e);
      if (v > 0) nt = std::min(nt, v);
    }
  }
  nt = std::max(1, std::min(nt, std::max(1, nbatch / 16)));
  if (nt == 1) {
    work(0, nbatch);
    return TFX_OK;
  }
  std::vector<std::thread> th;
  const int per = (nbatch + nt - 1) / nt;
  for (int t = 0; t < nt; ++t) {
    const int a = t * per, b = std::min(nbatch, a + per);
    if (a < b) th.emplace_back(work, a, b);
  }
  for (auto& x : th) x.join();
  return TFX_OK;
}
