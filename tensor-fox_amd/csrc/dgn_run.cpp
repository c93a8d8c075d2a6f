// The damped Gauss-Newton loop of GaussNewton.py:29-251 (with compute_step :254-312, the CG inner
// methods, the dogleg of :713-795 and the stopping rules) as one library call.  The per-iteration
// work is the library's device entry points on one stream (Gram products and regularization,
// MTTKRP, negative gradient, CG, update with equalization and transform, residual); the host keeps
// the reference's scalar control flow (gain ratio, damping, best point, stopping codes, dogleg
// trigger) in C++, reading 8 doubles per iteration through a page-locked buffer.  The CG caps are
// drawn here, one per iteration as the reference draws them, from the caller's MT19937 state
// (NumPy's legacy global generator: key and position, advanced in place), so that a loop that
// stops early consumes exactly the reference's draws.
//
// The call sequence, arguments and host arithmetic are those of the Python loop of the package
// (the same kernels in the same order), so both give bitwise identical results; the C++ loop only
// removes the interpreter and binding overhead between kernel launches.
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>
#include <string.h>

#include <algorithm>
#include <initializer_list>
#include <utility>

#include "tfxhip.h"

// library-internal (dgn.hip): up to four copies of n[k] doubles in one launch
int tfx_copies(tfx_ctx* ctx, int np, const double* const* src, double* const* dst, const int64_t* n,
               hipStream_t stream);

namespace {

struct Mt {
  uint32_t* key;
  int pos;
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
    if (pos >= 624) regen();
    uint32_t y = key[pos++];
    y ^= (y >> 11);
    y ^= (y << 7) & 0x9d2c5680u;
    y ^= (y << 15) & 0xefc60000u;
    y ^= (y >> 18);
    return y;
  }
  // randint(low, high) of the legacy generator for high - low - 1 < 2^32 - 1: masked rejection
  int64_t randint(int64_t low, int64_t high) {
    const uint32_t rng = (uint32_t)(high - low - 1);
    if (rng == 0) return low;
    uint32_t mask = rng;
    mask |= mask >> 1;
    mask |= mask >> 2;
    mask |= mask >> 4;
    mask |= mask >> 8;
    mask |= mask >> 16;
    uint32_t v;
    while ((v = (next32() & mask)) > rng) {
    }
    return low + (int64_t)v;
  }
};

// np.mean of a float64 slice: NumPy's pairwise summation (8 accumulators from 8 elements on,
// blocks of 128), divided by the count
double np_sum(const double* a, int64_t n) {
  if (n < 8) {
    double r = 0.0;
    for (int64_t i = 0; i < n; ++i) r += a[i];
    return r;
  }
  if (n <= 128) {
    double r[8];
    for (int k = 0; k < 8; ++k) r[k] = a[k];
    int64_t i = 8;
    for (; i < n - (n % 8); i += 8)
      for (int k = 0; k < 8; ++k) r[k] += a[i + k];
    double res = ((r[0] + r[1]) + (r[2] + r[3])) + ((r[4] + r[5]) + (r[6] + r[7]));
    for (; i < n; ++i) res += a[i];
    return res;
  }
  int64_t n2 = n / 2;
  n2 -= n2 % 8;
  return np_sum(a, n2) + np_sum(a + n2, n - n2);
}
double np_mean(const double* a, int64_t n) { return np_sum(a, n) / (double)n; }

double gain_ratio(double old_error, double error, double Tsize, double gain_den) {
  const double num = (Tsize * Tsize) * (old_error * old_error - error * error);
  if (gain_den != 0.0) return num / gain_den;
  if (num != 0.0) return num > 0.0 ? INFINITY : -INFINITY;
  return -INFINITY;
}

double update_damp(double damp, const double* sched, double gain, int it) {
  if (sched) return sched[it];
  if (gain < 0.25) return 3 * damp;
  if (gain > 0.75) return damp / 3;
  return damp;
}

double update_delta(double delta, double gain, double step) {
  if (gain > 0.75) return std::max(delta, 2 * step);
  const double sigma = 0.25 / (1 + exp(-14 * (gain - 0.75))) + 0.75;
  return std::min(sigma * delta, delta);
}

// dogleg step y_dl = a y + b g from [y.y, y.g, g.g, g.Jg] (GaussNewton.py:766-795)
void dogleg_coeffs(const double* d, double delta, double* a, double* b) {
  const double yy = d[0], yg = d[1], gg = d[2], gjg = d[3];
  const double alpha = gg / gjg;
  const double c1 = fabs(alpha) * sqrt(gg);
  const double c2 = alpha * yg - alpha * alpha * gg;
  const double c3 = yy - 2 * alpha * yg + alpha * alpha * gg;
  const double c4 = delta * delta - c1 * c1;
  const double m = c2 * c2 + c3 * c4;
  if (sqrt(yy) < delta) { *a = 1.0; *b = 0.0; return; }
  if (c1 >= delta) { *a = 0.0; *b = (delta / c1) * alpha; return; }
  if (m < 0) { *a = 1.0; *b = 0.0; return; }
  const double c5 = sqrt(m);
  const double beta = c2 <= 0 ? (-c2 + c5) / c3 : c4 / (c2 + c5);
  *a = beta;
  *b = alpha * (1 - beta);
}

int stop_code(int it, const double* errors, const double* steps, const double* improv, const double* grads,
              double tol, double tol_step, double tol_improv, double tol_grad, int cst, double error, double Tsize) {
  if (it <= 1) return -1;
  if (errors[it] < tol) return 0;
  if (steps[it] < tol_step) return 1;
  if (improv[it] < tol_improv) return 2;
  if (grads[it] < tol_grad) return 3;
  if (it > 2 * cst && it % cst == 0) {
    const double mean1 = np_mean(errors + it - 2 * cst, cst);
    const double mean2 = np_mean(errors + it - cst, cst);
    if (mean1 - mean2 <= tol_improv) return 4;
    const double mean3 = np_mean(improv + it - cst, cst);
    if (mean3 < 1e-3 * mean2) return 7;
  }
  if (error > std::max(1.0, Tsize * Tsize) / (1e-16 + tol)) return 6;
  return -1;
}

// Gram/Hadamard products and gradient of one point; two sets, so that the next iteration's can
// be formed while the current iteration's are still in use
struct GradSet {
  double *Gr, *P1, *P2, *gamma, *N, *grad;
};

struct Bufs {
  GradSet gs[2];
  double *y, *jtjg, *xa, *xb, *F, *best;
  double *dx[2], *dy[2], *dbx, *dby, *dbF, *d, *sx, *sF, *scal, *tscal;
};

int64_t carve(int L, const int64_t* dims, int R, double* base, Bufs* b) {
  int64_t n = 0;
  for (int l = 0; l < L; ++l) n += (int64_t)R * dims[l];
  const int64_t rr = (int64_t)R * R;
  const int64_t nal = (n + 31) & ~int64_t(31);
  double* p = base;
  auto take = [&](int64_t cnt) {
    double* q = p;
    if (p) p += (cnt + 31) & ~int64_t(31);
    return q;
  };
  int64_t used = 0;
  auto add = [&](int64_t cnt) { used += (cnt + 31) & ~int64_t(31); };
  for (GradSet& g : b->gs) {
    g.Gr = take(L * rr), add(L * rr);
    g.P1 = take(L * rr), add(L * rr);
    g.P2 = take((int64_t)L * L * rr), add((int64_t)L * L * rr);
    g.gamma = take((int64_t)L * R), add((int64_t)L * R);
    g.N = take(nal), add(nal);
    g.grad = take(nal), add(nal);
  }
  double** vecs[] = {&b->y, &b->jtjg, &b->xa, &b->xb, &b->F, &b->best, &b->dx[0], &b->dx[1],
                     &b->dy[0], &b->dy[1], &b->dbx, &b->dby, &b->dbF, &b->d, &b->sx, &b->sF};
  for (double** v : vecs) *v = take(nal), add(nal);
  b->scal = take(16), add(16);
  b->tscal = take(16), add(16);
  return used;
}

}  // namespace

extern "C" int64_t tfx_dgn_run_work(int L, const int64_t* dims, int R) {
  if (L < 2 || L > TFX_MAXL || !dims || R < 1) return -1;
  Bufs b;
  return carve(L, dims, R, nullptr, &b);
}

extern "C" int tfx_dgn_run(tfx_ctx* ctx, const void* T, int t_dtype, int prec, int L, const int64_t* dims, int R,
                           const double* F0, const tfx_dgn_opts* o, const int8_t* inner, const double* damp_sched,
                           uint32_t* mt_key, int* mt_pos, double Tsize, double* work, int64_t work_doubles,
                           double* F_best, double* hist, int* info, hipStream_t st) {
  if (!ctx || !T || !dims || !F0 || !o || !inner || !mt_key || !mt_pos || !work || !F_best || !hist || !info)
    return TFX_E_ARG;
  const int64_t need = tfx_dgn_run_work(L, dims, R);
  if (need < 0 || work_doubles < need) return TFX_E_ARG;
  Bufs b;
  carve(L, dims, R, work, &b);
  int64_t n = 0, sdims = 0;
  for (int l = 0; l < L; ++l) {
    n += (int64_t)R * dims[l];
    sdims += dims[l];
  }
  const int maxiter = o->maxiter;
  double* steps = hist;
  double* errors = hist + maxiter;
  double* improv = hist + 2 * (int64_t)maxiter;
  double* grads = hist + 3 * (int64_t)maxiter;
  for (int64_t i = 0; i < 4 * (int64_t)maxiter; ++i) hist[i] = 0.0;
  double tol = o->tol, tol_step = o->tol_step, tol_improv = o->tol_improv, tol_grad = o->tol_grad;
  if ((int64_t)R * sdims <= 100) tol = tol_step = tol_improv = tol_grad = 0.0;
  const int cst = 1 + (int)(maxiter / 10);
  const int symm1 = o->symm ? (o->factors_norm > 0 ? 1 : 2) : 0;   // the transform on shared arrays
  int symm_now = o->symm ? (o->factors_norm > 0 ? 1 : (o->aliased ? 2 : 1)) : 0;
  double* pin = nullptr;
  // coherent pinned memory: the copy kernel writes the scalars straight into it
  if (hipHostMalloc((void**)&pin, 64 * sizeof(double), hipHostMallocCoherent) != hipSuccess) return TFX_E_NOMEM;
  hipEvent_t ev_scal = nullptr;
  if (hipEventCreateWithFlags(&ev_scal, hipEventDisableTiming) != hipSuccess) {
    hipHostFree(pin);
    return TFX_E_HIP;
  }
  Mt mt{mt_key, *mt_pos};
  int rc = TFX_OK;
  double error = 1.0, best_error = INFINITY, damp = o->damp0;
  int stop = 5, it = 0, dogleg_calls = 0;
  int gcur = 0;          // the gradient set of the current iteration
  bool ahead = false;    // its Gram products and gradient were formed during the previous one
  double* x = b.xa;
  double* xs = b.xb;
#define RUN(call)            \
  do {                       \
    rc = (call);             \
    if (rc) goto done;       \
  } while (0)
#define HIPRUN(call)                      \
  do {                                    \
    if ((call) != hipSuccess) {           \
      rc = TFX_E_HIP;                     \
      goto done;                          \
    }                                     \
  } while (0)
  // copies in one launch each (hipMemcpyAsync costs a runtime blit per call)
  auto copies = [&](std::initializer_list<std::pair<const double*, double*>> pairs, int64_t cnt) {
    const double* src[4];
    double* dst[4];
    int64_t nn[4];
    int np = 0;
    for (const auto& pr : pairs) {
      src[np] = pr.first;
      dst[np] = pr.second;
      nn[np] = cnt;
      ++np;
    }
    return tfx_copies(ctx, np, src, dst, nn, st);
  };
  RUN(copies({{F0, x}, {F0, b.F}, {F0, b.best}}, n));
  for (it = 0; it < maxiter; ++it) {
    const double old_error = error;
    int cap;
    if (inner[it] == 0) {
      const int64_t lo = (int64_t)(1 + pow((double)it, 0.4)), hi = (int64_t)(2 + pow((double)it, 0.9));
      cap = 1 + (L - 2) * (int)trunc(o->cg_factor * (double)mt.randint(lo, hi));
    } else {
      cap = o->cg_maxiter;
    }
    const GradSet& g = b.gs[gcur];
    if (!ahead) {
      RUN(tfx_gramians(ctx, TFX_F64, L, dims, R, b.F, g.Gr, g.P1, g.P2, g.gamma, st));
      RUN(tfx_mttkrp_neg_grad(ctx, T, t_dtype, L, dims, b.F, TFX_F64, R, g.P1, g.N, g.grad, prec, st));
    }
    ahead = false;
    RUN(tfx_cg(ctx, TFX_F64, L, dims, R, b.F, g.P1, g.P2, g.gamma, damp, g.grad, cap, o->cg_tol, b.y, b.jtjg,
               b.scal, st));
    RUN(tfx_dgn_update(ctx, TFX_F64, L, dims, R, x, b.y, 1.0, g.grad, damp, 1, symm_now, o->factors_norm, xs, b.F,
                       b.scal + 2, st));
    symm_now = symm1;
    RUN(tfx_cpd_residual_mixed(ctx, T, t_dtype, L, dims, b.F, TFX_F64, R, b.scal + 6, st));
    RUN(copies({{b.scal, pin}}, 8));
    HIPRUN(hipEventRecord(ev_scal, st));
    // The next iteration's Gram products and gradient depend only on the updated factors: they
    // go into the other set now, so the device works while the host reads this iteration's
    // scalars.  A dogleg step changes the factors (they are formed again after it); after a stop
    // they are left unused.
    const GradSet& gn = b.gs[1 - gcur];
    if (it + 1 < maxiter) {
      RUN(tfx_gramians(ctx, TFX_F64, L, dims, R, b.F, gn.Gr, gn.P1, gn.P2, gn.gamma, st));
      RUN(tfx_mttkrp_neg_grad(ctx, T, t_dtype, L, dims, b.F, TFX_F64, R, gn.P1, gn.N, gn.grad, prec, st));
      ahead = true;
    }
    gcur = 1 - gcur;
    HIPRUN(hipEventSynchronize(ev_scal));   // the scalars only, not the work queued after them
    double* old_x = x;
    std::swap(x, xs);
    error = sqrt(pin[6]) / Tsize;
    double gain_den = pin[2], dx2 = pin[3];
    const double x2 = pin[4], gmax = pin[5];
    if (it > 3 && error > o->tol_jump * old_error) {
      // compute_dogleg_steps (GaussNewton.py:713-763) from x - y along the CG step and the gradient
      ++dogleg_calls;
      int cur = 0;
      RUN(tfx_axpby(ctx, TFX_F64, n, 1.0, x, -1.0, b.y, b.dx[0], st));
      RUN(copies({{b.y, b.dy[0]}, {b.dx[0], b.dbx}, {b.y, b.dby}, {b.F, b.dbF}}, n));
      double derr = error, dbest = error, gain = 1.0, delta = 1.0;
      int count = 0;
      while (gain > 0) {
        const double dold = derr;
        RUN(tfx_dogleg_dots(ctx, TFX_F64, n, b.dy[cur], g.grad, b.jtjg, b.tscal, st));
        RUN(copies({{b.tscal, pin + 16}}, 4));
        HIPRUN(hipStreamSynchronize(st));
        double ca, cb;
        dogleg_coeffs(pin + 16, delta, &ca, &cb);
        const int nx = 1 - cur;
        RUN(tfx_axpby(ctx, TFX_F64, n, ca, b.dy[cur], cb, g.grad, b.dy[nx], st));
        RUN(tfx_dgn_update(ctx, TFX_F64, L, dims, R, b.dx[cur], b.dy[nx], 1.0, g.grad, damp, 0, symm1,
                           o->factors_norm, b.dx[nx], b.F, b.tscal + 4, st));
        RUN(tfx_cpd_residual_mixed(ctx, T, t_dtype, L, dims, b.F, TFX_F64, R, b.tscal + 8, st));
        RUN(copies({{b.tscal + 4, pin + 24}}, 5));
        HIPRUN(hipStreamSynchronize(st));
        derr = sqrt(pin[28]) / Tsize;
        gain = gain_ratio(dold, derr, Tsize, pin[24]);
        if (derr < dold) {
          RUN(copies({{b.dx[nx], b.dbx}, {b.dy[nx], b.dby}, {b.F, b.dbF}}, n));
          dbest = derr;
        }
        delta = update_delta(delta, gain, sqrt(pin[25]));
        cur = nx;
        if (++count > 10) break;
      }
      error = dbest;
      RUN(copies({{b.dbF, b.F}, {b.dbx, x}, {b.dby, b.y}}, n));
      // x - old_x statistics for this iteration's gain ratio and step size
      RUN(tfx_axpby(ctx, TFX_F64, n, 1.0, x, -1.0, old_x, b.d, st));
      RUN(tfx_dgn_update(ctx, TFX_F64, L, dims, R, old_x, b.d, 1.0, g.grad, damp, 0, 0, 0.0, b.sx, b.sF,
                         b.tscal, st));
      RUN(copies({{b.tscal, pin + 32}}, 4));
      HIPRUN(hipStreamSynchronize(st));
      gain_den = pin[32];
      dx2 = pin[33];
      ahead = false;   // the factors moved: the next iteration forms its products again
    }
    const double gr = gain_ratio(old_error, error, Tsize, gain_den);
    damp = update_damp(damp, damp_sched, gr, it);
    if (error < best_error) {
      best_error = error;
      RUN(copies({{b.F, b.best}}, n));
    }
    errors[it] = error;
    steps[it] = sqrt(dx2) / sqrt(x2);
    grads[it] = gmax;
    improv[it] = it == 0 ? errors[it] : fabs(errors[it] - errors[it - 1]);
    const int code = stop_code(it, errors, steps, improv, grads, tol, tol_step, tol_improv, tol_grad, cst, error,
                               Tsize);
    if (code >= 0) {
      stop = code;
      break;
    }
  }
  if (it >= maxiter) it = maxiter - 1;
  RUN(copies({{b.best, F_best}}, n));
  HIPRUN(hipStreamSynchronize(st));
  info[0] = it + 1;
  info[1] = stop;
  info[2] = dogleg_calls;
done:
#undef RUN
#undef HIPRUN
  *mt_pos = mt.pos;
  hipHostFree(pin);
  hipEventDestroy(ev_scal);
  return rc;
}
