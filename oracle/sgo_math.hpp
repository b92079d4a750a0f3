// ORACLE — test infrastructure only. CPU restatement of singularity-eos's
// base numerics (robust arithmetic, NQT fast logs, 1-D root finding,
// compensated sums, small dense solves) used as the parity checker and as the
// reported host baseline. Nothing in the shipped library includes, links or
// calls this code; only tests/, __graft_entry__.smoke() and bench.py's
// cpu_baseline leg load it.
//
// Compiled with g++ -O3 -ffp-contract=off (no fused multiply-add, no
// fast-math) so that every floating-point operation happens in the order
// written here.
#pragma once

#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <limits>

#include "sgo_real.hpp"

namespace sgo {

// ---------------------------------------------------------------------------
// robust:: — singularity-eos/base/robust_utils.hpp:26-68
// ---------------------------------------------------------------------------
namespace robust {
inline constexpr Real SMALL() { return 10 * std::numeric_limits<Real>::min(); }
inline constexpr Real EPS() { return 10 * std::numeric_limits<Real>::epsilon(); }
// sgn(0) = +1 (robust_utils.hpp:56-63)
inline int sgn(const Real v) { return (Real(0) <= v) - (v < Real(0)); }
inline Real ratio(const Real a, const Real b) { return a / (b + sgn(b) * SMALL()); }
} // namespace robust

// ---------------------------------------------------------------------------
// error_utils::bad_value — singularity-eos/base/error_utils.hpp:28-73
// true when the value is not (zero, or normal and normal after *1e10)
// ---------------------------------------------------------------------------
inline bool bad_value(const Real v) {
  const bool good = (v == 0.0) || (std::isnormal(1.0e10 * v) && std::isnormal(v));
  return !good;
}

// ---------------------------------------------------------------------------
// FastMath NQT order-2, integer-aliased (the reference default build) —
// singularity-eos/base/fast-math/logs.hpp:113-125, 188-214, 282-292.
// Method: arXiv:2206.08957.
// ---------------------------------------------------------------------------
namespace fastmath {
inline std::int64_t as_int(double f) {
  std::int64_t i;
  std::memcpy(&i, &f, sizeof(i));
  return i;
}
inline double as_double(std::int64_t i) {
  double f;
  std::memcpy(&f, &i, sizeof(f));
  return f;
}
constexpr std::int64_t one = 1;
constexpr std::int64_t one_as_int = (one << 62) - (one << 52);
constexpr double scale_down = 2.22044604925031e-16;
constexpr double scale_up = (double)(one << 52);
constexpr std::int64_t mantissa_mask = (one << 52) - one;
constexpr std::int64_t low_mask = (one << 26) - 1;

inline double lg(const double x) {
  const std::int64_t x_as_int = as_int(x) - one_as_int;
  const std::int64_t frac_as_int = x_as_int & mantissa_mask;
  const std::int64_t frac_high = frac_as_int >> 26;
  const std::int64_t frac_low = frac_as_int & low_mask;
  const std::int64_t frac_squared = frac_high * frac_high + ((frac_high * frac_low) >> 25);
  return static_cast<double>(x_as_int + ((frac_as_int - frac_squared) / 3)) * scale_down;
}
inline double pow2(const double x) {
  constexpr std::int64_t a = 9007199254740992;  // 2 * 2^52
  constexpr double b = 67108864;                // 2^26
  constexpr std::int64_t c = 18014398509481984; // 4 * 2^52
  const std::int64_t x_as_int = static_cast<std::int64_t>(x * scale_up);
  const std::int64_t frac_as_int = x_as_int & mantissa_mask;
  const std::int64_t frac_sqrt =
      static_cast<std::int64_t>(b * std::sqrt(static_cast<double>(c - 3 * frac_as_int)));
  return as_double(x_as_int + a - frac_sqrt - frac_as_int + one_as_int);
}
inline double log10(const double x) { return 0.301029995663981195 * lg(x); }
inline double pow10(const double x) { return pow2(3.321928094887362626 * x); }
} // namespace fastmath

// exp and cbrt built from IEEE +,-,*,/ only. The reference calls std::exp and
// std::cbrt (eos_vinet.hpp:235-239); host and device math libraries differ in
// the last bit, so the oracle and the device both evaluate these two with the
// same published methods to stay bit-comparable (a CPU test bounds them
// against the host libm):
//  exp:  Cody-Waite reduction with fdlibm's two-part ln2, degree-13 Taylor.
//  cbrt: exponent split x = m 2^(3q+s), quadratic seed on [1,8), five Newton steps.
namespace pmath {
inline double pow2i(int k) { return fastmath::as_double(static_cast<std::int64_t>(k + 1023) << 52); }
inline double exp(double x) {
  if (!(x == x)) return x;
  if (x > 709.782712893384) return std::numeric_limits<double>::infinity();
  if (x < -745.1332191019412) return 0.0;
  const double INV_LN2 = 1.4426950408889634;
  const double LN2_HI = 6.93147180369123816490e-01;
  const double LN2_LO = 1.90821492927058770002e-10;
  const double kd = std::floor(x * INV_LN2 + 0.5);
  const double r = (x - kd * LN2_HI) - kd * LN2_LO;
  // 1/13!, 1/12!, ..., 1/2!, 1, 1
  static const double c[14] = {1.6059043836821613e-10, 2.08767569878681e-09,
                               2.505210838544172e-08,  2.755731922398589e-07,
                               2.7557319223985893e-06, 2.48015873015873e-05,
                               0.0001984126984126984,  0.001388888888888889,
                               0.008333333333333333,   0.041666666666666664,
                               0.16666666666666666,    0.5, 1.0, 1.0};
  double p = c[0];
  for (int i = 1; i < 14; ++i) p = p * r + c[i];
  int k = (int)kd;
  if (k > 1023) {
    p *= pow2i(1023);
    k -= 1023;
  } else if (k < -1022) {
    p *= pow2i(k + 54);
    k = -54;
  }
  return p * pow2i(k);
}
inline double cbrt(double x) {
  if (x == 0.0 || !(x == x) || !(std::fabs(x) <= std::numeric_limits<double>::max())) return x;
  double a = std::fabs(x);
  int adj = 0;
  if (a < std::numeric_limits<double>::min()) {
    a *= 18014398509481984.0; // 2^54
    adj = -18;
  }
  const std::int64_t bits = fastmath::as_int(a);
  const int e = (int)((bits >> 52) & 0x7ff) - 1023;
  const double m =
      fastmath::as_double((bits & fastmath::mantissa_mask) | (std::int64_t(1023) << 52));
  int q = e / 3, s = e % 3;
  if (s < 0) {
    s += 3;
    q -= 1;
  }
  const double mm = s == 0 ? m : (s == 1 ? 2.0 * m : 4.0 * m);
  double y = (-0.011593381976926909 * mm + 0.23742467276201396) * mm + 0.8208312163657417;
  for (int i = 0; i < 5; ++i) y = (2.0 * y + mm / (y * y)) / 3.0;
  y *= pow2i(q + adj);
  return x < 0 ? -y : y;
}
} // namespace pmath

#ifdef SGO_COUNT_FLOPS
// counted forms: NQT lg = 1 multiply (+ integer work), pow2 = 2 multiplies and
// a sqrt, log10/pow10 one more multiply; exp = 2 + floor + 4 (reduction) +
// 26 (Horner) + 1 (scale) = 34; cbrt = 4 (seed) + 5 x 5 (Newton) + 1 = 30
namespace fastmath {
inline Real lg(Real x) { ++g_flops; return Real(lg(val(x))); }
inline Real pow2(Real x) { g_flops += 3; return Real(pow2(val(x))); }
inline Real log10(Real x) { g_flops += 2; return Real(log10(val(x))); }
inline Real pow10(Real x) { g_flops += 4; return Real(pow10(val(x))); }
} // namespace fastmath
namespace pmath {
inline Real exp(Real x) { g_flops += 34; return Real(exp(val(x))); }
inline Real cbrt(Real x) { g_flops += 30; return Real(cbrt(val(x))); }
} // namespace pmath
#endif

// The reference's Vinet calls std::exp and std::cbrt
// (singularity-eos/eos/eos_vinet.hpp:235-239). The oracle, like the device
// code, calls the portable pmath routines, so that host and device agree bit
// for bit; built with -DSGO_LIBM_EXP_CBRT it calls the host libm instead, to
// measure what that substitution changes.
namespace vmath {
#ifdef SGO_LIBM_EXP_CBRT
inline double exp(double x) { return std::exp(x); }
inline double cbrt(double x) { return std::cbrt(x); }
#else
using pmath::cbrt;
using pmath::exp;
#endif
} // namespace vmath

// spiner_common::to_log / from_log — singularity-eos/eos/eos_spiner_common.hpp:52-57
inline Real to_log(const Real x, const Real offset) {
  return fastmath::log10(std::abs(std::max(x, -offset) + offset) + robust::EPS());
}
inline Real from_log(const Real lx, const Real offset) {
  return fastmath::pow10(lx) - offset;
}

// ---------------------------------------------------------------------------
// RootFinding1D — singularity-eos/base/root-finding-1d/root_finding.hpp
// ---------------------------------------------------------------------------
namespace root {
enum Status { SUCCESS = 0, FAIL = 1 };
constexpr int SECANT_NITER_MAX = 1000;
constexpr int BISECT_NITER_MAX = 1000;
constexpr int BISECT_REG_MAX = 1000;

// root_finding.hpp:114-116
inline bool check_bracket(const Real ya, const Real yb) { return ya * yb <= 0.0; }

// root_finding.hpp:118-151 (f is already shifted by the target)
template <typename F>
inline bool set_bracket(const F &f, Real &a, const Real guess, Real &b, Real &ya,
                        const Real yg, Real &yb) {
  constexpr int max_search_depth = 6;
  const Real dx = b - a;
  for (int level = 0; level < max_search_depth; level++) {
    const int nlev = 1 << level;
    for (int i = 0; i < nlev; i++) {
      const Real x = a + (i + 0.5) * dx;
      const Real yx = f(x);
      if (check_bracket(yx, yg)) {
        if (x < guess) {
          a = x;
          ya = yx;
          b = guess;
          yb = yg;
        } else {
          a = guess;
          ya = yg;
          b = x;
          yb = yx;
        }
        return true;
      }
    }
  }
  return false;
}

// root_finding.hpp:154-240. `iters` reports the loop count (for histograms).
template <typename F>
inline Status regula_falsi(const F &f, const Real ytarget, const Real guess, Real a,
                           Real b, const Real xtol, const Real ytol, Real &xroot,
                           int *iters = nullptr) {
  constexpr int max_iter = SECANT_NITER_MAX;
  auto func = [&](const Real x) { return f(x) - ytarget; };
  Real ya = func(a);
  Real yg = func(guess);
  Real yb;
  if (check_bracket(ya, yg)) {
    b = guess;
    yb = yg;
  } else {
    yb = func(b);
    if (check_bracket(yg, yb)) {
      a = guess;
      ya = yg;
    } else {
      if (!set_bracket(func, a, guess, b, ya, yg, yb)) {
        if (iters) *iters = 0;
        return FAIL;
      }
    }
  }
  const Real sign = (ya < 0 ? 1.0 : -1.0);
  ya *= sign;
  yb *= sign;
  int b1 = 0, b2 = 0, iteration_count = 0;
  while (b - a > 2.0 * xtol && (std::abs(ya) > ytol || std::abs(yb) > ytol) &&
         iteration_count < max_iter) {
    const Real c = (a * yb - b * ya) / (yb - ya);
    if (c == a) {
      b = a;
      continue;
    } else if (c == b) {
      a = b;
      continue;
    }
    const Real yc = sign * func(c);
    if (yc > 0.0) {
      b = c;
      yb = yc;
      b1++;
      ya *= (b1 > 1 ? 0.5 : 1.0);
      b2 = 0;
    } else if (yc < 0.0) {
      a = c;
      ya = yc;
      b2++;
      yb *= (b2 > 1 ? 0.5 : 1.0);
      b1 = 0;
    } else {
      a = c;
      b = c;
    }
    iteration_count++;
  }
  Status status = SUCCESS;
  if (iteration_count == max_iter) status = FAIL;
  if (iters) *iters = iteration_count;
  xroot = 0.5 * (a + b);
  return status;
}

// root_finding.hpp:430-510
template <typename F>
inline Status secant(const F &f, const Real ytarget, const Real xguess, const Real xmin,
                     const Real xmax, const Real xtol, const Real ytol, Real &xroot) {
  Real dx, x_last, y, yp, ym, dyNum, dyDen, dy;
  Real x = xguess;
  int iter = 0;
  for (iter = 0; iter < SECANT_NITER_MAX; ++iter) {
    x_last = x;
    dx = std::fabs(1.e-7 * x) + xtol;
    y = f(x) - ytarget;
    yp = f(x + dx);
    ym = f(x - dx);
    dyNum = yp - ym;
    dyDen = (2. * dx);
    dy = dyNum / dyDen;
    x -= y / dy;
    if (x < xmin) x = xmin;
    if (x > xmax) x = xmax;
    if (std::isnan(x) || std::isinf(x)) return FAIL;
    if (std::fabs(x - x_last) / (std::fabs(x) + xtol) < xtol) break;
  }
  ++iter;
  xroot = x;
  y = f(x);
  const Real frac_error = std::fabs(y - ytarget) / (std::fabs(y) + ytol);
  const bool secant_failed = (std::abs(x - x_last) > xtol ||
                              std::abs(frac_error) > ytol || std::isnan(x) ||
                              std::isinf(x));
  return secant_failed ? FAIL : SUCCESS;
}

// root_finding.hpp:512-600
template <typename F>
inline Status bisect(const F &f, const Real ytarget, const Real xguess, const Real xmin,
                     const Real xmax, const Real xtol, const Real ytol, Real &xroot) {
  Real xl, xr, fl, fr, dx;
  Real grow = 0.01;
  Real x = xguess;
  if (std::fabs(x) < xtol) x += 2. * xtol;
  for (int i = 0; i < BISECT_REG_MAX; ++i) {
    dx = std::fabs(grow * x);
    xl = x - dx;
    xr = x + dx;
    fl = f(xl) - ytarget;
    fr = f(xr) - ytarget;
    grow *= 1.1;
    if (fl * fr < 0.0 || xl < xmin || xr > xmax) break;
    if (i > BISECT_REG_MAX - 2) return FAIL;
  }
  if (xr > xmax) {
    xr = xmax;
    fr = f(xr) - ytarget;
  }
  if (xl < xmin) {
    xl = xmin;
    fl = f(xl) - ytarget;
  }
  if (fl * fr > 0) {
    xl = xmin;
    fl = f(xl) - ytarget;
    if (fl * fr > 0) {
      xr = xmax;
      fr = f(xr) - ytarget;
      if (fl * fr > 0) return FAIL;
    }
  }
  for (int i = 0; i < BISECT_NITER_MAX; ++i) {
    const Real xm = 0.5 * (xl + xr);
    const Real fm = f(xm) - ytarget;
    if (fl * fm <= 0) {
      xr = xm;
      fr = fm;
    } else {
      xl = xm;
      fl = fm;
    }
    if (xr - xl < xtol) {
      xroot = 0.5 * (xl + xr);
      return SUCCESS;
    }
  }
  xroot = 0.5 * (xl + xr);
  return FAIL;
}

// root_finding.hpp:312-373
template <typename F>
inline Status findRoot(const F &f, const Real ytarget, Real xguess, const Real xmin,
                       const Real xmax, const Real xtol, const Real ytol, Real &xroot) {
  const Real fmax = f(xmax);
  const Real errmax = std::fabs(fmax - ytarget) / (std::fabs(fmax) + ytol);
  if (errmax < ytol) {
    xroot = xmax;
    return SUCCESS;
  }
  const Real fmin = f(xmin);
  const Real errmin = std::fabs(fmin - ytarget) / (std::fabs(fmin) + ytol);
  if (errmin < ytol) {
    xroot = xmin;
    return SUCCESS;
  }
  if (xguess >= xmax) xguess = xmax - xtol;
  if (xguess <= xmin) xguess = xmin + xtol;
  Status status = secant(f, ytarget, xguess, xmin, xmax, xtol, ytol, xroot);
  if (status == SUCCESS) return status;
  status = bisect(f, ytarget, xguess, xmin, xmax, xtol, ytol, xroot);
  if (std::isnan(xroot) || std::isinf(xroot)) return FAIL;
  if (xroot < xmin) return FAIL;
  if (xroot > xmax) return FAIL;
  return status;
}
} // namespace root

// ---------------------------------------------------------------------------
// mix_impl helpers — singularity-eos/closure/mixed_cell_models.hpp:153-305
// ---------------------------------------------------------------------------
namespace mix {
// mixed_cell_models.hpp:155-158
inline bool isfinite_(const Real a) { return (a == a) && ((a == 0) || (a != 2 * a)); }

inline bool check_nans(const Real *a, const int n) {
  bool ok = true;
  for (int i = 0; i < n; ++i)
    if (!isfinite_(a[i])) ok = false;
  return ok;
}

// Kahan–Neumaier sum, mixed_cell_models.hpp:287-305
template <typename Op>
inline Real sum_neumaier_op(const Real *data, int n, int offset, int iskip, Op op) {
  Real sum = 0, c = 0;
  for (int i = 0; i < n; ++i) {
    if (iskip >= 0 && i == iskip) continue;
    const Real x = op(data[i + offset]);
    const Real t = sum + x;
    if (std::abs(sum) >= std::abs(x)) {
      c += (sum - t) + x;
    } else {
      c += (x - t) + sum;
    }
    sum = t;
  }
  return sum + c;
}
inline Real sum_neumaier(const Real *data, int n, int offset = 0, int iskip = -1) {
  return sum_neumaier_op(data, n, offset, iskip, [](Real x) { return x; });
}

// Column-pivoted Householder QR solve for n >= 3. The reference delegates to
// Eigen's ColPivHouseholderQR (mixed_cell_models.hpp:265-273); Eigen is not
// available here, so this is a restatement of the published algorithm
// (Golub & Van Loan 5.4.1: pivot on the largest remaining column norm, first
// index on ties; LAPACK xGEQP3 norm downdating with a sqrt(eps) recompute
// threshold; Householder sign convention beta = -sign(c0)*||x||; pivots whose
// squared norm falls below (eps*max_norm)^2/n*(n-k) are treated as zero and
// their unknowns set to 0). Matrix is row-major n x n.
// perm_out (optional) receives the pivot order: the original column at each
// position of R.
inline void qr_colpiv_solve(const int n, Real *a, Real *b, Real *x, int *perm_out = nullptr) {
  constexpr int NMAX = 128;  // neq <= 2 * 64 materials
  int perm[NMAX];
  Real cn_upd[NMAX], cn_dir[NMAX];
  for (int j = 0; j < n; ++j) {
    perm[j] = j;
    Real s = 0;
    for (int i = 0; i < n; ++i) s += a[i * n + j] * a[i * n + j];
    cn_dir[j] = std::sqrt(s);
    cn_upd[j] = cn_dir[j];
  }
  const Real eps = std::numeric_limits<Real>::epsilon();
  const Real norm_downdate_threshold = std::sqrt(eps);
  int nonzero_pivots = n;
  Real threshold_helper = 0;
  for (int k = 0; k < n; ++k) {
    // pivot: largest updated column norm among k..n-1 (first on ties)
    int big = k;
    Real bign = cn_upd[k];
    for (int j = k + 1; j < n; ++j) {
      if (cn_upd[j] > bign) {
        bign = cn_upd[j];
        big = j;
      }
    }
    const Real big_sq = bign * bign;
    if (k == 0) threshold_helper = big_sq * eps * eps / n;
    if (nonzero_pivots == n && big_sq < threshold_helper * (n - k)) nonzero_pivots = k;
    if (big != k) {
      for (int i = 0; i < n; ++i) std::swap(a[i * n + k], a[i * n + big]);
      std::swap(perm[k], perm[big]);
      std::swap(cn_upd[k], cn_upd[big]);
      std::swap(cn_dir[k], cn_dir[big]);
    }
    // Householder vector for column k, rows k..n-1 (Eigen makeHouseholder)
    Real tailsq = 0;
    for (int i = k + 1; i < n; ++i) tailsq += a[i * n + k] * a[i * n + k];
    const Real c0 = a[k * n + k];
    Real tau, beta;
    if (tailsq <= std::numeric_limits<Real>::min()) {
      tau = 0;
      beta = c0;
      for (int i = k + 1; i < n; ++i) a[i * n + k] = 0;
    } else {
      beta = std::sqrt(c0 * c0 + tailsq);
      if (c0 >= 0) beta = -beta;
      const Real den = c0 - beta;
      for (int i = k + 1; i < n; ++i) a[i * n + k] = a[i * n + k] / den;
      tau = (beta - c0) / beta;
    }
    a[k * n + k] = beta;
    // apply H = I - tau v v^T (v = [1, a[k+1..,k]]) to trailing columns
    // tmp = v_tail . col_tail + row0; row0 -= tau*tmp; tail -= (tau*v)*tmp
    for (int j = k + 1; j < n; ++j) {
      Real s = 0;
      for (int i = k + 1; i < n; ++i) s += a[i * n + k] * a[i * n + j];
      s += a[k * n + j];
      a[k * n + j] -= tau * s;
      for (int i = k + 1; i < n; ++i) a[i * n + j] -= (tau * a[i * n + k]) * s;
    }
    // the same reflector applied to the right-hand side (Q^T b)
    {
      Real s = 0;
      for (int i = k + 1; i < n; ++i) s += a[i * n + k] * b[i];
      s += b[k];
      b[k] -= tau * s;
      for (int i = k + 1; i < n; ++i) b[i] -= (tau * a[i * n + k]) * s;
    }
    // norm downdate
    for (int j = k + 1; j < n; ++j) {
      if (cn_upd[j] != 0) {
        Real t = std::abs(a[k * n + j]) / cn_upd[j];
        t = (1 + t) * (1 - t);
        if (t < 0) t = 0;
        const Real r = cn_upd[j] / cn_dir[j];
        const Real t2 = t * r * r;
        if (t2 <= norm_downdate_threshold) {
          Real s = 0;
          for (int i = k + 1; i < n; ++i) s += a[i * n + j] * a[i * n + j];
          cn_dir[j] = std::sqrt(s);
          cn_upd[j] = cn_dir[j];
        } else {
          cn_upd[j] *= std::sqrt(t);
        }
      }
    }
  }
  // solve with the leading nonzero_pivots x nonzero_pivots block of R
  // (column-oriented back substitution), zero the remaining unknowns
  const int rank = nonzero_pivots;
  Real z[NMAX];
  for (int i = 0; i < rank; ++i) z[i] = b[i];
  for (int i = rank - 1; i >= 0; --i) {
    z[i] = z[i] / a[i * n + i];
    for (int r = 0; r < i; ++r) z[r] -= z[i] * a[r * n + i];
  }
  for (int i = 0; i < n; ++i) x[i] = 0;
  for (int i = 0; i < rank; ++i) x[perm[i]] = z[i];
  if (perm_out)
    for (int i = 0; i < n; ++i) perm_out[i] = perm[i];
}

// solve_Ax_b_wscr — mixed_cell_models.hpp:176-278. a is row-major n x n,
// b is overwritten with x. Returns false on an all-zero row or non-finite x.
inline bool solve_Ax_b_wscr(const int n, Real *a, Real *b, Real *scr) {
  for (int row = 0; row < n; ++row) {
    Real maxabs = 0;
    for (int col = 0; col < n; ++col) maxabs = std::max(std::abs(a[row * n + col]), maxabs);
    if (maxabs == 0) return false;
    for (int col = 0; col < n; ++col) a[row * n + col] /= maxabs;
    b[row] /= maxabs;
  }
  if (n == 1) {
    b[0] = robust::ratio(b[0], a[0]);
  } else if (n == 2) {
    Real *x = scr;
    const Real det = a[0] * a[3] - a[1] * a[2];
    x[0] = robust::ratio(a[3] * b[0] - a[1] * b[1], det);
    x[1] = robust::ratio(a[0] * b[1] - a[2] * b[0], det);
    b[0] = x[0];
    b[1] = x[1];
  } else {
    Real x[128];
    qr_colpiv_solve(n, a, b, x);
    for (int i = 0; i < n; ++i) b[i] = x[i];
  }
  return check_nans(b, n);
}
} // namespace mix
} // namespace sgo
