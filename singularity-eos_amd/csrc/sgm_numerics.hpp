// Device (and host) numerics for the EOS / closure hot path on gfx950.
//
// Everything here is FP64 and compiled with -ffp-contract=off so that each
// multiply and add rounds exactly as written: the vector EOS calls, the table
// interpolation and the Newton closure reproduce the singularity-eos CPU
// arithmetic operation for operation (libm exp/cbrt/log aside).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#define SGM_HD __host__ __device__ __forceinline__

namespace sgm {
using Real = double;

// min/max with the exact semantics of std::min/std::max (first argument on
// ties and NaN handling of the comparison), not fmin/fmax.
SGM_HD Real smax(Real a, Real b) { return (a < b) ? b : a; }
SGM_HD Real smin(Real a, Real b) { return (b < a) ? b : a; }

// singularity-eos/base/robust_utils.hpp:26-68
namespace robust {
SGM_HD constexpr Real SMALL() { return 10 * 2.2250738585072014e-308; }
SGM_HD constexpr Real EPS() { return 10 * 2.220446049250313e-16; }
SGM_HD int sgn(Real v) { return (Real(0) <= v) - (v < Real(0)); }
SGM_HD Real ratio(Real a, Real b) { return a / (b + sgn(b) * SMALL()); }
} // namespace robust

SGM_HD bool is_normal(Real v) {
  const Real a = fabs(v);
  return a >= 2.2250738585072014e-308 && a <= 1.7976931348623157e308;
}
// singularity-eos/base/error_utils.hpp:28-73 (bad_value)
SGM_HD bool bad_value(Real v) {
  const bool good = (v == 0.0) || (is_normal(1.0e10 * v) && is_normal(v));
  return !good;
}
SGM_HD bool is_nan_or_inf(Real v) { return !(fabs(v) <= 1.7976931348623157e308); }

// NQT order-2 integer-aliased log2 / exp2 (singularity-eos/base/fast-math/logs.hpp:188-214,
// 282-292; arXiv:2206.08957). Integer arithmetic is exact on both host and device.
namespace nqt {
SGM_HD std::int64_t as_int(double f) { return __builtin_bit_cast(std::int64_t, f); }
SGM_HD double as_double(std::int64_t i) { return __builtin_bit_cast(double, i); }
constexpr std::int64_t one_as_int = (std::int64_t(1) << 62) - (std::int64_t(1) << 52);
constexpr double scale_down = 2.22044604925031e-16;
constexpr double scale_up = 4503599627370496.0;
constexpr std::int64_t mantissa_mask = (std::int64_t(1) << 52) - 1;
constexpr std::int64_t low_mask = (std::int64_t(1) << 26) - 1;

SGM_HD double lg(double x) {
  const std::int64_t xi = as_int(x) - one_as_int;
  const std::int64_t fr = xi & mantissa_mask;
  const std::int64_t hi = fr >> 26;
  const std::int64_t lo = fr & low_mask;
  const std::int64_t sq = hi * hi + ((hi * lo) >> 25);
  return static_cast<double>(xi + ((fr - sq) / 3)) * scale_down;
}
SGM_HD double pow2(double x) {
  const std::int64_t xi = static_cast<std::int64_t>(x * scale_up);
  const std::int64_t fr = xi & mantissa_mask;
  const std::int64_t fs = static_cast<std::int64_t>(
      67108864.0 * sqrt(static_cast<double>(std::int64_t(18014398509481984) - 3 * fr)));
  return as_double(xi + std::int64_t(9007199254740992) - fs - fr + one_as_int);
}
SGM_HD double log10(double x) { return 0.301029995663981195 * lg(x); }
SGM_HD double pow10(double x) { return pow2(3.321928094887362626 * x); }
} // namespace nqt

// Table coordinate transforms (singularity-eos/eos/eos_spiner_common.hpp:52-57)
SGM_HD Real to_log(Real x, Real offset) {
  return nqt::log10(fabs(smax(x, -offset) + offset) + robust::EPS());
}
SGM_HD Real from_log(Real lx, Real offset) { return nqt::pow10(lx) - offset; }

// ---------------------------------------------------------------------------
// 1-D root finding (singularity-eos/base/root-finding-1d/root_finding.hpp).
// Functors are called as f(x); every routine returns true on SUCCESS.
// ---------------------------------------------------------------------------
namespace root {
constexpr int NITER_MAX = 1000;
SGM_HD bool check_bracket(Real ya, Real yb) { return ya * yb <= 0.0; }

// regula_falsi (:154-240) including set_bracket (:118-151). g(x) = f(x) - target
template <typename F>
SGM_HD bool regula_falsi(const F &f, Real ytarget, Real guess, Real a, Real b, Real xtol,
                         Real ytol, Real &xroot, int *iters = nullptr) {
  Real ya = f(a) - ytarget;
  const Real yg = f(guess) - ytarget;
  Real yb;
  if (check_bracket(ya, yg)) {
    b = guess;
    yb = yg;
  } else {
    yb = f(b) - ytarget;
    if (check_bracket(yg, yb)) {
      a = guess;
      ya = yg;
    } else {
      // bracket search: midpoints of 2^level sub-intervals, 6 levels
      const Real dx = b - a;
      const Real a0 = a;
      bool found = false;
      for (int level = 0; level < 6 && !found; level++) {
        const int nlev = 1 << level;
        for (int i = 0; i < nlev; i++) {
          const Real x = a0 + (i + 0.5) * dx;
          const Real yx = f(x) - ytarget;
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
            found = true;
            break;
          }
        }
      }
      if (!found) {
        if (iters) *iters = 0;
        return false;
      }
    }
  }
  const Real sign = (ya < 0 ? 1.0 : -1.0);
  ya *= sign;
  yb *= sign;
  int b1 = 0, b2 = 0, it = 0;
  while (b - a > 2.0 * xtol && (fabs(ya) > ytol || fabs(yb) > ytol) && it < NITER_MAX) {
    const Real c = (a * yb - b * ya) / (yb - ya);
    if (c == a) {
      b = a;
      continue;
    } else if (c == b) {
      a = b;
      continue;
    }
    const Real yc = sign * (f(c) - ytarget);
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
    it++;
  }
  if (iters) *iters = it;
  xroot = 0.5 * (a + b);
  return it != NITER_MAX;
}

// secant (:430-510) with a central finite-difference slope
template <typename F>
SGM_HD bool secant(const F &f, Real ytarget, Real xguess, Real xmin, Real xmax, Real xtol,
                   Real ytol, Real &xroot) {
  Real x = xguess, x_last = xguess;
  for (int iter = 0; iter < NITER_MAX; ++iter) {
    x_last = x;
    const Real dx = fabs(1.e-7 * x) + xtol;
    const Real y = f(x) - ytarget;
    const Real yp = f(x + dx);
    const Real ym = f(x - dx);
    const Real dy = (yp - ym) / (2. * dx);
    x -= y / dy;
    if (x < xmin) x = xmin;
    if (x > xmax) x = xmax;
    if (is_nan_or_inf(x)) return false;
    if (fabs(x - x_last) / (fabs(x) + xtol) < xtol) break;
  }
  xroot = x;
  const Real y = f(x);
  const Real frac_error = fabs(y - ytarget) / (fabs(y) + ytol);
  return !(fabs(x - x_last) > xtol || fabs(frac_error) > ytol || is_nan_or_inf(x));
}

// bisect (:512-600) with the expanding bracket search
template <typename F>
SGM_HD bool bisect(const F &f, Real ytarget, Real xguess, Real xmin, Real xmax, Real xtol,
                   Real ytol, Real &xroot) {
  Real xl = 0, xr = 0, fl = 0, fr = 0;
  Real grow = 0.01;
  Real x = xguess;
  if (fabs(x) < xtol) x += 2. * xtol;
  for (int i = 0; i < NITER_MAX; ++i) {
    const Real dx = fabs(grow * x);
    xl = x - dx;
    xr = x + dx;
    fl = f(xl) - ytarget;
    fr = f(xr) - ytarget;
    grow *= 1.1;
    if (fl * fr < 0.0 || xl < xmin || xr > xmax) break;
    if (i > NITER_MAX - 2) return false;
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
      if (fl * fr > 0) return false;
    }
  }
  for (int i = 0; i < NITER_MAX; ++i) {
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
      return true;
    }
  }
  xroot = 0.5 * (xl + xr);
  return false;
}

// findRoot (:312-373): edge check, secant, then bisection
template <typename F>
SGM_HD bool find_root(const F &f, Real ytarget, Real xguess, Real xmin, Real xmax, Real xtol,
                      Real ytol, Real &xroot) {
  const Real fmax = f(xmax);
  if (fabs(fmax - ytarget) / (fabs(fmax) + ytol) < ytol) {
    xroot = xmax;
    return true;
  }
  const Real fmin = f(xmin);
  if (fabs(fmin - ytarget) / (fabs(fmin) + ytol) < ytol) {
    xroot = xmin;
    return true;
  }
  if (xguess >= xmax) xguess = xmax - xtol;
  if (xguess <= xmin) xguess = xmin + xtol;
  if (secant(f, ytarget, xguess, xmin, xmax, xtol, ytol, xroot)) return true;
  const bool ok = bisect(f, ytarget, xguess, xmin, xmax, xtol, ytol, xroot);
  if (is_nan_or_inf(xroot)) return false;
  if (xroot < xmin || xroot > xmax) return false;
  return ok;
}
} // namespace root

// ---------------------------------------------------------------------------
// exp and cbrt from IEEE +,-,*,/ only, so host and device produce the same
// bits (the Vinet model is the only caller on the hot path).
//  exp:  Cody-Waite reduction x = k ln2 + r with fdlibm's two-part ln2
//        (ln2_hi has 32 trailing zero bits, so k*ln2_hi is exact), |r| <= ln2/2,
//        degree-13 Taylor polynomial (truncation < 5e-18), scaling by 2^k.
//  cbrt: x = m 2^(3q+s), seed by a quadratic fit of cbrt on [1,8), five
//        Newton steps y <- (2y + a/y^2)/3, scaling by 2^q.
// Both are within a few ulp of the correctly rounded values.
// ---------------------------------------------------------------------------
namespace pmath {
SGM_HD double pow2i(int k) { // 2^k for -1022 <= k <= 1023
  return nqt::as_double(static_cast<std::int64_t>(k + 1023) << 52);
}
SGM_HD double exp(double x) {
  if (!(x == x)) return x;
  if (x > 709.782712893384) return __builtin_huge_val();
  if (x < -745.1332191019412) return 0.0;
  const double INV_LN2 = 1.4426950408889634;
  const double LN2_HI = 6.93147180369123816490e-01;
  const double LN2_LO = 1.90821492927058770002e-10;
  const double kd = floor(x * INV_LN2 + 0.5);
  const double r = (x - kd * LN2_HI) - kd * LN2_LO;
  double p = 1.6059043836821613e-10;
  p = p * r + 2.08767569878681e-09;
  p = p * r + 2.505210838544172e-08;
  p = p * r + 2.755731922398589e-07;
  p = p * r + 2.7557319223985893e-06;
  p = p * r + 2.48015873015873e-05;
  p = p * r + 0.0001984126984126984;
  p = p * r + 0.001388888888888889;
  p = p * r + 0.008333333333333333;
  p = p * r + 0.041666666666666664;
  p = p * r + 0.16666666666666666;
  p = p * r + 0.5;
  p = p * r + 1.0;
  p = p * r + 1.0;
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
SGM_HD double cbrt(double x) {
  if (x == 0.0 || !(x == x) || !(fabs(x) <= 1.7976931348623157e308)) return x;
  double a = fabs(x);
  int adj = 0;
  if (a < 2.2250738585072014e-308) {
    a *= 18014398509481984.0; // 2^54
    adj = -18;
  }
  const std::int64_t bits = nqt::as_int(a);
  const int e = (int)((bits >> 52) & 0x7ff) - 1023;
  const double m = nqt::as_double((bits & nqt::mantissa_mask) | (std::int64_t(1023) << 52));
  int q = e / 3, s = e % 3;
  if (s < 0) {
    s += 3;
    q -= 1;
  }
  const double mm = s == 0 ? m : (s == 1 ? 2.0 * m : 4.0 * m); // [1, 8)
  double y = (-0.011593381976926909 * mm + 0.23742467276201396) * mm + 0.8208312163657417;
#pragma unroll
  for (int i = 0; i < 5; ++i) y = (2.0 * y + mm / (y * y)) / 3.0;
  y *= pow2i(q + adj);
  return x < 0 ? -y : y;
}
} // namespace pmath

// Kahan–Neumaier summation over a register array
// (singularity-eos/closure/mixed_cell_models.hpp:287-305)
struct NeumaierSum {
  Real sum = 0, c = 0;
  SGM_HD void add(Real x) {
    const Real t = sum + x;
    if (fabs(sum) >= fabs(x)) {
      c += (sum - t) + x;
    } else {
      c += (x - t) + sum;
    }
    sum = t;
  }
  SGM_HD Real result() const { return sum + c; }
};

} // namespace sgm
