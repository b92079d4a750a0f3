// ORACLE — test infrastructure only. The scalar type of the restatement.
//
// Normally `double`. Built with -DSGO_COUNT_FLOPS it becomes CReal, a double
// that counts every floating-point +, -, *, / and every sqrt/cbrt/exp/log/floor
// (one each) in a thread-local counter: the algorithmic FP64 operation count of
// the closure, used as the denominator of the compute roofline. Comparisons,
// selections, copies and integer work are not counted.
#pragma once

#include <cmath>
#include <cstdint>
#include <limits>

namespace sgo {
#ifndef SGO_COUNT_FLOPS
using Real = double;
inline double val(double x) { return x; }
#else
inline thread_local std::uint64_t g_flops = 0;
struct CReal {
  double v = 0;
  constexpr CReal() = default;
  constexpr CReal(double x) : v(x) {}
  constexpr CReal(int x) : v(x) {}
  constexpr CReal(long x) : v((double)x) {}
  explicit operator double() const { return v; }
  CReal &operator+=(const CReal &o) { ++g_flops; v += o.v; return *this; }
  CReal &operator-=(const CReal &o) { ++g_flops; v -= o.v; return *this; }
  CReal &operator*=(const CReal &o) { ++g_flops; v *= o.v; return *this; }
  CReal &operator/=(const CReal &o) { ++g_flops; v /= o.v; return *this; }
  CReal operator-() const { return CReal(-v); }
};
inline CReal operator+(CReal a, CReal b) { ++g_flops; return CReal(a.v + b.v); }
inline CReal operator-(CReal a, CReal b) { ++g_flops; return CReal(a.v - b.v); }
inline CReal operator*(CReal a, CReal b) { ++g_flops; return CReal(a.v * b.v); }
inline CReal operator/(CReal a, CReal b) { ++g_flops; return CReal(a.v / b.v); }
#define SGO_CMP(op)                                                                              \
  inline bool operator op(CReal a, CReal b) { return a.v op b.v; }                                \
  inline bool operator op(CReal a, double b) { return a.v op b; }                                 \
  inline bool operator op(double a, CReal b) { return a op b.v; }                                 \
  inline bool operator op(CReal a, int b) { return a.v op b; }                                    \
  inline bool operator op(int a, CReal b) { return a op b.v; }
SGO_CMP(<)
SGO_CMP(>)
SGO_CMP(<=)
SGO_CMP(>=)
SGO_CMP(==)
SGO_CMP(!=)
#undef SGO_CMP
#define SGO_MIX(op)                                                                              \
  inline CReal operator op(CReal a, double b) { return a op CReal(b); }                          \
  inline CReal operator op(double a, CReal b) { return CReal(a) op b; }                          \
  inline CReal operator op(CReal a, int b) { return a op CReal(b); }                             \
  inline CReal operator op(int a, CReal b) { return CReal(a) op b; }                             \
  inline CReal operator op(CReal a, long b) { return a op CReal(b); }                            \
  inline CReal operator op(long a, CReal b) { return CReal(a) op b; }
SGO_MIX(+)
SGO_MIX(-)
SGO_MIX(*)
SGO_MIX(/)
#undef SGO_MIX
inline double val(CReal x) { return x.v; }
using Real = CReal;
#endif
} // namespace sgo

#ifdef SGO_COUNT_FLOPS
namespace std {
inline sgo::CReal abs(sgo::CReal x) { return sgo::CReal(std::fabs(x.v)); }
inline sgo::CReal fabs(sgo::CReal x) { return sgo::CReal(std::fabs(x.v)); }
inline sgo::CReal sqrt(sgo::CReal x) { ++sgo::g_flops; return sgo::CReal(std::sqrt(x.v)); }
inline sgo::CReal log(sgo::CReal x) { ++sgo::g_flops; return sgo::CReal(std::log(x.v)); }
inline sgo::CReal floor(sgo::CReal x) { ++sgo::g_flops; return sgo::CReal(std::floor(x.v)); }
inline sgo::CReal atan(sgo::CReal x) { ++sgo::g_flops; return sgo::CReal(std::atan(x.v)); }
inline bool isnan(sgo::CReal x) { return std::isnan(x.v); }
inline bool isinf(sgo::CReal x) { return std::isinf(x.v); }
inline bool isnormal(sgo::CReal x) { return std::isnormal(x.v); }
template <>
struct numeric_limits<sgo::CReal> : numeric_limits<double> {};
// std::max / std::min semantics for mixed double / CReal arguments
inline sgo::CReal max(sgo::CReal a, double b) { return (a.v < b) ? sgo::CReal(b) : a; }
inline sgo::CReal max(double a, sgo::CReal b) { return (a < b.v) ? b : sgo::CReal(a); }
inline sgo::CReal min(sgo::CReal a, double b) { return (b < a.v) ? sgo::CReal(b) : a; }
inline sgo::CReal min(double a, sgo::CReal b) { return (b.v < a) ? b : sgo::CReal(a); }
} // namespace std
#endif
