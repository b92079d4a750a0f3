// ORACLE — test infrastructure only. CPU restatement of the singularity-eos
// equations of state on the hot path: IdealGas, StiffGas, Gruneisen, Vinet and
// SpinerEOSDependsRhoT (with a restatement of spiner's regular-grid linear
// interpolation), plus a tagged-union stand-in for singularity::Variant.
// Never linked into the shipped library.
#pragma once

#include "sgo_math.hpp"

#include <vector>

namespace sgo {

// thermalqs bits — singularity-eos/base/constants.hpp:22-35
namespace thermalqs {
constexpr unsigned long none = 0;
constexpr unsigned long density = 1 << 0;
constexpr unsigned long specific_internal_energy = 1 << 1;
constexpr unsigned long pressure = 1 << 2;
constexpr unsigned long temperature = 1 << 3;
constexpr unsigned long specific_heat = 1 << 4;
constexpr unsigned long bulk_modulus = 1 << 5;
} // namespace thermalqs
constexpr Real ROOM_TEMPERATURE = 293; // constants.hpp:41
constexpr Real ATMOSPHERIC_PRESSURE = 1e6;
enum TableStatus { OnTable = 0, OffBottom = 1, OffTop = 2 }; // constants.hpp:39

// Error flags raised where the reference aborts (EOS_ERROR / PORTABLE_*THROW*).
// The oracle records them per call instead of terminating.
struct ErrFlag {
  int code = 0;
};
// Error bits, OR'd per point or cell (the conditions the reference aborts or
// throws on, plus the two solver bailouts of PTESolver).
enum ErrCode {
  ERR_NONE = 0,
  ERR_GRUN_MAXRHO = 1,
  ERR_GRUN_PT = 2,
  ERR_FILLEOS = 4,
  ERR_LINSOLVE = 8,
  ERR_SMALLSTEP = 16
};

// ---------------------------------------------------------------------------
// IdealGas — singularity-eos/eos/eos_ideal.hpp:38-262
// ---------------------------------------------------------------------------
struct IdealGas {
  Real Cv, gm1;
  static constexpr unsigned long preferred =
      thermalqs::density | thermalqs::specific_internal_energy | thermalqs::temperature |
      thermalqs::pressure; // eos_ideal.hpp:227-229
  static Real mymax(Real a, Real b) { return a > b ? a : b; } // MYMAX, eos_ideal.hpp:32
  Real T_re(Real rho, Real sie) const { return mymax(0.0, sie / Cv); }
  Real e_rT(Real rho, Real T) const { return mymax(0.0, Cv * T); }
  Real P_rT(Real rho, Real T) const { return mymax(0.0, gm1 * rho * Cv * T); }
  Real P_re(Real rho, Real sie) const { return mymax(0.0, gm1 * rho * sie); }
  Real cv() const { return Cv; }
  Real bmod_rT(Real rho, Real T) const { return mymax(0.0, (gm1 + 1) * gm1 * rho * Cv * T); }
  Real bmod_re(Real rho, Real sie) const { return mymax(0.0, (gm1 + 1) * gm1 * rho * sie); }
  void rhoe_PT(Real press, Real temp, Real &rho, Real &sie) const {
    sie = mymax(0.0, Cv * temp);
    rho = mymax(0.0, press / (gm1 * sie));
  }
  // eos_ideal.hpp:236-262
  void fill(Real &rho, Real &temp, Real &sie, Real &press, Real &cv_, Real &bmod,
            unsigned long output, ErrFlag &err) const {
    using namespace thermalqs;
    if (output & density && output & specific_internal_energy) {
      if (output & pressure || output & temperature) {
        err.code |= ERR_FILLEOS;
        return;
      }
      rhoe_PT(press, temp, rho, sie);
    }
    if (output & pressure && output & specific_internal_energy) {
      if (output & density || output & temperature) {
        err.code |= ERR_FILLEOS;
        return;
      }
      sie = e_rT(rho, temp);
    }
    if (output & temperature && output & specific_internal_energy) {
      sie = press / (gm1 * rho);
    }
    if (output & pressure) press = P_re(rho, sie);
    if (output & temperature) temp = T_re(rho, sie);
    if (output & bulk_modulus) bmod = bmod_re(rho, sie);
    if (output & specific_heat) cv_ = Cv;
  }
};

// ---------------------------------------------------------------------------
// StiffGas — singularity-eos/eos/eos_stiff.hpp:36-229
// ---------------------------------------------------------------------------
struct StiffGas {
  Real Cv, gm1, Pinf, qq;
  static constexpr unsigned long preferred =
      thermalqs::density | thermalqs::specific_internal_energy;
  Real T_re(Real rho, Real sie) const {
    return std::max(robust::SMALL(), robust::ratio(rho * (sie - qq) - Pinf, rho * Cv));
  }
  Real e_rT(Real rho, Real T) const {
    return std::max(qq, robust::ratio(rho * Cv * T + Pinf, rho) + qq);
  }
  Real P_rT(Real rho, Real T) const { return std::max(-Pinf, gm1 * rho * Cv * T - Pinf); }
  Real P_re(Real rho, Real sie) const {
    return std::max(-Pinf, gm1 * rho * (sie - qq) - (gm1 + 1.0) * Pinf);
  }
  Real bmod_rT(Real rho, Real T) const {
    return std::max(robust::SMALL(), gm1 * (gm1 + 1.0) * rho * Cv * T);
  }
  Real bmod_re(Real rho, Real sie) const {
    return std::max(robust::SMALL(), gm1 * (gm1 + 1.0) * (rho * (sie - qq) - Pinf));
  }
  void rhoe_PT(Real press, Real temp, Real &rho, Real &sie) const {
    sie = std::max(qq, robust::ratio(press + (gm1 + 1.0) * Pinf, press + Pinf) * Cv * temp +
                           qq);
    rho = std::max(robust::SMALL(), robust::ratio(press + Pinf, gm1 * Cv * temp));
  }
  // eos_stiff.hpp:203-229
  void fill(Real &rho, Real &temp, Real &sie, Real &press, Real &cv_, Real &bmod,
            unsigned long output, ErrFlag &err) const {
    using namespace thermalqs;
    if (output & density && output & specific_internal_energy) {
      if (output & pressure || output & temperature) {
        err.code |= ERR_FILLEOS;
        return;
      }
      rhoe_PT(press, temp, rho, sie);
    }
    if (output & pressure && output & specific_internal_energy) {
      if (output & density || output & temperature) {
        err.code |= ERR_FILLEOS;
        return;
      }
      sie = e_rT(rho, temp);
    }
    if (output & temperature && output & specific_internal_energy) {
      sie = robust::ratio(press + (gm1 + 1.0) * Pinf, gm1 * rho) + qq;
    }
    if (output & pressure) press = P_re(rho, sie);
    if (output & temperature) temp = T_re(rho, sie);
    if (output & bulk_modulus) bmod = bmod_re(rho, sie);
    if (output & specific_heat) cv_ = Cv;
  }
};

// ---------------------------------------------------------------------------
// Gruneisen (Steinberg) — singularity-eos/eos/eos_gruneisen.hpp:36-549
// ---------------------------------------------------------------------------
struct Gruneisen {
  Real C0, s1, s2, s3, G0, b, rho0, T0, P0, Cv, rho_max;
  static constexpr unsigned long preferred =
      thermalqs::density | thermalqs::specific_internal_energy;
  static constexpr Real RHOMAX_SAFETY = 0.99;

  // eos_gruneisen.hpp:231-329
  static Real ComputeRhoMax(Real s1, Real s2, Real s3, Real rho0) {
    auto poly = [=](Real eta) {
      return 1 - s1 * eta - s2 * (eta * eta) - s3 * ((eta * eta) * eta);
    };
    const Real EPS = robust::EPS();
    auto near0 = [&](Real v) { return std::abs(v) < EPS; };
    auto find_min_bounded = [](Real v1, Real v2, Real mn, Real mx) {
      const Real minval = std::min(v1, v2), maxval = std::max(v1, v2);
      if (minval > mn && minval < mx) return minval;
      if (maxval > mn && maxval < mx) return maxval;
      return Real(-std::numeric_limits<double>::infinity());
    };
    Real root = -1.;
    if (near0(s2) && near0(s3) && s1 > 1.) {
      root = 1. / s1;
    } else if (near0(s3) && !near0(s2)) {
      const Real disc = s1 * s1 + 4. * s2;
      if (disc < 0.) {
        root = -1.;
      } else {
        const Real r1 = (s1 + std::sqrt(disc)) / (-2. * s2);
        const Real r2 = (s1 - std::sqrt(disc)) / (-2. * s2);
        root = find_min_bounded(r1, r2, 0., 1.);
      }
    } else if (!near0(s3)) {
      const Real m2 = -s2;
      const Real disc = -18. * s3 * s2 * s1 - 4. * ((m2 * m2) * m2) + (s2 * s2) * (s1 * s1) -
                        4. * s3 * ((s1 * s1) * s1) - 27. * (s3 * s3);
      Real minbound = 0., maxbound = 1.;
      if (disc >= 0) {
        const Real e1 = (s2 + std::sqrt(s2 * s2 - 3. * s3 * s1)) / (-3. * s3);
        const Real e2 = (s2 - std::sqrt(s2 * s2 - 3. * s3 * s1)) / (-3. * s3);
        const Real mne = std::min(e1, e2), mxe = std::max(e1, e2);
        if (s3 < 0.) {
          minbound = std::max(mne, minbound);
          maxbound = std::min(std::max(mxe, minbound), maxbound);
        } else {
          if (mne > 0) maxbound = std::min(mne, maxbound);
        }
      }
      if (poly(minbound) * poly(maxbound) < 0.) {
        const Real xtol = 100 * EPS, ytol = 100. / 10 * EPS;
        root::regula_falsi(poly, 0., 0.001, minbound, maxbound, xtol, ytol, root);
      } else if (near0(poly(minbound))) {
        root = minbound;
      } else if (near0(poly(maxbound))) {
        root = maxbound;
      } else {
        root = -1.;
      }
    }
    if (root > 0. && root < 1.) return rho0 / ((1 - root) + EPS);
    return std::numeric_limits<Real>::infinity();
  }
  static Gruneisen make(Real C0, Real s1, Real s2, Real s3, Real G0, Real b, Real rho0,
                        Real T0, Real P0, Real Cv) {
    return Gruneisen{C0, s1, s2, s3, G0, b, rho0, T0, P0, Cv,
                     RHOMAX_SAFETY * ComputeRhoMax(s1, s2, s3, rho0)};
  }
  // eos_gruneisen.hpp:191-195
  Real Gamma(Real rho_in) const {
    const Real rho = std::min(rho_in, rho_max);
    return rho < rho0 ? G0 : G0 * rho0 / rho + b * (1 - rho0 / rho);
  }
  // eos_gruneisen.hpp:331-351
  Real dPres_drho_e(Real rho_in, Real sie) const {
    const Real rho = std::min(rho_in, rho_max);
    if (rho < rho0) {
      return (C0 * C0) + Gamma(rho) * sie;
    }
    const Real eta = 1 - rho0 / rho;
    const Real s = s1 * eta + s2 * (eta * eta) + s3 * ((eta * eta) * eta);
    const Real ds = s1 + 2 * s2 * eta + 3 * s3 * (eta * eta);
    const Real deta = rho0 / (rho * rho);
    const Real dGam = (b - G0) * deta;
    const Real oms = 1 - s;
    const Real P_H = P0 + (C0 * C0) * rho0 * eta / (oms * oms);
    const Real dP_H = (C0 * C0) * rho0 / (oms * oms) * deta * (1 + 2 * eta * ds / (1 - s));
    const Real E_H = (P_H + P0) * eta / rho0 / 2.;
    const Real dE_H = deta * (P_H + P0) / rho0 / 2. + eta / rho0 / 2 * dP_H;
    return dP_H + Gamma(rho) * (sie - E_H) + rho * dGam * (sie - E_H) -
           rho * Gamma(rho) * dE_H;
  }
  Real T_re(Real rho, Real sie) const { return T0 + sie / Cv; }
  Real e_rT(Real rho, Real T) const { return Cv * (T - T0); }
  // eos_gruneisen.hpp:353-371
  Real P_re(Real rho_in, Real sie) const {
    const Real rho = std::min(rho_in, rho_max);
    Real P_H, E_H;
    if (rho >= rho0) {
      const Real eta = 1 - rho0 / rho;
      const Real s = s1 * eta + s2 * (eta * eta) + s3 * ((eta * eta) * eta);
      const Real oms = 1 - s;
      P_H = P0 + (C0 * C0) * rho0 * eta / (oms * oms);
      E_H = (P_H + P0) * eta / rho0 / 2.;
    } else {
      P_H = P0 + (C0 * C0) * (rho - rho0);
      E_H = 0.;
    }
    return P_H + Gamma(rho) * rho * (sie - E_H);
  }
  // eos_gruneisen.hpp:385-400
  Real bmod_re(Real rho_in, Real sie) const {
    const Real rho = std::min(rho_in, rho_max);
    if (rho < rho0) {
      return rho * (C0 * C0) + G0 * (rho * sie + P_re(rho, sie));
    }
    const Real dPdr_e = dPres_drho_e(rho, sie);
    const Real dPde_r = rho * Gamma(rho);
    return rho * dPdr_e + P_re(rho, sie) / rho * dPde_r;
  }
  Real P_rT(Real rho_in, Real T) const {
    const Real rho = std::min(rho_in, rho_max);
    return P_re(rho, e_rT(rho, T));
  }
  Real bmod_rT(Real rho_in, Real T) const {
    const Real rho = std::min(rho_in, rho_max);
    return bmod_re(rho, e_rT(rho, T));
  }
  // eos_gruneisen.hpp:423-477
  Real MaxStableDensityAtTemperature(Real T, ErrFlag &err) const {
    const Real slope_at_max = dPres_drho_e(rho_max, e_rT(rho_max, T));
    const Real slope_at_ref = dPres_drho_e(rho0, e_rT(rho0, T));
    if (slope_at_max >= 0) return rho_max;
    if (slope_at_ref < 0) {
      err.code |= ERR_GRUN_MAXRHO;
      return rho_max;
    }
    auto dPdrho_T = [this, T](Real r) { return dPres_drho_e(r, e_rT(r, T)); };
    const Real lo = rho0, hi = std::min(rho_max, 1.0e4);
    const Real guess = (lo + hi) / 2.;
    Real r = hi; // defined value when the bracket search fails (the reference aborts)
    if (root::regula_falsi(dPdrho_T, 0., guess, lo, hi, 1.0e-8, 1.0e-8, r) != root::SUCCESS)
      err.code |= ERR_GRUN_MAXRHO;
    return r;
  }
  // eos_gruneisen.hpp:478-526
  void rhoe_PT(Real press, Real temp, Real &rho, Real &sie, ErrFlag &err) const {
    const Real Pref = P_rT(rho0, temp);
    Real lo, hi, guess, p_used = press;
    if (press < Pref) {
      lo = 0.;
      hi = 1.1 * rho0;
      guess = rho0;
    } else {
      lo = 0.9 * rho0;
      hi = MaxStableDensityAtTemperature(temp, err);
      const Real pres_max = P_rT(hi, temp);
      if (press > pres_max) p_used = pres_max;
      const Real slope = (hi - rho0) / (pres_max - Pref);
      guess = rho0 + slope * (p_used - Pref);
    }
    auto PofRatT = [this, temp](Real r) { return P_rT(r, temp); };
    if (root::regula_falsi(PofRatT, p_used, guess, lo, hi, 1.0e-8, 1.0e-8, rho) !=
        root::SUCCESS)
      err.code |= ERR_GRUN_PT;
    sie = e_rT(rho, temp);
  }
  // eos_gruneisen.hpp:527-549
  void fill(Real &rho_in, Real &temp, Real &sie, Real &press, Real &cv_, Real &bmod,
            unsigned long output, ErrFlag &err) const {
    using namespace thermalqs;
    const unsigned long input = ~output;
    if (temperature & input && pressure & input) {
      rhoe_PT(press, temp, rho_in, sie, err);
    } else if (density & output || specific_internal_energy & output) {
      err.code |= ERR_FILLEOS;
      return;
    }
    const Real rho = std::min(rho_in, rho_max);
    if (output & pressure) press = P_re(rho, sie);
    if (output & temperature) temp = T_re(rho, sie);
    if (output & bulk_modulus) bmod = bmod_re(rho, sie);
    if (output & specific_heat) cv_ = Cv;
  }
};

// ---------------------------------------------------------------------------
// Vinet — singularity-eos/eos/eos_vinet.hpp:38-420
// ---------------------------------------------------------------------------
struct Vinet {
  Real rho0, T0, B0, BP0, A0, Cv0, E0, S0;
  Real VIP[43], d2tod40[39];
  static constexpr unsigned long preferred = thermalqs::density | thermalqs::temperature;
  // eos_vinet.hpp:199-226
  static Vinet make(Real rho0, Real T0, Real B0, Real BP0, Real A0, Real Cv0, Real E0,
                    Real S0, const Real *d2to40) {
    Vinet v{rho0, T0, B0, BP0, A0, Cv0, E0, S0, {}, {}};
    for (int i = 0; i < 39; ++i) v.d2tod40[i] = d2to40[i];
    v.VIP[0] = robust::ratio(B0, rho0) + robust::ratio(A0 * A0 * B0 * B0 * T0, rho0 * rho0 * Cv0);
    v.VIP[1] = 3.0 / 2.0 * (BP0 - 1.0);
    v.VIP[2] = 1.0;
    v.VIP[3] = 0.0;
    for (int i = 4; i < 43; ++i) v.VIP[i] = v.d2tod40[i - 4];
    for (int i = 43 - 2; i >= 2; i--) v.VIP[i] = v.VIP[i] - (i) / v.VIP[1] * v.VIP[i + 1];
    return v;
  }
  // eos_vinet.hpp:228-292; out = {e, P, dPdrho, dPdT, dEdT, dEdrho, S, cs}
  void F_DT(Real rho, Real T, Real *out) const {
    constexpr int pref0vp = -2, pref0vip = 2, maxind = 40;
    Real sumP = 0.0, sumB = 0.0, sumE = 0.0;
    const Real x = vmath::cbrt(robust::ratio(rho0, rho));
    const Real x2inv = robust::ratio(1.0, x * x);
    const Real onemx = 1.0 - x;
    const Real etatimes1mx = VIP[1] * onemx;
    const Real expetatimes1mx = vmath::exp(etatimes1mx);
    for (int ind = maxind; ind >= 2; ind--) {
      sumP = d2tod40[pref0vp + ind] + onemx * sumP;
      sumB = d2tod40[pref0vp + ind] * ind + onemx * sumB;
      sumE = VIP[pref0vip + ind] + onemx * sumE;
    }
    for (int ind = 1; ind >= 0; ind--) sumE = VIP[pref0vip + ind] + onemx * sumE;
    sumP = 1.0 + sumP * (onemx * onemx);
    sumB = sumB * onemx;
    sumE = VIP[1] * onemx * sumE;
    Real energy = 9.0 * robust::ratio(B0, VIP[1] * VIP[1] * rho0) *
                      (VIP[pref0vip] - expetatimes1mx * (VIP[pref0vip] - sumE)) -
                  (A0 * B0) * (robust::ratio(T0, rho0) - robust::ratio(T0, rho));
    Real pressure = 3.0 * B0 * x2inv * onemx * expetatimes1mx * sumP;
    Real temp = (1.0 + onemx * (VIP[1] * x + 1.0)) * sumP + x * onemx * sumB;
    temp = robust::ratio(B0, rho0) * x * expetatimes1mx * temp;
    energy = energy + Cv0 * (T - T0) + E0;
    pressure = pressure + (A0 * B0) * (T - T0);
    const Real dpdrho = temp, dpdt = A0 * B0, dedt = Cv0;
    const Real dedrho = robust::ratio(pressure - T * (A0 * B0), rho * rho);
    Real entropy;
    if (T < 0.0) {
      entropy = 0.0;
    } else {
      entropy = (A0 * B0) * (robust::ratio(1.0, rho) - robust::ratio(1, rho0)) +
                Cv0 * std::log(robust::ratio(T, T0)) + S0;
    }
    Real cs = std::max(0.0, dpdrho + T * robust::ratio(dpdt * dpdt, rho * rho * Cv0));
    cs = std::sqrt(cs);
    out[0] = energy;
    out[1] = pressure;
    out[2] = dpdrho;
    out[3] = dpdt;
    out[4] = dedt;
    out[5] = dedrho;
    out[6] = entropy;
    out[7] = cs;
  }
  Real e_rT(Real rho, Real T) const {
    Real o[8];
    F_DT(rho, T, o);
    return o[0];
  }
  Real P_rT(Real rho, Real T) const {
    Real o[8];
    F_DT(rho, T, o);
    return o[1];
  }
  Real bmod_rT(Real rho, Real T) const {
    Real o[8];
    F_DT(rho, T, o);
    return o[7] * o[7] * rho;
  }
  Real T_re(Real rho, Real sie) const {
    Real o[8];
    F_DT(rho, T0, o);
    return robust::ratio(sie - o[0], Cv0) + T0;
  }
  Real P_re(Real rho, Real sie) const {
    const Real t = T_re(rho, sie);
    Real o[8];
    F_DT(rho, t, o);
    return o[1];
  }
  Real bmod_re(Real rho, Real sie) const {
    const Real t = T_re(rho, sie);
    Real o[8];
    F_DT(rho, t, o);
    return o[7] * o[7] * rho;
  }
  Real MinimumDensity() const { return vmath::cbrt(10 * robust::EPS()) * rho0; }
  // eos_vinet.hpp:373-395
  void fill(Real &rho, Real &temp, Real &sie, Real &press, Real &cv_, Real &bmod,
            unsigned long output, ErrFlag &err) const {
    using namespace thermalqs;
    const unsigned long input = ~output;
    if (density & output) {
      err.code |= ERR_FILLEOS;
      return;
    } else if (temperature & output && specific_internal_energy & output) {
      err.code |= ERR_FILLEOS;
      return;
    }
    if (specific_internal_energy & input) temp = T_re(rho, sie);
    Real o[8];
    F_DT(rho, temp, o);
    if (output & specific_internal_energy) sie = o[0];
    if (output & pressure) press = o[1];
    if (output & specific_heat) cv_ = o[4];
    if (output & bulk_modulus) bmod = o[7] * o[7] * rho;
  }
};

// ---------------------------------------------------------------------------
// spiner-equivalent regular grid and linear/bilinear interpolation. spiner is
// an external dependency (lanl/spiner >= 1.6, cmake/singularity-eos/spiner.cmake:17)
// absent from the reference tree; restated from its documented behaviour:
// uniform grid x_i = min + i*dx, dx = (max-min)/(N-1); index =
// floor((x-min)/dx) clamped to [0, N-2]; weights w1 = (x - x_ix)/dx, w0 = 1-w1.
// 2-D layout is [n_slow][n_fast] with the fast axis evaluated first
// (interpToReal(x_slow, x_fast), call sites eos_spiner_rho_temp.hpp:817,1318).
// ---------------------------------------------------------------------------
struct Grid1D {
  Real min = 0, max = 1, dx = 1, inv_dx = 1;
  int N = 2;
  static Grid1D make(Real mn, Real mx, int n) {
    Grid1D g;
    g.min = mn;
    g.max = mx;
    g.N = n;
    g.dx = (mx - mn) / ((Real)(n - 1));
    g.inv_dx = 1 / g.dx;
    return g;
  }
  Real x(int i) const { return min + i * dx; }
  void weights(Real xx, int &ix, Real &w0, Real &w1) const {
    ix = (int)val(std::floor((xx - min) * inv_dx));
    if (ix < 0) ix = 0;
    if (ix >= N - 1) ix = N - 2;
    const Real fl = ix * dx + min;
    w1 = inv_dx * (xx - fl);
    w0 = (1. - w1);
  }
};
struct Table1D {
  Grid1D g;
  std::vector<Real> d;
  Real interp(Real x) const {
    int ix;
    Real w0, w1;
    g.weights(x, ix, w0, w1);
    return w0 * d[ix] + w1 * d[ix + 1];
  }
};
struct Table2D {
  Grid1D gs, gf; // slow (lRho), fast (lT)
  std::vector<Real> d;
  Real at(int j, int i) const { return d[(size_t)j * gf.N + i]; }
  Real interp(Real xs, Real xf) const {
    int i1, i2;
    Real a0, a1, b0, b1;
    gf.weights(xf, i1, a0, a1);
    gs.weights(xs, i2, b0, b1);
    return (b0 * (a0 * at(i2, i1) + a1 * at(i2, i1 + 1)) +
            b1 * (a0 * at(i2 + 1, i1) + a1 * at(i2 + 1, i1 + 1)));
  }
};

// ---------------------------------------------------------------------------
// SpinerEOSDependsRhoT on in-memory tables —
// singularity-eos/eos/eos_spiner_rho_temp.hpp:70-1404 (HDF5 loading replaced
// by caller-provided fields; the load-time derived tables of
// loadDataboxes_ :515-651, fixBulkModulus_ :653-679, setlTColdCrit_ :681-740 and
// spiner_common::SetRhoPMin eos_spiner_common.hpp:59-109 are restated).
// ---------------------------------------------------------------------------
struct SpinerRhoT {
  static constexpr unsigned long preferred = thermalqs::density | thermalqs::temperature;
  static constexpr Real ROOT_THRESH = 1e-14, SOFT_THRESH = 1e-8, VAPOR_DPDR_THRESH = 1e-8;
  Table2D P_, sie_, bMod_, dPdRho_, dPdE_, dEdRho_, dEdT_;
  Table1D PCold_, sieCold_, bModCold_, dPdECold_, dEdTCold_, lTColdCrit_;
  Table1D PMax_, sielTMax_, dEdTMax_, gm1Max_, rho_at_pmin_;
  int numRho_ = 0, numT_ = 0;
  Real lRhoMin_, lRhoMax_, rhoMax_, lTMin_, lTMax_, TMax_, PMin_;
  Real lRhoOffset_ = 0, lTOffset_ = 0;
  bool reproducible_ = false, pmin_vapor_dome_ = false;

  Real lRho_(Real rho) const { return to_log(rho, lRhoOffset_); }
  Real lT_(Real T) const { return to_log(T, lTOffset_); }
  Real rho_(Real lRho) const {
    const Real r = from_log(lRho, lRhoOffset_);
    return r < 0 ? 0 : r;
  }
  Real T_(Real lT) const { return from_log(lT, lTOffset_); }

  // Build from 2-D fields [nRho][nT] (T fastest) and 1-D cold curves [nRho].
  void init(const Grid1D &gRho, const Grid1D &gT, const Real *P, const Real *sie,
            const Real *bMod, const Real *dPdRho, const Real *dPdE, const Real *dEdRho,
            const Real *dEdT, const Real *PCold, const Real *sieCold, const Real *bModCold,
            Real rhoOffset, Real TOffset) {
    numRho_ = gRho.N;
    numT_ = gT.N;
    const size_t n2 = (size_t)numRho_ * numT_;
    auto mk2 = [&](Table2D &t, const Real *src) {
      t.gs = gRho;
      t.gf = gT;
      t.d.assign(src, src + n2);
    };
    auto mk1 = [&](Table1D &t, const Grid1D &g, const Real *src) {
      t.g = g;
      if (src)
        t.d.assign(src, src + g.N);
      else
        t.d.assign(g.N, 0.0);
    };
    mk2(P_, P);
    mk2(sie_, sie);
    mk2(bMod_, bMod);
    mk2(dPdRho_, dPdRho);
    mk2(dPdE_, dPdE);
    mk2(dEdRho_, dEdRho);
    mk2(dEdT_, dEdT);
    mk1(PCold_, gRho, PCold);
    mk1(sieCold_, gRho, sieCold);
    mk1(bModCold_, gRho, bModCold);
    lRhoOffset_ = std::abs(rhoOffset);
    lTOffset_ = std::abs(TOffset);
    lRhoMin_ = gRho.min;
    lRhoMax_ = gRho.max;
    rhoMax_ = from_log(lRhoMax_, lRhoOffset_);
    lTMin_ = gT.min;
    lTMax_ = gT.max;
    TMax_ = from_log(lTMax_, lTOffset_);
    fixBulkModulus_();
    setlTColdCrit_();
    PMin_ = SetRhoPMin_();
    mk1(dPdECold_, gRho, nullptr);
    mk1(dEdTCold_, gRho, nullptr);
    for (int j = 0; j < numRho_; j++) {
      const Real lRho = bModCold_.g.x(j);
      const Real lT = lTColdCrit_.d[j];
      bModCold_.d[j] = bMod_.interp(lRho, lT);
      dPdECold_.d[j] = dPdE_.interp(lRho, lT);
      dEdTCold_.d[j] = dEdT_.interp(lRho, lT);
    }
    mk1(PMax_, gRho, nullptr);
    mk1(dEdTMax_, gRho, nullptr);
    mk1(gm1Max_, gRho, nullptr);
    mk1(sielTMax_, gRho, nullptr);
    for (int j = 0; j < numRho_; j++) {
      const Real lRho = PMax_.g.x(j);
      const Real rho = rho_(lRho);
      PMax_.d[j] = P_.at(j, numT_ - 1);
      dEdTMax_.d[j] = dEdT_.at(j, numT_ - 1);
      gm1Max_.d[j] = robust::ratio(dPdE_.at(j, numT_ - 1), rho);
      sielTMax_.d[j] = sie_.at(j, numT_ - 1);
    }
  }
  // eos_spiner_rho_temp.hpp:653-679
  void fixBulkModulus_() {
    for (int j = 0; j < numRho_; j++) {
      const Real lRho = bMod_.gs.x(j);
      const Real rho = rho_(lRho);
      for (int i = 0; i < numT_; i++) {
        const Real lT = bMod_.gf.x(i);
        const Real press = P_.interp(lRho, lT);
        const Real DPDR_E = dPdRho_.interp(lRho, lT);
        const Real DPDE_R = dPdE_.interp(lRho, lT);
        const Real DEDR_T = dEdRho_.interp(lRho, lT);
        const Real DPDR_T = DPDR_E + DPDE_R * DEDR_T;
        Real bm;
        if (DPDE_R > 0.0 && rho > 0.0) {
          bm = rho * DPDR_E + DPDE_R * (press / rho);
        } else if (rho > 0.0) {
          bm = std::max(rho * DPDR_T, 0.0);
        } else {
          bm = 0.0;
        }
        bMod_.d[(size_t)j * numT_ + i] = std::max(bm, std::abs(robust::EPS()));
      }
    }
  }
  // eos_spiner_rho_temp.hpp:681-740
  void setlTColdCrit_() {
    lTColdCrit_.g = bModCold_.g;
    lTColdCrit_.d.assign(numRho_, 0.0);
    for (int j = 0; j < numRho_; j++) {
      const Real lRho = bModCold_.g.x(j);
      const Real sieCold = sieCold_.interp(lRho);
      // the reference collects every crossing of sie(lRho_j, .) through the
      // cold-curve energy, then uses the last upward one unless it is the
      // first entry of that list (:703-722)
      int ncross = 0, idx_last_pos = -1, last_i = -1;
      for (int i = 0; i < numT_ - 1; i++) {
        const Real sc = sie_.at(j, i) - sieCold;
        const Real sn = sie_.at(j, i + 1) - sieCold;
        if (sc < 0 && sn > 0) {
          idx_last_pos = ncross;
          last_i = i;
          ncross++;
        } else if (sc > 0 && sn < 0) {
          ncross++;
        }
      }
      if (idx_last_pos <= 0) {
        lTColdCrit_.d[j] = lTMin_;
      } else {
        auto sieFunc = [&](Real x) { return sie_.interp(lRho, x); };
        const Real lTlower = bMod_.gf.x(last_i) - 1.0e-14;
        const Real lTupper = bMod_.gf.x(last_i + 1) + 1.0e-14;
        const Real lTGuess = 0.5 * (lTlower + lTupper);
        Real lT;
        if (root::regula_falsi(sieFunc, sieCold, lTGuess, lTlower, lTupper, ROOT_THRESH,
                               ROOT_THRESH, lT) != root::SUCCESS)
          lT = lTGuess;
        lTColdCrit_.d[j] = lT;
      }
      lTColdCrit_.d[j] = std::max(lTMin_, lTColdCrit_.d[j]);
    }
  }
  // spiner_common::SetRhoPMin, eos_spiner_common.hpp:59-109
  Real SetRhoPMin_() {
    Real PMin = std::numeric_limits<Real>::max();
    rho_at_pmin_.g = P_.gf;
    rho_at_pmin_.d.assign(numT_, 0.0);
    for (int i = 0; i < numT_; ++i) {
      Real PMin_at_T = std::numeric_limits<Real>::max();
      int jmax = 0;
      for (int j = 0; j < numRho_; ++j) {
        if (P_.at(j, i) < PMin_at_T) {
          PMin_at_T = P_.at(j, i);
          jmax = j;
        }
        if ((j > 1) && pmin_vapor_dome_) {
          const Real dP = P_.at(j, i) - P_.at(j - 2, i);
          const Real dr =
              from_log(P_.gs.x(j), lRhoOffset_) - from_log(P_.gs.x(j - 2), lRhoOffset_);
          if (robust::ratio(dP, dr) < VAPOR_DPDR_THRESH) jmax = j;
        }
      }
      if ((PMin_at_T > 0) && !pmin_vapor_dome_) {
        rho_at_pmin_.d[i] = 0;
        PMin_at_T = 0;
      } else {
        rho_at_pmin_.d[i] = from_log(P_.gs.x(jmax), lRhoOffset_);
      }
      PMin = std::min(PMin_at_T, PMin);
    }
    if (pmin_vapor_dome_) {
      for (int i = numT_ - 2; i >= 0; i--)
        if (rho_at_pmin_.d[i] < rho_at_pmin_.d[i + 1]) rho_at_pmin_.d[i] = rho_at_pmin_.d[i + 1];
    }
    return PMin;
  }

  // :1377-1392
  TableStatus locRhoSie(Real lRho, Real sie) const {
    const Real sielTMax = sielTMax_.interp(lRho);
    const Real sieCold = sieCold_.interp(lRho);
    if (sie >= sielTMax - SOFT_THRESH * std::abs(sielTMax)) return OffTop;
    if (sie <= sieCold + SOFT_THRESH * std::abs(sieCold)) return OffBottom;
    return OnTable;
  }
  // :1394-1404
  TableStatus locRhoT(Real lRho, Real lT) const {
    if (lT <= (1 + SOFT_THRESH) * lTMin_) return OffBottom;
    if (lT >= (1 - SOFT_THRESH) * lTMax_) return OffTop;
    return OnTable;
  }
  // :1192-1250; lambda may be null, else lambda[0]=lRho, lambda[1]=lT
  Real lTFromlRhoSie_(Real lRho, Real sie, TableStatus &where, Real *lambda,
                      int *iters = nullptr) const {
    Real lT;
    root::Status status = root::SUCCESS;
    where = locRhoSie(lRho, sie);
    if (iters) *iters = 0;
    if (where == OffBottom) {
      lT = lTMin_;
    } else if (where == OffTop) {
      const Real Cv = dEdTMax_.interp(lRho);
      const Real e0 = sielTMax_.interp(lRho);
      const Real T = TMax_ + robust::ratio(sie - e0, Cv);
      lT = lT_(T);
    } else {
      Real lTGuess = reproducible_ ? lTMin_ : 0.5 * (lTMin_ + lTMax_);
      if (lambda) {
        const Real c = lambda[1];
        if ((lTMin_ <= c) && (c <= lTMax_)) lTGuess = c;
      }
      auto sieFunc = [&](Real x) { return sie_.interp(lRho, x); };
      status = root::regula_falsi(sieFunc, sie, lTGuess, lTMin_, lTMax_, ROOT_THRESH,
                                  ROOT_THRESH, lT, iters);
      if (status != root::SUCCESS) lT = reproducible_ ? lTMin_ : lTGuess;
    }
    if (lambda) {
      lambda[0] = lRho;
      lambda[1] = lT;
    }
    return lT;
  }
  // :1135-1190
  Real lRhoFromPlT_(Real P, Real lT, TableStatus &where, Real *lambda) const {
    root::Status status;
    const Real rhopmin = lRho_(rho_at_pmin_.interp(lT));
    Real lRho;
    Real lRhoGuess = reproducible_ ? lRhoMax_ : 0.5 * (rhopmin + lRhoMax_);
    if (lambda) {
      const Real c = lambda[0];
      if ((rhopmin <= c) && (c <= lRhoMax_)) lRhoGuess = c;
    }
    if (lT <= lTMin_) {
      where = OffBottom;
      auto PFunc = [&](Real x) { return PCold_.interp(x); };
      status = root::regula_falsi(PFunc, P, lRhoGuess, rhopmin, lRhoMax_, ROOT_THRESH,
                                  ROOT_THRESH, lRho);
    } else if (lT >= lTMax_) {
      where = OffTop;
      auto PFunc = [&](Real x) { return gm1Max_.interp(x) * dEdTMax_.interp(x) * lT * x; };
      status = root::regula_falsi(PFunc, P, lRhoGuess, rhopmin, lRhoMax_, ROOT_THRESH,
                                  ROOT_THRESH, lRho);
    } else {
      where = OnTable;
      auto PFunc = [&](Real x) { return P_.interp(x, lT); };
      status = root::regula_falsi(PFunc, P, lRhoGuess, rhopmin, lRhoMax_, ROOT_THRESH,
                                  ROOT_THRESH, lRho);
    }
    if (status != root::SUCCESS) lRho = reproducible_ ? lRhoMax_ : lRhoGuess;
    if (lambda) {
      lambda[0] = lRho;
      lambda[1] = lT;
    }
    return lRho;
  }
  Real sieFromlRhoTlT_(Real lRho, Real T, Real lT, TableStatus where) const {
    if (where == OffBottom) return sieCold_.interp(lRho);
    if (where == OffTop) {
      const Real Cv = dEdTMax_.interp(lRho);
      const Real e0 = sielTMax_.interp(lRho);
      return e0 + Cv * (T - TMax_);
    }
    return sie_.interp(lRho, lT);
  }
  Real PFromRholRhoTlT_(Real rho, Real lRho, Real T, Real lT, TableStatus where) const {
    if (where == OffBottom) return PCold_.interp(lRho);
    if (where == OffTop) {
      const Real Cv = dEdTMax_.interp(lRho);
      const Real gm1 = gm1Max_.interp(lRho);
      const Real e0 = sielTMax_.interp(lRho);
      const Real e = e0 + Cv * (T - TMax_);
      return gm1 * rho * e;
    }
    return P_.interp(lRho, lT);
  }
  Real CvFromlRholT_(Real lRho, Real lT, TableStatus where) const {
    Real Cv;
    if (where == OffBottom)
      Cv = dEdTCold_.interp(lRho);
    else if (where == OffTop)
      Cv = dEdTMax_.interp(lRho);
    else
      Cv = dEdT_.interp(lRho, lT);
    return Cv > robust::EPS() ? Cv : robust::EPS();
  }
  Real bModFromRholRhoTlT_(Real rho, Real lRho, Real T, Real lT, TableStatus where) const {
    Real bm;
    if (where == OffBottom) {
      bm = bModCold_.interp(lRho);
    } else if (where == OffTop) {
      const Real Cv = dEdTMax_.interp(lRho);
      const Real gm1 = gm1Max_.interp(lRho);
      const Real e0 = sielTMax_.interp(lRho);
      const Real e = e0 + Cv * (T - TMax_);
      bm = (gm1 + 1) * gm1 * rho * e;
    } else {
      bm = bMod_.interp(lRho, lT);
    }
    return bm > robust::EPS() ? bm : robust::EPS();
  }
  void setlog(Real *lambda, Real lRho, Real lT) const {
    if (lambda) {
      lambda[0] = lRho;
      lambda[1] = lT;
    }
  }
  // scalar API (:777-860)
  Real T_re(Real rho, Real sie, Real *lambda) const {
    TableStatus w;
    const Real lRho = lRho_(rho);
    const Real lT = lTFromlRhoSie_(lRho, sie, w, lambda);
    return T_(lT);
  }
  Real e_rT(Real rho, Real T, Real *lambda) const {
    const Real lRho = lRho_(rho), lT = lT_(T);
    setlog(lambda, lRho, lT);
    return sieFromlRhoTlT_(lRho, T, lT, locRhoT(lRho, lT));
  }
  Real P_rT(Real rho, Real T, Real *lambda) const {
    const Real lRho = lRho_(rho), lT = lT_(T);
    setlog(lambda, lRho, lT);
    return PFromRholRhoTlT_(rho, lRho, T, lT, locRhoT(lRho, lT));
  }
  Real P_re(Real rho, Real sie, Real *lambda) const {
    TableStatus w;
    const Real lRho = lRho_(rho);
    const Real lT = lTFromlRhoSie_(lRho, sie, w, lambda);
    if (w == OffBottom) return PCold_.interp(lRho);
    if (w == OffTop) {
      const Real gm1 = gm1Max_.interp(lRho);
      return gm1 * rho * sie;
    }
    return P_.interp(lRho, lT);
  }
  Real cv_rT(Real rho, Real T, Real *lambda) const {
    const Real lRho = lRho_(rho), lT = lT_(T);
    setlog(lambda, lRho, lT);
    return CvFromlRholT_(lRho, lT, locRhoT(lRho, lT));
  }
  // GruneisenParamFromDensityTemperature — eos_spiner_rho_temp.hpp:877-895
  Real gruneisen_rT(Real rho, Real T, Real *lambda) const {
    const Real lRho = lRho_(rho), lT = lT_(T);
    setlog(lambda, lRho, lT);
    const TableStatus w = locRhoT(lRho, lT);
    if (w == OffBottom) return robust::ratio(std::abs(dPdECold_.interp(lRho)), std::abs(rho));
    if (w == OffTop) return gm1Max_.interp(lRho);
    return robust::ratio(std::abs(dPdE_.interp(lRho, lT)), std::abs(rho));
  }
  // the temperature of ValuesAtReferenceState (eos_spiner_rho_temp.hpp:671-677)
  Real TNormal() const {
    constexpr Real ROOM_TEMPERATURE = 293;  // base/constants.hpp:41
    const Real lTN = lT_(ROOM_TEMPERATURE);
    if (!(lTMin_ < lTN && lTN < lTMax_)) return T_(0.5 * (lTMin_ + lTMax_));
    return ROOM_TEMPERATURE;
  }
  Real cv_re(Real rho, Real sie, Real *lambda) const {
    TableStatus w;
    const Real lRho = lRho_(rho);
    const Real lT = lTFromlRhoSie_(lRho, sie, w, lambda);
    Real Cv;
    if (w == OffBottom)
      Cv = dEdTCold_.interp(lRho);
    else if (w == OffTop)
      Cv = dEdTMax_.interp(lRho);
    else
      Cv = dEdT_.interp(lRho, lT);
    return Cv > robust::EPS() ? Cv : robust::EPS();
  }
  Real bmod_rT(Real rho, Real T, Real *lambda) const {
    const Real lRho = lRho_(rho), lT = lT_(T);
    setlog(lambda, lRho, lT);
    return bModFromRholRhoTlT_(rho, lRho, T, lT, locRhoT(lRho, lT));
  }
  Real bmod_re(Real rho, Real sie, Real *lambda) const {
    TableStatus w;
    const Real lRho = lRho_(rho);
    const Real lT = lTFromlRhoSie_(lRho, sie, w, lambda);
    if (w == OffBottom) return bModCold_.interp(lRho);
    if (w == OffTop) {
      const Real gm1 = gm1Max_.interp(lRho);
      return (gm1 + 1) * gm1 * rho * sie;
    }
    return bMod_.interp(lRho, lT);
  }
  // :1006-1014
  void rhoe_PT(Real press, Real temp, Real &rho, Real &sie, Real *lambda) const {
    TableStatus w;
    const Real lT = lT_(temp);
    const Real lRho = lRhoFromPlT_(press, lT, w, lambda);
    rho = rho_(lRho);
    sie = e_rT(rho, temp, lambda);
  }
  // :1117-1123
  Real RhoPmin(Real temp) const {
    const Real lT = lT_(temp);
    if (lT <= lTMin_) return rho_at_pmin_.d[0];
    if (lT >= lTMax_) return 0.0;
    return rho_at_pmin_.interp(lT);
  }
  // FillEos :1052-1101 for the (rho, sie) and (rho, T) input modes
  void fill(Real &rho, Real &temp, Real &energy, Real &press, Real &cv, Real &bmod,
            unsigned long output, Real *lambda, ErrFlag &err) const {
    using namespace thermalqs;
    Real lRho, lT;
    TableStatus w;
    const unsigned long input = ~output;
    if (output == none) {
      err.code |= ERR_FILLEOS;
      return;
    }
    if (output & density) {
      if (input & pressure && input & temperature) {
        lT = lT_(temp);
        lRho = lRhoFromPlT_(press, lT, w, lambda);
        rho = rho_(lRho);
      } else {
        err.code |= ERR_FILLEOS;
        return;
      }
    } else {
      lRho = lRho_(rho);
    }
    if (output & temperature) {
      if (input & density && input & specific_internal_energy) {
        lT = lTFromlRhoSie_(lRho, energy, w, lambda);
      } else {
        err.code |= ERR_FILLEOS; // (rho, P) input mode is not on the hot path
        return;
      }
      temp = T_(lT);
    } else {
      lT = lT_(temp);
    }
    w = locRhoT(lRho, lT);
    if (output & specific_internal_energy) energy = sieFromlRhoTlT_(lRho, temp, lT, w);
    if (output & pressure) press = PFromRholRhoTlT_(rho, lRho, temp, lT, w);
    if (output & specific_heat) cv = CvFromlRholT_(lRho, lT, w);
    if (output & bulk_modulus) bmod = bModFromRholRhoTlT_(rho, lRho, temp, lT, w);
    setlog(lambda, lRho, lT);
  }
};

// ---------------------------------------------------------------------------
// Variant stand-in — singularity-eos/eos/eos_variant.hpp:151-868. Kind tags.
// Defaults from EosBase (eos_base.hpp:533-557) where a model has no override.
// ---------------------------------------------------------------------------
enum Kind { K_IDEAL = 0, K_STIFF = 1, K_GRUN = 2, K_VINET = 3, K_TABLE = 4 };

struct EOS {
  int kind = K_IDEAL;
  // ShiftedEOS / ScaledEOS (singularity-eos/eos/modifiers/shifted_eos.hpp,
  // scaled_eos.hpp) in the one order the reference's C interface builds,
  // ScaledEOS<ShiftedEOS<T>> (singularity_eos_init_utils.hpp:25-41): either
  // may be off.
  bool shifted = false, scaled = false;
  Real shift = 0.0, scale = 1.0, inv_scale = 1.0;
  IdealGas ig{};
  StiffGas sg{};
  Gruneisen gr{};
  Vinet vn{};
  const SpinerRhoT *tb = nullptr;

  unsigned long PreferredInput_b() const {
    switch (kind) {
    case K_IDEAL: return IdealGas::preferred;
    case K_STIFF: return StiffGas::preferred;
    case K_GRUN: return Gruneisen::preferred;
    case K_VINET: return Vinet::preferred;
    default: return SpinerRhoT::preferred;
    }
  }
  Real T_re_b(Real rho, Real sie, Real *lam) const {
    switch (kind) {
    case K_IDEAL: return ig.T_re(rho, sie);
    case K_STIFF: return sg.T_re(rho, sie);
    case K_GRUN: return gr.T_re(rho, sie);
    case K_VINET: return vn.T_re(rho, sie);
    default: return tb->T_re(rho, sie, lam);
    }
  }
  Real e_rT_b(Real rho, Real T, Real *lam) const {
    switch (kind) {
    case K_IDEAL: return ig.e_rT(rho, T);
    case K_STIFF: return sg.e_rT(rho, T);
    case K_GRUN: return gr.e_rT(rho, T);
    case K_VINET: return vn.e_rT(rho, T);
    default: return tb->e_rT(rho, T, lam);
    }
  }
  Real P_rT_b(Real rho, Real T, Real *lam) const {
    switch (kind) {
    case K_IDEAL: return ig.P_rT(rho, T);
    case K_STIFF: return sg.P_rT(rho, T);
    case K_GRUN: return gr.P_rT(rho, T);
    case K_VINET: return vn.P_rT(rho, T);
    default: return tb->P_rT(rho, T, lam);
    }
  }
  Real P_re_b(Real rho, Real sie, Real *lam) const {
    switch (kind) {
    case K_IDEAL: return ig.P_re(rho, sie);
    case K_STIFF: return sg.P_re(rho, sie);
    case K_GRUN: return gr.P_re(rho, sie);
    case K_VINET: return vn.P_re(rho, sie);
    default: return tb->P_re(rho, sie, lam);
    }
  }
  Real cv_rT_b(Real rho, Real T, Real *lam) const {
    switch (kind) {
    case K_IDEAL: return ig.Cv;
    case K_STIFF: return sg.Cv;
    case K_GRUN: return gr.Cv;
    case K_VINET: return vn.Cv0;
    default: return tb->cv_rT(rho, T, lam);
    }
  }
  Real cv_re_b(Real rho, Real sie, Real *lam) const {
    switch (kind) {
    case K_IDEAL: return ig.Cv;
    case K_STIFF: return sg.Cv;
    case K_GRUN: return gr.Cv;
    case K_VINET: return vn.Cv0;
    default: return tb->cv_re(rho, sie, lam);
    }
  }
  Real bmod_rT_b(Real rho, Real T, Real *lam) const {
    switch (kind) {
    case K_IDEAL: return ig.bmod_rT(rho, T);
    case K_STIFF: return sg.bmod_rT(rho, T);
    case K_GRUN: return gr.bmod_rT(rho, T);
    case K_VINET: return vn.bmod_rT(rho, T);
    default: return tb->bmod_rT(rho, T, lam);
    }
  }
  Real bmod_re_b(Real rho, Real sie, Real *lam) const {
    switch (kind) {
    case K_IDEAL: return ig.bmod_re(rho, sie);
    case K_STIFF: return sg.bmod_re(rho, sie);
    case K_GRUN: return gr.bmod_re(rho, sie);
    case K_VINET: return vn.bmod_re(rho, sie);
    default: return tb->bmod_re(rho, sie, lam);
    }
  }
  // GruneisenParamFromDensityTemperature: IdealGas / StiffGas gm1
  // (eos_ideal.hpp:140-144, eos_stiff.hpp:135-139), Gruneisen Gamma(min(rho,
  // rho_max)) (eos_gruneisen.hpp:134-138), Vinet A0 B0 / (Cv0 rho)
  // (eos_vinet.hpp:110-114), the table's dPdE / rho (see SpinerRhoT)
  Real gruneisen_rT_b(Real rho, Real T, Real *lam) const {
    switch (kind) {
    case K_IDEAL: return ig.gm1;
    case K_STIFF: return sg.gm1;
    case K_GRUN: return gr.Gamma(std::min(rho, gr.rho_max));
    case K_VINET: return robust::ratio(vn.A0 * vn.B0, vn.Cv0 * rho);
    default: return tb->gruneisen_rT(rho, T, lam);
    }
  }
  // the temperature of ValuesAtReferenceState: room temperature for IdealGas
  // and StiffGas (eos_ideal.hpp:224, eos_stiff.hpp:42), T0 for Gruneisen and
  // Vinet (eos_gruneisen.hpp:558, eos_vinet.hpp:408), TNormal for the table
  Real T_reference_b() const {
    switch (kind) {
    case K_IDEAL:
    case K_STIFF: return 293;
    case K_GRUN: return gr.T0;
    case K_VINET: return vn.T0;
    default: return tb->TNormal();
    }
  }
  // EosBase::MinInternalEnergyFromDensity default = e(rho, 0) (eos_base.hpp:598-603);
  // IdealGas returns 0 (eos_ideal.hpp:95-99), tables the cold curve (:822-827)
  Real min_e_b(Real rho) const {
    switch (kind) {
    case K_IDEAL: return 0.0;
    case K_STIFF: return sg.e_rT(rho, 0.);
    case K_GRUN: return gr.e_rT(std::min(rho, gr.rho_max), 0.);
    case K_VINET: return vn.e_rT(rho, 0.);
    default: return tb->sieCold_.interp(tb->lRho_(rho));
    }
  }
  Real MinimumDensity_b() const {
    if (kind == K_VINET) return vn.MinimumDensity();
    if (kind == K_TABLE) return tb->rho_(tb->lRhoMin_);
    return 0;
  }
  Real MinimumTemperature_b() const {
    if (kind == K_TABLE) return tb->T_(tb->lTMin_);
    return 0;
  }
  Real MaximumDensity_b() const {
    if (kind == K_GRUN) return gr.rho_max;
    if (kind == K_TABLE) return tb->rhoMax_;
    return 1e100;
  }
  Real MinimumPressure_b() const {
    if (kind == K_GRUN) return gr.P_re(0, 0);
    if (kind == K_VINET) return -1e100;
    if (kind == K_TABLE) return tb->PMin_;
    return 0;
  }
  // Gruneisen returns its maximum stable density here, exactly as the
  // reference does (eos_gruneisen.hpp:160-163).
  Real MaximumPressureAtTemperature_b(Real T, ErrFlag &err) const {
    if (kind == K_GRUN) return gr.MaxStableDensityAtTemperature(T, err);
    return 1e100;
  }
  Real RhoPmin_b(Real T) const {
    if (kind == K_TABLE) return tb->RhoPmin(T);
    return 0.0;
  }
  // EosBase::DensityEnergyFromPressureTemperature default (eos_base.hpp:630-671)
  // used by Vinet; the others override.
  void rhoe_PT_b(Real press, Real temp, Real *lam, Real &rho, Real &sie, ErrFlag &err) const {
    switch (kind) {
    case K_IDEAL: ig.rhoe_PT(press, temp, rho, sie); return;
    case K_STIFF: sg.rhoe_PT(press, temp, rho, sie); return;
    case K_GRUN: gr.rhoe_PT(press, temp, rho, sie, err); return;
    case K_TABLE: tb->rhoe_PT(press, temp, rho, sie, lam); return;
    default: break;
    }
    constexpr Real DEFAULT_RHO_GUESS = 12;
    const Real rhomin = std::max(RhoPmin_b(temp), MinimumDensity_b());
    const Real rhomax = MaximumDensity_b();
    auto PofRT = [&](Real r) { return P_rT_b(r, temp, lam); };
    Real rhoguess = rho;
    if ((rhoguess <= rhomin) || (rhoguess >= rhomax)) {
      if ((rhomin < DEFAULT_RHO_GUESS) && (DEFAULT_RHO_GUESS < rhomax))
        rhoguess = DEFAULT_RHO_GUESS;
      else
        rhoguess = 0.5 * (rhomin + rhomax);
    }
    root::findRoot(PofRT, press, rhoguess, rhomin, rhomax, robust::EPS(), robust::EPS(), rho);
    sie = e_rT_b(rho, temp, lam);
  }
  void fill_b(Real &rho, Real &temp, Real &sie, Real &press, Real &cv, Real &bmod,
            unsigned long output, Real *lam, ErrFlag &err) const {
    switch (kind) {
    case K_IDEAL: ig.fill(rho, temp, sie, press, cv, bmod, output, err); return;
    case K_STIFF: sg.fill(rho, temp, sie, press, cv, bmod, output, err); return;
    case K_GRUN: gr.fill(rho, temp, sie, press, cv, bmod, output, err); return;
    case K_VINET: vn.fill(rho, temp, sie, press, cv, bmod, output, err); return;
    default: tb->fill(rho, temp, sie, press, cv, bmod, output, lam, err); return;
    }
  }

  // ---- the material as the variant sees it: base model under the modifiers
  // ScaledEOS forwards (scale rho, sie / scale) to ShiftedEOS, which forwards
  // (rho, sie - shift) to the model; energies come back + shift, then x scale.
  Real rho_in(Real rho) const { return scaled ? scale * rho : rho; }
  Real sie_in(Real sie) const {
    Real s = scaled ? inv_scale * sie : sie;
    return shifted ? s - shift : s;
  }
  Real sie_out(Real e) const {
    if (shifted) e = e + shift;
    return scaled ? scale * e : e;
  }
  unsigned long PreferredInput() const { return PreferredInput_b(); }
  Real T_re(Real rho, Real sie, Real *lam) const { return T_re_b(rho_in(rho), sie_in(sie), lam); }
  // scaled_eos.hpp:80-84: the energy of ScaledEOS takes rho unscaled
  Real e_rT(Real rho, Real T, Real *lam) const { return sie_out(e_rT_b(rho, T, lam)); }
  Real P_rT(Real rho, Real T, Real *lam) const { return P_rT_b(rho_in(rho), T, lam); }
  Real P_re(Real rho, Real sie, Real *lam) const { return P_re_b(rho_in(rho), sie_in(sie), lam); }
  Real cv_rT(Real rho, Real T, Real *lam) const { return cv_rT_b(rho_in(rho), T, lam); }
  Real cv_re(Real rho, Real sie, Real *lam) const {
    return cv_re_b(rho_in(rho), sie_in(sie), lam);
  }
  Real bmod_rT(Real rho, Real T, Real *lam) const { return bmod_rT_b(rho_in(rho), T, lam); }
  Real bmod_re(Real rho, Real sie, Real *lam) const {
    return bmod_re_b(rho_in(rho), sie_in(sie), lam);
  }
  Real gruneisen_rT(Real rho, Real T, Real *lam) const {
    return gruneisen_rT_b(rho_in(rho), T, lam);
  }
  Real T_reference() const { return T_reference_b(); }
  // scaled_eos.hpp:88-91 (rho scaled, energy not), shifted_eos.hpp:81-84 (+ shift)
  Real min_e(Real rho) const {
    const Real e = min_e_b(rho_in(rho));
    return shifted ? e + shift : e;
  }
  Real MinimumDensity() const { return scaled ? inv_scale * MinimumDensity_b() : MinimumDensity_b(); }
  Real MinimumTemperature() const { return MinimumTemperature_b(); }
  Real MaximumDensity() const { return scaled ? inv_scale * MaximumDensity_b() : MaximumDensity_b(); }
  Real MinimumPressure() const { return MinimumPressure_b(); }
  Real MaximumPressureAtTemperature(Real T, ErrFlag &err) const {
    return MaximumPressureAtTemperature_b(T, err);
  }
  Real RhoPmin(Real T) const { return scaled ? inv_scale * RhoPmin_b(T) : RhoPmin_b(T); }
  // shifted_eos.hpp:399-403, scaled_eos.hpp:328-335
  void rhoe_PT(Real press, Real temp, Real *lam, Real &rho, Real &sie, ErrFlag &err) const {
    rhoe_PT_b(press, temp, lam, rho, sie, err);
    if (shifted) sie = sie + shift;
    if (scaled) {
      rho = rho * inv_scale;
      sie = sie * scale;
    }
  }
  // FillEos of ScaledEOS over ShiftedEOS (scaled_eos.hpp:148-166,
  // shifted_eos.hpp:142-158): by the model's preferred input, (rho, T) or
  // (rho, sie); any other preferred input (IdealGas's) is an error
  void fill(Real &rho, Real &temp, Real &sie, Real &press, Real &cv, Real &bmod,
            unsigned long output, Real *lam, ErrFlag &err) const {
    if (!shifted && !scaled) {
      fill_b(rho, temp, sie, press, cv, bmod, output, lam, err);
      return;
    }
    const unsigned long pref = PreferredInput_b();
    if (pref == (thermalqs::density | thermalqs::temperature)) {
      Real srho = rho_in(rho);
      Real &r = scaled ? srho : rho;
      fill_b(r, temp, sie, press, cv, bmod, output, lam, err);
      sie = sie_out(sie);
    } else if (pref == (thermalqs::density | thermalqs::specific_internal_energy)) {
      Real srho = rho_in(rho);
      Real &r = scaled ? srho : rho;
      Real senergy = sie_in(sie);
      fill_b(r, temp, senergy, press, cv, bmod, output, lam, err);
    } else {
      err.code |= ERR_FILLEOS;
    }
  }
};

} // namespace sgo
