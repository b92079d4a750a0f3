// Device equations of state for the hot path: IdealGas, StiffGas, Gruneisen,
// Vinet and a Spiner-style rho-T table, behind one POD descriptor (EosDev)
// that kernels stage per material and dispatch on with a switch. Every
// function restates the singularity-eos arithmetic for the model it names
// (file:line cited), in the same operation order.
#pragma once

#include "sgm_numerics.hpp"

// heavy paths (root-finding inversions, the Vinet energy/pressure kernel) are
// real calls rather than inlined into every unrolled material slot
#define SGM_NOINL __host__ __device__ __attribute__((noinline)) inline

namespace sgm {

// thermalqs bits (singularity-eos/base/constants.hpp:22-35)
namespace tq {
constexpr unsigned long density = 1 << 0;
constexpr unsigned long specific_internal_energy = 1 << 1;
constexpr unsigned long pressure = 1 << 2;
constexpr unsigned long temperature = 1 << 3;
constexpr unsigned long specific_heat = 1 << 4;
constexpr unsigned long bulk_modulus = 1 << 5;
} // namespace tq

enum EosKind : int { KIND_IDEAL = 0, KIND_STIFF = 1, KIND_GRUN = 2, KIND_VINET = 3, KIND_TABLE = 4 };

// Error codes written where the reference aborts or warns
enum ErrBits : int {
  ERRB_GRUN_MAXRHO = 1, // Gruneisen::MaxStableDensityAtTemperature failure
  ERRB_GRUN_PT = 2,     // Gruneisen P-T inversion root find failed
  ERRB_FILLEOS = 4,     // FillEos called with an unsupported output set
  ERRB_LINSOLVE = 8,    // PTE: singular or non-finite Newton step
  ERRB_SMALLSTEP = 16,  // PTE: step scale below threshold repeatedly
};

// Spiner-style rho-T table on regular log grids (NQT log10 coordinates).
// 2-D fields are [nRho][nT], T fastest; 1-D cold/edge curves over lRho;
// rho_at_pmin over lT. Fields and derived curves are built on the host
// (load-time derivations of singularity-eos/eos/eos_spiner_rho_temp.hpp:515-740).
struct TableDev {
  Real lRhoMin, lRhoMax, dlRho, inv_dlRho;
  Real lTMin, lTMax, dlT, inv_dlT;
  Real rhoMax, TMax, PMin, lRhoOffset, lTOffset;
  int nRho, nT, reproducible, pad;
  const Real *P, *sie, *bMod, *dEdT, *dPdE;
  const Real *PCold, *sieCold, *bModCold, *dPdECold, *dEdTCold;
  const Real *PMax, *sielTMax, *dEdTMax, *gm1Max;
  const Real *rho_at_pmin;
};

// 288-byte POD describing one material; the closure and vector kernels stage
// the registry into LDS, so everything an evaluation needs besides table and
// coefficient arrays is one LDS read away. p[] holds the model parameters:
//  IdealGas  {Cv, gm1}
//  StiffGas  {Cv, gm1, Pinf, qq}
//  Gruneisen {C0, s1, s2, s3, G0, b, rho0, T0, P0, Cv, rho_max}
//  Vinet     {rho0, T0, B0, BP0, A0, Cv0, E0, S0, VIP1, min_density, VIP2, VIP3}
// Vinet Horner coefficients (d2..d40 then VIP0..VIP42) live in `coef`;
// nvinet_skip counts the leading Horner steps that can be skipped because
// every coefficient they touch is exactly zero (the sums stay exactly 0
// there, so skipping them leaves every bit of the result unchanged). A table
// material carries its whole TableDev record (grid, scalars, array pointers).
// Modifiers (singularity-eos/eos/modifiers/shifted_eos.hpp, scaled_eos.hpp)
// in the one order the reference's C interface builds them,
// ScaledEOS<ShiftedEOS<model>> (singularity_eos_init_utils.hpp:25-41):
// MOD_SHIFTED and MOD_SCALED in `mods`, either may be off.
enum EosMods : int { MOD_SHIFTED = 1, MOD_SCALED = 2 };
struct alignas(16) EosDev {
  int kind;
  int nvinet_skip;
  int mods;
  int pad_;
  Real shift, scale, inv_scale, pad2_;
  union {
    struct {
      Real p[12];
      const Real *coef;
    } m;
    TableDev tab;
  };
};
static_assert(sizeof(EosDev) == 288, "EosDev is 288 bytes");

// Lambda cache for the table EOS: (lRho, lT) per point, used as root-find guesses
struct Lam {
  bool present = false;
  Real lRho = 0, lT = 0;
  SGM_HD void set(Real r, Real t) {
    if (present) {
      lRho = r;
      lT = t;
    }
  }
};

// ---------------------------------------------------------------------------
// regular-grid linear interpolation (spiner RegularGrid1D/DataBox semantics:
// index clamped to [0, N-2], weights from the cell's left node)
// ---------------------------------------------------------------------------
struct W1 {
  int ix;
  Real w0, w1;
};
SGM_HD W1 grid_weights(Real x, Real mn, Real dx, Real inv_dx, int N) {
  W1 w;
  const Real fi = floor((x - mn) * inv_dx);
  int ix;
  if (!(fi >= 0.0))
    ix = 0;
  else if (fi >= (Real)(N - 1))
    ix = N - 2;
  else
    ix = (int)fi;
  const Real fl = ix * dx + mn;
  w.w1 = inv_dx * (x - fl);
  w.w0 = (1. - w.w1);
  w.ix = ix;
  return w;
}
SGM_HD Real lerp_rho(const TableDev &t, const Real *f, Real lRho) {
  const W1 w = grid_weights(lRho, t.lRhoMin, t.dlRho, t.inv_dlRho, t.nRho);
  return w.w0 * f[w.ix] + w.w1 * f[w.ix + 1];
}
SGM_HD Real lerp_T(const TableDev &t, const Real *f, Real lT) {
  const W1 w = grid_weights(lT, t.lTMin, t.dlT, t.inv_dlT, t.nT);
  return w.w0 * f[w.ix] + w.w1 * f[w.ix + 1];
}
// bilinear with precomputed lRho weights (the fast-axis weights are
// evaluated first in the reference; the arithmetic below is identical)
SGM_HD Real bilerp_w(const TableDev &t, const Real *f, const W1 &wr, Real lT) {
  const W1 wt = grid_weights(lT, t.lTMin, t.dlT, t.inv_dlT, t.nT);
  const Real *r0 = f + (size_t)wr.ix * t.nT + wt.ix;
  const Real *r1 = r0 + t.nT;
  return (wr.w0 * (wt.w0 * r0[0] + wt.w1 * r0[1]) + wr.w1 * (wt.w0 * r1[0] + wt.w1 * r1[1]));
}
SGM_HD Real bilerp(const TableDev &t, const Real *f, Real lRho, Real lT) {
  const W1 wr = grid_weights(lRho, t.lRhoMin, t.dlRho, t.inv_dlRho, t.nRho);
  return bilerp_w(t, f, wr, lT);
}

// ---------------------------------------------------------------------------
// Table EOS (SpinerEOSDependsRhoT), singularity-eos/eos/eos_spiner_rho_temp.hpp
// ---------------------------------------------------------------------------
namespace table {
enum Where { OnTable = 0, OffBottom = 1, OffTop = 2 };
constexpr Real ROOT_THRESH = 1e-14, SOFT_THRESH = 1e-8;
SGM_HD Real lRho_(const TableDev &t, Real rho) { return to_log(rho, t.lRhoOffset); }
SGM_HD Real lT_(const TableDev &t, Real T) { return to_log(T, t.lTOffset); }
SGM_HD Real rho_(const TableDev &t, Real l) {
  const Real r = from_log(l, t.lRhoOffset);
  return r < 0 ? 0 : r;
}
SGM_HD Real T_(const TableDev &t, Real l) { return from_log(l, t.lTOffset); }

// :1377-1392
SGM_HD int loc_rho_sie(const TableDev &t, Real lRho, Real sie) {
  const Real sielTMax = lerp_rho(t, t.sielTMax, lRho);
  const Real sieCold = lerp_rho(t, t.sieCold, lRho);
  if (sie >= sielTMax - SOFT_THRESH * fabs(sielTMax)) return OffTop;
  if (sie <= sieCold + SOFT_THRESH * fabs(sieCold)) return OffBottom;
  return OnTable;
}
// :1394-1404
SGM_HD int loc_rho_T(const TableDev &t, Real lRho, Real lT) {
  if (lT <= (1 + SOFT_THRESH) * t.lTMin) return OffBottom;
  if (lT >= (1 - SOFT_THRESH) * t.lTMax) return OffTop;
  return OnTable;
}
// lTFromlRhoSie_ :1192-1250
struct InvT {
  Real lT;
  int where, iters;
  Lam lam;
};
// The table inversions (regula falsi over a grid column) run out of line in
// the closure kernels, where an inlined copy at every material's P(rho, e)
// call site would grow code that is already instruction-cache bound, and
// inline in the vector kernels (SGM_TABLE_INVERSE_INLINE), where an
// out-of-line call would save the callee's registers to scratch memory at
// every point. Same arithmetic either way.
#ifdef SGM_TABLE_INVERSE_INLINE
#define SGM_TABLE_INV SGM_HD
#else
#define SGM_TABLE_INV SGM_NOINL
#endif
SGM_TABLE_INV InvT lT_from_lRho_sie(const TableDev &t, Real lRho, Real sie, Lam lam) {
  InvT o;
  Real lT;
  const int where = loc_rho_sie(t, lRho, sie);
  int iters = 0;
  if (where == OffBottom) {
    lT = t.lTMin;
  } else if (where == OffTop) {
    const Real Cv = lerp_rho(t, t.dEdTMax, lRho);
    const Real e0 = lerp_rho(t, t.sielTMax, lRho);
    const Real T = t.TMax + robust::ratio(sie - e0, Cv);
    lT = lT_(t, T);
  } else {
    Real lTGuess = t.reproducible ? t.lTMin : 0.5 * (t.lTMin + t.lTMax);
    if (lam.present && (t.lTMin <= lam.lT) && (lam.lT <= t.lTMax)) lTGuess = lam.lT;
    const W1 wr = grid_weights(lRho, t.lRhoMin, t.dlRho, t.inv_dlRho, t.nRho);
    const Real *sief = t.sie;
    auto f = [&](Real x) { return bilerp_w(t, sief, wr, x); };
    if (!root::regula_falsi(f, sie, lTGuess, t.lTMin, t.lTMax, ROOT_THRESH, ROOT_THRESH, lT,
                            &iters))
      lT = t.reproducible ? t.lTMin : lTGuess;
  }
  lam.set(lRho, lT);
  o.lT = lT;
  o.where = where;
  o.iters = iters;
  o.lam = lam;
  return o;
}
// lRhoFromPlT_ :1135-1190
struct InvRho {
  Real lRho;
  int where;
  Lam lam;
};
SGM_TABLE_INV InvRho lRho_from_P_lT(const TableDev &t, Real P, Real lT, Lam lam) {
  const Real rhopmin = lRho_(t, lerp_T(t, t.rho_at_pmin, lT));
  Real lRho;
  int where;
  Real lRhoGuess = t.reproducible ? t.lRhoMax : 0.5 * (rhopmin + t.lRhoMax);
  if (lam.present && (rhopmin <= lam.lRho) && (lam.lRho <= t.lRhoMax)) lRhoGuess = lam.lRho;
  bool ok;
  if (lT <= t.lTMin) {
    where = OffBottom;
    auto f = [&](Real x) { return lerp_rho(t, t.PCold, x); };
    ok = root::regula_falsi(f, P, lRhoGuess, rhopmin, t.lRhoMax, ROOT_THRESH, ROOT_THRESH, lRho);
  } else if (lT >= t.lTMax) {
    where = OffTop;
    // prod_interp_1d(gm1Max_, dEdTMax_, lT) (eos_spiner_common.hpp:240-251)
    auto f = [&](Real x) {
      return lerp_rho(t, t.gm1Max, x) * lerp_rho(t, t.dEdTMax, x) * lT * x;
    };
    ok = root::regula_falsi(f, P, lRhoGuess, rhopmin, t.lRhoMax, ROOT_THRESH, ROOT_THRESH, lRho);
  } else {
    where = OnTable;
    const W1 wt = grid_weights(lT, t.lTMin, t.dlT, t.inv_dlT, t.nT);
    auto f = [&](Real x) {
      const W1 wr = grid_weights(x, t.lRhoMin, t.dlRho, t.inv_dlRho, t.nRho);
      const Real *r0 = t.P + (size_t)wr.ix * t.nT + wt.ix;
      const Real *r1 = r0 + t.nT;
      return (wr.w0 * (wt.w0 * r0[0] + wt.w1 * r0[1]) + wr.w1 * (wt.w0 * r1[0] + wt.w1 * r1[1]));
    };
    ok = root::regula_falsi(f, P, lRhoGuess, rhopmin, t.lRhoMax, ROOT_THRESH, ROOT_THRESH, lRho);
  }
  if (!ok) lRho = t.reproducible ? t.lRhoMax : lRhoGuess;
  lam.set(lRho, lT);
  InvRho o;
  o.lRho = lRho;
  o.where = where;
  o.lam = lam;
  return o;
}
SGM_HD Real sie_from(const TableDev &t, Real lRho, Real T, Real lT, int where) {
  if (where == OffBottom) return lerp_rho(t, t.sieCold, lRho);
  if (where == OffTop) {
    const Real Cv = lerp_rho(t, t.dEdTMax, lRho);
    const Real e0 = lerp_rho(t, t.sielTMax, lRho);
    return e0 + Cv * (T - t.TMax);
  }
  return bilerp(t, t.sie, lRho, lT);
}
SGM_HD Real P_from(const TableDev &t, Real rho, Real lRho, Real T, Real lT, int where) {
  if (where == OffBottom) return lerp_rho(t, t.PCold, lRho);
  if (where == OffTop) {
    const Real Cv = lerp_rho(t, t.dEdTMax, lRho);
    const Real gm1 = lerp_rho(t, t.gm1Max, lRho);
    const Real e0 = lerp_rho(t, t.sielTMax, lRho);
    const Real e = e0 + Cv * (T - t.TMax);
    return gm1 * rho * e;
  }
  return bilerp(t, t.P, lRho, lT);
}
SGM_HD Real cv_from(const TableDev &t, Real lRho, Real lT, int where) {
  Real Cv;
  if (where == OffBottom)
    Cv = lerp_rho(t, t.dEdTCold, lRho);
  else if (where == OffTop)
    Cv = lerp_rho(t, t.dEdTMax, lRho);
  else
    Cv = bilerp(t, t.dEdT, lRho, lT);
  return Cv > robust::EPS() ? Cv : robust::EPS();
}
SGM_HD Real bmod_from(const TableDev &t, Real rho, Real lRho, Real T, Real lT, int where) {
  Real bm;
  if (where == OffBottom) {
    bm = lerp_rho(t, t.bModCold, lRho);
  } else if (where == OffTop) {
    const Real Cv = lerp_rho(t, t.dEdTMax, lRho);
    const Real gm1 = lerp_rho(t, t.gm1Max, lRho);
    const Real e0 = lerp_rho(t, t.sielTMax, lRho);
    const Real e = e0 + Cv * (T - t.TMax);
    bm = (gm1 + 1) * gm1 * rho * e;
  } else {
    bm = bilerp(t, t.bMod, lRho, lT);
  }
  return bm > robust::EPS() ? bm : robust::EPS();
}
SGM_HD Real e_rT(const TableDev &t, Real rho, Real T, Lam &lam) {
  const Real lRho = lRho_(t, rho), lT = lT_(t, T);
  lam.set(lRho, lT);
  return sie_from(t, lRho, T, lT, loc_rho_T(t, lRho, lT));
}
SGM_HD Real P_rT(const TableDev &t, Real rho, Real T, Lam &lam) {
  const Real lRho = lRho_(t, rho), lT = lT_(t, T);
  lam.set(lRho, lT);
  return P_from(t, rho, lRho, T, lT, loc_rho_T(t, lRho, lT));
}
SGM_HD Real cv_rT(const TableDev &t, Real rho, Real T, Lam &lam) {
  const Real lRho = lRho_(t, rho), lT = lT_(t, T);
  lam.set(lRho, lT);
  return cv_from(t, lRho, lT, loc_rho_T(t, lRho, lT));
}
// GruneisenParamFromDensityTemperature (eos_spiner_rho_temp.hpp:877-895)
SGM_HD Real gruneisen_rT(const TableDev &t, Real rho, Real T, Lam &lam) {
  const Real lRho = lRho_(t, rho), lT = lT_(t, T);
  lam.set(lRho, lT);
  const int w = loc_rho_T(t, lRho, lT);
  if (w == OffBottom) return robust::ratio(fabs(lerp_rho(t, t.dPdECold, lRho)), fabs(rho));
  if (w == OffTop) return lerp_rho(t, t.gm1Max, lRho);
  return robust::ratio(fabs(bilerp(t, t.dPdE, lRho, lT)), fabs(rho));
}
// the temperature of ValuesAtReferenceState (eos_spiner_rho_temp.hpp:671-677)
SGM_HD Real T_normal(const TableDev &t) {
  const Real room = 293;  // ROOM_TEMPERATURE, base/constants.hpp:41
  const Real lTN = lT_(t, room);
  if (!(t.lTMin < lTN && lTN < t.lTMax)) return T_(t, 0.5 * (t.lTMin + t.lTMax));
  return room;
}
SGM_HD Real bmod_rT(const TableDev &t, Real rho, Real T, Lam &lam) {
  const Real lRho = lRho_(t, rho), lT = lT_(t, T);
  lam.set(lRho, lT);
  return bmod_from(t, rho, lRho, T, lT, loc_rho_T(t, lRho, lT));
}
SGM_HD Real T_re(const TableDev &t, Real rho, Real sie, Lam &lam) {
  const InvT iv = lT_from_lRho_sie(t, lRho_(t, rho), sie, lam);
  lam = iv.lam;
  return T_(t, iv.lT);
}
SGM_HD Real P_re(const TableDev &t, Real rho, Real sie, Lam &lam) {
  const Real lRho = lRho_(t, rho);
  const InvT iv = lT_from_lRho_sie(t, lRho, sie, lam);
  lam = iv.lam;
  const int w = iv.where;
  const Real lT = iv.lT;
  if (w == OffBottom) return lerp_rho(t, t.PCold, lRho);
  if (w == OffTop) return lerp_rho(t, t.gm1Max, lRho) * rho * sie;
  return bilerp(t, t.P, lRho, lT);
}
SGM_HD Real cv_re(const TableDev &t, Real rho, Real sie, Lam &lam) {
  const Real lRho = lRho_(t, rho);
  const InvT iv = lT_from_lRho_sie(t, lRho, sie, lam);
  lam = iv.lam;
  const int w = iv.where;
  const Real lT = iv.lT;
  Real Cv;
  if (w == OffBottom)
    Cv = lerp_rho(t, t.dEdTCold, lRho);
  else if (w == OffTop)
    Cv = lerp_rho(t, t.dEdTMax, lRho);
  else
    Cv = bilerp(t, t.dEdT, lRho, lT);
  return Cv > robust::EPS() ? Cv : robust::EPS();
}
SGM_HD Real bmod_re(const TableDev &t, Real rho, Real sie, Lam &lam) {
  const Real lRho = lRho_(t, rho);
  const InvT iv = lT_from_lRho_sie(t, lRho, sie, lam);
  lam = iv.lam;
  const int w = iv.where;
  const Real lT = iv.lT;
  if (w == OffBottom) return lerp_rho(t, t.bModCold, lRho);
  if (w == OffTop) {
    const Real gm1 = lerp_rho(t, t.gm1Max, lRho);
    return (gm1 + 1) * gm1 * rho * sie;
  }
  return bilerp(t, t.bMod, lRho, lT);
}
// :1117-1123
SGM_HD Real rho_pmin(const TableDev &t, Real temp) {
  const Real lT = lT_(t, temp);
  if (lT <= t.lTMin) return t.rho_at_pmin[0];
  if (lT >= t.lTMax) return 0.0;
  return lerp_T(t, t.rho_at_pmin, lT);
}
// :1006-1014
SGM_HD void rhoe_PT(const TableDev &t, Real P, Real T, Real &rho, Real &sie, Lam &lam) {
  const Real lT = lT_(t, T);
  const InvRho iv = lRho_from_P_lT(t, P, lT, lam);
  lam = iv.lam;
  rho = rho_(t, iv.lRho);
  sie = e_rT(t, rho, T, lam);
}
} // namespace table

// ---------------------------------------------------------------------------
// Analytic models
// ---------------------------------------------------------------------------
namespace ideal { // singularity-eos/eos/eos_ideal.hpp:59-207
SGM_HD Real mx(Real a, Real b) { return a > b ? a : b; } // MYMAX (:32)
SGM_HD Real T_re(const Real *p, Real rho, Real sie) { return mx(0.0, sie / p[0]); }
SGM_HD Real e_rT(const Real *p, Real rho, Real T) { return mx(0.0, p[0] * T); }
SGM_HD Real P_rT(const Real *p, Real rho, Real T) { return mx(0.0, p[1] * rho * p[0] * T); }
SGM_HD Real P_re(const Real *p, Real rho, Real sie) { return mx(0.0, p[1] * rho * sie); }
SGM_HD Real bmod_rT(const Real *p, Real rho, Real T) {
  return mx(0.0, (p[1] + 1) * p[1] * rho * p[0] * T);
}
SGM_HD Real bmod_re(const Real *p, Real rho, Real sie) {
  return mx(0.0, (p[1] + 1) * p[1] * rho * sie);
}
SGM_HD void rhoe_PT(const Real *p, Real P, Real T, Real &rho, Real &sie) {
  sie = mx(0.0, p[0] * T);
  rho = mx(0.0, P / (p[1] * sie));
}
} // namespace ideal

namespace stiff { // singularity-eos/eos/eos_stiff.hpp:63-185; p = {Cv, gm1, Pinf, qq}
SGM_HD Real T_re(const Real *p, Real rho, Real sie) {
  return smax(robust::SMALL(), robust::ratio(rho * (sie - p[3]) - p[2], rho * p[0]));
}
SGM_HD Real e_rT(const Real *p, Real rho, Real T) {
  return smax(p[3], robust::ratio(rho * p[0] * T + p[2], rho) + p[3]);
}
SGM_HD Real P_rT(const Real *p, Real rho, Real T) {
  return smax(-p[2], p[1] * rho * p[0] * T - p[2]);
}
SGM_HD Real P_re(const Real *p, Real rho, Real sie) {
  return smax(-p[2], p[1] * rho * (sie - p[3]) - (p[1] + 1.0) * p[2]);
}
SGM_HD Real bmod_rT(const Real *p, Real rho, Real T) {
  return smax(robust::SMALL(), p[1] * (p[1] + 1.0) * rho * p[0] * T);
}
SGM_HD Real bmod_re(const Real *p, Real rho, Real sie) {
  return smax(robust::SMALL(), p[1] * (p[1] + 1.0) * (rho * (sie - p[3]) - p[2]));
}
SGM_HD void rhoe_PT(const Real *p, Real P, Real T, Real &rho, Real &sie) {
  sie = smax(p[3], robust::ratio(P + (p[1] + 1.0) * p[2], P + p[2]) * p[0] * T + p[3]);
  rho = smax(robust::SMALL(), robust::ratio(P + p[2], p[1] * p[0] * T));
}
} // namespace stiff

namespace grun { // singularity-eos/eos/eos_gruneisen.hpp
// p = {C0, s1, s2, s3, G0, b, rho0, T0, P0, Cv, rho_max}
SGM_HD Real Gamma(const Real *p, Real rho_in) { // :191-195
  const Real rho = smin(rho_in, p[10]);
  return rho < p[6] ? p[4] : p[4] * p[6] / rho + p[5] * (1 - p[6] / rho);
}
SGM_HD Real dPres_drho_e(const Real *p, Real rho_in, Real sie) { // :331-351
  const Real C0 = p[0], s1 = p[1], s2 = p[2], s3 = p[3], G0 = p[4], b = p[5], rho0 = p[6],
             P0 = p[8];
  const Real rho = smin(rho_in, p[10]);
  if (rho < rho0) return (C0 * C0) + Gamma(p, rho) * sie;
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
  const Real G = Gamma(p, rho);
  return dP_H + G * (sie - E_H) + rho * dGam * (sie - E_H) - rho * G * dE_H;
}
SGM_HD Real T_re(const Real *p, Real rho, Real sie) { return p[7] + sie / p[9]; }
SGM_HD Real e_rT(const Real *p, Real rho, Real T) { return p[9] * (T - p[7]); }
SGM_HD Real P_re(const Real *p, Real rho_in, Real sie) { // :353-371
  const Real C0 = p[0], rho0 = p[6], P0 = p[8];
  const Real rho = smin(rho_in, p[10]);
  Real P_H, E_H;
  if (rho >= rho0) {
    const Real eta = 1 - rho0 / rho;
    const Real s = p[1] * eta + p[2] * (eta * eta) + p[3] * ((eta * eta) * eta);
    const Real oms = 1 - s;
    P_H = P0 + (C0 * C0) * rho0 * eta / (oms * oms);
    E_H = (P_H + P0) * eta / rho0 / 2.;
  } else {
    P_H = P0 + (C0 * C0) * (rho - rho0);
    E_H = 0.;
  }
  return P_H + Gamma(p, rho) * rho * (sie - E_H);
}
SGM_HD Real bmod_re(const Real *p, Real rho_in, Real sie) { // :385-400
  const Real rho = smin(rho_in, p[10]);
  if (rho < p[6]) return rho * (p[0] * p[0]) + p[4] * (rho * sie + P_re(p, rho, sie));
  const Real dPdr_e = dPres_drho_e(p, rho, sie);
  const Real dPde_r = rho * Gamma(p, rho);
  return rho * dPdr_e + P_re(p, rho, sie) / rho * dPde_r;
}
SGM_HD Real P_rT(const Real *p, Real rho_in, Real T) {
  const Real rho = smin(rho_in, p[10]);
  return P_re(p, rho, e_rT(p, rho, T));
}
SGM_HD Real bmod_rT(const Real *p, Real rho_in, Real T) {
  const Real rho = smin(rho_in, p[10]);
  return bmod_re(p, rho, e_rT(p, rho, T));
}
// :423-477
SGM_NOINL Real max_stable_density(const Real *p, Real T, int &err) {
  const Real rho_max = p[10], rho0 = p[6];
  const Real slope_at_max = dPres_drho_e(p, rho_max, e_rT(p, rho_max, T));
  const Real slope_at_ref = dPres_drho_e(p, rho0, e_rT(p, rho0, T));
  if (slope_at_max >= 0) return rho_max;
  if (slope_at_ref < 0) {
    err |= ERRB_GRUN_MAXRHO;
    return rho_max;
  }
  auto f = [&](Real r) { return dPres_drho_e(p, r, e_rT(p, r, T)); };
  const Real lo = rho0, hi = smin(rho_max, 1.0e4);
  Real r = hi;
  if (!root::regula_falsi(f, 0., (lo + hi) / 2., lo, hi, 1.0e-8, 1.0e-8, r))
    err |= ERRB_GRUN_MAXRHO;
  return r;
}
// :478-526
SGM_NOINL void rhoe_PT(const Real *p, Real P, Real T, Real &rho, Real &sie, int &err) {
  const Real rho0 = p[6];
  const Real Pref = P_rT(p, rho0, T);
  Real lo, hi, guess, p_used = P;
  if (P < Pref) {
    lo = 0.;
    hi = 1.1 * rho0;
    guess = rho0;
  } else {
    lo = 0.9 * rho0;
    hi = max_stable_density(p, T, err);
    const Real pres_max = P_rT(p, hi, T);
    if (P > pres_max) p_used = pres_max;
    const Real slope = (hi - rho0) / (pres_max - Pref);
    guess = rho0 + slope * (p_used - Pref);
  }
  auto f = [&](Real r) { return P_rT(p, r, T); };
  if (!root::regula_falsi(f, p_used, guess, lo, hi, 1.0e-8, 1.0e-8, rho)) err |= ERRB_GRUN_PT;
  sie = e_rT(p, rho, T);
}
} // namespace grun

namespace vinet { // singularity-eos/eos/eos_vinet.hpp:228-352
// p = {rho0, T0, B0, BP0, A0, Cv0, E0, S0, VIP1, min_density}; coef = d2..d40 (39) then
// VIP[0..42] (43). Returns energy and pressure (and optionally the sound speed^2 term).
struct EP {
  Real e, P, cs2;
};
template <bool WITH_CS>
SGM_HD EP F_DT(const Real *p, const Real *coef, int nskip, Real rho, Real T) {
  const Real rho0 = p[0], T0 = p[1], B0 = p[2], A0 = p[4], Cv0 = p[5], E0 = p[6];
  const Real *d2tod40 = coef;
  const Real *VIP = coef + 39;
  const Real VIP1 = p[8];
  Real sumP = 0.0, sumB = 0.0, sumE = 0.0;
  const Real x = pmath::cbrt(robust::ratio(rho0, rho));
  const Real x2inv = robust::ratio(1.0, x * x);
  const Real onemx = 1.0 - x;
  const Real etatimes1mx = VIP1 * onemx;
  const Real expetatimes1mx = pmath::exp(etatimes1mx);
  for (int ind = 40 - nskip; ind >= 2; ind--) {
    sumP = d2tod40[ind - 2] + onemx * sumP;
    sumB = d2tod40[ind - 2] * ind + onemx * sumB;
    sumE = VIP[2 + ind] + onemx * sumE;
  }
  // VIP[3], VIP[2] are kept in p[11], p[10] (the registry entry, in LDS)
  sumE = p[11] + onemx * sumE;
  sumE = p[10] + onemx * sumE;
  sumP = 1.0 + sumP * (onemx * onemx);
  sumB = sumB * onemx;
  sumE = VIP1 * onemx * sumE;
  Real energy = 9.0 * robust::ratio(B0, VIP1 * VIP1 * rho0) *
                    (p[10] - expetatimes1mx * (p[10] - sumE)) -
                (A0 * B0) * (robust::ratio(T0, rho0) - robust::ratio(T0, rho));
  Real pressure = 3.0 * B0 * x2inv * onemx * expetatimes1mx * sumP;
  EP out;
  out.cs2 = 0;
  if (WITH_CS) {
    Real tmp = (1.0 + onemx * (VIP1 * x + 1.0)) * sumP + x * onemx * sumB;
    tmp = robust::ratio(B0, rho0) * x * expetatimes1mx * tmp;
    const Real dpdt = A0 * B0;
    out.cs2 = smax(0.0, tmp + T * robust::ratio(dpdt * dpdt, rho * rho * Cv0));
  }
  energy = energy + Cv0 * (T - T0) + E0;
  pressure = pressure + (A0 * B0) * (T - T0);
  out.e = energy;
  out.P = pressure;
  return out;
}
SGM_HD Real e_rT(const Real *p, const Real *c, int ns, Real rho, Real T) {
  return F_DT<false>(p, c, ns, rho, T).e;
}
SGM_HD Real P_rT(const Real *p, const Real *c, int ns, Real rho, Real T) {
  return F_DT<false>(p, c, ns, rho, T).P;
}
SGM_HD Real bmod_rT(const Real *p, const Real *c, int ns, Real rho, Real T) {
  const Real cs = sqrt(F_DT<true>(p, c, ns, rho, T).cs2);
  return cs * cs * rho;
}
SGM_HD Real T_re(const Real *p, const Real *c, int ns, Real rho, Real sie) {
  const EP o = F_DT<false>(p, c, ns, rho, p[1]);
  return robust::ratio(sie - o.e, p[5]) + p[1];
}
SGM_HD Real P_re(const Real *p, const Real *c, int ns, Real rho, Real sie) {
  return F_DT<false>(p, c, ns, rho, T_re(p, c, ns, rho, sie)).P;
}
SGM_HD Real bmod_re(const Real *p, const Real *c, int ns, Real rho, Real sie) {
  return bmod_rT(p, c, ns, rho, T_re(p, c, ns, rho, sie));
}
} // namespace vinet

// ---------------------------------------------------------------------------
// Variant-style dispatch over EosDev (singularity-eos/eos/eos_variant.hpp)
// ---------------------------------------------------------------------------
SGM_HD unsigned long base_preferred_input(const EosDev &e) {
  switch (e.kind) {
  case KIND_IDEAL:
    return tq::density | tq::specific_internal_energy | tq::temperature | tq::pressure;
  case KIND_STIFF:
  case KIND_GRUN: return tq::density | tq::specific_internal_energy;
  default: return tq::density | tq::temperature;
  }
}
SGM_HD Real base_e_rT(const EosDev &e, Real rho, Real T, Lam &lam) {
  switch (e.kind) {
  case KIND_IDEAL: return ideal::e_rT(e.m.p, rho, T);
  case KIND_STIFF: return stiff::e_rT(e.m.p, rho, T);
  case KIND_GRUN: return grun::e_rT(e.m.p, rho, T);
  case KIND_VINET: return vinet::e_rT(e.m.p, e.m.coef, e.nvinet_skip, rho, T);
  default: return table::e_rT(e.tab, rho, T, lam);
  }
}
// e(rho, T) and, for models whose preferred input is (rho, T) (Vinet, table),
// P(rho, T) at the same point from one evaluation: one Vinet F_DT, one table
// index and weight computation. The values equal separate eos_e_rT and
// eos_P_rT calls bit for bit. For (rho, e)-preferred models P_T is 0 and the
// caller evaluates P(rho, e).
SGM_HD Real base_e_rT_with_P(const EosDev &e, Real rho, Real T, Lam &lam, Real &P_T,
                             Real rhoP) {
  // rhoP: the density of the pressure (a scaled material's P sits at scale x
  // rho, its energy at rho); where it differs from rho the model is
  // evaluated a second time there, by the same code (a loop of one or two
  // passes, kept rolled)
  switch (e.kind) {
  case KIND_VINET: {
    Real en = 0.0, r = rho;
#pragma clang loop unroll(disable)
    for (int pass = 0; pass < 2; ++pass) {
      const vinet::EP o = vinet::F_DT<false>(e.m.p, e.m.coef, e.nvinet_skip, r, T);
      if (pass == 0) en = o.e;
      P_T = o.P;
      if (rhoP == rho) break;
      r = rhoP;
    }
    return en;
  }
  case KIND_TABLE: {
    const TableDev &t = e.tab;
    const Real lT = table::lT_(t, T);
    Real en = 0.0, r = rho;
#pragma clang loop unroll(disable)
    for (int pass = 0; pass < 2; ++pass) {
      const Real lRho = table::lRho_(t, r);
      lam.set(lRho, lT);
      const int w = table::loc_rho_T(t, lRho, lT);
      if (pass == 0) en = table::sie_from(t, lRho, T, lT, w);
      P_T = table::P_from(t, r, lRho, T, lT, w);
      if (rhoP == rho) break;
      r = rhoP;
    }
    return en;
  }
  default: P_T = 0.0; return base_e_rT(e, rho, T, lam);
  }
}
SGM_HD Real base_P_rT(const EosDev &e, Real rho, Real T, Lam &lam) {
  switch (e.kind) {
  case KIND_IDEAL: return ideal::P_rT(e.m.p, rho, T);
  case KIND_STIFF: return stiff::P_rT(e.m.p, rho, T);
  case KIND_GRUN: return grun::P_rT(e.m.p, rho, T);
  case KIND_VINET: return vinet::P_rT(e.m.p, e.m.coef, e.nvinet_skip, rho, T);
  default: return table::P_rT(e.tab, rho, T, lam);
  }
}
// e(rho, T) and P(rho, T) for every model: the shared evaluation where there
// is one (Vinet, table), separate closed forms otherwise
SGM_HD Real base_e_and_P_rT(const EosDev &e, Real rho, Real T, Lam &lam, Real &P) {
  if (e.kind == KIND_VINET || e.kind == KIND_TABLE) return base_e_rT_with_P(e, rho, T, lam, P, rho);
  P = base_P_rT(e, rho, T, lam);
  return base_e_rT(e, rho, T, lam);
}
SGM_HD Real base_P_re(const EosDev &e, Real rho, Real sie, Lam &lam) {
  switch (e.kind) {
  case KIND_IDEAL: return ideal::P_re(e.m.p, rho, sie);
  case KIND_STIFF: return stiff::P_re(e.m.p, rho, sie);
  case KIND_GRUN: return grun::P_re(e.m.p, rho, sie);
  case KIND_VINET: return vinet::P_re(e.m.p, e.m.coef, e.nvinet_skip, rho, sie);
  default: return table::P_re(e.tab, rho, sie, lam);
  }
}
SGM_HD Real base_T_re(const EosDev &e, Real rho, Real sie, Lam &lam) {
  switch (e.kind) {
  case KIND_IDEAL: return ideal::T_re(e.m.p, rho, sie);
  case KIND_STIFF: return stiff::T_re(e.m.p, rho, sie);
  case KIND_GRUN: return grun::T_re(e.m.p, rho, sie);
  case KIND_VINET: return vinet::T_re(e.m.p, e.m.coef, e.nvinet_skip, rho, sie);
  default: return table::T_re(e.tab, rho, sie, lam);
  }
}
SGM_HD Real base_cv_rT(const EosDev &e, Real rho, Real T, Lam &lam) {
  switch (e.kind) {
  case KIND_IDEAL:
  case KIND_STIFF: return e.m.p[0];
  case KIND_GRUN: return e.m.p[9];
  case KIND_VINET: return e.m.p[5];
  default: return table::cv_rT(e.tab, rho, T, lam);
  }
}
SGM_HD Real base_cv_re(const EosDev &e, Real rho, Real sie, Lam &lam) {
  switch (e.kind) {
  case KIND_IDEAL:
  case KIND_STIFF: return e.m.p[0];
  case KIND_GRUN: return e.m.p[9];
  case KIND_VINET: return e.m.p[5];
  default: return table::cv_re(e.tab, rho, sie, lam);
  }
}
SGM_HD Real base_bmod_rT(const EosDev &e, Real rho, Real T, Lam &lam) {
  switch (e.kind) {
  case KIND_IDEAL: return ideal::bmod_rT(e.m.p, rho, T);
  case KIND_STIFF: return stiff::bmod_rT(e.m.p, rho, T);
  case KIND_GRUN: return grun::bmod_rT(e.m.p, rho, T);
  case KIND_VINET: return vinet::bmod_rT(e.m.p, e.m.coef, e.nvinet_skip, rho, T);
  default: return table::bmod_rT(e.tab, rho, T, lam);
  }
}
// GruneisenParamFromDensityTemperature: IdealGas / StiffGas gm1
// (eos_ideal.hpp:140-144, eos_stiff.hpp:135-139), Gruneisen Gamma(min(rho,
// rho_max)) (eos_gruneisen.hpp:134-138), Vinet A0 B0 / (Cv0 rho)
// (eos_vinet.hpp:110-114), the table's dPdE / rho
SGM_HD Real base_gruneisen_rT(const EosDev &e, Real rho, Real T, Lam &lam) {
  switch (e.kind) {
  case KIND_IDEAL:
  case KIND_STIFF: return e.m.p[1];
  case KIND_GRUN: return grun::Gamma(e.m.p, smin(rho, e.m.p[10]));
  case KIND_VINET: return robust::ratio(e.m.p[4] * e.m.p[2], e.m.p[5] * rho);
  default: return table::gruneisen_rT(e.tab, rho, T, lam);
  }
}
// the temperature of ValuesAtReferenceState: room temperature for IdealGas
// and StiffGas (eos_ideal.hpp:224, eos_stiff.hpp:42), T0 for Gruneisen and
// Vinet (eos_gruneisen.hpp:558, eos_vinet.hpp:408), TNormal for the table
SGM_HD Real base_T_reference(const EosDev &e) {
  switch (e.kind) {
  case KIND_IDEAL:
  case KIND_STIFF: return 293;
  case KIND_GRUN: return e.m.p[7];
  case KIND_VINET: return e.m.p[1];
  default: return table::T_normal(e.tab);
  }
}
SGM_HD Real base_bmod_re(const EosDev &e, Real rho, Real sie, Lam &lam) {
  switch (e.kind) {
  case KIND_IDEAL: return ideal::bmod_re(e.m.p, rho, sie);
  case KIND_STIFF: return stiff::bmod_re(e.m.p, rho, sie);
  case KIND_GRUN: return grun::bmod_re(e.m.p, rho, sie);
  case KIND_VINET: return vinet::bmod_re(e.m.p, e.m.coef, e.nvinet_skip, rho, sie);
  default: return table::bmod_re(e.tab, rho, sie, lam);
  }
}
// EosBase::MinInternalEnergyFromDensity = e(rho, 0) (eos_base.hpp:598-603);
// IdealGas 0 (eos_ideal.hpp:95-99); Gruneisen clamps rho (:373-377); table cold curve.
SGM_HD Real base_min_e(const EosDev &e, Real rho) {
  Lam none;
  switch (e.kind) {
  case KIND_IDEAL: return 0.0;
  case KIND_STIFF: return stiff::e_rT(e.m.p, rho, 0.);
  case KIND_GRUN: return grun::e_rT(e.m.p, smin(rho, e.m.p[10]), 0.);
  case KIND_VINET: return vinet::e_rT(e.m.p, e.m.coef, e.nvinet_skip, rho, 0.);
  default: return lerp_rho(e.tab, e.tab.sieCold, table::lRho_(e.tab, rho));
  }
  (void)none;
}
SGM_HD Real base_min_density(const EosDev &e) {
  if (e.kind == KIND_VINET) return e.m.p[9];
  if (e.kind == KIND_TABLE) return table::rho_(e.tab, e.tab.lRhoMin);
  return 0;
}
SGM_HD Real base_min_temperature(const EosDev &e) {
  if (e.kind == KIND_TABLE) return table::T_(e.tab, e.tab.lTMin);
  return 0;
}
SGM_HD Real base_max_density(const EosDev &e) {
  if (e.kind == KIND_GRUN) return e.m.p[10];
  if (e.kind == KIND_TABLE) return e.tab.rhoMax;
  return 1e100;
}
SGM_HD Real base_min_pressure(const EosDev &e) {
  if (e.kind == KIND_GRUN) return grun::P_re(e.m.p, 0, 0);
  if (e.kind == KIND_VINET) return -1e100;
  if (e.kind == KIND_TABLE) return e.tab.PMin;
  return 0;
}
// Gruneisen returns its maximum stable density here, as the reference does
// (eos_gruneisen.hpp:160-163).
SGM_HD Real base_max_pressure_at_T(const EosDev &e, Real T, int &err) {
  if (e.kind == KIND_GRUN) return grun::max_stable_density(e.m.p, T, err);
  return 1e100;
}
SGM_HD Real base_rho_pmin(const EosDev &e, Real T) {
  if (e.kind == KIND_TABLE) return table::rho_pmin(e.tab, T);
  return 0.0;
}
// DensityEnergyFromPressureTemperature; rho is the in/out guess. The
// EosBase default (eos_base.hpp:630-671, findRoot) serves Vinet.
SGM_HD void base_rhoe_PT(const EosDev &e, Real P, Real T, Real &rho, Real &sie, Lam &lam,
                           int &err) {
  switch (e.kind) {
  case KIND_IDEAL: ideal::rhoe_PT(e.m.p, P, T, rho, sie); return;
  case KIND_STIFF: stiff::rhoe_PT(e.m.p, P, T, rho, sie); return;
  case KIND_GRUN: grun::rhoe_PT(e.m.p, P, T, rho, sie, err); return;
  case KIND_TABLE: table::rhoe_PT(e.tab, P, T, rho, sie, lam); return;
  default: break;
  }
  const Real rhomin = smax(0.0, e.m.p[9]);
  const Real rhomax = 1e100;
  auto f = [&](Real r) { return vinet::P_rT(e.m.p, e.m.coef, e.nvinet_skip, r, T); };
  Real rhoguess = rho;
  if ((rhoguess <= rhomin) || (rhoguess >= rhomax)) {
    if ((rhomin < 12.0) && (12.0 < rhomax))
      rhoguess = 12.0;
    else
      rhoguess = 0.5 * (rhomin + rhomax);
  }
  root::find_root(f, P, rhoguess, rhomin, rhomax, robust::EPS(), robust::EPS(), rho);
  sie = vinet::e_rT(e.m.p, e.m.coef, e.nvinet_skip, rho, T);
}

// FillEos for the (rho, sie) -> {T, P, cv, bmod} and (rho, T) -> {sie, P, cv, bmod}
// input modes (eos_ideal.hpp:236-262, eos_stiff.hpp:203-229, eos_gruneisen.hpp:527-549,
// eos_vinet.hpp:373-395, eos_spiner_rho_temp.hpp:1052-1101) plus (P, T) -> {rho, sie}.
// FillEos per model (the bodies of each model's FillEos, eos_base.hpp and
// the model headers), one function per kind so that a kernel built for one
// kind holds that model's code alone
SGM_HD void fill_gas(const EosDev &e, Real &rho, Real &temp, Real &sie, Real &press,
                     Real &cv, Real &bmod, unsigned long output, Lam &, int &err) {
  using namespace tq;
  const bool st = e.kind == KIND_STIFF;
  if (output & density && output & specific_internal_energy) {
    if (output & pressure || output & temperature) {
      err |= ERRB_FILLEOS;
      return;
    }
    if (st)
      stiff::rhoe_PT(e.m.p, press, temp, rho, sie);
    else
      ideal::rhoe_PT(e.m.p, press, temp, rho, sie);
  }
  if (output & pressure && output & specific_internal_energy) {
    if (output & density || output & temperature) {
      err |= ERRB_FILLEOS;
      return;
    }
    sie = st ? stiff::e_rT(e.m.p, rho, temp) : ideal::e_rT(e.m.p, rho, temp);
  }
  if (output & temperature && output & specific_internal_energy) {
    sie = st ? robust::ratio(press + (e.m.p[1] + 1.0) * e.m.p[2], e.m.p[1] * rho) + e.m.p[3]
             : press / (e.m.p[1] * rho);
  }
  if (output & pressure) press = st ? stiff::P_re(e.m.p, rho, sie) : ideal::P_re(e.m.p, rho, sie);
  if (output & temperature) temp = st ? stiff::T_re(e.m.p, rho, sie) : ideal::T_re(e.m.p, rho, sie);
  if (output & bulk_modulus)
    bmod = st ? stiff::bmod_re(e.m.p, rho, sie) : ideal::bmod_re(e.m.p, rho, sie);
  if (output & specific_heat) cv = e.m.p[0];
}
SGM_HD void fill_grun(const EosDev &e, Real &rho, Real &temp, Real &sie, Real &press,
                     Real &cv, Real &bmod, unsigned long output, Lam &lam, int &err) {
  using namespace tq;
  const unsigned long input = ~output;
  if (temperature & input && pressure & input) {
    grun::rhoe_PT(e.m.p, press, temp, rho, sie, err);
  } else if (density & output || specific_internal_energy & output) {
    err |= ERRB_FILLEOS;
    return;
  }
  const Real r = smin(rho, e.m.p[10]);
  if (output & pressure) press = grun::P_re(e.m.p, r, sie);
  if (output & temperature) temp = grun::T_re(e.m.p, r, sie);
  if (output & bulk_modulus) bmod = grun::bmod_re(e.m.p, r, sie);
  if (output & specific_heat) cv = e.m.p[9];
  return;
}
SGM_HD void fill_vinet(const EosDev &e, Real &rho, Real &temp, Real &sie, Real &press,
                     Real &cv, Real &bmod, unsigned long output, Lam &lam, int &err) {
  using namespace tq;
  const unsigned long input = ~output;
  if (density & output) {
    err |= ERRB_FILLEOS;
    return;
  } else if (temperature & output && specific_internal_energy & output) {
    err |= ERRB_FILLEOS;
    return;
  }
  if (specific_internal_energy & input)
    temp = vinet::T_re(e.m.p, e.m.coef, e.nvinet_skip, rho, sie);
  const bool need_cs = (output & bulk_modulus) != 0;
  vinet::EP o = need_cs ? vinet::F_DT<true>(e.m.p, e.m.coef, e.nvinet_skip, rho, temp)
                        : vinet::F_DT<false>(e.m.p, e.m.coef, e.nvinet_skip, rho, temp);
  if (output & specific_internal_energy) sie = o.e;
  if (output & pressure) press = o.P;
  if (output & specific_heat) cv = e.m.p[5];
  if (output & bulk_modulus) {
    const Real cs = sqrt(o.cs2);
    bmod = cs * cs * rho;
  }
  return;
}
SGM_HD void fill_table(const EosDev &e, Real &rho, Real &temp, Real &sie, Real &press,
                     Real &cv, Real &bmod, unsigned long output, Lam &lam, int &err) {
  using namespace tq;
  const unsigned long input = ~output;
  const TableDev &t = e.tab;
  Real lRho, lT;
  int w;
  if (output == 0) {
    err |= ERRB_FILLEOS;
    return;
  }
  if (output & density) {
    if (input & pressure && input & temperature) {
      lT = table::lT_(t, temp);
      const table::InvRho iv = table::lRho_from_P_lT(t, press, lT, lam);
      lam = iv.lam;
      lRho = iv.lRho;
      rho = table::rho_(t, lRho);
    } else {
      err |= ERRB_FILLEOS;
      return;
    }
  } else {
    lRho = table::lRho_(t, rho);
  }
  if (output & temperature) {
    if (input & density && input & specific_internal_energy) {
      const table::InvT iv = table::lT_from_lRho_sie(t, lRho, sie, lam);
      lam = iv.lam;
      lT = iv.lT;
    } else {
      err |= ERRB_FILLEOS;
      return;
    }
    temp = table::T_(t, lT);
  } else {
    lT = table::lT_(t, temp);
  }
  w = table::loc_rho_T(t, lRho, lT);
  if (output & specific_internal_energy) sie = table::sie_from(t, lRho, temp, lT, w);
  if (output & pressure) press = table::P_from(t, rho, lRho, temp, lT, w);
  if (output & specific_heat) cv = table::cv_from(t, lRho, lT, w);
  if (output & bulk_modulus) bmod = table::bmod_from(t, rho, lRho, temp, lT, w);
  lam.set(lRho, lT);
}
SGM_HD void base_fill(const EosDev &e, Real &rho, Real &temp, Real &sie, Real &press, Real &cv,
                     Real &bmod, unsigned long output, Lam &lam, int &err) {
  switch (e.kind) {
  case KIND_IDEAL:
  case KIND_STIFF: fill_gas(e, rho, temp, sie, press, cv, bmod, output, lam, err); return;
  case KIND_GRUN: fill_grun(e, rho, temp, sie, press, cv, bmod, output, lam, err); return;
  case KIND_VINET: fill_vinet(e, rho, temp, sie, press, cv, bmod, output, lam, err); return;
  default: fill_table(e, rho, temp, sie, press, cv, bmod, output, lam, err); return;
  }
}
// base_fill for a material known to be of kind KIND
template <int KIND>
SGM_HD void base_fill_k(const EosDev &e, Real &rho, Real &temp, Real &sie, Real &press, Real &cv,
                        Real &bmod, unsigned long output, Lam &lam, int &err) {
  if constexpr (KIND == KIND_IDEAL || KIND == KIND_STIFF)
    fill_gas(e, rho, temp, sie, press, cv, bmod, output, lam, err);
  else if constexpr (KIND == KIND_GRUN)
    fill_grun(e, rho, temp, sie, press, cv, bmod, output, lam, err);
  else if constexpr (KIND == KIND_VINET)
    fill_vinet(e, rho, temp, sie, press, cv, bmod, output, lam, err);
  else
    fill_table(e, rho, temp, sie, press, cv, bmod, output, lam, err);
}

// ---------------------------------------------------------------------------
// The material as the variant sees it: the model under its modifiers.
// ScaledEOS forwards (scale rho, sie / scale) to ShiftedEOS, which forwards
// (rho, sie - shift) to the model; energies come back + shift, then x scale
// (shifted_eos.hpp:64-158, 399-403; scaled_eos.hpp:66-166, 328-356).
// Each call is one model evaluation with its inputs and outputs mapped (an
// exact pass-through without modifiers), inlined like the models; nothing is
// called out of line, so no caller state has to leave registers. A scaled
// material evaluates its energy at rho but its pressure at scale x rho, so
// the fused e/P evaluation of a (rho, T) model runs its model code a second
// time there.
// ---------------------------------------------------------------------------
// the modifier fields are read only on a modified material's path (a
// branch, not a select), so an unmodified material's evaluation reads one
// int more than the model's own
SGM_HD int mods_of(const EosDev &e) {
  int m = e.mods;
#if defined(__HIP_DEVICE_COMPILE__)
  asm volatile("" : "+v"(m));
#endif
  return m;
}
SGM_HD Real mod_rho(const EosDev &e, Real rho) {
  if (__builtin_expect((mods_of(e) & MOD_SCALED) != 0, 0)) rho = e.scale * rho;
  return rho;
}
SGM_HD Real mod_sie(const EosDev &e, Real sie) {
  const int m = mods_of(e);
  if (__builtin_expect(m != 0, 0)) {
    if (m & MOD_SCALED) sie = e.inv_scale * sie;
    if (m & MOD_SHIFTED) sie = sie - e.shift;
  }
  return sie;
}
SGM_HD Real mod_energy_out(const EosDev &e, Real en) {
  const int m = mods_of(e);
  if (__builtin_expect(m != 0, 0)) {
    if (m & MOD_SHIFTED) en = en + e.shift;
    if (m & MOD_SCALED) en = e.scale * en;
  }
  return en;
}
SGM_HD Real mod_rho_bound(const EosDev &e, Real r) {
  if (__builtin_expect((mods_of(e) & MOD_SCALED) != 0, 0)) r = e.inv_scale * r;
  return r;
}
SGM_HD unsigned long preferred_input(const EosDev &e) { return base_preferred_input(e); }
// scaled_eos.hpp:80-84: ScaledEOS's energy takes rho unscaled
SGM_HD Real eos_e_rT(const EosDev &e, Real rho, Real T, Lam &lam) {
  return mod_energy_out(e, base_e_rT(e, rho, T, lam));
}
SGM_HD Real eos_P_rT(const EosDev &e, Real rho, Real T, Lam &lam) {
  return base_P_rT(e, mod_rho(e, rho), T, lam);
}
SGM_HD Real eos_e_rT_with_P(const EosDev &e, Real rho, Real T, Lam &lam, Real &P_T) {
  return mod_energy_out(e, base_e_rT_with_P(e, rho, T, lam, P_T, mod_rho(e, rho)));
}
SGM_HD Real eos_e_and_P_rT(const EosDev &e, Real rho, Real T, Lam &lam, Real &P) {
  if (e.kind == KIND_VINET || e.kind == KIND_TABLE) return eos_e_rT_with_P(e, rho, T, lam, P);
  P = eos_P_rT(e, rho, T, lam);
  return eos_e_rT(e, rho, T, lam);
}
SGM_HD Real eos_P_re(const EosDev &e, Real rho, Real sie, Lam &lam) {
  return base_P_re(e, mod_rho(e, rho), mod_sie(e, sie), lam);
}
SGM_HD Real eos_T_re(const EosDev &e, Real rho, Real sie, Lam &lam) {
  return base_T_re(e, mod_rho(e, rho), mod_sie(e, sie), lam);
}
SGM_HD Real eos_cv_rT(const EosDev &e, Real rho, Real T, Lam &lam) {
  return base_cv_rT(e, mod_rho(e, rho), T, lam);
}
SGM_HD Real eos_cv_re(const EosDev &e, Real rho, Real sie, Lam &lam) {
  return base_cv_re(e, mod_rho(e, rho), mod_sie(e, sie), lam);
}
SGM_HD Real eos_bmod_rT(const EosDev &e, Real rho, Real T, Lam &lam) {
  return base_bmod_rT(e, mod_rho(e, rho), T, lam);
}
SGM_HD Real eos_gruneisen_rT(const EosDev &e, Real rho, Real T, Lam &lam) {
  return base_gruneisen_rT(e, mod_rho(e, rho), T, lam);
}
SGM_HD Real eos_T_reference(const EosDev &e) { return base_T_reference(e); }
SGM_HD Real eos_bmod_re(const EosDev &e, Real rho, Real sie, Lam &lam) {
  return base_bmod_re(e, mod_rho(e, rho), mod_sie(e, sie), lam);
}
// scaled_eos.hpp:88-91 (rho scaled, energy not), shifted_eos.hpp:81-84 (+ shift)
SGM_HD Real eos_min_e(const EosDev &e, Real rho) {
  const Real en = base_min_e(e, mod_rho(e, rho));
  return (e.mods & MOD_SHIFTED) ? en + e.shift : en;
}
SGM_HD Real eos_min_density(const EosDev &e) { return mod_rho_bound(e, base_min_density(e)); }
SGM_HD Real eos_min_temperature(const EosDev &e) { return base_min_temperature(e); }
SGM_HD Real eos_max_density(const EosDev &e) { return mod_rho_bound(e, base_max_density(e)); }
SGM_HD Real eos_min_pressure(const EosDev &e) { return base_min_pressure(e); }
SGM_HD Real eos_max_pressure_at_T(const EosDev &e, Real T, int &err) {
  return base_max_pressure_at_T(e, T, err);
}
SGM_HD Real eos_rho_pmin(const EosDev &e, Real T) { return mod_rho_bound(e, base_rho_pmin(e, T)); }
// the model's P-T inversion out of line, its arguments and results by value
// (a reference to the caller's state would keep that state out of registers)
struct RhoSie {
  Real rho, sie;
  Lam lam;
  int err;
};
SGM_NOINL RhoSie base_rhoe_PT_v(const EosDev *e, Real P, Real T, Real rho, Lam lam) {
  RhoSie r;
  r.rho = rho;
  r.sie = 0.0;
  r.err = 0;
  base_rhoe_PT(*e, P, T, r.rho, r.sie, lam, r.err);
  r.lam = lam;
  return r;
}
SGM_HD void eos_rhoe_PT(const EosDev &e, Real P, Real T, Real &rho, Real &sie, Lam &lam,
                        int &err) {
  const RhoSie v = base_rhoe_PT_v(&e, P, T, rho, lam);
  rho = v.rho;
  sie = v.sie;
  lam = v.lam;
  err |= v.err;
  if (e.mods & MOD_SHIFTED) sie = sie + e.shift;
  if (e.mods & MOD_SCALED) {
    rho = rho * e.inv_scale;
    sie = sie * e.scale;
  }
}
// FillEos of ScaledEOS over ShiftedEOS (scaled_eos.hpp:148-166,
// shifted_eos.hpp:142-158): by the model's preferred input, (rho, T) or
// (rho, sie); a modified IdealGas (whose preferred input is neither) is the
// reference's EOS_ERROR, here ERRB_FILLEOS
// eos_fill for a material of kind KIND (KIND < 0: any kind)
template <int KIND>
SGM_HD void eos_fill_k(const EosDev &e, Real &rho, Real &temp, Real &sie, Real &press, Real &cv,
                       Real &bmod, unsigned long output, Lam &lam, int &err) {
  auto fill = [&](Real &r, Real &s) {
    if constexpr (KIND < 0)
      base_fill(e, r, temp, s, press, cv, bmod, output, lam, err);
    else
      base_fill_k<KIND>(e, r, temp, s, press, cv, bmod, output, lam, err);
  };
  if (!e.mods) {
    fill(rho, sie);
    return;
  }
  const unsigned long pref = base_preferred_input(e);
  const bool rhoT = pref == (tq::density | tq::temperature);
  if (!rhoT && pref != (tq::density | tq::specific_internal_energy)) {
    err |= ERRB_FILLEOS;
    return;
  }
  Real srho = mod_rho(e, rho);
  Real senergy = rhoT ? sie : mod_sie(e, sie);
  fill(srho, senergy);
  if (!(e.mods & MOD_SCALED)) rho = srho;  // an unscaled model fills the caller's rho
  if (rhoT) sie = mod_energy_out(e, senergy);
}
SGM_HD void eos_fill(const EosDev &e, Real &rho, Real &temp, Real &sie, Real &press, Real &cv,
                     Real &bmod, unsigned long output, Lam &lam, int &err) {
  eos_fill_k<-1>(e, rho, temp, sie, press, cv, bmod, output, lam, err);
}


// The EOS calls of the closure solvers, selected at compile time: Ev<true>
// applies the modifiers (the eos_* functions above), Ev<false> calls the
// models alone (the base_* functions) for kernels that only ever see
// unmodified materials, whose hot loops then carry no modifier code.
template <bool MODS>
struct Ev;
template <>
struct Ev<true> {
  static SGM_HD unsigned long preferred_input(const EosDev &e) { return sgm::preferred_input(e); }
  static SGM_HD Real P_rT(const EosDev &e, Real r, Real T, Lam &l) { return eos_P_rT(e, r, T, l); }
  static SGM_HD Real P_re(const EosDev &e, Real r, Real s, Lam &l) { return eos_P_re(e, r, s, l); }
  static SGM_HD Real T_re(const EosDev &e, Real r, Real s, Lam &l) { return eos_T_re(e, r, s, l); }
  static SGM_HD Real cv_rT(const EosDev &e, Real r, Real T, Lam &l) { return eos_cv_rT(e, r, T, l); }
  static SGM_HD Real cv_re(const EosDev &e, Real r, Real s, Lam &l) { return eos_cv_re(e, r, s, l); }
  static SGM_HD Real e_rT(const EosDev &e, Real r, Real T, Lam &l) { return eos_e_rT(e, r, T, l); }
  static SGM_HD Real e_rT_with_P(const EosDev &e, Real r, Real T, Lam &l, Real &P) {
    return eos_e_rT_with_P(e, r, T, l, P);
  }
  static SGM_HD Real e_and_P_rT(const EosDev &e, Real r, Real T, Lam &l, Real &P) {
    return eos_e_and_P_rT(e, r, T, l, P);
  }
  static SGM_HD Real max_density(const EosDev &e) { return eos_max_density(e); }
  static SGM_HD Real max_pressure_at_T(const EosDev &e, Real T, int &err) {
    return eos_max_pressure_at_T(e, T, err);
  }
  static SGM_HD Real min_pressure(const EosDev &e) { return eos_min_pressure(e); }
  static SGM_HD Real min_temperature(const EosDev &e) { return eos_min_temperature(e); }
  static SGM_HD Real rho_pmin(const EosDev &e, Real T) { return eos_rho_pmin(e, T); }
  static SGM_HD void rhoe_PT(const EosDev &e, Real P, Real T, Real &r, Real &s, Lam &l, int &err) {
    eos_rhoe_PT(e, P, T, r, s, l, err);
  }
  static SGM_HD bool modified(const EosDev &e) { return e.mods != 0; }
};
template <>
struct Ev<false> {
  static SGM_HD unsigned long preferred_input(const EosDev &e) { return base_preferred_input(e); }
  static SGM_HD Real P_rT(const EosDev &e, Real r, Real T, Lam &l) { return base_P_rT(e, r, T, l); }
  static SGM_HD Real P_re(const EosDev &e, Real r, Real s, Lam &l) { return base_P_re(e, r, s, l); }
  static SGM_HD Real T_re(const EosDev &e, Real r, Real s, Lam &l) { return base_T_re(e, r, s, l); }
  static SGM_HD Real cv_rT(const EosDev &e, Real r, Real T, Lam &l) { return base_cv_rT(e, r, T, l); }
  static SGM_HD Real cv_re(const EosDev &e, Real r, Real s, Lam &l) { return base_cv_re(e, r, s, l); }
  static SGM_HD Real e_rT(const EosDev &e, Real r, Real T, Lam &l) { return base_e_rT(e, r, T, l); }
  static SGM_HD Real e_rT_with_P(const EosDev &e, Real r, Real T, Lam &l, Real &P) {
    return base_e_rT_with_P(e, r, T, l, P, r);
  }
  static SGM_HD Real e_and_P_rT(const EosDev &e, Real r, Real T, Lam &l, Real &P) {
    return base_e_and_P_rT(e, r, T, l, P);
  }
  static SGM_HD Real max_density(const EosDev &e) { return base_max_density(e); }
  static SGM_HD Real max_pressure_at_T(const EosDev &e, Real T, int &err) {
    return base_max_pressure_at_T(e, T, err);
  }
  static SGM_HD Real min_pressure(const EosDev &e) { return base_min_pressure(e); }
  static SGM_HD Real min_temperature(const EosDev &e) { return base_min_temperature(e); }
  static SGM_HD Real rho_pmin(const EosDev &e, Real T) { return base_rho_pmin(e, T); }
  static SGM_HD void rhoe_PT(const EosDev &e, Real P, Real T, Real &r, Real &s, Lam &l, int &err) {
    const RhoSie v = base_rhoe_PT_v(&e, P, T, r, l);
    r = v.rho;
    s = v.sie;
    l = v.lam;
    err |= v.err;
  }
  static SGM_HD bool modified(const EosDev &) { return false; }
};

} // namespace sgm
