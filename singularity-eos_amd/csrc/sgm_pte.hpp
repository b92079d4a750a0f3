// Register-resident pressure-temperature-equilibrium closure for one mixed
// cell per lane, with the material count NM fixed at compile time so every
// per-material array lives in VGPRs. Restates
// singularity-eos/closure/mixed_cell_models.hpp (PTESolverBase :329-786,
// PTESolverRhoT :873-1193, PTESolverPT :1198-1446, PTESolverFixedT
// :1451-1669, PTESolverFixedP :1674-1921, PTESolver :2188-2291) in
// the same operation order. Indices that the reference computes at run time
// (the excluded material `ms`, column pivots) are resolved with unrolled
// selects instead of dynamic register indexing.
#pragma once

#include "sgm_eos.hpp"

#include <type_traits>

namespace sgm {

// MixParams (mixed_cell_models.hpp:43-78) as a flat POD for kernel arguments
struct MixParamsDev {
  Real derivative_eps;
  Real pte_rel_tolerance_p, pte_rel_tolerance_t, pte_rel_tolerance_e, pte_rel_tolerance_v;
  Real pte_abs_tolerance_p, pte_abs_tolerance_t;
  Real pte_slow_convergence_thresh;
  Real pte_rel_tolerance_p_sufficient, pte_rel_tolerance_t_sufficient;
  Real pte_rel_tolerance_e_sufficient, pte_rel_tolerance_v_sufficient;
  Real pte_abs_tolerance_p_sufficient, pte_abs_tolerance_t_sufficient;
  Real line_search_alpha, line_search_fac, vfrac_safety_fac, temperature_limit;
  Real default_tguess, min_dtde, pte_small_step_thresh, pte_max_dpdv;
  int pte_max_iter_per_mat, line_search_max_iter, pte_small_step_tries;
  int pte_reduced_system_exclude_idx, iterate_t_guess, pad;
};

struct CellStatus {
  int converged;
  int niter;
  int err;
  Real residual;
};

// --- unrolled selects -------------------------------------------------------
// A value the optimizer cannot look through. Selecting between two array
// elements by a run-time index is otherwise rewritten into one load from a
// run-time address, which pins the whole per-cell state in scratch memory
// instead of registers; selecting between opaque values keeps it in VGPRs.
SGM_HD Real opaque(Real x) {
#if defined(__HIP_DEVICE_COMPILE__)
  asm("" : "+v"(x));
#endif
  return x;
}
SGM_HD Real sel(bool c, Real a, Real b) { return c ? opaque(a) : opaque(b); }
template <int N>
SGM_HD Real pick(const Real (&a)[N], int i) {
  Real r = opaque(a[0]);
#pragma unroll
  for (int k = 1; k < N; ++k) r = sel(i == k, a[k], r);
  return r;
}
template <int N>
SGM_HD void put(Real (&a)[N], int i, Real v) {
#pragma unroll
  for (int k = 0; k < N; ++k) a[k] = sel(i == k, v, a[k]);
}

// --- small dense solve (mixed_cell_models.hpp:176-278) ----------------------
// Column-pivoted Householder QR for N >= 3 (the reference's Eigen
// colPivHouseholderQr). Pivot on the largest updated column norm (first on
// ties), LAPACK-style norm downdating with a sqrt(eps) recompute threshold,
// beta = -sign(c0)||x||, pivots with squared norm below (max*eps)^2/N*(N-k)
// treated as zero.
template <int N>
SGM_HD void qr_colpiv_solve(Real (&a)[N * N], Real (&b)[N], Real (&x)[N]) {
  int perm[N];
  Real cu[N], cd[N];
#pragma unroll
  for (int j = 0; j < N; ++j) {
    perm[j] = j;
    Real s = 0;
#pragma unroll
    for (int i = 0; i < N; ++i) s += a[i * N + j] * a[i * N + j];
    cd[j] = sqrt(s);
    cu[j] = cd[j];
  }
  const Real eps = 2.220446049250313e-16;
  const Real norm_downdate_threshold = sqrt(eps);
  int nzp = N;
  Real th = 0;
#pragma unroll
  for (int k = 0; k < N; ++k) {
    int big = k;
    Real bign = cu[k];
#pragma unroll
    for (int j = k + 1; j < N; ++j)
      if (cu[j] > bign) {
        bign = cu[j];
        big = j;
      }
    const Real big_sq = bign * bign;
    if (k == 0) th = big_sq * eps * eps / N;
    if (nzp == N && big_sq < th * (N - k)) nzp = k;
#pragma unroll
    for (int j = k + 1; j < N; ++j) {
      if (big == j) {
#pragma unroll
        for (int i = 0; i < N; ++i) {
          const Real t = a[i * N + k];
          a[i * N + k] = a[i * N + j];
          a[i * N + j] = t;
        }
        const int tp = perm[k];
        perm[k] = perm[j];
        perm[j] = tp;
        Real t = cu[k];
        cu[k] = cu[j];
        cu[j] = t;
        t = cd[k];
        cd[k] = cd[j];
        cd[j] = t;
      }
    }
    Real tailsq = 0;
#pragma unroll
    for (int i = k + 1; i < N; ++i) tailsq += a[i * N + k] * a[i * N + k];
    const Real c0 = a[k * N + k];
    Real tau, beta;
    if (tailsq <= 2.2250738585072014e-308) {
      tau = 0;
      beta = c0;
#pragma unroll
      for (int i = k + 1; i < N; ++i) a[i * N + k] = 0;
    } else {
      beta = sqrt(c0 * c0 + tailsq);
      if (c0 >= 0) beta = -beta;
      const Real den = c0 - beta;
#pragma unroll
      for (int i = k + 1; i < N; ++i) a[i * N + k] = a[i * N + k] / den;
      tau = (beta - c0) / beta;
    }
    a[k * N + k] = beta;
#pragma unroll
    for (int j = k + 1; j < N; ++j) {
      Real s = 0;
#pragma unroll
      for (int i = k + 1; i < N; ++i) s += a[i * N + k] * a[i * N + j];
      s += a[k * N + j];
      a[k * N + j] -= tau * s;
#pragma unroll
      for (int i = k + 1; i < N; ++i) a[i * N + j] -= (tau * a[i * N + k]) * s;
    }
    {
      Real s = 0;
#pragma unroll
      for (int i = k + 1; i < N; ++i) s += a[i * N + k] * b[i];
      s += b[k];
      b[k] -= tau * s;
#pragma unroll
      for (int i = k + 1; i < N; ++i) b[i] -= (tau * a[i * N + k]) * s;
    }
#pragma unroll
    for (int j = k + 1; j < N; ++j) {
      if (cu[j] != 0) {
        Real t = fabs(a[k * N + j]) / cu[j];
        t = (1 + t) * (1 - t);
        if (t < 0) t = 0;
        const Real r = cu[j] / cd[j];
        const Real t2 = t * r * r;
        if (t2 <= norm_downdate_threshold) {
          Real s = 0;
#pragma unroll
          for (int i = k + 1; i < N; ++i) s += a[i * N + j] * a[i * N + j];
          cd[j] = sqrt(s);
          cu[j] = cd[j];
        } else {
          cu[j] *= sqrt(t);
        }
      }
    }
  }
  Real z[N];
#pragma unroll
  for (int i = 0; i < N; ++i) z[i] = b[i];
#pragma unroll
  for (int i = N - 1; i >= 0; --i) {
    if (i < nzp) {
      z[i] = z[i] / a[i * N + i];
#pragma unroll
      for (int r = 0; r < i; ++r) z[r] -= z[i] * a[r * N + i];
    }
  }
#pragma unroll
  for (int j = 0; j < N; ++j) x[j] = 0;
#pragma unroll
  for (int i = 0; i < N; ++i) {
    const Real zi = (i < nzp) ? z[i] : 0.0;
#pragma unroll
    for (int j = 0; j < N; ++j)
      if (perm[i] == j) x[j] = zi;
  }
}

SGM_HD bool isfinite_ref(Real a) { return (a == a) && ((a == 0) || (a != 2 * a)); }

// solve_Ax_b_wscr: Jacobi row scaling, then direct (N=1), Cramer (N=2) or
// pivoted QR (N>=3). b is overwritten with x. false on a zero row or non-finite x.
template <int N>
SGM_HD bool solve_small(Real (&a)[N * N], Real (&b)[N]) {
#pragma unroll
  for (int row = 0; row < N; ++row) {
    Real maxabs = 0;
#pragma unroll
    for (int c = 0; c < N; ++c) maxabs = smax(fabs(a[row * N + c]), maxabs);
    if (maxabs == 0) return false;
#pragma unroll
    for (int c = 0; c < N; ++c) a[row * N + c] /= maxabs;
    b[row] /= maxabs;
  }
  if constexpr (N == 1) {
    b[0] = robust::ratio(b[0], a[0]);
  } else if constexpr (N == 2) {
    const Real det = a[0] * a[3] - a[1] * a[2];
    const Real x0 = robust::ratio(a[3] * b[0] - a[1] * b[1], det);
    const Real x1 = robust::ratio(a[0] * b[1] - a[2] * b[0], det);
    b[0] = x0;
    b[1] = x1;
  } else {
    Real x[N];
    qr_colpiv_solve<N>(a, b, x);
#pragma unroll
    for (int i = 0; i < N; ++i) b[i] = x[i];
  }
  bool ok = true;
#pragma unroll
  for (int i = 0; i < N; ++i)
    if (!isfinite_ref(b[i])) ok = false;
  return ok;
}

// lane k's x, k the same on every lane (two v_readlane)
__device__ __forceinline__ Real lane_value(Real x, int k) {
  const unsigned long long u = __builtin_bit_cast(unsigned long long, x);
  const unsigned lo = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)u, k);
  const unsigned hi = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)(u >> 32), k);
  return __builtin_bit_cast(Real, ((unsigned long long)hi << 32) | lo);
}

// solve_small<N> by a whole wave whose lanes hold identical copies of a and
// b: lane j < N keeps column j of the scaled matrix in registers and applies
// the reflections and norm downdates to it; pivoting swaps entries of a
// wave-uniform position -> lane map instead of moving columns; the pivot
// search, each Householder vector, the reflection of b and the back
// substitution read the columns they need with v_readlane and run alike on
// every lane. Every value is formed by the operations of solve_small in
// their order, so b (overwritten with x on every lane) and the result are
// bit-identical to it. Needs every lane of the wave active.
template <int N>
__device__ bool solve_small_wave(const Real (&a)[N * N], Real (&b)[N], int lane) {
  static_assert(N >= 3, "the wave form covers the QR sizes");
  const int j = lane < N ? lane : N - 1;
  Real rmax[N];
  bool zero = false;
#pragma unroll
  for (int row = 0; row < N; ++row) {
    Real maxabs = 0;
#pragma unroll
    for (int c = 0; c < N; ++c) maxabs = smax(fabs(a[row * N + c]), maxabs);
    rmax[row] = maxabs;
    zero = zero || maxabs == 0;
  }
  if (zero) return false;
  Real col[N];
#pragma unroll
  for (int i = 0; i < N; ++i) {
    Real v = opaque(a[i * N]);
#pragma unroll
    for (int c = 1; c < N; ++c) v = sel(j == c, a[i * N + c], v);
    col[i] = v / rmax[i];
    b[i] /= rmax[i];
  }
  Real cu, cd;
  {
    Real s = 0;
#pragma unroll
    for (int i = 0; i < N; ++i) s += col[i] * col[i];
    cd = sqrt(s);
    cu = cd;
  }
  int L[N];  // position -> lane (= the column's original index)
#pragma unroll
  for (int p = 0; p < N; ++p) L[p] = p;
  bool done = false;  // this lane's column is final (an R column)
  const Real eps = 2.220446049250313e-16;
  const Real norm_downdate_threshold = sqrt(eps);
  int nzp = N;
  Real th = 0;
#pragma unroll
  for (int k = 0; k < N; ++k) {
    int big = k;
    Real bign = lane_value(cu, L[k]);
#pragma unroll
    for (int p = k + 1; p < N; ++p) {
      const Real c = lane_value(cu, L[p]);
      if (c > bign) {
        bign = c;
        big = p;
      }
    }
    big = __builtin_amdgcn_readfirstlane(big);
    const Real big_sq = bign * bign;
    if (k == 0) th = big_sq * eps * eps / N;
    if (nzp == N && big_sq < th * (N - k)) nzp = k;
#pragma unroll
    for (int p = k + 1; p < N; ++p) {
      if (big == p) {
        const int t = L[k];
        L[k] = L[p];
        L[p] = t;
      }
    }
    const int lk = L[k];
    Real v[N];
    Real tailsq = 0;
#pragma unroll
    for (int i = k + 1; i < N; ++i) {
      v[i] = lane_value(col[i], lk);
      tailsq += v[i] * v[i];
    }
    const Real c0 = lane_value(col[k], lk);
    Real tau, beta;
    if (tailsq <= 2.2250738585072014e-308) {
      tau = 0;
      beta = c0;
#pragma unroll
      for (int i = k + 1; i < N; ++i) v[i] = 0;
    } else {
      beta = sqrt(c0 * c0 + tailsq);
      if (c0 >= 0) beta = -beta;
      const Real den = c0 - beta;
#pragma unroll
      for (int i = k + 1; i < N; ++i) v[i] = v[i] / den;
      tau = (beta - c0) / beta;
    }
    const bool mine = lane == lk;
    const bool refl = lane < N && !done && !mine;
    if (mine) {
      col[k] = beta;
#pragma unroll
      for (int i = k + 1; i < N; ++i) col[i] = v[i];
    }
    if (refl) {
      Real s = 0;
#pragma unroll
      for (int i = k + 1; i < N; ++i) s += v[i] * col[i];
      s += col[k];
      col[k] -= tau * s;
#pragma unroll
      for (int i = k + 1; i < N; ++i) col[i] -= (tau * v[i]) * s;
    }
    {
      Real s = 0;
#pragma unroll
      for (int i = k + 1; i < N; ++i) s += v[i] * b[i];
      s += b[k];
      b[k] -= tau * s;
#pragma unroll
      for (int i = k + 1; i < N; ++i) b[i] -= (tau * v[i]) * s;
    }
    if (refl && cu != 0) {
      Real t = fabs(col[k]) / cu;
      t = (1 + t) * (1 - t);
      if (t < 0) t = 0;
      const Real r = cu / cd;
      const Real t2 = t * r * r;
      if (t2 <= norm_downdate_threshold) {
        Real s = 0;
#pragma unroll
        for (int i = k + 1; i < N; ++i) s += col[i] * col[i];
        cd = sqrt(s);
        cu = cd;
      } else {
        cu *= sqrt(t);
      }
    }
    done = done || mine;
  }
  Real z[N];
#pragma unroll
  for (int i = 0; i < N; ++i) z[i] = b[i];
#pragma unroll
  for (int i = N - 1; i >= 0; --i) {
    if (i < nzp) {
      z[i] = z[i] / lane_value(col[i], L[i]);
#pragma unroll
      for (int r = 0; r < i; ++r) z[r] -= z[i] * lane_value(col[r], L[i]);
    }
  }
#pragma unroll
  for (int c = 0; c < N; ++c) b[c] = 0;
#pragma unroll
  for (int i = 0; i < N; ++i) {
    const Real zi = (i < nzp) ? z[i] : 0.0;
#pragma unroll
    for (int c = 0; c < N; ++c)
      if (L[i] == c) b[c] = zi;
  }
  bool ok = true;
#pragma unroll
  for (int i = 0; i < N; ++i)
    if (!isfinite_ref(b[i])) ok = false;
  return ok;
}

// The closure kernels stage the material registry into LDS (dynamic shared
// memory) once per block; cells name their materials by registry index, so
// every EOS parameter read is an LDS access.
// Host compilations of this header (debugging tools that step the closure
// code on the CPU) install a host-pointer registry here; the library runs the
// closure on the device only and never sets it.
inline const EosDev *&host_registry() {
  static const EosDev *r = nullptr;
  return r;
}
SGM_HD const EosDev &registry_entry(int i) {
#if defined(__HIP_DEVICE_COMPILE__)
  extern __shared__ EosDev sgm_lds_registry[];
  return sgm_lds_registry[i];
#else
  static const EosDev none{};
  const EosDev *r = host_registry();
  return r ? r[i] : none;
#endif
}

// GetSieCvFromTAndPreferred (mixed_cell_models.hpp:599-620) for one material
template <bool MODS>
SGM_HD void sie_cv_from_T_preferred(const EosDev &e, Lam &l, Real r, Real Pin, Real T, Real &s,
                                    Real &cv) {
  using Ev = sgm::Ev<MODS>;
  if (Ev::preferred_input(e) & tq::pressure) {
    // IdealGas (modified or not) is the only model here whose preferred
    // input includes pressure: MinimumPressure() = 0,
    // MaximumPressureAtTemperature() = 1e100 (eos_base.hpp:544-549), its
    // closed-form P-T inversion (eos_ideal.hpp:201-207) under the
    // modifiers' DensityEnergyFromPressureTemperature (shifted_eos.hpp:399-403,
    // scaled_eos.hpp:328-335), inlined: no call takes this cell's state
    // by reference
    const Real Pmin = 0.0, Pmax = 1e100;
    const Real Pguess = bad_value(Pin) ? 0.5 * (Pmin + Pmax) : smax(Pmin, smin(Pmax, Pin));
    ideal::rhoe_PT(e.m.p, Pguess, T, r, s);
    if (Ev::modified(e)) {
      if (e.mods & MOD_SHIFTED) s = s + e.shift;
      if (e.mods & MOD_SCALED) {
        r = r * e.inv_scale;
        s = s * e.scale;
      }
    }
    cv = Ev::cv_re(e, r, s, l);
  } else {
    s = Ev::e_rT(e, r, T, l);
    cv = Ev::cv_rT(e, r, T, l);
  }
}

// ---------------------------------------------------------------------------
// Shared base state (PTESolverBase)
// ---------------------------------------------------------------------------
template <int NM, int NEQ, bool MODS = true>
struct PteBase {
  // the EOS calls: with or without the materials' modifiers
  using Ev = sgm::Ev<MODS>;
  unsigned char eidx[NM];  // registry index of each material
  SGM_HD const EosDev &E(int m) const { return registry_entry(eidx[m]); }
  Lam lam[NM];
  Real rho[NM], vfrac[NM], sie[NM], press[NM], u[NM], rhobar[NM];
  Real jac[NEQ * NEQ], dx[NEQ], res[NEQ];
  Real vfrac_total, sie_total, rho_total, uscale, utotal_scale, Tnorm;
  int err;
  MixParamsDev P;

  // :388-399
  SGM_HD void InitRhoBarandRho() {
    NeumaierSum s;
#pragma unroll
    for (int m = 0; m < NM; ++m) {
      rhobar[m] = rho[m] * vfrac[m];
      s.add(rhobar[m]);
    }
    rho_total = s.result();
  }
  SGM_HD void NormalizeVfrac() {
    NeumaierSum s;
#pragma unroll
    for (int m = 0; m < NM; ++m) s.add(vfrac[m]);
    const Real f = robust::ratio(vfrac_total, s.result());
#pragma unroll
    for (int m = 0; m < NM; ++m) vfrac[m] *= f;
  }
  // :423-498; returns true on FAIL
  SGM_HD bool SetVfracFromT(Real T) {
    constexpr Real min_dv = 1e-12;
    constexpr Real safety = 0.9;
    Real vmax[NM];
    Real vmax_sum = 0;
#pragma unroll
    for (int m = 0; m < NM; ++m) {
      vmax[m] = smin(safety * robust::ratio(rhobar[m], Ev::rho_pmin(E(m), T)), 1.0);
      if (vfrac[m] <= 0) vfrac[m] = vmax[m];
      vmax_sum += vmax[m];
    }
    if (vmax_sum < vfrac_total) return true;
    NormalizeVfrac();
    for (int it = 0; it < NM; ++it) {
      Real remaining = 0;
      int n_uncapped = NM;
#pragma unroll
      for (int m = 0; m < NM; ++m) {
        const Real vnew = smin(vmax[m], vfrac[m]);
        const Real dv = vfrac[m] - vnew;
        vfrac[m] -= dv;
        remaining += dv;
        n_uncapped -= (dv > min_dv);
      }
      if (fabs(remaining) < min_dv) break;
#pragma unroll
      for (int m = 0; m < NM; ++m)
        // ratio(Real, size_t) in the reference adds SMALL<size_t>() == 0: a plain division
        if ((vmax[m] - vfrac[m]) > min_dv) vfrac[m] += remaining / (Real)n_uncapped;
    }
    NormalizeVfrac();
    return false;
  }
  // :583-597
  SGM_HD Real PressureFromPreferred(int m, Real r, Real T, Real s) {
    const EosDev &e = E(m);
    if (Ev::preferred_input(e) & tq::specific_internal_energy) return Ev::P_re(e, r, s, lam[m]);
    return Ev::P_rT(e, r, T, lam[m]);
  }
  // the same, with P(rho, T) already evaluated beside e (eos_e_rT_with_P)
  SGM_HD Real PressureFromPreferredP(int m, Real r, Real s, Real P_T) {
    const EosDev &e = E(m);
    if (Ev::preferred_input(e) & tq::specific_internal_energy) return Ev::P_re(e, r, s, lam[m]);
    return P_T;
  }
  // :599-620
  SGM_HD void SieCvFromTAndPreferred(int m, Real r, Real Pin, Real T, Real &s, Real &cv) {
    sie_cv_from_T_preferred<MODS>(E(m), lam[m], r, Pin, T, s, cv);
  }
  // :500-581; Tin are the caller's input temperatures
  SGM_HD Real GetTguess(const Real (&Tin)[NM]) {
    Real Tguess;
    if (Tnorm > 0.0) {
      Tguess = Tnorm;
    } else {
      Tguess = P.default_tguess;
#pragma unroll
      for (int m = 0; m < NM; ++m) Tguess = smax(Tguess, Tin[m]);
    }
    bool bad = false;
    for (int i = 0; i < 3; i++) {
      if (P.iterate_t_guess) {
        Real dudt = 0, usum = 0;
#pragma unroll
        for (int m = 0; m < NM; ++m) {
          const Real rho_max = Ev::max_density(E(m));
          Real s, cv;
          SieCvFromTAndPreferred(m, smin(rho[m], rho_max), press[m], Tguess, s, cv);
          usum += rhobar[m] * s;
          dudt += rhobar[m] * cv;
        }
        const Real Tn = Tguess - robust::ratio((usum - utotal_scale * uscale), dudt);
        Tguess = smin(P.temperature_limit, smax(Tguess, Tn));
      }
#pragma unroll
      for (int m = 0; m < NM; ++m) Tguess = smax(Ev::min_temperature(E(m)), Tguess);
      bad = SetVfracFromT(Tguess);
      if (bad) {
        Tguess *= 10.0;
        continue;
      }
#pragma unroll
      for (int m = 0; m < NM; ++m) rho[m] = robust::ratio(rhobar[m], vfrac[m]);
      bad = false;
#pragma unroll
      for (int m = 0; m < NM; ++m) {
        if (!bad && rho[m] < Ev::rho_pmin(E(m), Tguess)) {
          bad = true;
          Tguess *= 10.0;
        }
      }
      if (!bad) break;
    }
    return Tguess;
  }
  // :627-662
  SGM_HD void InitBase(const Real (&Tin)[NM]) {
    InitRhoBarandRho();
    const Real utotal = rho_total * sie_total;
    uscale = smax(fabs(utotal), 1.0e-14);
    utotal_scale = robust::ratio(utotal, uscale);
    const Real Tguess = GetTguess(Tin);
    Tnorm = Tguess;
#pragma unroll
    for (int m = 0; m < NM; m++) {
      Real P_T;
      sie[m] = Ev::e_rT_with_P(E(m), rho[m], Tguess, lam[m], P_T);
      if ((Ev::preferred_input(E(m)) & tq::pressure) && !bad_value(press[m])) {
        press[m] = robust::ratio(press[m], uscale);
      } else {
        press[m] = robust::ratio(PressureFromPreferredP(m, rho[m], sie[m], P_T), uscale);
      }
    }
#pragma unroll
    for (int m = 0; m < NM; ++m) u[m] = sie[m] * robust::ratio(rhobar[m], uscale);
  }
  // Every field of the solver state that lives from one Newton iteration to
  // the next (the Jacobian, the step and the line-search copies are rewritten
  // before they are read; the mixing parameters P come from the kernel
  // arguments), in a fixed order: a cell can be saved between iterations and
  // resumed bit for bit.
  template <class F>
  SGM_HD constexpr void VisitBase(F &f) {
#pragma unroll
    for (int m = 0; m < NM; ++m) {
      f(eidx[m]);
      f(lam[m].present);
      f(lam[m].lRho);
      f(lam[m].lT);
      f(rho[m]);
      f(vfrac[m]);
      f(sie[m]);
      f(press[m]);
      f(u[m]);
      f(rhobar[m]);
    }
#pragma unroll
    for (int i = 0; i < NEQ; ++i) f(res[i]);
    f(vfrac_total);
    f(sie_total);
    f(rho_total);
    f(uscale);
    f(utotal_scale);
    f(Tnorm);
    f(err);
  }
  SGM_HD Real ResidualNorm() const {
    NeumaierSum s;
#pragma unroll
    for (int i = 0; i < NEQ; ++i) s.add(res[i] * res[i]);
    return 0.5 * s.result();
  }
  SGM_HD bool Solve() {
#pragma unroll
    for (int i = 0; i < NEQ; ++i) dx[i] = res[i];
    return solve_small<NEQ>(jac, dx);
  }
  // Solve() by a whole wave holding identical copies of the cell
  __device__ __forceinline__ bool SolveWave(int lane) {
#pragma unroll
    for (int i = 0; i < NEQ; ++i) dx[i] = res[i];
    if constexpr (NEQ < 3) {
      return solve_small<NEQ>(jac, dx);
    } else {
      return solve_small_wave<NEQ>(jac, dx, lane);
    }
  }
};

// ---------------------------------------------------------------------------
// PTESolverRhoT (:873-1193). Unknowns (T, alpha_m for m != ms).
// ---------------------------------------------------------------------------
template <int NM, bool MODS = true>
struct PteRhoT : PteBase<NM, NM, MODS> {
  using Ev = sgm::Ev<MODS>;
  using B = PteBase<NM, NM, MODS>;
  using B::dx;
  using B::E;
  using B::jac;
  using B::lam;
  using B::P;
  using B::press;
  using B::res;
  using B::rho;
  using B::rhobar;
  using B::sie;
  using B::Tnorm;
  using B::u;
  using B::uscale;
  using B::utotal_scale;
  using B::vfrac;
  static constexpr bool kStoresTemp = true;
  Real vtemp[NM];
  Real Tequil, Ttemp;
  int ms;

  SGM_HD void SetCellInputs(Real vfrac_tot, Real sie_tot, Real Tnorm_in) {
    B::vfrac_total = vfrac_tot;
    B::sie_total = sie_tot;
    Tnorm = Tnorm_in;
  }
  // material index of reduced-system row/column k (1..NM-1)
  SGM_HD int mat_of(int k) const { return (k <= ms) ? k - 1 : k; }

  SGM_HD void Residual() {
    NeumaierSum s;
#pragma unroll
    for (int m = 0; m < NM; ++m) s.add(u[m]);
    res[0] = utotal_scale - s.result();
    const Real pms = pick<NM>(press, ms);
#pragma unroll
    for (int k = 1; k < NM; ++k) {
      const Real pm = sel(k <= ms, press[k - 1], press[k]);
      res[k] = pms - pm;
    }
  }
  // :947-984
  SGM_HD Real Init(const Real (&Tin)[NM]) {
    B::InitBase(Tin);
    if (P.pte_reduced_system_exclude_idx >= 0) {
      ms = P.pte_reduced_system_exclude_idx;
    } else {
      ms = 0;
      Real xmax = 0;
#pragma unroll
      for (int m = 0; m < NM; ++m) {
        const Real x = robust::ratio(rhobar[m], B::rho_total);
        if (x > xmax) {
          xmax = x;
          ms = m;
        }
      }
    }
    Residual();
    Tequil = 1.0;
    return B::ResidualNorm();
  }
  // :986-1008
  SGM_HD void Fixup() {
    if (P.pte_reduced_system_exclude_idx < 0) {
      Real amax = 0;
      ms = 0;
#pragma unroll
      for (int m = 0; m < NM; ++m)
        if (vfrac[m] > amax) {
          amax = vfrac[m];
          ms = m;
        }
      Residual();
    }
  }
  // :1021-1040 -> (converged, acceptable)
  SGM_HD void CheckPTE(bool &conv, bool &ok) const {
    Real mean_p = 0;
#pragma unroll
    for (int m = 0; m < NM; ++m) mean_p += vfrac[m] * press[m];
    NeumaierSum s;
#pragma unroll
    for (int k = 1; k < NM; ++k) s.add(res[k] * res[k]);
    const Real error_p = sqrt(s.result());
    const Real error_u = fabs(res[0]);
    const Real amp = fabs(mean_p);
    const bool cp = (error_p < P.pte_rel_tolerance_p * amp || error_p < P.pte_abs_tolerance_p);
    const bool cu = (error_u < P.pte_rel_tolerance_e);
    const bool ap = (error_p < P.pte_rel_tolerance_p_sufficient * amp ||
                     error_p < P.pte_abs_tolerance_p_sufficient);
    const bool au = (error_u < P.pte_rel_tolerance_e_sufficient);
    conv = cp && cu;
    ok = ap && au;
  }
  // One material's finite-difference derivatives of the Jacobian (the body
  // of the reference's material loop, :1042-1080): two fused e/P
  // evaluations, which also update that material's lambda cache.
  SGM_HD void MaterialDerivs(const EosDev &e, Lam &l, Real vf, Real rb, Real rh, Real pr,
                             Real um, Real Tcur, Real Tpert, Real dT, Real &dpdv, Real &dedv,
                             Real &dpdT, Real &dedT) const {
    const Real dv = (vf < 0.5 ? 1.0 : -1.0) * vf * P.derivative_eps;
    const Real vf_pert = vf + dv;
    const Real rho_pert = robust::ratio(rb, vf_pert);
    Real P_T;
    Real e_pert = Ev::e_rT_with_P(e, rho_pert, Tcur, l, P_T);
    Real p_pert = robust::ratio(PressureFromPreferredE(e, l, rho_pert, e_pert, P_T), uscale);
    dpdv = smin(robust::ratio((p_pert - pr), dv), -fabs(P.pte_max_dpdv));
    dedv = robust::ratio(rb * robust::ratio(e_pert, uscale) - um, dv);
    e_pert = Ev::e_rT_with_P(e, rh, Tpert, l, P_T);
    p_pert = robust::ratio(PressureFromPreferredE(e, l, rh, e_pert, P_T), uscale);
    dpdT = robust::ratio((p_pert - pr), dT);
    dedT = robust::ratio(rb * robust::ratio(e_pert, uscale) - um, dT);
  }
  SGM_HD static Real PressureFromPreferredE(const EosDev &e, Lam &l, Real r, Real s, Real P_T) {
    if (Ev::preferred_input(e) & tq::specific_internal_energy) return Ev::P_re(e, r, s, l);
    return P_T;
  }
  // :1042-1105
  SGM_HD void Jacobian() {
    Real dpdv[NM], dedv[NM], dpdT[NM];
    Real dedT_sum = 0.0;
    const Real Tcur = Tnorm * Tequil;
    const Real dT = Tequil * P.derivative_eps;
    const Real Tpert = Tnorm * (Tequil + dT);
#pragma unroll
    for (int m = 0; m < NM; m++) {
      Real dedT;
      MaterialDerivs(E(m), lam[m], vfrac[m], rhobar[m], rho[m], press[m], u[m], Tcur, Tpert, dT,
                     dpdv[m], dedv[m], dpdT[m], dedT);
      dedT_sum += dedT;
    }
    AssembleJacobian(dpdv, dedv, dpdT, dedT_sum);
  }
  // Jacobian() by a group of G >= NM lanes holding identical copies of the
  // cell (lanes base .. base+G-1, this lane g): lane g evaluates material g's
  // derivatives, then every lane gathers them, the lambda caches included, in
  // material order. The values, the order of the dedT sum and the caches are
  // those of Jacobian().
  template <int G>
  __device__ void JacobianGroup(int g, int base) {
    static_assert(G >= NM, "one lane per material");
    const Real Tcur = Tnorm * Tequil;
    const Real dT = Tequil * P.derivative_eps;
    const Real Tpert = Tnorm * (Tequil + dT);
    const int m = g < NM ? g : 0;
    int ei = B::eidx[0];
    Real lr = lam[0].lRho, lt = lam[0].lT;
#pragma unroll
    for (int k = 1; k < NM; ++k) {
      ei = (m == k) ? (int)B::eidx[k] : ei;
      lr = sel(m == k, lam[k].lRho, lr);
      lt = sel(m == k, lam[k].lT, lt);
    }
    Lam l;
    l.present = lam[0].present;
    l.lRho = lr;
    l.lT = lt;
    Real my_dpdv, my_dedv, my_dpdT, my_dedT;
    MaterialDerivs(registry_entry(ei), l, pick<NM>(vfrac, m), pick<NM>(rhobar, m),
                   pick<NM>(rho, m), pick<NM>(press, m), pick<NM>(u, m), Tcur, Tpert, dT,
                   my_dpdv, my_dedv, my_dpdT, my_dedT);
    Real dpdv[NM], dedv[NM], dpdT[NM];
    Real dedT_sum = 0.0;
#pragma unroll
    for (int k = 0; k < NM; ++k) {
      dpdv[k] = __shfl(my_dpdv, base + k, 64);
      dedv[k] = __shfl(my_dedv, base + k, 64);
      dpdT[k] = __shfl(my_dpdT, base + k, 64);
      dedT_sum += __shfl(my_dedT, base + k, 64);
      lam[k].lRho = __shfl(l.lRho, base + k, 64);
      lam[k].lT = __shfl(l.lT, base + k, 64);
    }
    AssembleJacobian(dpdv, dedv, dpdT, dedT_sum);
  }
  // Jacobian() by a whole wave holding identical copies of the cell
  // (pte_iterate_wide): lane 8 t + m evaluates material m's volume-fraction
  // derivatives (t even) or temperature derivatives (t odd), the two fused
  // e/P evaluations of MaterialDerivs, written as one code path so that the
  // wave runs each material kind's evaluation once. Every lane then gathers
  // the derivatives in material order and each lambda cache from the
  // temperature lane, whose evaluation MaterialDerivs makes last (the
  // evaluations write the caches and never read them). The values, the order
  // of the dedT sum and the caches are those of Jacobian().
  __device__ __forceinline__ void JacobianWide(int lane) {
    const Real Tcur = Tnorm * Tequil;
    const Real dT = Tequil * P.derivative_eps;
    const Real Tpert = Tnorm * (Tequil + dT);
    const int m = (lane & 7) < NM ? (lane & 7) : NM - 1;
    const bool tpart = (lane & 8) != 0;
    int ei = B::eidx[0];
    Lam l = lam[0];
#pragma unroll
    for (int k = 1; k < NM; ++k) {
      ei = (m == k) ? (int)B::eidx[k] : ei;
      l.lRho = sel(m == k, lam[k].lRho, l.lRho);
      l.lT = sel(m == k, lam[k].lT, l.lT);
    }
    const EosDev &e = registry_entry(ei);
    const Real vf = pick<NM>(vfrac, m), rb = pick<NM>(rhobar, m), rh = pick<NM>(rho, m),
               pr = pick<NM>(press, m), um = pick<NM>(u, m);
    const Real dv = (vf < 0.5 ? 1.0 : -1.0) * vf * P.derivative_eps;
    const Real rho_pert = robust::ratio(rb, vf + dv);
    const Real rx = tpart ? rh : rho_pert;
    const Real Tx = tpart ? Tpert : Tcur;
    const Real delta = tpart ? dT : dv;
    Real P_T;
    const Real e_pert = Ev::e_rT_with_P(e, rx, Tx, l, P_T);
    const Real p_pert = robust::ratio(PressureFromPreferredE(e, l, rx, e_pert, P_T), uscale);
    Real dp = robust::ratio((p_pert - pr), delta);
    if (!tpart) dp = smin(dp, -fabs(P.pte_max_dpdv));
    const Real de = robust::ratio(rb * robust::ratio(e_pert, uscale) - um, delta);
    Real dpdv[NM], dedv[NM], dpdT[NM];
    Real dedT_sum = 0.0;
#pragma unroll
    for (int k = 0; k < NM; ++k) {
      dpdv[k] = __shfl(dp, k, 64);
      dedv[k] = __shfl(de, k, 64);
      dpdT[k] = __shfl(dp, 8 + k, 64);
      dedT_sum += __shfl(de, 8 + k, 64);
      lam[k].lRho = __shfl(l.lRho, 8 + k, 64);
      lam[k].lT = __shfl(l.lT, 8 + k, 64);
    }
    AssembleJacobian(dpdv, dedv, dpdT, dedT_sum);
  }
  SGM_HD void AssembleJacobian(const Real (&dpdv)[NM], const Real (&dedv)[NM],
                               const Real (&dpdT)[NM], Real dedT_sum) {
    const Real dedv_s = pick<NM>(dedv, ms), dpdT_s = pick<NM>(dpdT, ms),
               dpdv_s = pick<NM>(dpdv, ms);
    jac[0] = dedT_sum;
#pragma unroll
    for (int k = 1; k < NM; ++k) {
      const Real dedv_k = sel(k <= ms, dedv[k - 1], dedv[k]);
      jac[k] = dedv_k - dedv_s;
    }
#pragma unroll
    for (int r = 1; r < NM; ++r) {
      const Real dpdT_r = sel(r <= ms, dpdT[r - 1], dpdT[r]);
      const Real dpdv_r = sel(r <= ms, dpdv[r - 1], dpdv[r]);
      jac[r * NM + 0] = dpdT_r - dpdT_s;
#pragma unroll
      for (int c = 1; c < NM; ++c) jac[r * NM + c] = (c == r) ? dpdv_r + dpdv_s : dpdv_s;
    }
  }
  // :1107-1151 (including the reference's idx bookkeeping, which does not
  // advance for materials skipped in the second pass)
  SGM_HD Real ScaleDx() {
    Real scale = 1.0;
    NeumaierSum s;
#pragma unroll
    for (int i = 0; i < NM - 1; ++i) s.add(dx[1 + i]);
    const Real dalphaskip = -s.result();
#pragma unroll
    for (int m = 0; m < NM; ++m) {
      const Real mydx = sel(m == ms, dalphaskip, sel(m < ms, dx[m + 1], dx[m]));
      if (scale * mydx < -P.vfrac_safety_fac * vfrac[m])
        scale = smin(scale, fabs(P.vfrac_safety_fac * robust::ratio(vfrac[m], mydx)));
    }
    const Real Tnew = Tequil + scale * dx[0];
    int idx = 1;
#pragma unroll
    for (int m = 0; m < NM; m++) {
      const Real rho_min =
          smax(Ev::rho_pmin(E(m), Tnorm * Tequil), Ev::rho_pmin(E(m), Tnorm * Tnew));
      const Real alpha_max = smin(1.0, robust::ratio(rhobar[m], rho_min));
      if (!(alpha_max < vfrac[m])) {
        Real mydx;
        if (m == ms) {
          mydx = dalphaskip;
        } else {
          mydx = pick<NM>(dx, idx);
          idx++;
        }
        if (scale * mydx > 0.5 * (alpha_max - vfrac[m]))
          scale = smin(robust::ratio(0.5 * (alpha_max - vfrac[m]), mydx), scale);
      }
    }
    if (scale * dx[0] < -0.95 * Tequil) scale = smin(scale, robust::ratio(-0.95 * Tequil, dx[0]));
#pragma unroll
    for (int i = 0; i < NM; ++i) dx[i] *= scale;
    return scale;
  }
  // :1155-1187
  SGM_HD Real TestUpdate(Real scale, bool cache_state) {
    if (cache_state) {
      Ttemp = Tequil;
#pragma unroll
      for (int m = 0; m < NM; ++m) vtemp[m] = vfrac[m];
    }
    Tequil = Ttemp + scale * dx[0];
#pragma unroll
    for (int m = 0; m < NM; ++m) {
      const Real dxm = sel(m < ms, dx[m + 1], dx[m]);
      if (m != ms) vfrac[m] = vtemp[m] + scale * dxm;
    }
    NeumaierSum s;
#pragma unroll
    for (int m = 0; m < NM; ++m)
      if (m != ms) s.add(vfrac[m]);
    put<NM>(vfrac, ms, 1 - s.result());
    const Real Tc = Tnorm * Tequil;
#pragma unroll
    for (int m = 0; m < NM; ++m) {
      rho[m] = robust::ratio(rhobar[m], vfrac[m]);
      Real P_T;
      Real um = rhobar[m] * Ev::e_rT_with_P(E(m), rho[m], Tc, lam[m], P_T);
      sie[m] = robust::ratio(um, rhobar[m]);
      u[m] = robust::ratio(um, uscale);
      press[m] = robust::ratio(B::PressureFromPreferredP(m, rho[m], sie[m], P_T), uscale);
    }
    Residual();
    return B::ResidualNorm();
  }
  // TestUpdate by a whole wave (pte_iterate_wide): lane 8 t + m takes the
  // trial's new temperature and volume fractions like every lane, then
  // evaluates material m only; the eight lanes of each t then gather the
  // materials' densities, energies, pressures and lambda caches in material
  // order and form the residual. Each group of eight lanes can run a trial
  // of its own scale. The values, the order of every sum and the caches are
  // those of TestUpdate.
  __device__ __forceinline__ Real TestUpdateWide(Real scale, bool cache_state, int lane) {
    if (cache_state) {
      Ttemp = Tequil;
#pragma unroll
      for (int m = 0; m < NM; ++m) vtemp[m] = vfrac[m];
    }
    Tequil = Ttemp + scale * dx[0];
#pragma unroll
    for (int m = 0; m < NM; ++m) {
      const Real dxm = sel(m < ms, dx[m + 1], dx[m]);
      if (m != ms) vfrac[m] = vtemp[m] + scale * dxm;
    }
    NeumaierSum s;
#pragma unroll
    for (int m = 0; m < NM; ++m)
      if (m != ms) s.add(vfrac[m]);
    put<NM>(vfrac, ms, 1 - s.result());
    const Real Tc = Tnorm * Tequil;
    const int m = (lane & 7) < NM ? (lane & 7) : NM - 1;
    int ei = B::eidx[0];
    Lam l = lam[0];
#pragma unroll
    for (int k = 1; k < NM; ++k) {
      ei = (m == k) ? (int)B::eidx[k] : ei;
      l.lRho = sel(m == k, lam[k].lRho, l.lRho);
      l.lT = sel(m == k, lam[k].lT, l.lT);
    }
    const EosDev &e = registry_entry(ei);
    const Real rb = pick<NM>(rhobar, m);
    const Real rh = robust::ratio(rb, pick<NM>(vfrac, m));
    Real P_T;
    const Real um = rb * Ev::e_rT_with_P(e, rh, Tc, l, P_T);
    const Real se = robust::ratio(um, rb);
    const Real us = robust::ratio(um, uscale);
    const Real pr = robust::ratio(PressureFromPreferredE(e, l, rh, se, P_T), uscale);
    const int b = lane & ~7;
#pragma unroll
    for (int k = 0; k < NM; ++k) {
      rho[k] = __shfl(rh, b + k, 64);
      sie[k] = __shfl(se, b + k, 64);
      u[k] = __shfl(us, b + k, 64);
      press[k] = __shfl(pr, b + k, 64);
      lam[k].lRho = __shfl(l.lRho, b + k, 64);
      lam[k].lT = __shfl(l.lT, b + k, 64);
    }
    Residual();
    return B::ResidualNorm();
  }
  SGM_HD Real Tout() const { return Tequil * Tnorm; }
  template <class F>
  SGM_HD constexpr void Visit(F &f) {
    B::VisitBase(f);
    f(Tequil);
    f(ms);
  }
  SGM_HD Real Pout(int m) const { return press[m] * uscale; }
};

// ---------------------------------------------------------------------------
// PTESolverPT (:1198-1446). Unknowns (T, P); 2x2 system.
// ---------------------------------------------------------------------------
template <int NM, bool MODS = true>
struct PtePT : PteBase<NM, 2, MODS> {
  using Ev = sgm::Ev<MODS>;
  using B = PteBase<NM, 2, MODS>;
  using B::dx;
  using B::E;
  using B::err;
  using B::jac;
  using B::lam;
  using B::P;
  using B::press;
  using B::res;
  using B::rho;
  using B::rhobar;
  using B::sie;
  using B::Tnorm;
  using B::u;
  using B::uscale;
  using B::utotal_scale;
  using B::vfrac;
  using B::vfrac_total;
  static constexpr bool kStoresTemp = true;
  Real Pequil, Tequil, Ptemp, Ttemp;

  SGM_HD void SetCellInputs(Real vfrac_tot, Real sie_tot, Real Tnorm_in) {
    vfrac_total = vfrac_tot;
    B::sie_total = sie_tot;
    Tnorm = Tnorm_in;
  }
  SGM_HD void Residual() {
    NeumaierSum sv, se;
#pragma unroll
    for (int m = 0; m < NM; ++m) sv.add(vfrac[m]);
#pragma unroll
    for (int m = 0; m < NM; ++m) se.add(u[m]);
    res[0] = vfrac_total - sv.result();
    res[1] = utotal_scale - se.result();
  }
  // :1258-1297
  SGM_HD Real Init(const Real (&Tin)[NM]) {
    B::InitBase(Tin);
    Residual();
    Pequil = 0;
#pragma unroll
    for (int m = 0; m < NM; ++m) Pequil += fabs(press[m]) * vfrac[m];
    NeumaierSum s;
#pragma unroll
    for (int m = 0; m < NM; ++m) s.add(vfrac[m]);
    Pequil /= s.result();
    Tequil = 1;
#pragma unroll
    for (int m = 0; m < NM; ++m) {
      Ev::rhoe_PT(E(m), Pequil * uscale, Tequil * Tnorm, rho[m], sie[m], lam[m], err);
      vfrac[m] = robust::ratio(rhobar[m], rho[m]);
      u[m] = robust::ratio(sie[m] * rhobar[m], uscale);
    }
    Residual();
    return B::ResidualNorm();
  }
  SGM_HD void Fixup() {}
  SGM_HD void CheckPTE(bool &conv, bool &ok) const {
    const Real ev = fabs(res[0]), eu = fabs(res[1]);
    conv = (ev < P.pte_rel_tolerance_v) && (eu < P.pte_rel_tolerance_e);
    ok = (eu < P.pte_rel_tolerance_e_sufficient) && (ev < P.pte_rel_tolerance_v_sufficient);
  }
  // :1336-1385
  SGM_HD void Jacobian() {
    Real dudT = 0.0, dudP = 0.0, rbT = 0.0, rbP = 0.0;
    const Real dp = -Pequil * P.derivative_eps;
    const Real dT = Tequil * P.derivative_eps;
#pragma unroll
    for (int m = 0; m < NM; ++m) {
      Real r_pert = rho[m];
      Real e_pert = sie[m];
      Ev::rhoe_PT(E(m), uscale * (Pequil + dp), Tnorm * Tequil, r_pert, e_pert, lam[m], err);
      const Real drdp = robust::ratio(r_pert - rho[m], dp);
      const Real dudp = robust::ratio(robust::ratio(rhobar[m] * e_pert, uscale) - u[m], dp);
      rbP += robust::ratio(rhobar[m], rho[m] * rho[m]) * drdp;
      dudP += dudp;
      Ev::rhoe_PT(E(m), uscale * Pequil, Tnorm * (Tequil + dT), r_pert, e_pert, lam[m], err);
      const Real drdT = robust::ratio(r_pert - rho[m], dT);
      const Real dudTm = robust::ratio(robust::ratio(rhobar[m] * e_pert, uscale) - u[m], dT);
      rbT += robust::ratio(rhobar[m], rho[m] * rho[m]) * drdT;
      dudT += dudTm;
    }
    jac[0] = -rbT;
    jac[1] = -rbP;
    jac[2] = dudT;
    jac[3] = dudP;
  }
  // :1387-1409
  SGM_HD Real ScaleDx() {
    Real scale = 1.0;
    if (scale * dx[0] < -0.95 * Tequil) scale = robust::ratio(-0.95 * Tequil, dx[0]);
#pragma unroll
    for (int m = 0; m < NM; ++m) {
      const Real Pmin = Ev::min_pressure(E(m));
      scale = smin(fabs(scale),
                   fabs(0.95 * robust::ratio(robust::ratio(Pmin, uscale) - Pequil, dx[1])));
    }
#pragma unroll
    for (int m = 0; m < NM; ++m) {
      const Real Ttest = (Tequil + scale * dx[0]) * Tnorm;
      const Real Pmax = Ev::max_pressure_at_T(E(m), Ttest, err);
      scale = smin(fabs(scale),
                   fabs(0.95 * robust::ratio(robust::ratio(Pmax, uscale) - Pequil, dx[0])));
    }
    dx[0] *= scale;
    dx[1] *= scale;
    return scale;
  }
  // :1413-1435
  SGM_HD Real TestUpdate(Real scale, bool cache_state) {
    if (cache_state) {
      Ttemp = Tequil;
      Ptemp = Pequil;
    }
    Tequil = Ttemp + scale * dx[0];
    Pequil = Ptemp + scale * dx[1];
#pragma unroll
    for (int m = 0; m < NM; ++m) {
      Ev::rhoe_PT(E(m), Pequil * uscale, Tequil * Tnorm, rho[m], sie[m], lam[m], err);
      vfrac[m] = robust::ratio(rhobar[m], rho[m]);
      u[m] = robust::ratio(sie[m] * rhobar[m], uscale);
    }
    Residual();
    return B::ResidualNorm();
  }
  SGM_HD Real Tout() const { return Tequil * Tnorm; }
  template <class F>
  SGM_HD constexpr void Visit(F &f) {
    B::VisitBase(f);
    f(Pequil);
    f(Tequil);
  }
  SGM_HD Real Pout(int) const { return Pequil * uscale; }
};

// ---------------------------------------------------------------------------
// PTESolverFixedT (:1451-1669). Temperature given; unknowns the volume
// fractions; residual = volume closure + neighbouring pressure differences.
// The solver never writes temperatures (its Finalize scales them by Tnorm=1).
// ---------------------------------------------------------------------------
template <int NM, bool MODS = true>
struct PteFixedT : PteBase<NM, NM, MODS> {
  using Ev = sgm::Ev<MODS>;
  using B = PteBase<NM, NM, MODS>;
  using B::dx;
  using B::E;
  using B::jac;
  using B::lam;
  using B::P;
  using B::press;
  using B::res;
  using B::rho;
  using B::rhobar;
  using B::sie;
  using B::Tnorm;
  using B::u;
  using B::uscale;
  using B::vfrac;
  using B::vfrac_total;
  static constexpr bool kStoresTemp = false;
  Real vtemp[NM];
  Real Tequil;

  // constructor (:1496-1509): base sie_tot = 1, Tnorm = 1 after the base call
  SGM_HD void SetCellInputs(Real vfrac_tot, Real T_true, Real) {
    vfrac_total = vfrac_tot;
    B::sie_total = 1.0;
    Tequil = T_true;
    Tnorm = 1.0;
  }
  SGM_HD void Residual() {
    NeumaierSum s;
#pragma unroll
    for (int m = 0; m < NM; ++m) s.add(vfrac[m]);
    res[0] = vfrac_total - s.result();
#pragma unroll
    for (int m = 0; m < NM - 1; ++m) res[1 + m] = press[m] - press[m + 1];
  }
  // :1523-1551 (Tin unused: the temperature is the cell input)
  SGM_HD Real Init(const Real (&)[NM]) {
    Tnorm = 1.0;
    B::InitRhoBarandRho();
    B::SetVfracFromT(Tequil);
    uscale = 0.0;
#pragma unroll
    for (int m = 0; m < NM; m++) {
      rho[m] = robust::ratio(rhobar[m], vfrac[m]);
      Real P_T;
      sie[m] = Ev::e_and_P_rT(E(m), rho[m], Tequil, lam[m], P_T);
      uscale += fabs(sie[m]) * rho[m];
      press[m] = P_T;
    }
    uscale = smax(uscale, 1.0e-14);
#pragma unroll
    for (int m = 0; m < NM; ++m) {
      press[m] = robust::ratio(press[m], uscale);
      u[m] = sie[m] * robust::ratio(rhobar[m], uscale);
    }
    Residual();
    return B::ResidualNorm();
  }
  SGM_HD void Fixup() {}
  // :1562-1581
  SGM_HD void CheckPTE(bool &conv, bool &ok) const {
    Real mean_p = vfrac[0] * press[0];
    Real error_p = 0;
#pragma unroll
    for (int m = 1; m < NM; ++m) {
      mean_p += vfrac[m] * press[m];
      error_p += res[m] * res[m];
    }
    error_p = sqrt(error_p);
    const Real error_v = fabs(res[0]);
    const Real amp = fabs(mean_p);
    conv = (error_p < P.pte_rel_tolerance_p * amp || error_p < P.pte_abs_tolerance_p) &&
           (error_v < P.pte_rel_tolerance_v);
    ok = (error_p < P.pte_rel_tolerance_p_sufficient * amp ||
          error_p < P.pte_abs_tolerance_p_sufficient) &&
         (error_v < P.pte_rel_tolerance_v_sufficient);
  }
  // :1583-1609
  SGM_HD void Jacobian() {
    Real dpdv[NM];
#pragma unroll
    for (int m = 0; m < NM; m++) {
      const Real dv = (vfrac[m] < 0.5 ? 1.0 : -1.0) * vfrac[m] * P.derivative_eps;
      const Real vf_pert = vfrac[m] + dv;
      const Real rho_pert = robust::ratio(rhobar[m], vf_pert);
      const Real p_pert = robust::ratio(Ev::P_rT(E(m), rho_pert, Tequil, lam[m]), uscale);
      dpdv[m] = smin(robust::ratio((p_pert - press[m]), dv), -fabs(P.pte_max_dpdv));
    }
#pragma unroll
    for (int i = 0; i < NM * NM; ++i) jac[i] = 0.0;
#pragma unroll
    for (int m = 0; m < NM; ++m) jac[m] = 1.0;
#pragma unroll
    for (int m = 0; m < NM - 1; m++) {
      jac[(m + 1) * NM + m] = -dpdv[m];
      jac[(m + 1) * NM + m + 1] = dpdv[m + 1];
    }
  }
  // :1611-1641
  SGM_HD Real ScaleDx() {
    Real scale = 1.0;
#pragma unroll
    for (int m = 0; m < NM; ++m)
      if (scale * dx[m] < -P.vfrac_safety_fac * vfrac[m])
        scale = robust::ratio(-P.vfrac_safety_fac * vfrac[m], dx[m]);
#pragma unroll
    for (int m = 0; m < NM; m++) {
      const Real alpha_max = robust::ratio(rhobar[m], Ev::rho_pmin(E(m), Tequil));
      // already unstable (alpha_max < vfrac): no limit, as the reference
      if (!(alpha_max < vfrac[m]) && scale * dx[m] > 0.5 * (alpha_max - vfrac[m]))
        scale = robust::ratio(0.5 * (alpha_max - vfrac[m]), dx[m]);
    }
#pragma unroll
    for (int i = 0; i < NM; ++i) dx[i] *= scale;
    return scale;
  }
  // :1643-1664
  SGM_HD Real TestUpdate(Real scale, bool cache_state) {
    if (cache_state) {
#pragma unroll
      for (int m = 0; m < NM; ++m) vtemp[m] = vfrac[m];
    }
#pragma unroll
    for (int m = 0; m < NM; ++m) {
      vfrac[m] = vtemp[m] + scale * dx[m];
      rho[m] = robust::ratio(rhobar[m], vfrac[m]);
      Real P_T;
      const Real um = rhobar[m] * Ev::e_and_P_rT(E(m), rho[m], Tequil, lam[m], P_T);
      sie[m] = robust::ratio(um, rhobar[m]);
      u[m] = robust::ratio(um, uscale);
      press[m] = robust::ratio(P_T, uscale);
    }
    Residual();
    return B::ResidualNorm();
  }
  SGM_HD Real Tout() const { return Tequil; }
  template <class F>
  SGM_HD constexpr void Visit(F &f) {
    B::VisitBase(f);
    f(Tequil);
  }
  SGM_HD Real Pout(int m) const { return press[m] * uscale; }
};

// ---------------------------------------------------------------------------
// PTESolverFixedP (:1674-1921). Pressure given; unknowns the volume fractions
// and the temperature; residual = each material's pressure against P, then
// the volume closure.
// ---------------------------------------------------------------------------
template <int NM, bool MODS = true>
struct PteFixedP : PteBase<NM, NM + 1, MODS> {
  using Ev = sgm::Ev<MODS>;
  using B = PteBase<NM, NM + 1, MODS>;
  using B::dx;
  using B::E;
  using B::jac;
  using B::lam;
  using B::P;
  using B::press;
  using B::res;
  using B::rho;
  using B::rhobar;
  using B::sie;
  using B::Tnorm;
  using B::u;
  using B::uscale;
  using B::utotal_scale;
  using B::vfrac;
  using B::vfrac_total;
  static constexpr bool kStoresTemp = true;
  static constexpr int NEQ = NM + 1;
  Real vtemp[NM];
  Real Tequil, Ttemp, Pequil;

  // constructor (:1721-1733): base sie_tot = 1, Tnorm = 0. The reference's
  // GetTguess reads uscale and utotal_scale (:522-537) before Init assigns
  // them, so their value there is indeterminate; both are zero here, which
  // makes that Newton step T - (sum rhobar e) / (sum rhobar cv).
  SGM_HD void SetCellInputs(Real vfrac_tot, Real P_in, Real) {
    vfrac_total = vfrac_tot;
    B::sie_total = 1.0;
    Pequil = P_in;
    Tnorm = 0.0;
    uscale = 0.0;
    utotal_scale = 0.0;
  }
  // the PreferredInput pressure with the energy looked up first
  // (GetPressureFromPreferred(..., do_e_lookup = true), :583-597)
  SGM_HD Real PressureLookup(int m, Real r, Real T) {
    const EosDev &e = E(m);
    if (Ev::preferred_input(e) & tq::specific_internal_energy)
      return Ev::P_re(e, r, Ev::e_rT(e, r, T, lam[m]), lam[m]);
    return Ev::P_rT(e, r, T, lam[m]);
  }
  SGM_HD void Residual() {
    NeumaierSum s;
#pragma unroll
    for (int m = 0; m < NM; ++m) s.add(vfrac[m]);
    const Real pt = robust::ratio(Pequil, uscale);
#pragma unroll
    for (int m = 0; m < NM; ++m) res[m] = pt - press[m];
    res[NM] = vfrac_total - s.result();
  }
  // :1747-1779
  SGM_HD Real Init(const Real (&Tin)[NM]) {
    B::InitRhoBarandRho();
    const Real Tguess = B::GetTguess(Tin);
    Tnorm = Tguess;
    uscale = smax(fabs(Pequil), 1.0e-14);
#pragma unroll
    for (int m = 0; m < NM; m++) {
      Real P_T;
      sie[m] = Ev::e_and_P_rT(E(m), rho[m], Tguess, lam[m], P_T);
      u[m] = robust::ratio(sie[m] * rhobar[m], uscale);
      press[m] = robust::ratio(P_T, uscale);
    }
    Residual();
    Tequil = 1.0;  // temp[0], set to 1 above in the reference
    Ttemp = Tequil;
    return B::ResidualNorm();
  }
  SGM_HD void Fixup() {}
  // :1790-1808
  SGM_HD void CheckPTE(bool &conv, bool &ok) const {
    Real error_p = 0;
#pragma unroll
    for (int m = 0; m < NM; ++m) error_p += res[m] * res[m];
    error_p = sqrt(error_p);
    error_p *= uscale;
    const Real error_v = fabs(res[NM]);
    conv = (error_p < P.pte_rel_tolerance_p || error_p < P.pte_abs_tolerance_p) &&
           (error_v < P.pte_rel_tolerance_v);
    ok = (error_p < P.pte_rel_tolerance_p_sufficient ||
          error_p < P.pte_abs_tolerance_p_sufficient) &&
         (error_v < P.pte_rel_tolerance_v_sufficient);
  }
  // :1809-1850
  SGM_HD void Jacobian() {
    Real dpdv[NM], dpdT[NM];
    const Real dT = Tequil * P.derivative_eps;
#pragma unroll
    for (int m = 0; m < NM; m++) {
      const Real dv = (vfrac[m] < 0.5 ? 1.0 : -1.0) * vfrac[m] * P.derivative_eps;
      const Real vf_pert = vfrac[m] + dv;
      const Real rho_pert = robust::ratio(rhobar[m], vf_pert);
      Real p_pert = robust::ratio(PressureLookup(m, rho_pert, Tnorm * Tequil), uscale);
      dpdv[m] = smin(robust::ratio((p_pert - press[m]), dv), -fabs(P.pte_max_dpdv));
      p_pert = robust::ratio(PressureLookup(m, rho[m], Tnorm * (Tequil + dT)), uscale);
      dpdT[m] = robust::ratio((p_pert - press[m]), dT);
    }
#pragma unroll
    for (int i = 0; i < NEQ * NEQ; ++i) jac[i] = 0.0;
#pragma unroll
    for (int m = 0; m < NM; m++) {
      jac[m * NEQ + m] = dpdv[m];
      jac[m * NEQ + NM] = dpdT[m];
    }
#pragma unroll
    for (int m = 0; m < NM; ++m) jac[NM * NEQ + m] = 1.0;
  }
  // :1851-1887
  SGM_HD Real ScaleDx() {
    Real scale = 1.0;
#pragma unroll
    for (int m = 0; m < NM; ++m)
      if (scale * dx[m] < -P.vfrac_safety_fac * vfrac[m])
        scale = -P.vfrac_safety_fac * robust::ratio(vfrac[m], dx[m]);
    const Real Tnew = Tequil + scale * dx[NM];
#pragma unroll
    for (int m = 0; m < NM; m++) {
      const Real rho_min =
          smax(Ev::rho_pmin(E(m), Tnorm * Tequil), Ev::rho_pmin(E(m), Tnorm * Tnew));
      const Real alpha_max = robust::ratio(rhobar[m], rho_min);
      if (!(alpha_max < vfrac[m]) && scale * dx[m] > 0.5 * (alpha_max - vfrac[m]))
        scale = 0.5 * robust::ratio(alpha_max - vfrac[m], dx[m]);
    }
    if (scale * dx[NM] < -0.95 * Tequil) scale = -0.95 * robust::ratio(Tequil, dx[NM]);
#pragma unroll
    for (int i = 0; i < NEQ; ++i) dx[i] *= scale;
    return scale;
  }
  // :1889-1917
  SGM_HD Real TestUpdate(Real scale, bool cache_state) {
    if (cache_state) {
      Ttemp = Tequil;
#pragma unroll
      for (int m = 0; m < NM; ++m) vtemp[m] = vfrac[m];
    }
    Tequil = Ttemp + scale * dx[NM];
    const Real Tc = Tnorm * Tequil;
#pragma unroll
    for (int m = 0; m < NM; ++m) {
      vfrac[m] = vtemp[m] + scale * dx[m];
      rho[m] = robust::ratio(rhobar[m], vfrac[m]);
      Real P_T;
      const Real um = rhobar[m] * Ev::e_and_P_rT(E(m), rho[m], Tc, lam[m], P_T);
      press[m] = robust::ratio(P_T, uscale);
      sie[m] = robust::ratio(um, rhobar[m]);
      u[m] = robust::ratio(um, uscale);
    }
    Residual();
    return B::ResidualNorm();
  }
  SGM_HD Real Tout() const { return Tequil * Tnorm; }
  template <class F>
  SGM_HD constexpr void Visit(F &f) {
    B::VisitBase(f);
    f(Tequil);
    f(Pequil);
  }
  SGM_HD Real Pout(int m) const { return press[m] * uscale; }
};

// ---------------------------------------------------------------------------
// PTESolverRhoU (:1924-2183, with the base's GetIdealPTE / TryIdealPTE,
// :672-763). Unknowns: the volume fractions and scaled energy densities u_m of
// every material (NEQ = 2 NM); residual: volume and energy closure, pressure
// and temperature differences of neighbouring materials; temperatures from
// T(rho, e) of each EOS. Scratch of the reference that lives only within one
// iteration (dpdv, dpde) stays local to Jacobian; dtdv / dtde carry from
// Jacobian to ScaleDx in the same iteration.
// ---------------------------------------------------------------------------
template <int NM, bool MODS = true>
struct PteRhoU : PteBase<NM, 2 * NM, MODS> {
  using Ev = sgm::Ev<MODS>;
  using B = PteBase<NM, 2 * NM, MODS>;
  using B::dx;
  using B::E;
  using B::jac;
  using B::lam;
  using B::P;
  using B::press;
  using B::res;
  using B::rho;
  using B::rhobar;
  using B::sie;
  using B::Tnorm;
  using B::u;
  using B::uscale;
  using B::utotal_scale;
  using B::vfrac;
  using B::vfrac_total;
  static constexpr bool kStoresTemp = true;
  static constexpr bool kPerMaterialT = true;
  static constexpr int NEQ = 2 * NM;
  Real temp[NM];
  Real vtemp[NM], utemp[NM];
  Real dtdv[NM], dtde[NM];

  SGM_HD void SetCellInputs(Real vfrac_tot, Real sie_tot, Real Tnorm_in) {
    vfrac_total = vfrac_tot;
    B::sie_total = sie_tot;
    Tnorm = Tnorm_in;
  }
  // :2014-2026
  SGM_HD void Residual() {
    NeumaierSum vs, es;
#pragma unroll
    for (int m = 0; m < NM; ++m) {
      vs.add(vfrac[m]);
      es.add(u[m]);
    }
    res[0] = vfrac_total - vs.result();
    res[1] = utotal_scale - es.result();
#pragma unroll
    for (int m = 0; m < NM - 1; ++m) res[2 + m] = press[m + 1] - press[m];
#pragma unroll
    for (int j = 0; j < NM - 1; ++j) res[NM + 1 + j] = temp[j + 1] - temp[j];
  }
  // :672-687
  SGM_HD void GetIdealPTE(Real &Pideal, Real &Tideal) const {
    Real rhoBsum = 0.0;
    Real Asum = 0.0;
#pragma unroll
    for (int m = 0; m < NM; m++) {
      Asum += vfrac[m] * robust::ratio(press[m], temp[m]);
      rhoBsum += rho[m] * vfrac[m] * robust::ratio(sie[m], temp[m]);
    }
    Asum *= uscale / Tnorm;
    rhoBsum /= Tnorm;
    Tideal = uscale / rhoBsum / Tnorm;
    Pideal = robust::ratio(Tnorm * Tideal * Asum, uscale * vfrac_total);
  }
  // :691-763
  SGM_HD void TryIdealPTE() {
    Real Pideal, Tideal;
    GetIdealPTE(Pideal, Tideal);
    Real etemp[NM], ptemp[NM], vt[NM], rtemp[NM], rs[NEQ];
#pragma unroll
    for (int m = 0; m < NM; ++m) {
      etemp[m] = u[m];
      ptemp[m] = press[m];
      vt[m] = vfrac[m];
      rtemp[m] = rho[m];
    }
    Real res_norm_old = 0.0;
#pragma unroll
    for (int i = 0; i < NEQ; ++i) {
      rs[i] = res[i];
      res_norm_old += rs[i] * rs[i];
    }
    const Real alpha = robust::ratio(Pideal, Tideal);
    bool abort = false;
#pragma unroll
    for (int m = 0; m < NM; ++m) {
      if (!abort) {
        vfrac[m] *= robust::ratio(press[m], (temp[m] * alpha));
        if (Tnorm * Tideal < 0 ||
            robust::ratio(rhobar[m], vfrac[m]) < Ev::rho_pmin(E(m), Tnorm * Tideal)) {
          // restore the volume fractions changed so far and keep the guess
#pragma unroll
          for (int n = 0; n < NM; n++)
            if (n <= m) vfrac[n] = vt[n];
          abort = true;
        }
      }
    }
    if (abort) return;
#pragma unroll
    for (int m = 0; m < NM; ++m) {
      rho[m] = robust::ratio(rhobar[m], vfrac[m]);
      Real P_T;
      const Real sie_m = Ev::e_rT_with_P(E(m), rho[m], Tnorm * Tideal, lam[m], P_T);
      u[m] = rhobar[m] * robust::ratio(sie_m, uscale);
      press[m] = robust::ratio(B::PressureFromPreferredP(m, rho[m], sie_m, P_T), uscale);
    }
    Residual();
    Real res_norm_new = 0.0;
#pragma unroll
    for (int i = 0; i < NEQ; i++) res_norm_new += res[i] * res[i];
    if (res_norm_new > res_norm_old) {
#pragma unroll
      for (int m = 0; m < NM; ++m) {
        vfrac[m] = vt[m];
        rho[m] = rtemp[m];
        u[m] = etemp[m];
        press[m] = ptemp[m];
      }
#pragma unroll
      for (int i = 0; i < NEQ; ++i) res[i] = rs[i];
    } else {
#pragma unroll
      for (int m = 0; m < NM; ++m) {
        temp[m] = Tideal;
        sie[m] = uscale * robust::ratio(u[m], rhobar[m]);
      }
    }
  }
  // :2005-2011 (InitBase sets every scaled temperature to 1, :641)
  SGM_HD Real Init(const Real (&Tin)[NM]) {
    B::InitBase(Tin);
#pragma unroll
    for (int m = 0; m < NM; ++m) temp[m] = 1.0;
    Residual();
    TryIdealPTE();
    return B::ResidualNorm();
  }
  SGM_HD void Fixup() {}  // the base's hook (:346)
  // :2029-2055 -> (converged, acceptable)
  SGM_HD void CheckPTE(bool &conv, bool &ok) const {
    Real mean_p = vfrac[0] * press[0];
    Real mean_t = rhobar[0] * temp[0];
    Real error_p = 0.0;
    Real error_t = 0.0;
#pragma unroll
    for (int m = 1; m < NM; ++m) {
      mean_p += vfrac[m] * press[m];
      mean_t += rhobar[m] * temp[m];
      error_p += res[m + 1] * res[m + 1];
      error_t += res[m + NM] * res[m + NM];
    }
    mean_t = robust::ratio(mean_t, B::rho_total);
    error_p = sqrt(error_p);
    error_t = sqrt(error_t);
    const bool cp = (error_p < P.pte_rel_tolerance_p * fabs(mean_p) ||
                     error_p < P.pte_abs_tolerance_p);
    const bool ct = (error_t < P.pte_rel_tolerance_t * mean_t ||
                     error_t < P.pte_abs_tolerance_t);
    const bool ap = (error_p < P.pte_rel_tolerance_p_sufficient * fabs(mean_p) ||
                     error_p < P.pte_abs_tolerance_p_sufficient);
    const bool at = (error_t < P.pte_rel_tolerance_t_sufficient * mean_t ||
                     error_t < P.pte_abs_tolerance_t_sufficient);
    conv = cp && ct;
    ok = ap && at;
  }
  // :2058-2121
  SGM_HD void Jacobian() {
    Real dpdv[NM], dpde[NM];
#pragma unroll
    for (int m = 0; m < NM; m++) {
      const Real dv = (vfrac[m] < 0.5 ? 1.0 : -1.0) * vfrac[m] * P.derivative_eps;
      const Real vf_pert = vfrac[m] + dv;
      const Real rho_pert = robust::ratio(rhobar[m], vf_pert);
      Real t_pert = robust::ratio(Ev::T_re(E(m), rho_pert, sie[m], lam[m]), Tnorm);
      Real p_pert =
          robust::ratio(B::PressureFromPreferred(m, rho_pert, Tnorm * t_pert, sie[m]), uscale);
      dpdv[m] = smin(robust::ratio(p_pert - press[m], dv), -fabs(P.pte_max_dpdv));
      dtdv[m] = robust::ratio(t_pert - temp[m], dv);
      const Real de = fabs(u[m]) * P.derivative_eps;
      const Real e_pert = robust::ratio(u[m] + de, rhobar[m]);
      t_pert = robust::ratio(Ev::T_re(E(m), rho[m], uscale * e_pert, lam[m]), Tnorm);
      p_pert = robust::ratio(
          B::PressureFromPreferred(m, rho[m], Tnorm * t_pert, uscale * e_pert), uscale);
      dpde[m] = robust::ratio(p_pert - press[m], de);
      dtde[m] = robust::ratio(t_pert - temp[m], de);
      if (fabs(dtde[m]) < P.min_dtde) dtde[m] = P.derivative_eps;
    }
    constexpr int n = NM, N = NEQ;
#pragma unroll
    for (int i = 0; i < N * N; ++i) jac[i] = 0.0;
#pragma unroll
    for (int m = 0; m < n; ++m) {
      jac[m] = 1.0;
      jac[2 * n + n + m] = 1.0;
    }
    jac[2 * N] = dpdv[0];
    jac[n * N + n - 1] = -dpdv[n - 1];
    jac[(n + 1) * N] = dtdv[0];
    jac[(2 * n - 1) * N + n - 1] = -dtdv[n - 1];
    jac[2 * N + n] = dpde[0];
    jac[n * N + 2 * n - 1] = -dpde[n - 1];
    jac[(n + 1) * N + n] = dtde[0];
    jac[(2 * n - 1) * N + 2 * n - 1] = -dtde[n - 1];
#pragma unroll
    for (int m = 1; m < n - 1; ++m) {
      jac[(1 + m) * N + m] = -dpdv[m];
      jac[(2 + m) * N + m] = dpdv[m];
      jac[(n + m) * N + m] = -dtdv[m];
      jac[(n + m + 1) * N + m] = dtdv[m];
      jac[(1 + m) * N + n + m] = -dpde[m];
      jac[(2 + m) * N + n + m] = dpde[m];
      jac[(n + m) * N + n + m] = -dtde[m];
      jac[(n + m + 1) * N + n + m] = dtde[m];
    }
  }
  // :2124-2158
  SGM_HD Real ScaleDx() {
    Real scale = 1.0;
#pragma unroll
    for (int m = 0; m < NM; ++m) {
      if (scale * dx[m] < -0.1 * vfrac[m]) scale = -0.1 * robust::ratio(vfrac[m], dx[m]);
      const Real dt = (dtdv[m] * dx[m] + dtde[m] * dx[m + NM]);
      if (scale * dt < -0.1 * temp[m]) scale = -0.1 * robust::ratio(temp[m], dt);
      const Real tt = temp[m] + scale * dt;
      const Real rho_min =
          smax(Ev::rho_pmin(E(m), Tnorm * temp[m]), Ev::rho_pmin(E(m), Tnorm * tt));
      const Real alpha_max = robust::ratio(rhobar[m], rho_min);
      if (!(alpha_max < vfrac[m])) {
        if (scale * dx[m] > 0.5 * (alpha_max - vfrac[m]))
          scale = 0.5 * robust::ratio(alpha_max - vfrac[m], dx[m]);
      }
    }
#pragma unroll
    for (int i = 0; i < NEQ; ++i) dx[i] *= scale;
    return scale;
  }
  // :2162-2181
  SGM_HD Real TestUpdate(Real scale, bool cache_state) {
    if (cache_state) {
#pragma unroll
      for (int m = 0; m < NM; ++m) {
        vtemp[m] = vfrac[m];
        utemp[m] = u[m];
      }
    }
#pragma unroll
    for (int m = 0; m < NM; ++m) {
      vfrac[m] = vtemp[m] + scale * dx[m];
      rho[m] = robust::ratio(rhobar[m], vfrac[m]);
      u[m] = utemp[m] + scale * dx[NM + m];
      sie[m] = uscale * robust::ratio(u[m], rhobar[m]);
      temp[m] = robust::ratio(Ev::T_re(E(m), rho[m], sie[m], lam[m]), Tnorm);
      press[m] =
          robust::ratio(B::PressureFromPreferred(m, rho[m], Tnorm * temp[m], sie[m]), uscale);
    }
    Residual();
    return B::ResidualNorm();
  }
  SGM_HD Real Tout() const { return temp[0] * Tnorm; }
  SGM_HD Real TempOut(int m) const { return temp[m] * Tnorm; }
  template <class F>
  SGM_HD constexpr void Visit(F &f) {
    B::VisitBase(f);
#pragma unroll
    for (int m = 0; m < NM; ++m) f(temp[m]);
  }
  SGM_HD Real Pout(int m) const { return press[m] * uscale; }
};

// ---------------------------------------------------------------------------
// PTESolver driver (:2188-2291)
// ---------------------------------------------------------------------------
// The driver is split into begin / one iteration / end so a lane can retire
// its cell and start the next one between iterations (the batched kernels);
// pte_solve chains them for one cell. The iteration body is the reference's
// loop body, in its order.
struct DriverState {
  Real err;
  int niter, small_step_iters, max_iter;
  // SolverStatus::max_line_niter and slow_convergence_detected
  int max_line_niter;
  bool converged, close_enough, slow;
};

template <class F>
SGM_HD constexpr void VisitDriver(DriverState &d, CellStatus &st, F &f) {
  // (fixed_point_probe skips words 1 and 9 of this order: the counters)
  f(d.err);
  f(d.niter);
  f(d.small_step_iters);
  f(d.max_iter);
  f(d.converged);
  f(d.close_enough);
  f(d.max_line_niter);
  f(d.slow);
  f(st.converged);
  f(st.niter);
  f(st.err);
  f(st.residual);
}

// A cell's whole solver and driver state as FP64 words (integers and flags
// are exact in a double): the batched kernels park cells that are still
// iterating after a set number of Newton iterations and a tail kernel
// resumes them. ParkWords bounds the word count of every formulation.
constexpr int ParkWords(int nm, int neq) { return 10 * nm + neq + 24; }
template <class Sys>
SGM_HD void ParkSave(Sys &s, DriverState &d, CellStatus &st, double *w) {
  int k = 0;
  auto put = [&](auto &x) { w[k++] = (double)x; };
  s.Visit(put);
  VisitDriver(d, st, put);
}
template <class Sys>
SGM_HD void ParkLoad(Sys &s, DriverState &d, CellStatus &st, const double *w) {
  int k = 0;
  auto get = [&](auto &x) { x = (typename std::remove_reference<decltype(x)>::type)w[k++]; };
  s.Visit(get);
  VisitDriver(d, st, get);
}
template <class Sys>
constexpr int ParkWordCount() {
  Sys s{};
  DriverState d{};
  CellStatus st{};
  int k = 0;
  auto count = [&](auto &) { ++k; };
  s.Visit(count);
  VisitDriver(d, st, count);
  return k;
}

// the loop state before the first pass; err: the residual of the initial state
SGM_HD void driver_start(DriverState &d, Real err, int max_iter) {
  d.err = err;
  d.converged = false;
  d.close_enough = false;
  d.max_iter = max_iter;
  d.small_step_iters = 0;
  d.niter = 0;
  d.max_line_niter = 0;
  d.slow = false;
}
SGM_HD void status_start(CellStatus &st) {
  st.converged = 0;
  st.niter = 0;
  // set by pte_end; defined here so that a parked state (and the fixed-point
  // probe's comparison of it) never holds an indeterminate word
  st.err = 0;
  st.residual = 0;
}

template <int NM, class S>
SGM_HD void pte_begin(S &s, const Real (&Tin)[NM], DriverState &d, CellStatus &st) {
  status_start(st);
  const Real err = s.Init(Tin);
  driver_start(d, err, NM * s.P.pte_max_iter_per_mat);
}

// one pass of `for (niter = 0; niter < max_iter; ++niter)`; true once the
// loop has ended (break or exhausted). Any solver with the formulation
// interface (CheckPTE, Jacobian, Solve, ScaleDx, TestUpdate, Fixup, err, P):
// the register-resident ones below and the one-wave-per-cell ones.
template <class S>
SGM_HD bool pte_step(S &s, DriverState &d, CellStatus &st) {
  const MixParamsDev &p = s.P;
  if (d.niter >= d.max_iter) return true;
  st.niter = d.niter;
  s.CheckPTE(d.converged, d.close_enough);
  if (d.converged) return true;
  s.Jacobian();
  if (!s.Solve()) {
    d.converged = false;
    s.err |= ERRB_LINSOLVE;
    return true;
  }
  Real scale = smin(1.0, s.ScaleDx());
  if (fabs(scale) < p.pte_small_step_thresh) d.small_step_iters++;
  if (d.small_step_iters >= p.pte_small_step_tries) {
    d.converged = false;
    s.err |= ERRB_SMALLSTEP;
    return true;
  }
  const Real gradfdx = -2.0 * scale * d.err;
  const Real err_old = d.err;
  // The reference's line search -- the full step; if it does not decrease
  // the residual enough, the half step, a quadratic-fit scale and up to
  // line_search_max_iter backtracking steps -- written as one loop around a
  // single TestUpdate call, so the (large, per-material) update code exists
  // once in the kernel instead of three times. The sequence of trial scales,
  // the accepted state and err are those of the reference's nested form.
  scale = 1.0;
  Real err = 0, err_full = 0;
  int stage = 0, li = 0;
  for (;;) {
    const Real e = s.TestUpdate(scale, stage == 0);
    if (stage == 0) {  // the full step
      err = err_full = e;
      if (!(e > err_old + p.line_search_alpha * gradfdx)) break;
      stage = 1;
      scale = 0.5;
    } else if (stage == 1) {  // the half step
      scale = 0.5;
      if (e < err_full && e < err_old)
        scale = 0.75 + 0.5 * robust::ratio(e - err_full, err_full - 2.0 * e + err_old);
      if (p.line_search_max_iter <= 0) break;  // the half step's state, err of the full step
      stage = 2;
    } else {  // backtracking
      err = e;
      if (e < err_old + p.line_search_alpha * scale * gradfdx) break;
      scale *= p.line_search_fac;
      if (li > d.max_line_niter) d.max_line_niter = li;
      if (++li >= p.line_search_max_iter) break;
    }
  }
  d.err = err;
  s.Fixup();
  if (((err > p.pte_slow_convergence_thresh * err_old) || d.niter == d.max_iter - 1) &&
      d.close_enough) {
    d.converged = true;
    d.slow = true;
    return true;
  }
  ++d.niter;
  return d.niter >= d.max_iter;
}
template <int NM, class S>
SGM_HD bool pte_iterate(S &s, DriverState &d, CellStatus &st) {
  return pte_step(s, d, st);
}

// Every lane of a group takes lane `src`'s copy of the cell state.
struct GroupTake {
  int src;
  template <class T>
  __device__ __forceinline__ void operator()(T &x) const {
    x = (T)__shfl((double)x, src, 64);
  }
};
template <class S>
__device__ __forceinline__ void group_take(S &s, DriverState &d, CellStatus &st, int src) {
  GroupTake take{src};
  s.Visit(take);
  VisitDriver(d, st, take);
}

// Fixed points of the Newton loop. A cell that stalls at the roundoff floor
// often maps its whole iteration-live state onto itself bit for bit; every
// further pass of the loop body then repeats that state until the iteration
// cap, where the last pass also applies the cap's slow-convergence test.
// fixed_point_probe compares the cell's state (every word Visit and
// VisitDriver see, the two iteration counters aside) with the copy in w,
// then stores the current state there (the layout of ParkSave, so a park
// record is a valid copy of the state it holds); true when the two are
// identical. A loop that finds its state unchanged by one pass may skip to
// the last pass: the skipped passes would reproduce the same state, so the
// state, counters and status after the last pass are those of the full loop.
struct FixedPointProbe {
  double *w;
  int k;
  bool same;
  int skip0, skip1;  // word positions not compared (the iteration counters)
  template <class T>
  SGM_HD void operator()(T &x) {
    const double v = (double)x;
    same = same && (k == skip0 || k == skip1 ||
                    __builtin_bit_cast(unsigned long long, v) ==
                        __builtin_bit_cast(unsigned long long, w[k]));
    w[k++] = v;
  }
};
// the number of words Visit writes for a formulation's solver state
template <class Sys>
constexpr int SolverWordCount() {
  Sys s{};
  int k = 0;
  auto count = [&](auto &) { ++k; };
  s.Visit(count);
  return k;
}
template <class Sys>
SGM_HD bool fixed_point_probe(Sys &s, DriverState &d, CellStatus &st, double *w) {
  // VisitDriver's order: d.err, d.niter, ..., st.converged, st.niter, ...
  constexpr int base = SolverWordCount<Sys>();
  FixedPointProbe f{w, 0, true, base + 1, base + 9};
  s.Visit(f);
  VisitDriver(d, st, f);
  return f.same;
}
// after a pass whose state matched the probe copy: resume at the last pass
SGM_HD void skip_to_last_pass(DriverState &d) {
  if (d.niter < d.max_iter - 1) d.niter = d.max_iter - 1;
}

// Whether one TestUpdate's result depends only on the state before the line
// search and its own scale (so trial steps can run on separate lanes):
// RhoT, FixedT and FixedP evaluate each material at the trial state; PT's
// P-T inversions start from the previous trial's densities.
template <class S>
struct PureTrialSteps : std::true_type {};
template <int NM, bool MODS>
struct PureTrialSteps<PtePT<NM, MODS>> : std::false_type {};

// The Jacobian of a group's cell: RhoT spreads its per-material derivative
// evaluations over the group's lanes; the other formulations evaluate it on
// every lane.
template <int G, class S>
__device__ __forceinline__ void group_jacobian(S &s, int, int) {
  s.Jacobian();
}
template <int G, int NM, bool MODS>
__device__ __forceinline__ void group_jacobian(PteRhoT<NM, MODS> &s, int g, int base) {
  if constexpr (G >= NM)
    s.template JacobianGroup<G>(g, base);
  else
    s.Jacobian();
}

// One pass of the PTESolver loop body (pte_iterate) by a group of G lanes
// (lanes base .. base+G-1, this lane g) that hold identical copies of one
// cell. Everything up to the line search runs on every lane; the line
// search's trial updates run side by side: lane 1 the half step and every
// other lane the full step, then lane k the k-th scale of the backtracking
// loop. The accepted trial is the one the sequential loop would stop at, and
// every lane takes its state, so the result is bit-identical to pte_iterate.
// Needs G >= max(2, line_search_max_iter); every branch below depends only
// on values shared by the whole group, so its lanes stay converged.
template <int NM, int G, class S>
__device__ __forceinline__ bool pte_iterate_group(S &s, DriverState &d, CellStatus &st, int g,
                                                  int base) {
  const MixParamsDev &p = s.P;
  if (d.niter >= d.max_iter) return true;
  st.niter = d.niter;
  s.CheckPTE(d.converged, d.close_enough);
  if (d.converged) return true;
  group_jacobian<G>(s, g, base);
  if (!s.Solve()) {
    d.converged = false;
    s.err |= ERRB_LINSOLVE;
    return true;
  }
  Real scale = smin(1.0, s.ScaleDx());
  if (fabs(scale) < p.pte_small_step_thresh) d.small_step_iters++;
  if (d.small_step_iters >= p.pte_small_step_tries) {
    d.converged = false;
    s.err |= ERRB_SMALLSTEP;
    return true;
  }
  const Real gradfdx = -2.0 * scale * d.err;
  const Real err_old = d.err;
  // trial 1: the full step (lane 1: the half step)
  const Real e1 = s.TestUpdate(g == 1 ? 0.5 : 1.0, true);
  const Real err_full = __shfl(e1, base, 64);
  const Real err_mid = __shfl(e1, base + 1, 64);
  const bool search = err_full > err_old + p.line_search_alpha * gradfdx;
  Real sc0 = 0.5;
  if (err_mid < err_full && err_mid < err_old)
    sc0 = 0.75 + 0.5 * robust::ratio(err_mid - err_full, err_full - 2.0 * err_mid + err_old);
  // trial 2: lane k the k-th scale of the backtracking loop
  Real ek = 0;
  if (search && g < p.line_search_max_iter) {
    Real sk = sc0;
    for (int k = 0; k < g; ++k) sk *= p.line_search_fac;
    ek = s.TestUpdate(sk, false);
  }
  Real ks[G];
#pragma unroll
  for (int k = 0; k < G; ++k) ks[k] = __shfl(ek, base + k, 64);
  int src = 0;
  Real err = err_full;
  if (search) {
    if (p.line_search_max_iter <= 0) {
      src = 1;  // the state of the half step stays, with the full step's error
    } else {
      int chosen = p.line_search_max_iter - 1;
      Real sc = sc0;
      bool found = false;
#pragma unroll
      for (int k = 0; k < G; ++k) {
        if (!found && k < p.line_search_max_iter) {
          if (ks[k] < err_old + p.line_search_alpha * sc * gradfdx) {
            chosen = k;
            found = true;
          }
          sc *= p.line_search_fac;
        }
      }
      src = chosen;
      err = pick<G>(ks, chosen);
      // the sequential loop records every trial index that failed
      const int failed = found ? chosen - 1 : p.line_search_max_iter - 1;
      if (failed > d.max_line_niter) d.max_line_niter = failed;
    }
  }
  group_take(s, d, st, base + src);
  d.err = err;
  s.Fixup();
  if (((err > p.pte_slow_convergence_thresh * err_old) || d.niter == d.max_iter - 1) &&
      d.close_enough) {
    d.converged = true;
    d.slow = true;
    return true;
  }
  ++d.niter;
  return d.niter >= d.max_iter;
}

// One pass of the PTESolver loop body by a whole wave whose 64 lanes hold
// identical copies of one RhoT cell, lane = 8 t + m: the Jacobian's
// derivative evaluations and every trial's material evaluations run one
// material per lane (JacobianWide, TestUpdateWide), so the wave evaluates
// each material kind once per stage instead of each material in turn; the
// full and the half step run as t = 0 and t = 1, then the backtracking
// scales as t = 0 .. line_search_max_iter - 1. The accepted trial is the one
// the sequential loop would stop at, and every lane takes its state, so the
// result is bit-identical to pte_iterate. Needs line_search_max_iter <= 8
// and NM <= 8, and a converged wave (every lane active); every branch
// depends on values that are equal on all lanes.
template <int NM, class S>
__device__ __forceinline__ bool pte_iterate_wide(S &s, DriverState &d, CellStatus &st, int lane) {
  const MixParamsDev &p = s.P;
  const int t = lane >> 3;
  if (d.niter >= d.max_iter) return true;
  st.niter = d.niter;
  s.CheckPTE(d.converged, d.close_enough);
  if (d.converged) return true;
  s.JacobianWide(lane);
  if (!s.SolveWave(lane)) {
    d.converged = false;
    s.err |= ERRB_LINSOLVE;
    return true;
  }
  Real scale = smin(1.0, s.ScaleDx());
  if (fabs(scale) < p.pte_small_step_thresh) d.small_step_iters++;
  if (d.small_step_iters >= p.pte_small_step_tries) {
    d.converged = false;
    s.err |= ERRB_SMALLSTEP;
    return true;
  }
  const Real gradfdx = -2.0 * scale * d.err;
  const Real err_old = d.err;
  // The two trial stages around one (rolled) TestUpdateWide. Stage 0: the
  // full step (t = 1: the half step). Stage 1, when the full step does not
  // decrease the residual enough: lanes of t the t-th scale of the
  // backtracking loop (scales past line_search_max_iter are computed and
  // unused; every lane takes part in the gathers).
  Real err_full = 0, err_mid = 0, sc0 = 0.5;
  bool search = false;
  Real ks[8];
#pragma clang loop unroll(disable)
  for (int stage = 0; stage < 2; ++stage) {
    Real sk = t == 1 ? 0.5 : 1.0;
    if (stage == 1) {
      sk = sc0;
      for (int k = 0; k < t; ++k) sk *= p.line_search_fac;
    }
    const Real e = s.TestUpdateWide(sk, stage == 0, lane);
    if (stage == 1) {
#pragma unroll
      for (int k = 0; k < 8; ++k) ks[k] = __shfl(e, 8 * k, 64);
      break;
    }
    err_full = __shfl(e, 0, 64);
    err_mid = __shfl(e, 8, 64);
    // a scalar: the loop's exit is uniform
    search = __builtin_amdgcn_readfirstlane(
        err_full > err_old + p.line_search_alpha * gradfdx ? 1 : 0);
    if (err_mid < err_full && err_mid < err_old)
      sc0 = 0.75 + 0.5 * robust::ratio(err_mid - err_full, err_full - 2.0 * err_mid + err_old);
    if (!(search && p.line_search_max_iter > 0)) break;
  }
  int src = 0;
  Real err = err_full;
  if (search) {
    if (p.line_search_max_iter <= 0) {
      src = 1;  // the state of the half step stays, with the full step's error
    } else {
      int chosen = p.line_search_max_iter - 1;
      Real sc = sc0;
      bool found = false;
#pragma unroll
      for (int k = 0; k < 8; ++k) {
        if (!found && k < p.line_search_max_iter) {
          if (ks[k] < err_old + p.line_search_alpha * sc * gradfdx) {
            chosen = k;
            found = true;
          }
          sc *= p.line_search_fac;
        }
      }
      src = chosen;
      err = pick<8>(ks, chosen);
      const int failed = found ? chosen - 1 : p.line_search_max_iter - 1;
      if (failed > d.max_line_niter) d.max_line_niter = failed;
    }
  }
  group_take(s, d, st, 8 * src);
  d.err = err;
  s.Fixup();
  if (((err > p.pte_slow_convergence_thresh * err_old) || d.niter == d.max_iter - 1) &&
      d.close_enough) {
    d.converged = true;
    d.slow = true;
    return true;
  }
  ++d.niter;
  return d.niter >= d.max_iter;
}

// When a loop that probes for fixed points (fixed_point_probe) needs the
// full comparison: a pass can only have repeated the state when it left the
// residual norm d.err unchanged, which wandering cells almost never do. After
// each pass q, ProbeGate::due is true when d.err equals the previous pass's;
// only then does the loop run the probe (compare with the copy, then store
// the state as the new copy), and it may skip when the probe matches and the
// copy was stored after pass q - 1 (a copy from an earlier pass would find a
// periodic orbit, not a fixed point). A fixed point is then found one pass
// after the first pass that repeats d.err, and the copy is written only at
// those passes instead of after every pass.
struct ProbeGate {
  unsigned long long last = 0;  // d.err of the state after the previous pass
  int copy_at = 0;              // pass whose state the copy holds (0: the state at load)
  int q = 0;                    // passes run
  SGM_HD void start(Real err) { last = __builtin_bit_cast(unsigned long long, err); }
  // after a pass: whether to run the probe
  SGM_HD bool due(Real err) {
    ++q;
    const unsigned long long b = __builtin_bit_cast(unsigned long long, err);
    const bool eq = b == last;
    last = b;
    return eq;
  }
  // after the probe that due() asked for: whether the loop may skip
  SGM_HD bool skip(bool same) {
    const bool ok = same && copy_at == q - 1;
    copy_at = q;
    return ok;
  }
};

template <int NM, class S>
SGM_HD void pte_end(S &s, const DriverState &d, CellStatus &st) {
  s.Fixup();
  st.converged = d.converged ? 1 : 0;
  st.residual = d.err;
  st.err = s.err;
}

template <int NM, class S>
SGM_HD CellStatus pte_solve(S &s, const Real (&Tin)[NM]) {
  CellStatus st;
  DriverState d;
  pte_begin<NM>(s, Tin, d, st);
  while (!pte_iterate<NM>(s, d, st)) {
  }
  pte_end<NM>(s, d, st);
  return st;
}

} // namespace sgm
