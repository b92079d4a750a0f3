// ORACLE — test infrastructure only. CPU restatement of the singularity-eos
// pressure-temperature-equilibrium closure: MixParams, SolverStatus,
// PTESolverBase, PTESolverRhoT, PTESolverPT, PTESolverFixedT,
// PTESolverFixedP, PTESolverRhoU, the PTESolver Newton driver,
// ApproxTemperatureFromRhoMatU and MixUtils. Follows
// singularity-eos/closure/mixed_cell_models.hpp line by line in algorithm;
// the per-cell scratch layout mirrors the reference's AssignIncrement carving.
// Never linked into the shipped library.
#pragma once

#include "sgo_eos.hpp"

#include <utility>

namespace sgo {

// mixed_cell_models.hpp:43-78
struct MixParams {
  bool verbose = false;
  bool iterate_t_guess = true;
  Real derivative_eps = 3.0e-6;
  Real pte_rel_tolerance_p = 1.0e-6;
  Real pte_rel_tolerance_t = 1.0e-4;
  Real pte_rel_tolerance_e = 1.0e-6;
  Real pte_rel_tolerance_v = 1.0e-6;
  Real pte_abs_tolerance_p = 1.0e-6;
  Real pte_abs_tolerance_t = 1.0e-4;
  Real pte_slow_convergence_thresh = 0.99;
  Real pte_rel_tolerance_p_sufficient = 1.e2 * 1.0e-6;
  Real pte_rel_tolerance_t_sufficient = 10 * 1.0e-4;
  Real pte_rel_tolerance_e_sufficient = 1.e2 * 1.0e-6;
  Real pte_rel_tolerance_v_sufficient = 1.e2 * 1.0e-6;
  Real pte_abs_tolerance_p_sufficient = 1.e2 * 1.0e-6;
  Real pte_abs_tolerance_t_sufficient = 10 * 1.0e-4;
  long pte_max_iter_per_mat = 128;
  Real line_search_alpha = 1.e-2;
  long line_search_max_iter = 6;
  Real line_search_fac = 0.5;
  Real vfrac_safety_fac = 0.95;
  Real temperature_limit = 1.0e15;
  Real default_tguess = 300.;
  Real min_dtde = 1.0e-16;
  long pte_small_step_tries = 2;
  Real pte_small_step_thresh = 1e-16;
  Real pte_max_dpdv = -1e-8;
  long pte_reduced_system_exclude_idx = -1;
};

// mixed_cell_models.hpp:80-87 (+ an error code for conditions the reference
// aborts on; 0 = none)
struct SolverStatus {
  bool converged = false;
  bool slow_convergence_detected = false;
  long max_niter = 0;
  long max_line_niter = 0;
  long small_step_iters = 0;
  Real residual = 0;
  int err = 0;
};

inline int PTESolverRhoTRequiredScratch(int nmat) { return nmat * nmat + 4 * nmat + 6 * nmat; }
inline int PTESolverPTRequiredScratch(int nmat) { return 4 + 8 + 2 * nmat; }

// ---------------------------------------------------------------------------
// PTESolverBase — mixed_cell_models.hpp:329-786
// lambda: per-material lambda slots (stride 3), may be null
// ---------------------------------------------------------------------------
struct PTEBase {
  MixParams params_;
  int nmat, neq;
  long niter = 0;
  Real vfrac_total, sie_total;
  const EOS *eos;
  Real *rho, *vfrac, *sie, *temp, *press;
  Real *lambda;
  Real *jacobian, *dx, *sol_scratch, *residual, *u, *rhobar;
  Real rho_total = 0, uscale = 0, utotal_scale = 0, Tnorm;
  ErrFlag err;

  Real *lam(int m) const { return lambda ? lambda + 3 * m : nullptr; }

  PTEBase(int nmats, int neqs, const EOS *eos_, Real vfrac_tot, Real sie_tot, Real *rho_,
          Real *vfrac_, Real *sie_, Real *temp_, Real *press_, Real *lambda_, Real *&scratch,
          Real Tnorm_, const MixParams &p)
      : params_(p), nmat(nmats), neq(neqs), vfrac_total(vfrac_tot), sie_total(sie_tot),
        eos(eos_), rho(rho_), vfrac(vfrac_), sie(sie_), temp(temp_), press(press_),
        lambda(lambda_), Tnorm(Tnorm_) {
    jacobian = inc(scratch, neq * neq);
    dx = inc(scratch, neq);
    sol_scratch = inc(scratch, 2 * neq);
    residual = inc(scratch, neq);
    u = inc(scratch, nmat);
    rhobar = inc(scratch, nmat);
  }
  static Real *inc(Real *&s, int n) {
    Real *p = s;
    s += n;
    return p;
  }
  bool Solve() const {
    for (int m = 0; m < neq; m++) dx[m] = residual[m];
    return mix::solve_Ax_b_wscr(neq, jacobian, dx, sol_scratch);
  }
  // :388-399
  void InitRhoBarandRho() {
    for (int m = 0; m < nmat; ++m) rhobar[m] = rho[m] * vfrac[m];
    rho_total = mix::sum_neumaier(rhobar, nmat);
  }
  // :401-407
  void NormalizeVfrac() const {
    const Real vsum = mix::sum_neumaier(vfrac, nmat);
    for (int m = 0; m < nmat; ++m) vfrac[m] *= robust::ratio(vfrac_total, vsum);
  }
  // :423-498 (returns true on FAIL)
  bool SetVfracFromT(const Real T) {
    constexpr Real min_dv = 1e-12;
    constexpr Real safety = 0.9;
    auto get_vfrac_max = [&](int m) {
      const Real rho_min = eos[m].RhoPmin(T);
      return std::min(safety * robust::ratio(rhobar[m], rho_min), 1.0);
    };
    auto get_dv = [&](int m) {
      const Real vmax = get_vfrac_max(m);
      const Real vnew = std::min(vmax, vfrac[m]);
      return vfrac[m] - vnew;
    };
    Real vfrac_max_sum = 0;
    for (int m = 0; m < nmat; ++m) {
      const Real vmax = get_vfrac_max(m);
      if (vfrac[m] <= 0) vfrac[m] = vmax;
      vfrac_max_sum += vmax;
    }
    if (vfrac_max_sum < vfrac_total) return true;
    NormalizeVfrac();
    for (int it = 0; it < nmat; ++it) {
      Real remaining = 0;
      int n_uncapped = nmat;
      for (int m = 0; m < nmat; ++m) {
        const Real dv = get_dv(m);
        vfrac[m] -= dv;
        remaining += dv;
        n_uncapped -= (dv > min_dv);
      }
      if (std::abs(remaining) < min_dv) break;
      for (int m = 0; m < nmat; ++m) {
        const bool uncapped = ((get_vfrac_max(m) - vfrac[m]) > min_dv);
        // robust::ratio(Real, size_t): SMALL<size_t>() is 0, so this is a plain division
        if (uncapped) vfrac[m] += remaining / (Real)n_uncapped;
      }
    }
    NormalizeVfrac();
    return false;
  }
  // :583-597
  Real GetPressureFromPreferred(const EOS &e, Real r, Real T, Real s, Real *lm,
                                bool do_e_lookup) const {
    if (e.PreferredInput() & thermalqs::specific_internal_energy) {
      if (do_e_lookup) s = e.e_rT(r, T, lm);
      return e.P_re(r, s, lm);
    }
    return e.P_rT(r, T, lm);
  }
  // :599-620
  void GetSieCvFromTAndPreferred(const EOS &e, Real r, Real P, Real T, Real *lm, Real &s,
                                 Real &cv) {
    if (e.PreferredInput() & thermalqs::pressure) {
      const Real Pmin = e.MinimumPressure();
      const Real Pmax = e.MaximumPressureAtTemperature(T, err);
      Real Pguess;
      if (bad_value(P))
        Pguess = 0.5 * (Pmin + Pmax);
      else
        Pguess = std::max(Pmin, std::min(Pmax, P));
      e.rhoe_PT(Pguess, T, lm, r, s, err);
      cv = e.cv_re(r, s, lm);
    } else {
      s = e.e_rT(r, T, lm);
      cv = e.cv_rT(r, T, lm);
    }
  }
  // :500-581
  Real GetTguess() {
    Real Tguess;
    if (Tnorm > 0.0) {
      Tguess = Tnorm;
    } else {
      Tguess = params_.default_tguess;
      for (int m = 0; m < nmat; ++m) Tguess = std::max(Tguess, temp[m]);
    }
    auto newton_step = [&](Real T) {
      Real dudt = 0, usum = 0;
      for (int m = 0; m < nmat; ++m) {
        const Real rho_max = eos[m].MaximumDensity();
        Real s, cv;
        GetSieCvFromTAndPreferred(eos[m], std::min(rho[m], rho_max), press[m], T, lam(m), s,
                                  cv);
        usum += rhobar[m] * s;
        dudt += rhobar[m] * cv;
      }
      return T - robust::ratio((usum - utotal_scale * uscale), dudt);
    };
    const Real Tfactor = 10.0;
    bool bad_vfrac_guess = false;
    for (int i = 0; i < 3; i++) {
      if (params_.iterate_t_guess) {
        Tguess = std::min(params_.temperature_limit, std::max(Tguess, newton_step(Tguess)));
      }
      for (int m = 0; m < nmat; ++m) Tguess = std::max(eos[m].MinimumTemperature(), Tguess);
      bad_vfrac_guess = SetVfracFromT(Tguess);
      if (bad_vfrac_guess) {
        Tguess *= Tfactor;
        continue;
      }
      for (int m = 0; m < nmat; ++m) rho[m] = robust::ratio(rhobar[m], vfrac[m]);
      bad_vfrac_guess = false;
      for (int m = 0; m < nmat; ++m) {
        const Real rho_min = eos[m].RhoPmin(Tguess);
        if (rho[m] < rho_min) {
          bad_vfrac_guess = true;
          Tguess *= Tfactor;
          break;
        }
      }
      if (!bad_vfrac_guess) break;
    }
    return Tguess;
  }
  // :627-662
  void InitBase() {
    InitRhoBarandRho();
    const Real utotal = rho_total * sie_total;
    uscale = std::max(std::abs(utotal), 1.0e-14);
    utotal_scale = robust::ratio(utotal, uscale);
    const Real Tguess = GetTguess();
    Tnorm = Tguess;
    for (int m = 0; m < nmat; m++) {
      temp[m] = 1.0;
      sie[m] = eos[m].e_rT(rho[m], Tguess, lam(m));
      const unsigned long pref = eos[m].PreferredInput();
      if ((pref & thermalqs::pressure) && !bad_value(press[m])) {
        press[m] = robust::ratio(press[m], uscale);
      } else {
        press[m] = robust::ratio(
            GetPressureFromPreferred(eos[m], rho[m], Tguess, sie[m], lam(m), false), uscale);
      }
    }
    for (int m = 0; m < nmat; ++m) u[m] = sie[m] * robust::ratio(rhobar[m], uscale);
  }
  Real ResidualNorm() const {
    return 0.5 * mix::sum_neumaier_op(residual, neq, 0, -1, [](Real x) { return x * x; });
  }
  // :327-329 base Finalize
  void FinalizeBase() {
    for (int m = 0; m < nmat; m++) {
      temp[m] *= Tnorm;
      u[m] *= uscale;
      press[m] *= uscale;
    }
  }
};

// ---------------------------------------------------------------------------
// PTESolverRhoT — mixed_cell_models.hpp:873-1193
// ---------------------------------------------------------------------------
struct PTERhoT : PTEBase {
  Real *dpdv, *dedv, *dpdT, *vtemp;
  Real Tequil = 0, Ttemp = 0;
  int ms = 0;
  PTERhoT(int nmat_, const EOS *eos_, Real vfrac_tot, Real sie_tot, Real *rho_, Real *vfrac_,
          Real *sie_, Real *temp_, Real *press_, Real *lambda_, Real *scratch, Real Tnorm_,
          const MixParams &p)
      : PTEBase(nmat_, nmat_, eos_, vfrac_tot, sie_tot, rho_, vfrac_, sie_, temp_, press_,
                lambda_, scratch, Tnorm_, p) {
    dpdv = inc(scratch, nmat);
    dedv = inc(scratch, nmat);
    dpdT = inc(scratch, nmat);
    vtemp = inc(scratch, nmat);
  }
  // :947-984
  Real Init() {
    InitBase();
    if (params_.pte_reduced_system_exclude_idx >= 0) {
      ms = (int)params_.pte_reduced_system_exclude_idx;
    } else {
      ms = 0;
      Real xmax = 0;
      for (int m = 0; m < nmat; ++m) {
        const Real x = robust::ratio(rhobar[m], rho_total);
        if (x > xmax) {
          xmax = x;
          ms = m;
        }
      }
    }
    Residual();
    Tequil = temp[0];
    return ResidualNorm();
  }
  // :986-1008
  void Fixup() {
    if (params_.pte_reduced_system_exclude_idx < 0) {
      Real alphamax = 0;
      ms = 0;
      for (int m = 0; m < nmat; ++m) {
        if (vfrac[m] > alphamax) {
          alphamax = vfrac[m];
          ms = m;
        }
      }
      Residual();
    }
  }
  // :1010-1019
  void Residual() const {
    const Real esum = mix::sum_neumaier(u, nmat);
    residual[0] = utotal_scale - esum;
    int ires = 1;
    for (int m = 0; m < nmat; ++m) {
      if (m == ms) continue;
      residual[ires++] = press[ms] - press[m];
    }
  }
  // :1021-1040
  std::pair<bool, bool> CheckPTE() const {
    Real mean_p = 0;
    for (int m = 0; m < nmat; ++m) mean_p += vfrac[m] * press[m];
    const Real error_p =
        std::sqrt(mix::sum_neumaier_op(residual, neq - 1, 1, -1, [](Real x) { return x * x; }));
    const Real error_u = std::abs(residual[0]);
    const bool converged_p = (error_p < params_.pte_rel_tolerance_p * std::abs(mean_p) ||
                              error_p < params_.pte_abs_tolerance_p);
    const bool converged_u = (error_u < params_.pte_rel_tolerance_e);
    const bool acceptable_p =
        (error_p < params_.pte_rel_tolerance_p_sufficient * std::abs(mean_p) ||
         error_p < params_.pte_abs_tolerance_p_sufficient);
    const bool acceptable_u = (error_u < params_.pte_rel_tolerance_e_sufficient);
    return {converged_p && converged_u, acceptable_p && acceptable_u};
  }
  // :1042-1105
  void Jacobian() const {
    Real dedT_sum = 0.0;
    for (int m = 0; m < nmat; m++) {
      const Real dv = (vfrac[m] < 0.5 ? 1.0 : -1.0) * vfrac[m] * params_.derivative_eps;
      const Real vf_pert = vfrac[m] + dv;
      const Real rho_pert = robust::ratio(rhobar[m], vf_pert);
      Real e_pert = eos[m].e_rT(rho_pert, Tnorm * Tequil, lam(m));
      Real p_pert = robust::ratio(
          GetPressureFromPreferred(eos[m], rho_pert, Tnorm * Tequil, e_pert, lam(m), false),
          uscale);
      dpdv[m] = std::min(robust::ratio((p_pert - press[m]), dv), -std::abs(params_.pte_max_dpdv));
      dedv[m] = robust::ratio(rhobar[m] * robust::ratio(e_pert, uscale) - u[m], dv);
      const Real dT = Tequil * params_.derivative_eps;
      e_pert = eos[m].e_rT(rho[m], Tnorm * (Tequil + dT), lam(m));
      p_pert = robust::ratio(GetPressureFromPreferred(eos[m], rho[m], Tnorm * (Tequil + dT),
                                                      e_pert, lam(m), false),
                             uscale);
      dpdT[m] = robust::ratio((p_pert - press[m]), dT);
      dedT_sum += robust::ratio(rhobar[m] * robust::ratio(e_pert, uscale) - u[m], dT);
    }
    for (int i = 0; i < neq * neq; ++i) jacobian[i] = 0.0;
    jacobian[0] = dedT_sum;
    int col = 1;
    for (int m = 0; m < nmat; ++m) {
      if (m == ms) continue;
      jacobian[col++] = dedv[m] - dedv[ms];
    }
    int row = 1;
    for (int m = 0; m < nmat; m++) {
      if (m == ms) continue;
      jacobian[row * neq + 0] = dpdT[m] - dpdT[ms];
      int c = 1;
      for (int mm = 0; mm < nmat; ++mm) {
        if (mm == ms) continue;
        jacobian[row * neq + c] = (mm == m) ? dpdv[m] + dpdv[ms] : dpdv[ms];
        c++;
      }
      row++;
    }
  }
  // :1107-1151
  Real ScaleDx() const {
    Real scale = 1.0;
    const Real dalphaskip = -mix::sum_neumaier(dx, nmat - 1, 1);
    int idx = 1;
    for (int m = 0; m < nmat; ++m) {
      const Real mydx = (m == ms) ? dalphaskip : dx[idx++];
      if (scale * mydx < -params_.vfrac_safety_fac * vfrac[m]) {
        scale = std::min(scale, std::abs(params_.vfrac_safety_fac * robust::ratio(vfrac[m], mydx)));
      }
    }
    const Real Tnew = Tequil + scale * dx[0];
    idx = 1;
    for (int m = 0; m < nmat; m++) {
      const Real rho_min = std::max(eos[m].RhoPmin(Tnorm * Tequil), eos[m].RhoPmin(Tnorm * Tnew));
      const Real alpha_max = std::min(1.0, robust::ratio(rhobar[m], rho_min));
      if (alpha_max < vfrac[m]) continue;
      const Real mydx = (m == ms) ? dalphaskip : dx[idx++];
      if (scale * mydx > 0.5 * (alpha_max - vfrac[m])) {
        scale = std::min(robust::ratio(0.5 * (alpha_max - vfrac[m]), mydx), scale);
      }
    }
    if (scale * dx[0] < -0.95 * Tequil) {
      scale = std::min(scale, robust::ratio(-0.95 * Tequil, dx[0]));
    }
    for (int i = 0; i < neq; ++i) dx[i] *= scale;
    return scale;
  }
  // :1155-1187
  Real TestUpdate(const Real scale, bool cache_state = false) {
    if (cache_state) {
      Ttemp = Tequil;
      for (int m = 0; m < nmat; ++m) vtemp[m] = vfrac[m];
    }
    Tequil = Ttemp + scale * dx[0];
    int idx = 1;
    for (int m = 0; m < nmat; ++m) {
      if (ms == m) continue;
      vfrac[m] = vtemp[m] + scale * dx[idx++];
    }
    vfrac[ms] = mix::sum_neumaier(vfrac, nmat, 0, ms);
    vfrac[ms] = 1 - vfrac[ms];
    for (int m = 0; m < nmat; ++m) {
      rho[m] = robust::ratio(rhobar[m], vfrac[m]);
      u[m] = rhobar[m] * eos[m].e_rT(rho[m], Tnorm * Tequil, lam(m));
      sie[m] = robust::ratio(u[m], rhobar[m]);
      u[m] = robust::ratio(u[m], uscale);
      temp[m] = Tequil;
      press[m] = robust::ratio(
          GetPressureFromPreferred(eos[m], rho[m], Tnorm * Tequil, sie[m], lam(m), false),
          uscale);
    }
    Residual();
    return ResidualNorm();
  }
  void Finalize() { FinalizeBase(); }
};

// ---------------------------------------------------------------------------
// PTESolverPT — mixed_cell_models.hpp:1198-1446
// ---------------------------------------------------------------------------
struct PTEPT : PTEBase {
  Real Pequil = 0, Tequil = 0, Ptemp = 0, Ttemp = 0;
  PTEPT(int nmat_, const EOS *eos_, Real vfrac_tot, Real sie_tot, Real *rho_, Real *vfrac_,
        Real *sie_, Real *temp_, Real *press_, Real *lambda_, Real *scratch, Real Tnorm_,
        const MixParams &p)
      : PTEBase(nmat_, 2, eos_, vfrac_tot, sie_tot, rho_, vfrac_, sie_, temp_, press_, lambda_,
                scratch, Tnorm_, p) {}
  // :1258-1297
  Real Init() {
    InitBase();
    Residual();
    Pequil = 0;
    for (int m = 0; m < nmat; ++m) Pequil += std::abs(press[m]) * vfrac[m];
    const Real vsum = mix::sum_neumaier(vfrac, nmat);
    Pequil /= vsum;
    Tequil = 1;
    for (int m = 0; m < nmat; ++m) {
      eos[m].rhoe_PT(Pequil * uscale, Tequil * Tnorm, lam(m), rho[m], sie[m], err);
      vfrac[m] = robust::ratio(rhobar[m], rho[m]);
      u[m] = robust::ratio(sie[m] * rhobar[m], uscale);
      temp[m] = Tequil;
    }
    Residual();
    return ResidualNorm();
  }
  // :1299-1312
  void Finalize() {
    for (int m = 0; m < nmat; m++) {
      press[m] = Pequil * uscale;
      temp[m] = Tequil * Tnorm;
      u[m] *= uscale;
    }
  }
  void Fixup() {}
  // :1314-1320
  void Residual() const {
    const Real vsum = mix::sum_neumaier(vfrac, nmat);
    const Real esum = mix::sum_neumaier(u, nmat);
    residual[0] = vfrac_total - vsum;
    residual[1] = utotal_scale - esum;
  }
  // :1322-1334
  std::pair<bool, bool> CheckPTE() const {
    const Real error_v = std::abs(residual[0]);
    const Real error_u = std::abs(residual[1]);
    const bool converged_u = error_u < params_.pte_rel_tolerance_e;
    const bool converged_v = error_v < params_.pte_rel_tolerance_v;
    const bool acceptable_u = error_u < params_.pte_rel_tolerance_e_sufficient;
    const bool acceptable_v = error_v < params_.pte_rel_tolerance_v_sufficient;
    return {converged_v && converged_u, acceptable_u && acceptable_v};
  }
  // :1336-1385
  void Jacobian() {
    Real dudT_P_sum = 0.0, dudP_T_sum = 0.0, rb_dT = 0.0, rb_dP = 0.0;
    for (int m = 0; m < nmat; ++m) {
      Real r_pert = rho[m];
      Real e_pert = sie[m];
      const Real dp = -Pequil * params_.derivative_eps;
      eos[m].rhoe_PT(uscale * (Pequil + dp), Tnorm * Tequil, lam(m), r_pert, e_pert, err);
      const Real drdp = robust::ratio(r_pert - rho[m], dp);
      const Real dudp = robust::ratio(robust::ratio(rhobar[m] * e_pert, uscale) - u[m], dp);
      rb_dP += robust::ratio(rhobar[m], rho[m] * rho[m]) * drdp;
      dudP_T_sum += dudp;
      const Real dT = Tequil * params_.derivative_eps;
      eos[m].rhoe_PT(uscale * Pequil, Tnorm * (Tequil + dT), lam(m), r_pert, e_pert, err);
      const Real drdT = robust::ratio(r_pert - rho[m], dT);
      const Real dudT = robust::ratio(robust::ratio(rhobar[m] * e_pert, uscale) - u[m], dT);
      rb_dT += robust::ratio(rhobar[m], rho[m] * rho[m]) * drdT;
      dudT_P_sum += dudT;
    }
    jacobian[0] = -rb_dT;
    jacobian[1] = -rb_dP;
    jacobian[2] = dudT_P_sum;
    jacobian[3] = dudP_T_sum;
  }
  // :1387-1409
  Real ScaleDx() {
    Real scale = 1.0;
    if (scale * dx[0] < -0.95 * Tequil) scale = robust::ratio(-0.95 * Tequil, dx[0]);
    auto bounded = [&](Real Pbound, Real delta) {
      return robust::ratio(robust::ratio(Pbound, uscale) - Pequil, delta);
    };
    for (int m = 0; m < nmat; ++m) {
      const Real Pmin = eos[m].MinimumPressure();
      scale = std::min(std::abs(scale), std::abs(0.95 * bounded(Pmin, dx[1])));
    }
    for (int m = 0; m < nmat; ++m) {
      const Real Ttest = (Tequil + scale * dx[0]) * Tnorm;
      const Real Pmax = eos[m].MaximumPressureAtTemperature(Ttest, err);
      scale = std::min(std::abs(scale), std::abs(0.95 * bounded(Pmax, dx[0])));
    }
    for (int i = 0; i < neq; ++i) dx[i] *= scale;
    return scale;
  }
  // :1413-1435
  Real TestUpdate(const Real scale, bool cache_state = false) {
    if (cache_state) {
      Ttemp = Tequil;
      Ptemp = Pequil;
    }
    Tequil = Ttemp + scale * dx[0];
    Pequil = Ptemp + scale * dx[1];
    for (int m = 0; m < nmat; ++m) {
      eos[m].rhoe_PT(Pequil * uscale, Tequil * Tnorm, lam(m), rho[m], sie[m], err);
      vfrac[m] = robust::ratio(rhobar[m], rho[m]);
      u[m] = robust::ratio(sie[m] * rhobar[m], uscale);
      temp[m] = Tequil;
      press[m] = Pequil;
    }
    Residual();
    return ResidualNorm();
  }
};

// ---------------------------------------------------------------------------
// PTESolverFixedT — mixed_cell_models.hpp:1451-1669. Temperature given;
// unknowns the nmat volume fractions; residual: volume closure and nmat-1
// neighbouring pressure differences.
// ---------------------------------------------------------------------------
inline int PTESolverFixedTRequiredScratch(int nmat) {
  return nmat * nmat + 4 * nmat + 2 * nmat + 2 * nmat;  // :1451-1457
}

struct PTEFixedT : PTEBase {
  Real *dpdv, *vtemp;
  Real Tequil = 0, Ttemp = 0;
  // :1496-1509: the base gets sie_tot = 1 and Tnorm = T_true, then Tnorm = 1
  PTEFixedT(int nmat_, const EOS *eos_, Real vfrac_tot, Real T_true, Real *rho_, Real *vfrac_,
            Real *sie_, Real *temp_, Real *press_, Real *lambda_, Real *scratch,
            const MixParams &p)
      : PTEBase(nmat_, nmat_, eos_, vfrac_tot, 1.0, rho_, vfrac_, sie_, temp_, press_, lambda_,
                scratch, T_true, p) {
    dpdv = inc(scratch, nmat);
    vtemp = inc(scratch, nmat);
    Tequil = T_true;
    Ttemp = T_true;
    Tnorm = 1.0;
  }
  // :1523-1551
  Real Init() {
    Tnorm = 1.0;
    InitRhoBarandRho();
    SetVfracFromT(Tequil);
    uscale = 0.0;
    for (int m = 0; m < nmat; m++) {
      rho[m] = robust::ratio(rhobar[m], vfrac[m]);
      sie[m] = eos[m].e_rT(rho[m], Tequil, lam(m));
      uscale += std::abs(sie[m]) * rho[m];
      press[m] = eos[m].P_rT(rho[m], Tequil, lam(m));
    }
    uscale = std::max(uscale, 1.0e-14);
    for (int m = 0; m < nmat; ++m) {
      press[m] = robust::ratio(press[m], uscale);
      u[m] = sie[m] * robust::ratio(rhobar[m], uscale);
    }
    Residual();
    return ResidualNorm();
  }
  void Finalize() { FinalizeBase(); }
  void Fixup() {}
  // :1553-1560
  void Residual() const {
    const Real vsum = mix::sum_neumaier(vfrac, nmat);
    residual[0] = vfrac_total - vsum;
    for (int m = 0; m < nmat - 1; ++m) residual[1 + m] = press[m] - press[m + 1];
  }
  // :1562-1581
  std::pair<bool, bool> CheckPTE() const {
    Real mean_p = vfrac[0] * press[0];
    Real error_p = 0;
    for (int m = 1; m < nmat; ++m) {
      mean_p += vfrac[m] * press[m];
      error_p += residual[m] * residual[m];
    }
    error_p = std::sqrt(error_p);
    const Real error_v = std::abs(residual[0]);
    const bool converged_p = (error_p < params_.pte_rel_tolerance_p * std::abs(mean_p) ||
                              error_p < params_.pte_abs_tolerance_p);
    const bool converged_v = error_v < params_.pte_rel_tolerance_v;
    const bool acceptable_p =
        (error_p < params_.pte_rel_tolerance_p_sufficient * std::abs(mean_p) ||
         error_p < params_.pte_abs_tolerance_p_sufficient);
    const bool acceptable_v = error_v < params_.pte_rel_tolerance_v_sufficient;
    return {converged_p && converged_v, acceptable_p && acceptable_v};
  }
  // :1583-1609
  void Jacobian() {
    for (int m = 0; m < nmat; m++) {
      const Real dv = (vfrac[m] < 0.5 ? 1.0 : -1.0) * vfrac[m] * params_.derivative_eps;
      const Real vf_pert = vfrac[m] + dv;
      const Real rho_pert = robust::ratio(rhobar[m], vf_pert);
      const Real p_pert = robust::ratio(eos[m].P_rT(rho_pert, Tequil, lam(m)), uscale);
      dpdv[m] = std::min(robust::ratio((p_pert - press[m]), dv), -std::abs(params_.pte_max_dpdv));
    }
    for (int i = 0; i < neq * neq; ++i) jacobian[i] = 0.0;
    for (int m = 0; m < nmat; ++m) jacobian[m] = 1.0;
    for (int m = 0; m < nmat - 1; m++) {
      jacobian[(m + 1) * neq + m] = -dpdv[m];
      jacobian[(m + 1) * neq + m + 1] = dpdv[m + 1];
    }
  }
  // :1611-1641
  Real ScaleDx() {
    Real scale = 1.0;
    for (int m = 0; m < nmat; ++m) {
      if (scale * dx[m] < -params_.vfrac_safety_fac * vfrac[m])
        scale = robust::ratio(-params_.vfrac_safety_fac * vfrac[m], dx[m]);
    }
    for (int m = 0; m < nmat; m++) {
      const Real rho_min = eos[m].RhoPmin(Tequil);
      const Real alpha_max = robust::ratio(rhobar[m], rho_min);
      if (alpha_max < vfrac[m]) continue;  // already unstable: no limit (as the reference)
      if (scale * dx[m] > 0.5 * (alpha_max - vfrac[m]))
        scale = robust::ratio(0.5 * (alpha_max - vfrac[m]), dx[m]);
    }
    for (int i = 0; i < neq; ++i) dx[i] *= scale;
    return scale;
  }
  // :1643-1664
  Real TestUpdate(const Real scale, bool cache_state = false) {
    if (cache_state)
      for (int m = 0; m < nmat; ++m) vtemp[m] = vfrac[m];
    for (int m = 0; m < nmat; ++m) {
      vfrac[m] = vtemp[m] + scale * dx[m];
      rho[m] = robust::ratio(rhobar[m], vfrac[m]);
      u[m] = rhobar[m] * eos[m].e_rT(rho[m], Tequil, lam(m));
      sie[m] = robust::ratio(u[m], rhobar[m]);
      u[m] = robust::ratio(u[m], uscale);
      press[m] = robust::ratio(eos[m].P_rT(rho[m], Tequil, lam(m)), uscale);
    }
    Residual();
    return ResidualNorm();
  }
};

// ---------------------------------------------------------------------------
// PTESolverFixedP — mixed_cell_models.hpp:1674-1921. Pressure given;
// unknowns the nmat volume fractions and the temperature; residual: each
// material's pressure against P, then the volume closure.
// ---------------------------------------------------------------------------
inline int PTESolverFixedPRequiredScratch(int nmat) {
  const int neq = nmat + 1;  // :1674-1680
  return neq * neq + 4 * neq + 2 * nmat + 3 * nmat;
}

struct PTEFixedP : PTEBase {
  Real *dpdv, *dpdT, *vtemp;
  Real Tequil = 0, Ttemp = 0, Pequil = 0;
  // :1721-1733: the base gets sie_tot = 1 and Tnorm = 0
  PTEFixedP(int nmat_, const EOS *eos_, Real vfrac_tot, Real P, Real *rho_, Real *vfrac_,
            Real *sie_, Real *temp_, Real *press_, Real *lambda_, Real *scratch,
            const MixParams &p)
      : PTEBase(nmat_, nmat_ + 1, eos_, vfrac_tot, 1.0, rho_, vfrac_, sie_, temp_, press_,
                lambda_, scratch, 0.0, p) {
    dpdv = inc(scratch, nmat);
    dpdT = inc(scratch, nmat);
    vtemp = inc(scratch, nmat);
    Pequil = P;
    // The reference's Init calls GetTguess, whose Newton step reads uscale and
    // utotal_scale (:522-537) before anything has assigned them: their value
    // there is indeterminate. This restatement fixes both at zero, which makes
    // that step T - (sum rhobar e) / (sum rhobar cv).
    uscale = 0.0;
    utotal_scale = 0.0;
  }
  // :1747-1779
  Real Init() {
    InitRhoBarandRho();
    const Real Tguess = GetTguess();
    Tnorm = Tguess;
    uscale = std::max(std::abs(Pequil), 1.0e-14);
    for (int m = 0; m < nmat; m++) {
      temp[m] = 1.0;
      sie[m] = eos[m].e_rT(rho[m], Tguess, lam(m));
    }
    for (int m = 0; m < nmat; ++m) {
      u[m] = robust::ratio(sie[m] * rhobar[m], uscale);
      press[m] = robust::ratio(eos[m].P_rT(rho[m], Tguess, lam(m)), uscale);
    }
    Residual();
    Tequil = temp[0];
    Ttemp = Tequil;
    return ResidualNorm();
  }
  void Finalize() { FinalizeBase(); }
  void Fixup() {}
  // :1781-1788
  void Residual() const {
    const Real vsum = mix::sum_neumaier(vfrac, nmat);
    for (int m = 0; m < nmat; ++m) residual[m] = robust::ratio(Pequil, uscale) - press[m];
    residual[nmat] = vfrac_total - vsum;
  }
  // :1790-1808
  std::pair<bool, bool> CheckPTE() const {
    Real error_p = 0;
    for (int m = 0; m < nmat; ++m) error_p += residual[m] * residual[m];
    error_p = std::sqrt(error_p);
    error_p *= uscale;
    const Real error_v = std::abs(residual[neq - 1]);
    const bool converged_p =
        (error_p < params_.pte_rel_tolerance_p || error_p < params_.pte_abs_tolerance_p);
    const bool converged_v = error_v < params_.pte_rel_tolerance_v;
    const bool acceptable_p = (error_p < params_.pte_rel_tolerance_p_sufficient ||
                               error_p < params_.pte_abs_tolerance_p_sufficient);
    const bool acceptable_v = error_v < params_.pte_rel_tolerance_v_sufficient;
    return {converged_p && converged_v, acceptable_p && acceptable_v};
  }
  // :1809-1850
  void Jacobian() {
    for (int m = 0; m < nmat; m++) {
      const Real dv = (vfrac[m] < 0.5 ? 1.0 : -1.0) * vfrac[m] * params_.derivative_eps;
      const Real vf_pert = vfrac[m] + dv;
      const Real rho_pert = robust::ratio(rhobar[m], vf_pert);
      Real p_pert = robust::ratio(
          GetPressureFromPreferred(eos[m], rho_pert, Tnorm * Tequil, 0.0, lam(m), true), uscale);
      dpdv[m] = std::min(robust::ratio((p_pert - press[m]), dv), -std::abs(params_.pte_max_dpdv));
      const Real dT = Tequil * params_.derivative_eps;
      p_pert = robust::ratio(
          GetPressureFromPreferred(eos[m], rho[m], Tnorm * (Tequil + dT), 0.0, lam(m), true),
          uscale);
      dpdT[m] = robust::ratio((p_pert - press[m]), dT);
    }
    for (int i = 0; i < neq * neq; ++i) jacobian[i] = 0.0;
    for (int m = 0; m < nmat; m++) {
      jacobian[m * neq + m] = dpdv[m];
      jacobian[m * neq + nmat] = dpdT[m];
    }
    for (int m = 0; m < neq - 1; ++m) jacobian[(neq - 1) * neq + m] = 1.0;
  }
  // :1851-1887
  Real ScaleDx() {
    Real scale = 1.0;
    for (int m = 0; m < nmat; ++m) {
      if (scale * dx[m] < -params_.vfrac_safety_fac * vfrac[m])
        scale = -params_.vfrac_safety_fac * robust::ratio(vfrac[m], dx[m]);
    }
    const Real Tnew = Tequil + scale * dx[nmat];
    for (int m = 0; m < nmat; m++) {
      const Real rho_min =
          std::max(eos[m].RhoPmin(Tnorm * Tequil), eos[m].RhoPmin(Tnorm * Tnew));
      const Real alpha_max = robust::ratio(rhobar[m], rho_min);
      if (alpha_max < vfrac[m]) continue;  // already unstable: no limit (as the reference)
      if (scale * dx[m] > 0.5 * (alpha_max - vfrac[m]))
        scale = 0.5 * robust::ratio(alpha_max - vfrac[m], dx[m]);
    }
    if (scale * dx[nmat] < -0.95 * Tequil) scale = -0.95 * robust::ratio(Tequil, dx[nmat]);
    for (int i = 0; i < neq; ++i) dx[i] *= scale;
    return scale;
  }
  // :1889-1917
  Real TestUpdate(const Real scale, bool cache_state = false) {
    if (cache_state) {
      Ttemp = Tequil;
      for (int m = 0; m < nmat; ++m) vtemp[m] = vfrac[m];
    }
    Tequil = Ttemp + scale * dx[nmat];
    for (int m = 0; m < nmat; ++m) {
      vfrac[m] = vtemp[m] + scale * dx[m];
      rho[m] = robust::ratio(rhobar[m], vfrac[m]);
      u[m] = rhobar[m] * eos[m].e_rT(rho[m], Tnorm * Tequil, lam(m));
      press[m] = robust::ratio(eos[m].P_rT(rho[m], Tnorm * Tequil, lam(m)), uscale);
      sie[m] = robust::ratio(u[m], rhobar[m]);
      u[m] = robust::ratio(u[m], uscale);
      temp[m] = Tequil;
    }
    Residual();
    return ResidualNorm();
  }
};

// ---------------------------------------------------------------------------
// PTESolverRhoU — mixed_cell_models.hpp:1924-2183 (with the base's
// GetIdealPTE / TryIdealPTE, :672-763). Unknowns: the volume fractions and
// the scaled energy densities u_m of every material (neq = 2 nmat); residual:
// volume and energy closure, pressure and temperature differences of
// neighbouring materials. Temperatures come from T(rho, e) of each EOS.
// ---------------------------------------------------------------------------
inline int PTESolverRhoURequiredScratch(int nmat) {
  const int neq = 2 * nmat;
  return neq * neq + 4 * neq + 8 * nmat;
}

struct PTERhoU : PTEBase {
  Real *dpdv, *dtdv, *dpde, *dtde, *vtemp, *utemp;
  PTERhoU(int nmat_, const EOS *eos_, Real vfrac_tot, Real sie_tot, Real *rho_, Real *vfrac_,
          Real *sie_, Real *temp_, Real *press_, Real *lambda_, Real *scratch, Real Tnorm_,
          const MixParams &p)
      : PTEBase(nmat_, 2 * nmat_, eos_, vfrac_tot, sie_tot, rho_, vfrac_, sie_, temp_, press_,
                lambda_, scratch, Tnorm_, p) {
    dpdv = inc(scratch, nmat);
    dtdv = inc(scratch, nmat);
    dpde = inc(scratch, nmat);
    dtde = inc(scratch, nmat);
    vtemp = inc(scratch, nmat);
    utemp = inc(scratch, nmat);
  }
  // :672-687 (scaled ideal-gas estimate of the equilibrium P and T)
  void GetIdealPTE(Real &Pideal, Real &Tideal) const {
    Real rhoBsum = 0.0;
    Real Asum = 0.0;
    for (int m = 0; m < nmat; m++) {
      Asum += vfrac[m] * robust::ratio(press[m], temp[m]);
      rhoBsum += rho[m] * vfrac[m] * robust::ratio(sie[m], temp[m]);
    }
    Asum *= uscale / Tnorm;
    rhoBsum /= Tnorm;
    Tideal = uscale / rhoBsum / Tnorm;
    Pideal = robust::ratio(Tnorm * Tideal * Asum, uscale * vfrac_total);
  }
  // :691-763: replaces the initial guess by the ideal-gas estimate when that
  // lowers the residual (the reference keeps its copies in the Jacobian
  // scratch; here they are arrays of their own)
  void TryIdealPTE() {
    Real Pideal, Tideal;
    GetIdealPTE(Pideal, Tideal);
    std::vector<Real> etemp(nmat), ptemp(nmat), vt(nmat), rtemp(nmat), res(neq);
    for (int m = 0; m < nmat; ++m) {
      etemp[m] = u[m];
      ptemp[m] = press[m];
      vt[m] = vfrac[m];
      rtemp[m] = rho[m];
    }
    Real res_norm_old = 0.0;
    for (int m = 0; m < neq; ++m) {
      res[m] = residual[m];
      res_norm_old += res[m] * res[m];
    }
    const Real alpha = robust::ratio(Pideal, Tideal);
    for (int m = 0; m < nmat; ++m) {
      vfrac[m] *= robust::ratio(press[m], (temp[m] * alpha));
      if (Tnorm * Tideal < 0 ||
          robust::ratio(rhobar[m], vfrac[m]) < eos[m].RhoPmin(Tnorm * Tideal)) {
        for (int n = 0; n <= m; n++) vfrac[n] = vt[n];
        return;
      }
    }
    for (int m = 0; m < nmat; ++m) {
      rho[m] = robust::ratio(rhobar[m], vfrac[m]);
      const Real sie_m = eos[m].e_rT(rho[m], Tnorm * Tideal, lam(m));
      u[m] = rhobar[m] * robust::ratio(sie_m, uscale);
      press[m] = robust::ratio(
          GetPressureFromPreferred(eos[m], rho[m], Tnorm * Tideal, sie_m, lam(m), false), uscale);
    }
    Residual();
    Real res_norm_new = 0.0;
    for (int m = 0; m < neq; m++) res_norm_new += residual[m] * residual[m];
    if (res_norm_new > res_norm_old) {
      for (int m = 0; m < nmat; ++m) {
        vfrac[m] = vt[m];
        rho[m] = rtemp[m];
        u[m] = etemp[m];
        press[m] = ptemp[m];
      }
      for (int m = 0; m < neq; ++m) residual[m] = res[m];
    } else {
      for (int m = 0; m < nmat; ++m) {
        temp[m] = Tideal;
        sie[m] = uscale * robust::ratio(u[m], rhobar[m]);
      }
    }
  }
  // :2005-2011
  Real Init() {
    InitBase();
    Residual();
    TryIdealPTE();
    return ResidualNorm();
  }
  // :2014-2026
  void Residual() const {
    const Real vsum = mix::sum_neumaier(vfrac, nmat);
    const Real esum = mix::sum_neumaier(u, nmat);
    residual[0] = vfrac_total - vsum;
    residual[1] = utotal_scale - esum;
    for (int m = 0; m < nmat - 1; ++m) residual[2 + m] = press[m + 1] - press[m];
    for (int m = nmat + 1; m < neq; m++) residual[m] = temp[m - nmat] - temp[m - nmat - 1];
  }
  // :2029-2055
  std::pair<bool, bool> CheckPTE() const {
    Real mean_p = vfrac[0] * press[0];
    Real mean_t = rhobar[0] * temp[0];
    Real error_p = 0.0;
    Real error_t = 0.0;
    for (int m = 1; m < nmat; ++m) {
      mean_p += vfrac[m] * press[m];
      mean_t += rhobar[m] * temp[m];
      error_p += residual[m + 1] * residual[m + 1];
      error_t += residual[m + nmat] * residual[m + nmat];
    }
    mean_t = robust::ratio(mean_t, rho_total);
    error_p = std::sqrt(error_p);
    error_t = std::sqrt(error_t);
    const bool converged_p = (error_p < params_.pte_rel_tolerance_p * std::abs(mean_p) ||
                              error_p < params_.pte_abs_tolerance_p);
    const bool converged_t = (error_t < params_.pte_rel_tolerance_t * mean_t ||
                              error_t < params_.pte_abs_tolerance_t);
    const bool acceptable_p =
        (error_p < params_.pte_rel_tolerance_p_sufficient * std::abs(mean_p) ||
         error_p < params_.pte_abs_tolerance_p_sufficient);
    const bool acceptable_t = (error_t < params_.pte_rel_tolerance_t_sufficient * mean_t ||
                               error_t < params_.pte_abs_tolerance_t_sufficient);
    return {converged_p && converged_t, acceptable_p && acceptable_t};
  }
  // :2058-2121
  void Jacobian() const {
    for (int m = 0; m < nmat; m++) {
      const Real dv = (vfrac[m] < 0.5 ? 1.0 : -1.0) * vfrac[m] * params_.derivative_eps;
      const Real vf_pert = vfrac[m] + dv;
      const Real rho_pert = robust::ratio(rhobar[m], vf_pert);
      Real t_pert = robust::ratio(eos[m].T_re(rho_pert, sie[m], lam(m)), Tnorm);
      Real p_pert = robust::ratio(
          GetPressureFromPreferred(eos[m], rho_pert, Tnorm * t_pert, sie[m], lam(m), false),
          uscale);
      dpdv[m] = std::min(robust::ratio(p_pert - press[m], dv), -std::abs(params_.pte_max_dpdv));
      dtdv[m] = robust::ratio(t_pert - temp[m], dv);
      const Real de = std::abs(u[m]) * params_.derivative_eps;
      const Real e_pert = robust::ratio(u[m] + de, rhobar[m]);
      t_pert = robust::ratio(eos[m].T_re(rho[m], uscale * e_pert, lam(m)), Tnorm);
      p_pert = robust::ratio(GetPressureFromPreferred(eos[m], rho[m], Tnorm * t_pert,
                                                      uscale * e_pert, lam(m), false),
                             uscale);
      dpde[m] = robust::ratio(p_pert - press[m], de);
      dtde[m] = robust::ratio(t_pert - temp[m], de);
      if (std::abs(dtde[m]) < params_.min_dtde) dtde[m] = params_.derivative_eps;
    }
    const int n = nmat, N = 2 * nmat;
    for (int i = 0; i < neq * neq; ++i) jacobian[i] = 0.0;
    for (int m = 0; m < n; ++m) {
      jacobian[m] = 1.0;
      jacobian[2 * n + n + m] = 1.0;
    }
    jacobian[2 * N] = dpdv[0];
    jacobian[n * N + n - 1] = -dpdv[n - 1];
    jacobian[(n + 1) * N] = dtdv[0];
    jacobian[(2 * n - 1) * N + n - 1] = -dtdv[n - 1];
    jacobian[2 * N + n] = dpde[0];
    jacobian[n * N + 2 * n - 1] = -dpde[n - 1];
    jacobian[(n + 1) * N + n] = dtde[0];
    jacobian[(2 * n - 1) * N + 2 * n - 1] = -dtde[n - 1];
    for (int m = 1; m < n - 1; ++m) {
      jacobian[(1 + m) * N + m] = -dpdv[m];
      jacobian[(2 + m) * N + m] = dpdv[m];
      jacobian[(n + m) * N + m] = -dtdv[m];
      jacobian[(n + m + 1) * N + m] = dtdv[m];
      jacobian[(1 + m) * N + n + m] = -dpde[m];
      jacobian[(2 + m) * N + n + m] = dpde[m];
      jacobian[(n + m) * N + n + m] = -dtde[m];
      jacobian[(n + m + 1) * N + n + m] = dtde[m];
    }
  }
  // :2124-2158
  Real ScaleDx() const {
    Real scale = 1.0;
    for (int m = 0; m < nmat; ++m) {
      if (scale * dx[m] < -0.1 * vfrac[m]) scale = -0.1 * robust::ratio(vfrac[m], dx[m]);
      const Real dt = (dtdv[m] * dx[m] + dtde[m] * dx[m + nmat]);
      if (scale * dt < -0.1 * temp[m]) scale = -0.1 * robust::ratio(temp[m], dt);
      const Real tt = temp[m] + scale * dt;
      const Real rho_min =
          std::max(eos[m].RhoPmin(Tnorm * temp[m]), eos[m].RhoPmin(Tnorm * tt));
      const Real alpha_max = robust::ratio(rhobar[m], rho_min);
      if (alpha_max < vfrac[m]) continue;
      if (scale * dx[m] > 0.5 * (alpha_max - vfrac[m]))
        scale = 0.5 * robust::ratio(alpha_max - vfrac[m], dx[m]);
    }
    for (int i = 0; i < neq; ++i) dx[i] *= scale;
    return scale;
  }
  // :2162-2181
  Real TestUpdate(const Real scale, bool cache_state = false) {
    if (cache_state) {
      for (int m = 0; m < nmat; ++m) {
        vtemp[m] = vfrac[m];
        utemp[m] = u[m];
      }
    }
    for (int m = 0; m < nmat; ++m) {
      vfrac[m] = vtemp[m] + scale * dx[m];
      rho[m] = robust::ratio(rhobar[m], vfrac[m]);
      u[m] = utemp[m] + scale * dx[nmat + m];
      sie[m] = uscale * robust::ratio(u[m], rhobar[m]);
      temp[m] = robust::ratio(eos[m].T_re(rho[m], sie[m], lam(m)), Tnorm);
      press[m] = robust::ratio(
          GetPressureFromPreferred(eos[m], rho[m], Tnorm * temp[m], sie[m], lam(m), false),
          uscale);
    }
    Residual();
    return ResidualNorm();
  }
  void Fixup() {}  // the base's hook (:346), not overridden by RhoU
  void Finalize() { FinalizeBase(); }
};

// ---------------------------------------------------------------------------
// PTESolver driver — mixed_cell_models.hpp:2188-2291
// ---------------------------------------------------------------------------
template <class System>
inline SolverStatus PTESolver(System &s) {
  SolverStatus status;
  Real &err = status.residual;
  bool &converged = status.converged;
  bool close_enough = false;
  err = s.Init();
  const MixParams &params = s.params_;
  converged = false;
  const long pte_max_iter = s.nmat * params.pte_max_iter_per_mat;
  long &niter = s.niter;
  long &small_step_iters = status.small_step_iters;
  for (niter = 0; niter < pte_max_iter; ++niter) {
    status.max_niter = std::max(status.max_niter, niter);
    auto check = s.CheckPTE();
    converged = check.first;
    close_enough = check.second;
    if (converged) break;
    s.Jacobian();
    const bool success = s.Solve();
    if (!success) {
      converged = false;
      s.err.code |= ERR_LINSOLVE;
      break;
    }
    Real scale = std::min(1.0, s.ScaleDx());
    if (std::abs(scale) < params.pte_small_step_thresh) small_step_iters++;
    if (small_step_iters >= params.pte_small_step_tries) {
      converged = false;
      s.err.code |= ERR_SMALLSTEP;
      break;
    }
    const Real gradfdx = -2.0 * scale * err;
    scale = 1.0;
    const Real err_old = err;
    err = s.TestUpdate(scale, true);
    if (err > err_old + params.line_search_alpha * gradfdx) {
      scale = 0.5;
      const Real err_mid = s.TestUpdate(scale);
      if (err_mid < err && err_mid < err_old) {
        scale = 0.75 + 0.5 * robust::ratio(err_mid - err, err - 2.0 * err_mid + err_old);
      }
      for (long line_iter = 0; line_iter < params.line_search_max_iter; line_iter++) {
        err = s.TestUpdate(scale);
        if (err < err_old + params.line_search_alpha * scale * gradfdx) break;
        scale *= params.line_search_fac;
        status.max_line_niter = std::max(line_iter, status.max_line_niter);
      }
    }
    s.Fixup();
    if (((err > params.pte_slow_convergence_thresh * err_old) || niter == pte_max_iter - 1) &&
        close_enough) {
      converged = true;
      status.slow_convergence_detected = true;
      break;
    }
  }
  s.Fixup();
  s.Finalize();
  status.err = s.err.code;
  return status;
}

// ---------------------------------------------------------------------------
// ApproxTemperatureFromRhoMatU — mixed_cell_models.hpp:790-847
// ---------------------------------------------------------------------------
inline Real ApproxTemperatureFromRhoMatU(int nmat, const EOS *eos, Real u_tot, const Real *rho,
                                         const Real *vfrac, Real Tguess = 0.0,
                                         Real *lambda = nullptr) {
  constexpr Real minimum_temperature = 1.e-9;
  constexpr Real maximum_temperature = 1.e9;
  auto ufunc = [&](Real T) {
    Real usum = 0.0;
    for (int m = 0; m < nmat; m++)
      usum += rho[m] * vfrac[m] * eos[m].e_rT(rho[m], T, lambda ? lambda + 3 * m : nullptr);
    return usum;
  };
  Real ulo = ufunc(minimum_temperature);
  if (u_tot < ulo) return minimum_temperature;
  Real uhi = ufunc(maximum_temperature);
  if (u_tot > uhi) return maximum_temperature;
  Real lTlo = fastmath::lg(minimum_temperature);
  Real lThi = fastmath::lg(maximum_temperature);
  if (Tguess > minimum_temperature && Tguess < maximum_temperature) {
    const Real ug = ufunc(Tguess);
    if (ug < u_tot) {
      lTlo = fastmath::lg(Tguess);
      ulo = ug;
    } else {
      lThi = fastmath::lg(Tguess);
      uhi = ug;
    }
  }
  int iter = 0;
  constexpr int max_iter = 10;
  while (lThi - lTlo > 0.01 && iter < max_iter) {
    const Real lT = 0.5 * (lTlo + lThi);
    const Real uT = ufunc(fastmath::pow2(lT));
    if (uT < u_tot) {
      lTlo = lT;
      ulo = uT;
    } else {
      lThi = lT;
      uhi = uT;
    }
    iter++;
  }
  const Real alpha = robust::ratio((u_tot - ulo), (uhi - ulo));
  return fastmath::pow2((1.0 - alpha) * lTlo + alpha * lThi);
}

// ---------------------------------------------------------------------------
// MixUtils — mixed_cell_models.hpp:93-150 (returns false where the reference
// PORTABLE_REQUIREs)
// ---------------------------------------------------------------------------
inline bool EnforceMassVolumesSum(int nmat, Real tot_vol, Real *rho, Real *vfracs) {
  int imax = 0;
  Real fmax = -1;
  for (int m = 0; m < nmat; ++m) {
    if (!(vfracs[m] > 0)) return false;
    if (vfracs[m] > fmax) {
      fmax = vfracs[m];
      imax = m;
    }
  }
  Real vnew = tot_vol;
  for (int m = 0; m < nmat; ++m) {
    if (m == imax) continue;
    vnew -= vfracs[m];
  }
  if (!(vnew > 0) || !(vnew <= tot_vol)) return false;
  rho[imax] *= robust::ratio(vfracs[imax], vnew);
  vfracs[imax] = vnew;
  return true;
}
inline bool EnforceEnergiesSum(int nmat, Real tot_rho, Real tot_sie, const Real *rhos,
                               const Real *vfracs, Real *sies) {
  const Real tot_u = tot_sie * tot_rho;
  int imax = 0;
  Real maxc = 0;
  for (int m = 0; m < nmat; ++m) {
    const Real um = rhos[m] * vfracs[m] * sies[m];
    if (std::abs(um) > maxc) {
      imax = m;
      maxc = std::abs(um);
    }
  }
  if (!(maxc > 0)) return false;
  Real unew = tot_u;
  for (int m = 0; m < nmat; ++m) {
    if (m == imax) continue;
    unew -= rhos[m] * vfracs[m] * sies[m];
  }
  sies[imax] = robust::ratio(unew, rhos[imax] * vfracs[imax]);
  return true;
}

} // namespace sgo
