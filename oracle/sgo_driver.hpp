// ORACLE — test infrastructure only. CPU restatement of singularity-eos's
// get_sg_eos closure driver: eos/get_sg_eos.cpp:35-280 (input modes, scratch,
// copy-back), get_sg_eos_functors.hpp:28-139 (init_functor: mass-fraction
// normalisation and cutoff, PTE inputs) and :141-388 (final_functor: cell
// averages and per-material bulk modulus, heat capacity, Gruneisen
// parameter), and the four per-cell bodies get_sg_eos_rho_e.cpp:22-90,
// get_sg_eos_rho_t.cpp:22-113, get_sg_eos_rho_p.cpp:22-118,
// get_sg_eos_p_t.cpp:21-95. Per-material fraction arrays are column-major
// (Kokkos LayoutLeft): x(i, m) = x[i + m * cell_dim].
//
// One deliberate difference: the reference points each cell's lambda cache
// into the PTE solver's own scratch (get_sg_eos_rho_e.cpp:47-48 and the
// same offset in the other modes), where for enough materials it overlaps the
// solver's last arrays; here the cache has storage of its own, zeroed per
// cell as the reference zeroes its scratch. The two agree whenever no
// material writes the cache (only tabulated EOS do).
// Never linked into the shipped library.
#pragma once

#include "sgo_pte.hpp"

#include <vector>

namespace sgo {

constexpr Real EV2K = 1.160451930280894026e4;  // get_sg_eos.hpp:68

struct SgEosArgs {
  int nmat = 0, ncell = 0, cell_dim = 0, input = 1;
  const int *eos_offsets = nullptr;  // [nmat], 1-based indices into pool
  const EOS *pool = nullptr;
  const int *offsets = nullptr;      // [ncell], 1-based cell indices
  double *press = nullptr, *pmax = nullptr, *vol = nullptr, *spvol = nullptr;
  double *sie = nullptr, *temp = nullptr, *bmod = nullptr, *dpde = nullptr, *cv = nullptr;
  double *frac_mass = nullptr, *frac_vol = nullptr, *frac_ie = nullptr;
  double *frac_bmod = nullptr, *frac_dpde = nullptr, *frac_cv = nullptr;
  double mass_frac_cutoff = 0;
  MixParams params{};
};

// per-cell working storage (the reference's per-thread scratch views)
struct SgEosScratch {
  std::vector<int> idxs, mats;
  std::vector<Real> vfrac, sie, temp, press, rho, lambda, solver;
  explicit SgEosScratch(int nmat)
      : idxs(nmat), mats(nmat), vfrac(nmat), sie(nmat), temp(nmat), press(nmat), rho(nmat),
        lambda(3 * nmat), solver(PTESolverFixedPRequiredScratch(nmat) +
                                 PTESolverRhoTRequiredScratch(nmat) + 16) {}
};

// input modes (get_sg_eos.cpp:65-71)
enum SgInput { SG_RHO_T = -3, SG_RHO_P = -2, SG_P_T = -1, SG_NORM_RHO_E = 0, SG_RHO_E = 1 };

// one cell; returns the SolverStatus error bits of its PTE solve (0 for pure
// cells) and sets *converged
inline int sg_eos_cell(const SgEosArgs &a, int iloop, SgEosScratch &w, bool *converged) {
  const int nmat = a.nmat;
  const long cd = a.cell_dim;
  const int i = a.offsets[iloop] - 1;
  auto FM = [&](int m) -> double & { return a.frac_mass[i + m * cd]; };
  auto FV = [&](int m) -> double & { return a.frac_vol[i + m * cd]; };
  auto FI = [&](int m) -> double & { return a.frac_ie[i + m * cd]; };
  const EOS *pool = a.pool;
  for (Real &x : w.lambda) x = 0.0;
  for (Real &x : w.solver) x = 0.0;
  // The reference leaves the per-thread density scratch as the previous cell
  // left it where a mode does not set it (the P-T mode's initial density
  // guesses, get_sg_eos_p_t.cpp:72-75); here it starts at zero.
  for (Real &x : w.rho) x = 0.0;
  Real mass_sum = 0.0;
  int npte = 0;
  Real vfrac_sum = 0.0;
  int err = 0;
  bool pte_converged = true;
  Real t_mult = 0, s_mult = 0, p_mult = 0;  // init multipliers
  Real ft = 0, fs = 0, fp = 0;              // final multipliers
  const bool p_t = a.input == SG_P_T;
  if (!p_t) {
    if (a.input == SG_RHO_T) t_mult = 1;
    else if (a.input == SG_RHO_P) p_mult = 1;
    else s_mult = 1;
    // init_functor (get_sg_eos_functors.hpp:64-113)
    for (int m = 0; m < nmat; ++m) {
      mass_sum += FM(m);
      w.idxs[m] = a.eos_offsets[m] - 1;
      FV(m) = FM(m);
      FI(m) = a.sie[i] * FM(m) * s_mult;
    }
    for (int m = 0; m < nmat; ++m) FM(m) /= mass_sum;
    npte = 0;
    for (int m = 0; m < nmat; ++m) {
      if (FM(m) > a.mass_frac_cutoff) {
        w.idxs[npte] = a.eos_offsets[m] - 1;
        w.mats[npte] = m;
        npte += 1;
      }
      w.vfrac[m] = 0.0;
      w.sie[m] = 0.0;
      w.temp[m] = 0.0;
      w.press[m] = 0.0;
    }
    vfrac_sum = 0.;
    for (int mp = 0; mp < npte; ++mp) {
      const int m = w.mats[mp];
      w.vfrac[mp] = FM(m);
      vfrac_sum += w.vfrac[mp];
      w.rho[mp] = FM(m) / a.spvol[i] / w.vfrac[mp];
      w.temp[mp] = a.temp[i] * EV2K * t_mult;
      w.press[mp] = a.press[i] * p_mult;
      w.sie[mp] = a.sie[i] * FM(m) * s_mult;
    }
  }
  std::vector<EOS> eos_cell(nmat);
  for (int mp = 0; mp < nmat; ++mp) eos_cell[mp] = pool[w.idxs[mp]];
  Real *lam = w.lambda.data();
  if (a.input == SG_RHO_E || a.input == SG_NORM_RHO_E) {
    // get_sg_eos_rho_e.cpp:48-81
    if (npte > 1) {
      PTERhoT s(npte, eos_cell.data(), vfrac_sum, a.sie[i], w.rho.data(), w.vfrac.data(),
                w.sie.data(), w.temp.data(), w.press.data(), lam, w.solver.data(), 0.0,
                a.params);
      const SolverStatus st = PTESolver(s);
      pte_converged = st.converged;
      err = st.err;
    } else {
      w.temp[0] = eos_cell[0].T_re(w.rho[0], w.sie[0], lam);
      w.press[0] = eos_cell[0].P_rT(w.rho[0], w.temp[0], lam);
    }
    ft = 1.0, fs = 0.0, fp = 1.0;
  } else if (a.input == SG_RHO_T) {
    // get_sg_eos_rho_t.cpp:52-96
    Real sie_tot_true = 0.0;
    if (npte > 1) {
      PTEFixedT s(npte, eos_cell.data(), vfrac_sum, w.temp[0], w.rho.data(), w.vfrac.data(),
                  w.sie.data(), w.temp.data(), w.press.data(), lam, w.solver.data(), a.params);
      const SolverStatus st = PTESolver(s);
      pte_converged = st.converged;
      err = st.err;
      for (int mp = 0; mp < npte; ++mp) sie_tot_true += w.sie[mp] * FM(w.mats[mp]);
    } else {
      w.sie[0] = eos_cell[0].e_rT(w.rho[0], w.temp[0], lam);
      sie_tot_true = w.sie[0];
      w.press[0] = eos_cell[0].P_rT(w.rho[0], w.temp[0], lam);
    }
    a.sie[i] = sie_tot_true;
    ft = 0.0, fs = 0.0, fp = 1.0;
  } else if (a.input == SG_RHO_P) {
    // get_sg_eos_rho_p.cpp:52-103
    Real sie_tot_true = 0.0;
    if (npte > 1) {
      PTEFixedP s(npte, eos_cell.data(), vfrac_sum, w.press[0], w.rho.data(), w.vfrac.data(),
                  w.sie.data(), w.temp.data(), w.press.data(), lam, w.solver.data(), a.params);
      const SolverStatus st = PTESolver(s);
      pte_converged = st.converged;
      err = st.err;
      for (int mp = 0; mp < npte; ++mp) sie_tot_true += w.sie[mp] * FM(w.mats[mp]);
    } else {
      const EOS &e = eos_cell[0];
      auto p_from_t = [&](Real t) { return e.P_rT(w.rho[0], t, lam); };
      const Real r_temp = e.T_reference();
      Real temp_i = 0;
      root::findRoot(p_from_t, a.press[i], r_temp, 1.e-10 * r_temp, 1.e10 * r_temp, 1.e-12,
                     1.e-12, temp_i);
      w.sie[0] = e.e_rT(w.rho[0], temp_i, lam);
      sie_tot_true = w.sie[0];
      w.temp[0] = temp_i;
    }
    a.sie[i] = sie_tot_true;
    ft = 1.0, fs = 0.0, fp = 0.0;
  } else {
    // P-T input, get_sg_eos_p_t.cpp:40-90: every material, no PTE solve
    for (int m = 0; m < nmat; ++m) {
      mass_sum += FM(m);
      w.idxs[m] = a.eos_offsets[m] - 1;
      w.mats[m] = m;
      w.temp[m] = a.temp[i] * EV2K;
      w.press[m] = a.press[i];
    }
    for (int m = 0; m < nmat; ++m) FM(m) /= mass_sum;
    for (int mp = 0; mp < nmat; ++mp) eos_cell[mp] = pool[w.idxs[mp]];
    Real vfrac_tot = 0.0, sie_tot = 0.0;
    ErrFlag ef;
    for (int m = 0; m < nmat; ++m) {
      eos_cell[m].rhoe_PT(w.press[m], w.temp[m], lam + 3 * m, w.rho[m], w.sie[m], ef);
      w.vfrac[m] = FM(m) / w.rho[m] * mass_sum;
      vfrac_tot += w.vfrac[m];
      sie_tot += w.sie[m] * FM(m);
    }
    a.sie[i] = sie_tot;
    a.vol[i] = vfrac_tot;
    a.spvol[i] = a.vol[i] / mass_sum;
    for (int m = 0; m < nmat; ++m) w.vfrac[m] /= vfrac_tot;
    npte = nmat;
    err = ef.code;
    ft = 0.0, fs = 0.0, fp = 0.0;
  }
  // final_functor (get_sg_eos_functors.hpp:206-282)
  const bool do_t = ft == 1.0, do_s = fs == 1.0, do_p = fp == 1.0;
  if (do_t) a.temp[i] = 0.0;
  if (do_p) a.press[i] = 0.0;
  if (do_s) a.sie[i] = 0.0;
  a.bmod[i] = 0.0;
  a.cv[i] = 0.0;
  a.dpde[i] = 0.0;
  const bool do_frac_bmod = a.frac_bmod != nullptr;
  const bool do_frac_dpde = a.frac_dpde != nullptr;
  const bool do_frac_cv = a.frac_dpde != nullptr;  // as get_sg_eos.cpp:61-62 decides
  for (int mp = 0; mp < npte; ++mp) {
    const int m = w.mats[mp];
    const EOS &e = eos_cell[mp];
    Real *lm = lam + 3 * mp;
    a.press[i] += w.press[mp] * w.vfrac[mp] * fp;
    a.temp[i] += w.temp[mp] * w.vfrac[mp] * ft;
    const Real ie_m = w.sie[mp] * FM(m) * mass_sum;
    a.sie[i] += ie_m * fs;
    FI(m) = ie_m;
    FV(m) = w.vfrac[mp] * a.vol[i];
    const Real bmod_m = e.bmod_rT(w.rho[mp], w.temp[mp], lm);
    a.bmod[i] += bmod_m * w.vfrac[mp];
    const Real cv_m = EV2K * e.cv_rT(w.rho[mp], w.temp[mp], lm);
    a.cv[i] += cv_m * FM(m);
    const Real dpde_m = e.gruneisen_rT(w.rho[mp], w.temp[mp], lm);
    a.dpde[i] += dpde_m * w.vfrac[mp];
    if (do_frac_bmod) a.frac_bmod[i + m * cd] = bmod_m;
    if (do_frac_cv) a.frac_cv[i + m * cd] = cv_m;
    if (do_frac_dpde) a.frac_dpde[i + m * cd] = dpde_m;
  }
  if (do_t) a.temp[i] /= EV2K;
  if (do_s) a.sie[i] /= mass_sum;
  for (int m = 0; m < nmat; ++m) FM(m) *= mass_sum;
  // get_sg_eos_rho_e.cpp:84 (every mode)
  a.pmax[i] = a.press[i] > a.pmax[i] ? a.press[i] : a.pmax[i];
  if (converged) *converged = pte_converged;
  return err;
}

}  // namespace sgo
