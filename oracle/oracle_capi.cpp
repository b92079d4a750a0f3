// ORACLE — test infrastructure only. extern "C" surface of the CPU
// restatement, loaded with ctypes by tests/ (as the parity checker) and by
// bench.py's cpu_baseline leg (as the host baseline, OpenMP over cells in the
// shape of singularity-eos/eos/get_sg_eos_rho_e.cpp:30-88: one solver object per
// cell, per-thread scratch, schedule(dynamic,256)).
#include "sgo_driver.hpp"
#include "sgo_pte.hpp"

#include <omp.h>

#include <vector>

using namespace sgo;

namespace {
// MixParams packed as doubles in this order (see params_from_array)
MixParams params_from_array(const double *a) {
  MixParams p;
  if (!a) return p;
  p.derivative_eps = a[0];
  p.pte_rel_tolerance_p = a[1];
  p.pte_rel_tolerance_t = a[2];
  p.pte_rel_tolerance_e = a[3];
  p.pte_rel_tolerance_v = a[4];
  p.pte_abs_tolerance_p = a[5];
  p.pte_abs_tolerance_t = a[6];
  p.pte_slow_convergence_thresh = a[7];
  p.pte_rel_tolerance_p_sufficient = a[8];
  p.pte_rel_tolerance_t_sufficient = a[9];
  p.pte_rel_tolerance_e_sufficient = a[10];
  p.pte_rel_tolerance_v_sufficient = a[11];
  p.pte_abs_tolerance_p_sufficient = a[12];
  p.pte_abs_tolerance_t_sufficient = a[13];
  p.pte_max_iter_per_mat = (long)a[14];
  p.line_search_alpha = a[15];
  p.line_search_max_iter = (long)a[16];
  p.line_search_fac = a[17];
  p.vfrac_safety_fac = a[18];
  p.temperature_limit = a[19];
  p.default_tguess = a[20];
  p.min_dtde = a[21];
  p.pte_small_step_tries = (long)a[22];
  p.pte_small_step_thresh = a[23];
  p.pte_max_dpdv = a[24];
  p.pte_reduced_system_exclude_idx = (long)a[25];
  p.iterate_t_guess = a[26] != 0.0;
  p.verbose = false;
  return p;
}
int threads_or_default(int n) { return n > 0 ? n : omp_get_max_threads(); }
} // namespace

extern "C" {

// ---- EOS handles ----------------------------------------------------------
void *sgo_eos_ideal(double gm1, double Cv) {
  auto *e = new EOS();
  e->kind = K_IDEAL;
  e->ig = IdealGas{Cv, gm1};
  return e;
}
void *sgo_eos_stiff(double gm1, double Cv, double Pinf, double qq) {
  auto *e = new EOS();
  e->kind = K_STIFF;
  e->sg = StiffGas{Cv, gm1, Pinf, qq};
  return e;
}
void *sgo_eos_gruneisen(double C0, double s1, double s2, double s3, double G0, double b,
                        double rho0, double T0, double P0, double Cv, double rho_max) {
  auto *e = new EOS();
  e->kind = K_GRUN;
  e->gr = Gruneisen::make(C0, s1, s2, s3, G0, b, rho0, T0, P0, Cv);
  if (rho_max > 0) e->gr.rho_max = rho_max;
  return e;
}
double sgo_gruneisen_compute_rho_max(double s1, double s2, double s3, double rho0) {
  return Gruneisen::ComputeRhoMax(s1, s2, s3, rho0);
}
void *sgo_eos_vinet(double rho0, double T0, double B0, double BP0, double A0, double Cv0,
                    double E0, double S0, const double *d2to40) {
  auto *e = new EOS();
  e->kind = K_VINET;
  e->vn = Vinet::make(rho0, T0, B0, BP0, A0, Cv0, E0, S0, d2to40);
  return e;
}
void *sgo_table_create(double lRhoMin, double lRhoMax, int nRho, double lTMin, double lTMax,
                       int nT, const double *P, const double *sie, const double *bMod,
                       const double *dPdRho, const double *dPdE, const double *dEdRho,
                       const double *dEdT, const double *PCold, const double *sieCold,
                       const double *bModCold, double rhoOffset, double TOffset) {
  auto *t = new SpinerRhoT();
  t->init(Grid1D::make(lRhoMin, lRhoMax, nRho), Grid1D::make(lTMin, lTMax, nT), P, sie, bMod,
          dPdRho, dPdE, dEdRho, dEdT, PCold, sieCold, bModCold, rhoOffset, TOffset);
  return t;
}
void sgo_table_free(void *t) { delete static_cast<SpinerRhoT *>(t); }
void *sgo_eos_table(void *t) {
  auto *e = new EOS();
  e->kind = K_TABLE;
  e->tb = static_cast<SpinerRhoT *>(t);
  return e;
}
void sgo_eos_free(void *e) { delete static_cast<EOS *>(e); }
// ScaledEOS<ShiftedEOS<T>> as init_sg_* builds it from enabled[0] (scaled),
// enabled[1] (shifted), vals[0] (scale), vals[1] (shift)
// (singularity_eos_init_utils.hpp:25-41, 75-76)
void sgo_eos_set_modifiers(void *h, int scaled, int shifted, double scale, double shift) {
  EOS *e = static_cast<EOS *>(h);
  e->scaled = scaled != 0;
  e->shifted = shifted != 0;
  e->scale = scaled ? scale : 1.0;
  e->inv_scale = scaled ? 1. / scale : 1.0;
  e->shift = shifted ? shift : 0.0;
}

// Export a derived table: 0 PMax 1 sielTMax 2 dEdTMax 3 gm1Max 4 rho_at_pmin
// 5 lTColdCrit 6 bModCold 7 dPdECold 8 dEdTCold 9 bMod (2-D). Returns length.
long sgo_table_export(void *tv, int which, double *out) {
  auto *t = static_cast<SpinerRhoT *>(tv);
  const std::vector<Real> *v = nullptr;
  switch (which) {
  case 0: v = &t->PMax_.d; break;
  case 1: v = &t->sielTMax_.d; break;
  case 2: v = &t->dEdTMax_.d; break;
  case 3: v = &t->gm1Max_.d; break;
  case 4: v = &t->rho_at_pmin_.d; break;
  case 5: v = &t->lTColdCrit_.d; break;
  case 6: v = &t->bModCold_.d; break;
  case 7: v = &t->dPdECold_.d; break;
  case 8: v = &t->dEdTCold_.d; break;
  case 9: v = &t->bMod_.d; break;
  default: return -1;
  }
  if (out) std::copy(v->begin(), v->end(), out);
  return (long)v->size();
}
double sgo_table_scalar(void *tv, int which) {
  auto *t = static_cast<SpinerRhoT *>(tv);
  switch (which) {
  case 0: return t->PMin_;
  case 1: return t->rhoMax_;
  case 2: return t->TMax_;
  default: return 0;
  }
}
// Bilinear interpolation of field (0 P, 1 sie, 2 bMod, 3 dEdT, 4 dPdE) at
// (lRho, lT) points.
void sgo_table_interp(void *tv, int field, const double *lRho, const double *lT, long n,
                      double *out) {
  auto *t = static_cast<SpinerRhoT *>(tv);
  const Table2D *f = field == 0   ? &t->P_
                     : field == 1 ? &t->sie_
                     : field == 2 ? &t->bMod_
                     : field == 3 ? &t->dEdT_
                                  : &t->dPdE_;
  for (long i = 0; i < n; ++i) out[i] = f->interp(lRho[i], lT[i]);
}
// Inverse lT(lRho, sie) with regula-falsi iteration counts and table status.
void sgo_table_lT_from_sie(void *tv, const double *lRho, const double *sie, long n,
                           double *lT, int *iters, int *where) {
  auto *t = static_cast<SpinerRhoT *>(tv);
  for (long i = 0; i < n; ++i) {
    TableStatus w;
    int it = 0;
    lT[i] = t->lTFromlRhoSie_(lRho[i], sie[i], w, nullptr, &it);
    if (iters) iters[i] = it;
    if (where) where[i] = (int)w;
  }
}

// ---- scalar utilities -----------------------------------------------------
void sgo_nqt(int which, const double *x, double *y, long n) {
  for (long i = 0; i < n; ++i) {
    switch (which) {
    case 0: y[i] = fastmath::lg(x[i]); break;
    case 1: y[i] = fastmath::pow2(x[i]); break;
    case 2: y[i] = fastmath::log10(x[i]); break;
    case 3: y[i] = fastmath::pow10(x[i]); break;
    case 5: y[i] = pmath::exp(x[i]); break;
    case 6: y[i] = pmath::cbrt(x[i]); break;
    default: y[i] = to_log(x[i], 0.0); break;
    }
  }
}
void sgo_nqt_constants(long long *ints, double *dbls) {
  ints[0] = fastmath::one_as_int;
  ints[1] = fastmath::mantissa_mask;
  ints[2] = fastmath::low_mask;
  dbls[0] = fastmath::scale_down;
  dbls[1] = fastmath::scale_up;
}
// root of scale*atan(x-shift)+offset (test_math_utils.cpp:95-141)
int sgo_regula_falsi_atan(double shift, double scale, double offset, double guess, double a,
                          double b, double xtol, double ytol, double *root, int *iters) {
  auto f = [=](double x) { return scale * std::atan(x - shift) + offset; };
  return (int)root::regula_falsi(f, 0.0, guess, a, b, xtol, ytol, *root, iters);
}
// findRoot on a cubic x^3 - c (secant then bisection)
int sgo_findroot_cubic(double c, double guess, double xmin, double xmax, double xtol,
                       double ytol, double *root) {
  auto f = [](double x) { return x * x * x; };
  return (int)root::findRoot(f, c, guess, xmin, xmax, xtol, ytol, *root);
}
int sgo_solve(int n, double *a, double *b) {
  double scr[16];
  return mix::solve_Ax_b_wscr(n, a, b, scr) ? 1 : 0;
}
// the pivoted QR alone (no row scaling), n >= 1: a (row-major) is left
// holding R in its upper triangle, x the solution, perm the pivot order
void sgo_qr_colpiv(int n, double *a, double *b, double *x, int *perm) {
  mix::qr_colpiv_solve(n, a, b, x, perm);
}
double sgo_sum_neumaier(const double *d, int n, int offset, int iskip) {
  return mix::sum_neumaier(d, n, offset, iskip);
}
int sgo_enforce_mass_volumes(int nmat, double tot_vol, double *rho, double *vfracs) {
  return EnforceMassVolumesSum(nmat, tot_vol, rho, vfracs) ? 1 : 0;
}
int sgo_enforce_energies(int nmat, double tot_rho, double tot_sie, const double *rhos,
                         const double *vfracs, double *sies) {
  return EnforceEnergiesSum(nmat, tot_rho, tot_sie, rhos, vfracs, sies) ? 1 : 0;
}
double sgo_approx_temperature(int nmat, void **eos, double u_tot, const double *rho,
                              const double *vfrac, double Tguess) {
  std::vector<EOS> e(nmat);
  for (int m = 0; m < nmat; ++m) e[m] = *static_cast<EOS *>(eos[m]);
  return ApproxTemperatureFromRhoMatU(nmat, e.data(), u_tot, rho, vfrac, Tguess);
}

// ---- vector EOS calls (EosBase vector loops, eos_base.hpp:169-230) ---------
// func: 0 P(rho,e) 1 T(rho,e) 2 e(rho,T) 3 P(rho,T) 4 cv(rho,e) 5 cv(rho,T)
//       6 bmod(rho,e) 7 bmod(rho,T) 8 MinInternalEnergy(rho)
void sgo_vec_eval(void *ev, int func, const double *in1, const double *in2, double *out,
                  long n, double *lambda, int lambda_stride, int nthreads) {
  const EOS &e = *static_cast<EOS *>(ev);
#pragma omp parallel for schedule(static) num_threads(threads_or_default(nthreads))
  for (long i = 0; i < n; ++i) {
    Real *lm = lambda ? lambda + (long)lambda_stride * i : nullptr;
    switch (func) {
    case 0: out[i] = e.P_re(in1[i], in2[i], lm); break;
    case 1: out[i] = e.T_re(in1[i], in2[i], lm); break;
    case 2: out[i] = e.e_rT(in1[i], in2[i], lm); break;
    case 3: out[i] = e.P_rT(in1[i], in2[i], lm); break;
    case 4: out[i] = e.cv_re(in1[i], in2[i], lm); break;
    case 5: out[i] = e.cv_rT(in1[i], in2[i], lm); break;
    case 6: out[i] = e.bmod_re(in1[i], in2[i], lm); break;
    case 7: out[i] = e.bmod_rT(in1[i], in2[i], lm); break;
    default: out[i] = e.min_e(in1[i]); break;
    }
  }
}
// vector FillEos (eos_base.hpp:503-531); errs may be null
void sgo_vec_fill(void *ev, double *rho, double *temp, double *sie, double *press, double *cv,
                  double *bmod, long n, unsigned long output, double *lambda,
                  int lambda_stride, int *errs, int nthreads) {
  const EOS &e = *static_cast<EOS *>(ev);
#pragma omp parallel for schedule(static) num_threads(threads_or_default(nthreads))
  for (long i = 0; i < n; ++i) {
    Real *lm = lambda ? lambda + (long)lambda_stride * i : nullptr;
    ErrFlag err;
    e.fill(rho[i], temp[i], sie[i], press[i], cv[i], bmod[i], output, lm, err);
    if (errs) errs[i] = err.code;
  }
}
// DensityEnergyFromPressureTemperature; rho is the in/out initial guess
void sgo_vec_rhoe_PT(void *ev, const double *P, const double *T, double *rho, double *sie,
                     long n, double *lambda, int lambda_stride, int *errs, int nthreads) {
  const EOS &e = *static_cast<EOS *>(ev);
#pragma omp parallel for schedule(static) num_threads(threads_or_default(nthreads))
  for (long i = 0; i < n; ++i) {
    Real *lm = lambda ? lambda + (long)lambda_stride * i : nullptr;
    ErrFlag err;
    e.rhoe_PT(P[i], T[i], lm, rho[i], sie[i], err);
    if (errs) errs[i] = err.code;
  }
}

// ---- Gruneisen-sampled rho-T table (sesame2spiner-style node sampling) -----
// Fields [nRho][nT], T fastest; cold curve = lowest isotherm.
void sgo_build_gruneisen_table(void *grv, double lRhoMin, double lRhoMax, int nRho,
                               double lTMin, double lTMax, int nT, double *P, double *sie,
                               double *bMod, double *dPdRho, double *dPdE, double *dEdRho,
                               double *dEdT, double *PCold, double *sieCold, double *bModCold) {
  const Gruneisen &g = static_cast<EOS *>(grv)->gr;
  const Grid1D gr = Grid1D::make(lRhoMin, lRhoMax, nRho);
  const Grid1D gt = Grid1D::make(lTMin, lTMax, nT);
  for (int j = 0; j < nRho; ++j) {
    const Real rho = from_log(gr.x(j), 0.0);
    for (int i = 0; i < nT; ++i) {
      const Real T = from_log(gt.x(i), 0.0);
      const size_t k = (size_t)j * nT + i;
      const Real e = g.e_rT(rho, T);
      P[k] = g.P_rT(rho, T);
      sie[k] = e;
      bMod[k] = g.bmod_rT(rho, T);
      dPdRho[k] = g.dPres_drho_e(rho, e);
      dPdE[k] = g.Gamma(rho) * std::min(rho, g.rho_max);
      dEdRho[k] = 0.0;
      dEdT[k] = g.Cv;
    }
    PCold[j] = P[(size_t)j * nT];
    sieCold[j] = sie[(size_t)j * nT];
    bModCold[j] = bMod[(size_t)j * nT];
  }
}

// ---- PTE closure over a CSR cell list ------------------------------------
// method 0 = PTESolverRhoT, 1 = PTESolverPT, 2 = PTESolverFixedT,
// 3 = PTESolverFixedP, 4 = PTESolverRhoU. offsets[ncell+1]; mat_eos[nnz]
// indexes eos_handles; per-(cell,mat) arrays are in/out; lambda is nnz*3 or
// null; Tnorm may be null (0 = no guess; methods 0, 1 and 4 only). cell_in[c]
// is the cell's specific internal energy (methods 0, 1, 4), its temperature
// (2) or its pressure (3).
// Returns number of unconverged cells.
long sgo_pte_solve(int method, long ncell, const int *offsets, const int *mat_eos,
                   void **eos_handles, int neos, double *rho, double *vfrac, double *sie,
                   double *temp, double *press, const double *vfrac_tot, const double *cell_in,
                   const double *Tnorm, const double *params_arr, double *lambda,
                   int *out_converged, int *out_niter, double *out_residual, int *out_err,
                   int nthreads) {
  const MixParams params = params_from_array(params_arr);
  std::vector<EOS> pool(neos);
  for (int i = 0; i < neos; ++i) pool[i] = *static_cast<EOS *>(eos_handles[i]);
  long nfail = 0;
#pragma omp parallel num_threads(threads_or_default(nthreads)) reduction(+ : nfail)
  {
    // cells of up to 64 materials (the device's one-wave-per-cell limit)
    constexpr int kMaxMats = 64;
    std::vector<Real> scratch(PTESolverFixedPRequiredScratch(kMaxMats) +
                              PTESolverRhoTRequiredScratch(kMaxMats) +
                              PTESolverRhoURequiredScratch(kMaxMats));
    std::vector<EOS> eos_cell(kMaxMats);
#pragma omp for schedule(dynamic, 256)
    for (long c = 0; c < ncell; ++c) {
      const int off = offsets[c];
      const int nmat = offsets[c + 1] - off;
      if (nmat < 1 || nmat > kMaxMats) {
        if (out_converged) out_converged[c] = 0;
        if (out_niter) out_niter[c] = 0;
        if (out_residual) out_residual[c] = 0;
        if (out_err) out_err[c] = 32;  // malformed cell
        nfail += 1;
        continue;
      }
      for (int m = 0; m < nmat; ++m) eos_cell[m] = pool[mat_eos[off + m]];
      Real *lam = lambda ? lambda + 3L * off : nullptr;
      const Real tn = Tnorm ? Tnorm[c] : 0.0;
      SolverStatus st;
      if (method == 0) {
        PTERhoT s(nmat, eos_cell.data(), vfrac_tot[c], cell_in[c], rho + off, vfrac + off, sie + off,
                  temp + off, press + off, lam, scratch.data(), tn, params);
        st = PTESolver(s);
      } else if (method == 1) {
        PTEPT s(nmat, eos_cell.data(), vfrac_tot[c], cell_in[c], rho + off, vfrac + off, sie + off,
                temp + off, press + off, lam, scratch.data(), tn, params);
        st = PTESolver(s);
      } else if (method == 4) {
        PTERhoU s(nmat, eos_cell.data(), vfrac_tot[c], cell_in[c], rho + off, vfrac + off, sie + off,
                  temp + off, press + off, lam, scratch.data(), tn, params);
        st = PTESolver(s);
      } else if (method == 2) {
        PTEFixedT s(nmat, eos_cell.data(), vfrac_tot[c], cell_in[c], rho + off, vfrac + off, sie + off,
                    temp + off, press + off, lam, scratch.data(), params);
        st = PTESolver(s);
      } else {
        PTEFixedP s(nmat, eos_cell.data(), vfrac_tot[c], cell_in[c], rho + off, vfrac + off, sie + off,
                    temp + off, press + off, lam, scratch.data(), params);
        st = PTESolver(s);
      }
      if (out_converged) out_converged[c] = st.converged ? 1 : 0;
      if (out_niter) out_niter[c] = (int)st.max_niter;
      if (out_residual) out_residual[c] = st.residual;
      if (out_err) out_err[c] = st.err;
      nfail += st.converged ? 0 : 1;
    }
  }
  return nfail;
}

// ---- get_sg_eos driver --------------------------------------------------
// eos_handles[neos] is the material pool; eos_offsets[nmat] are 1-based pool
// indices; offsets[ncell] are 1-based cell indices; per-material arrays are
// column-major [cell_dim x nmat]. out_err / out_conv: per loop index, may be
// null. Returns the number of unconverged PTE solves.
long sgo_get_sg_eos(int nmat, int ncell, int cell_dim, int input, const int *eos_offsets,
                    void **eos_handles, int neos, const int *offsets, double *press,
                    double *pmax, double *vol, double *spvol, double *sie, double *temp,
                    double *bmod, double *dpde, double *cv, double *frac_mass,
                    double *frac_vol, double *frac_ie, double *frac_bmod, double *frac_dpde,
                    double *frac_cv, double mass_frac_cutoff, const double *params_arr,
                    int *out_err, int *out_conv, int nthreads) {
  std::vector<EOS> pool(neos);
  for (int i = 0; i < neos; ++i) pool[i] = *static_cast<EOS *>(eos_handles[i]);
  SgEosArgs a;
  a.nmat = nmat;
  a.ncell = ncell;
  a.cell_dim = cell_dim;
  a.input = input;
  a.eos_offsets = eos_offsets;
  a.pool = pool.data();
  a.offsets = offsets;
  a.press = press;
  a.pmax = pmax;
  a.vol = vol;
  a.spvol = spvol;
  a.sie = sie;
  a.temp = temp;
  a.bmod = bmod;
  a.dpde = dpde;
  a.cv = cv;
  a.frac_mass = frac_mass;
  a.frac_vol = frac_vol;
  a.frac_ie = frac_ie;
  a.frac_bmod = frac_bmod;
  a.frac_dpde = frac_dpde;
  a.frac_cv = frac_cv;
  a.mass_frac_cutoff = mass_frac_cutoff;
  a.params = params_from_array(params_arr);
  long nfail = 0;
#pragma omp parallel num_threads(threads_or_default(nthreads)) reduction(+ : nfail)
  {
    SgEosScratch w(nmat);
#pragma omp for schedule(dynamic, 64)
    for (int c = 0; c < ncell; ++c) {
      bool conv = true;
      const int e = sg_eos_cell(a, c, w, &conv);
      if (out_err) out_err[c] = e;
      if (out_conv) out_conv[c] = conv ? 1 : 0;
      nfail += conv ? 0 : 1;
    }
  }
  return nfail;
}

int sgo_max_threads() { return omp_get_max_threads(); }
} // extern "C"
