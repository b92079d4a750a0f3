// C ABI: context and material registry, vector EOS calls, plain PTE calls,
// MixParams defaults and device info. Argument checks happen here, on the
// host, before anything is launched.
#include "sgm_internal.hpp"
#include "sgm_approx.hpp"

#include <cmath>
#include <cstring>

using namespace sgm;

static_assert(sgm::ERRB_INPUT == SGM_ERRB_INPUT, "status bit of a malformed cell");

namespace sgm {
bool build_table_host(const sgm_table_desc &in, TableHost &T);
hipError_t upload_table(TableHost &T);
} // namespace sgm

namespace sgm {
int sync_registry(sgm_ctx *ctx) {
  if (ctx->host_only) return SGM_OK;
  if (!ctx->dirty) return SGM_OK;
  if (hipMemcpy(ctx->d_eos, ctx->host_eos.data(), sizeof(EosDev) * ctx->nmat,
                hipMemcpyHostToDevice) != hipSuccess)
    return SGM_E_HIP;
  ctx->dirty = false;
  return SGM_OK;
}

MixParamsDev to_dev(const sgm_mix_params *p) {
  sgm_mix_params q;
  if (p)
    q = *p;
  else
    sgm_mix_params_default(&q);
  MixParamsDev d;
  static_assert(sizeof(MixParamsDev) == sizeof(sgm_mix_params), "MixParams layout");
  std::memcpy(&d, &q, sizeof(d));
  return d;
}
} // namespace sgm

namespace {

int check_mat(const sgm_ctx *ctx, int m) {
  if (!ctx) return SGM_E_ARG;
  if (m < 0 || m >= ctx->nmat || !ctx->is_set[m]) return SGM_E_MAT;
  return SGM_OK;
}

// every entry point that launches work starts here
int check_launch(const sgm_ctx *ctx, int m) {
  const int r = check_mat(ctx, m);
  if (r != SGM_OK) return r;
  return ctx->host_only ? SGM_E_NOGPU : SGM_OK;
}

int set_common(sgm_ctx *ctx, int m, const EosDev &e) {
  if (!ctx) return SGM_E_ARG;
  if (m < 0 || m >= ctx->nmat) return SGM_E_MAT;
  std::lock_guard<std::mutex> g(ctx->mu);
  ctx->host_eos[m] = e;
  ctx->host_eval[m] = e;
  ctx->is_set[m] = 1;
  ctx->dirty = true;
  return sync_registry(ctx);
}

// singularity-eos/eos/eos_gruneisen.hpp:231-329 (ComputeRhoMax), host-side init
Real compute_rho_max(Real s1, Real s2, Real s3, Real rho0) {
  auto poly = [=](Real eta) { return 1 - s1 * eta - s2 * (eta * eta) - s3 * ((eta * eta) * eta); };
  const Real EPS = robust::EPS();
  auto near0 = [&](Real v) { return std::fabs(v) < EPS; };
  Real root = -1.;
  if (near0(s2) && near0(s3) && s1 > 1.) {
    root = 1. / s1;
  } else if (near0(s3) && !near0(s2)) {
    const Real disc = s1 * s1 + 4. * s2;
    if (disc >= 0.) {
      const Real r1 = (s1 + std::sqrt(disc)) / (-2. * s2);
      const Real r2 = (s1 - std::sqrt(disc)) / (-2. * s2);
      const Real mn = smin(r1, r2), mx = smax(r1, r2);
      if (mn > 0. && mn < 1.)
        root = mn;
      else if (mx > 0. && mx < 1.)
        root = mx;
      else
        root = -INFINITY;
    }
  } else if (!near0(s3)) {
    const Real m2 = -s2;
    const Real disc = -18. * s3 * s2 * s1 - 4. * ((m2 * m2) * m2) + (s2 * s2) * (s1 * s1) -
                      4. * s3 * ((s1 * s1) * s1) - 27. * (s3 * s3);
    Real lo = 0., hi = 1.;
    if (disc >= 0) {
      const Real e1 = (s2 + std::sqrt(s2 * s2 - 3. * s3 * s1)) / (-3. * s3);
      const Real e2 = (s2 - std::sqrt(s2 * s2 - 3. * s3 * s1)) / (-3. * s3);
      const Real mne = smin(e1, e2), mxe = smax(e1, e2);
      if (s3 < 0.) {
        lo = smax(mne, lo);
        hi = smin(smax(mxe, lo), hi);
      } else if (mne > 0) {
        hi = smin(mne, hi);
      }
    }
    if (poly(lo) * poly(hi) < 0.) {
      root::regula_falsi(poly, 0., 0.001, lo, hi, 100 * EPS, 100. / 10 * EPS, root);
    } else if (near0(poly(lo))) {
      root = lo;
    } else if (near0(poly(hi))) {
      root = hi;
    } else {
      root = -1.;
    }
  }
  if (root > 0. && root < 1.) return rho0 / ((1 - root) + EPS);
  return INFINITY;
}

} // namespace

namespace sgm {
Real gruneisen_compute_rho_max(Real s1, Real s2, Real s3, Real rho0) {
  return compute_rho_max(s1, s2, s3, rho0);
}
} // namespace sgm

extern "C" {

const char *sgm_version(void) { return "sgm 0.1 (gfx950)"; }

void sgm_mix_params_default(sgm_mix_params *p) {
  if (!p) return;
  std::memset(p, 0, sizeof(*p));
  p->derivative_eps = 3.0e-6;
  p->pte_rel_tolerance_p = 1.0e-6;
  p->pte_rel_tolerance_t = 1.0e-4;
  p->pte_rel_tolerance_e = 1.0e-6;
  p->pte_rel_tolerance_v = 1.0e-6;
  p->pte_abs_tolerance_p = 1.0e-6;
  p->pte_abs_tolerance_t = 1.0e-4;
  p->pte_slow_convergence_thresh = 0.99;
  p->pte_rel_tolerance_p_sufficient = 1.e2 * 1.0e-6;
  p->pte_rel_tolerance_t_sufficient = 10 * 1.0e-4;
  p->pte_rel_tolerance_e_sufficient = 1.e2 * 1.0e-6;
  p->pte_rel_tolerance_v_sufficient = 1.e2 * 1.0e-6;
  p->pte_abs_tolerance_p_sufficient = 1.e2 * 1.0e-6;
  p->pte_abs_tolerance_t_sufficient = 10 * 1.0e-4;
  p->line_search_alpha = 1.e-2;
  p->line_search_fac = 0.5;
  p->vfrac_safety_fac = 0.95;
  p->temperature_limit = 1.0e15;
  p->default_tguess = 300.;
  p->min_dtde = 1.0e-16;
  p->pte_small_step_thresh = 1e-16;
  p->pte_max_dpdv = -1e-8;
  p->pte_max_iter_per_mat = 128;
  p->line_search_max_iter = 6;
  p->pte_small_step_tries = 2;
  p->pte_reduced_system_exclude_idx = -1;
  p->iterate_t_guess = 1;
}

void sgm_mix_params_tight(sgm_mix_params *p) {
  sgm_mix_params_default(p);
  if (!p) return;
  p->pte_rel_tolerance_p = 1e-10;
  p->pte_rel_tolerance_e = 1e-10;
  p->pte_rel_tolerance_t = 1e-10;
  p->pte_rel_tolerance_p_sufficient = 1e-10;
  p->pte_rel_tolerance_e_sufficient = 1e-10;
  p->pte_rel_tolerance_t_sufficient = 1e-10;
  p->pte_abs_tolerance_p = 0;
  p->pte_abs_tolerance_t = 0;
  p->pte_abs_tolerance_p_sufficient = 0;
  p->pte_abs_tolerance_t_sufficient = 0;
  p->pte_reduced_system_exclude_idx = -1;
}

int sgm_device_info(int *ncu, int *arch_gfx, double *hbm_gbs) {
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || n < 1) return SGM_E_NOGPU;
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, 0) != hipSuccess) return SGM_E_HIP;
  if (ncu) *ncu = prop.multiProcessorCount;
  if (arch_gfx) {
    int g = 0;
    if (std::strncmp(prop.gcnArchName, "gfx", 3) == 0) g = (int)std::strtol(prop.gcnArchName + 3, nullptr, 16);
    *arch_gfx = g;
  }
  if (hbm_gbs) *hbm_gbs = 2.0 * prop.memoryClockRate * 1e3 * (prop.memoryBusWidth / 8) / 1e9;
  return SGM_OK;
}

int sgm_init(int nmat, sgm_ctx **out) {
  if (!out || nmat < 1 || nmat > MAX_REGISTRY) return SGM_E_ARG;
  *out = nullptr;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return SGM_E_NOGPU;
  auto *ctx = new sgm_ctx();
  hipGetDevice(&ctx->device);
  ctx->nmat = nmat;
  ctx->host_eos.assign(nmat, EosDev{});
  ctx->host_eval.assign(nmat, EosDev{});
  ctx->is_set.assign(nmat, 0);
  ctx->tables.resize(nmat);
  ctx->vinets.resize(nmat);
  if (hipMalloc(&ctx->d_eos, sizeof(EosDev) * nmat) != hipSuccess) {
    delete ctx;
    return SGM_E_HIP;
  }
  ctx->dirty = true;
  sync_registry(ctx);
  *out = ctx;
  return SGM_OK;
}

int sgm_init_host(int nmat, sgm_ctx **out) {
  if (!out || nmat < 1 || nmat > MAX_REGISTRY) return SGM_E_ARG;
  auto *ctx = new sgm_ctx();
  ctx->host_only = true;
  ctx->device = -1;
  ctx->nmat = nmat;
  ctx->host_eos.assign(nmat, EosDev{});
  ctx->host_eval.assign(nmat, EosDev{});
  ctx->is_set.assign(nmat, 0);
  ctx->tables.resize(nmat);
  ctx->vinets.resize(nmat);
  *out = ctx;
  return SGM_OK;
}

int sgm_finalize(sgm_ctx *ctx) {
  if (!ctx) return SGM_E_ARG;
  if (ctx->host_only) {
    delete ctx;
    return SGM_OK;
  }
  hipDeviceSynchronize();
  sgm_release_driver_cache(ctx);
  for (auto &t : ctx->tables)
    if (t)
      for (void *p : t->d_arrays) hipFree(p);
  for (auto &v : ctx->vinets)
    if (v && v->d_coef) hipFree(v->d_coef);
  if (ctx->d_eos) hipFree(ctx->d_eos);
  delete ctx;
  return SGM_OK;
}

int sgm_release_driver_cache(sgm_ctx *ctx) {
  if (!ctx) return SGM_E_ARG;
  std::lock_guard<std::mutex> g(ctx->drv_mu);
  if (ctx->drv_plan) sgm_plan_destroy(ctx->drv_plan);
  if (ctx->drv_cells) (void)hipFree(ctx->drv_cells);
  if (ctx->drv_entries) (void)hipFree(ctx->drv_entries);
  ctx->drv_plan = nullptr;
  ctx->drv_cells = ctx->drv_entries = nullptr;
  ctx->drv_cells_bytes = ctx->drv_entries_bytes = 0;
  return SGM_OK;
}

int sgm_driver_cache_info(sgm_ctx *ctx, long *out) {
  if (!ctx || !out) return SGM_E_ARG;
  std::lock_guard<std::mutex> g(ctx->drv_mu);
  out[0] = ctx->drv_plan_builds;
  out[1] = ctx->drv_plan_reuses;
  out[2] = (long)(ctx->drv_cells_bytes + ctx->drv_entries_bytes);
  out[3] = ctx->drv_plan ? 1 : 0;
  return SGM_OK;
}

int sgm_set_modifiers(sgm_ctx *ctx, int m, const int *enabled, const double *vals) {
  const int r = check_mat(ctx, m);
  if (r != SGM_OK) return r;
  if (!enabled || !vals) return SGM_E_ARG;
  if (enabled[2] == 1 || enabled[3] == 1) return SGM_E_KIND;  // ramps: not on this path
  const bool scaled = enabled[0] == 1, shifted = enabled[1] == 1;
  // the reference's checks (scaled_eos.hpp:55-62, shifted_eos.hpp:56-59)
  if (scaled && (!(std::fabs(vals[0]) > 0) || std::isnan(vals[0]) ||
                 !(std::fabs(1. / vals[0]) > 0)))
    return SGM_E_ARG;
  if (shifted && std::isnan(vals[1])) return SGM_E_ARG;
  std::lock_guard<std::mutex> g(ctx->mu);
  for (EosDev *e : {&ctx->host_eos[m], &ctx->host_eval[m]}) {
    e->mods = (shifted ? MOD_SHIFTED : 0) | (scaled ? MOD_SCALED : 0);
    e->shift = shifted ? vals[1] : 0.0;
    e->scale = scaled ? vals[0] : 1.0;
    e->inv_scale = scaled ? 1. / vals[0] : 1.0;
  }
  ctx->dirty = true;
  return sync_registry(ctx);
}

int sgm_get_kind(const sgm_ctx *ctx, int m) {
  const int r = check_mat(ctx, m);
  if (r != SGM_OK) return r;
  return ctx->host_eos[m].kind;
}

int sgm_set_IdealGas(sgm_ctx *ctx, int m, double gm1, double Cv) {
  if (!(Cv >= 0) || !(gm1 >= 0)) return SGM_E_ARG; // eos_ideal.hpp:66-71
  EosDev e{};
  e.kind = KIND_IDEAL;
  e.m.p[0] = Cv;
  e.m.p[1] = gm1;
  return set_common(ctx, m, e);
}

int sgm_set_StiffGas(sgm_ctx *ctx, int m, double gm1, double Cv, double Pinf, double qq) {
  if (!(Cv >= 0) || !(gm1 >= 0)) return SGM_E_ARG; // eos_stiff.hpp:69-72
  EosDev e{};
  e.kind = KIND_STIFF;
  e.m.p[0] = Cv;
  e.m.p[1] = gm1;
  e.m.p[2] = Pinf;
  e.m.p[3] = qq;
  return set_common(ctx, m, e);
}

int sgm_set_Gruneisen(sgm_ctx *ctx, int m, double C0, double s1, double s2, double s3, double G0,
                      double b, double rho0, double T0, double P0, double Cv, double rho_max) {
  EosDev e{};
  e.kind = KIND_GRUN;
  const double p[11] = {C0, s1, s2, s3, G0, b, rho0, T0, P0, Cv,
                        rho_max > 0 ? rho_max : 0.99 * compute_rho_max(s1, s2, s3, rho0)};
  for (int i = 0; i < 11; ++i) e.m.p[i] = p[i];
  return set_common(ctx, m, e);
}

int sgm_set_Vinet(sgm_ctx *ctx, int m, double rho0, double T0, double B0, double BP0, double A0,
                  double Cv0, double E0, double S0, const double *d2to40) {
  if (!ctx || !d2to40) return SGM_E_ARG;
  if (m < 0 || m >= ctx->nmat) return SGM_E_MAT;
  // eos_vinet.hpp:172-196
  if (rho0 < 0 || T0 < 0 || B0 < 0 || BP0 < 1.0 || A0 < 0 || Cv0 < 0) return SGM_E_ARG;
  auto v = std::make_unique<VinetHost>();
  v->coef.assign(39 + 43, 0.0);
  Real *d = v->coef.data();
  Real *VIP = d + 39;
  for (int i = 0; i < 39; ++i) d[i] = d2to40[i];
  // InitializeVinet (eos_vinet.hpp:199-226)
  VIP[0] = robust::ratio(B0, rho0) + robust::ratio(A0 * A0 * B0 * B0 * T0, rho0 * rho0 * Cv0);
  VIP[1] = 3.0 / 2.0 * (BP0 - 1.0);
  VIP[2] = 1.0;
  VIP[3] = 0.0;
  for (int i = 4; i < 43; ++i) VIP[i] = d[i - 4];
  for (int i = 43 - 2; i >= 2; i--) VIP[i] = VIP[i] - (i) / VIP[1] * VIP[i + 1];
  // leading Horner steps (ind = 40, 39, ...) whose coefficients are all zero
  int nskip = 0;
  for (int ind = 40; ind >= 2; --ind) {
    if (d[ind - 2] == 0.0 && VIP[2 + ind] == 0.0)
      ++nskip;
    else
      break;
  }
  if (!ctx->host_only) {
    if (hipMalloc(&v->d_coef, v->coef.size() * sizeof(Real)) != hipSuccess) return SGM_E_HIP;
    if (hipMemcpy(v->d_coef, v->coef.data(), v->coef.size() * sizeof(Real),
                  hipMemcpyHostToDevice) != hipSuccess)
      return SGM_E_HIP;
  }
  EosDev e{};
  e.kind = KIND_VINET;
  const double p[12] = {rho0, T0, B0, BP0, A0, Cv0, E0, S0, VIP[1],
                        pmath::cbrt(10 * robust::EPS()) * rho0, VIP[2], VIP[3]};
  for (int i = 0; i < 12; ++i) e.m.p[i] = p[i];
  e.nvinet_skip = nskip;
  e.m.coef = v->d_coef;
  EosDev eh = e;
  eh.m.coef = v->coef.data();
  {
    std::lock_guard<std::mutex> g(ctx->mu);
    if (ctx->vinets[m] && ctx->vinets[m]->d_coef) hipFree(ctx->vinets[m]->d_coef);
    ctx->vinets[m] = std::move(v);
    ctx->host_eos[m] = e;
    ctx->host_eval[m] = eh;
    ctx->is_set[m] = 1;
    ctx->dirty = true;
    return sync_registry(ctx);
  }
}

int sgm_set_TableRhoT(sgm_ctx *ctx, int m, const sgm_table_desc *desc) {
  if (!ctx || !desc) return SGM_E_ARG;
  if (m < 0 || m >= ctx->nmat) return SGM_E_MAT;
  auto T = std::make_unique<TableHost>();
  if (!build_table_host(*desc, *T)) return SGM_E_ARG;
  if (!ctx->host_only && upload_table(*T) != hipSuccess) return SGM_E_HIP;
  EosDev e{};
  e.kind = KIND_TABLE;
  e.tab = T->ddev;
  EosDev eh = e;
  eh.tab = T->hdev;
  std::lock_guard<std::mutex> g(ctx->mu);
  if (ctx->tables[m])
    for (void *p : ctx->tables[m]->d_arrays) hipFree(p);
  ctx->tables[m] = std::move(T);
  ctx->host_eos[m] = e;
  ctx->host_eval[m] = eh;
  ctx->is_set[m] = 1;
  ctx->dirty = true;
  return sync_registry(ctx);
}

long sgm_table_query(const sgm_ctx *ctx, int m, int which, double *out) {
  if (check_mat(ctx, m) != SGM_OK) return SGM_E_MAT;
  if (ctx->host_eos[m].kind != KIND_TABLE || !ctx->tables[m]) return SGM_E_KIND;
  const TableHost &T = *ctx->tables[m];
  const std::vector<Real> *v = nullptr;
  switch (which) {
  case 0: v = &T.PMax; break;
  case 1: v = &T.sielTMax; break;
  case 2: v = &T.dEdTMax; break;
  case 3: v = &T.gm1Max; break;
  case 4: v = &T.rho_at_pmin; break;
  case 5: v = &T.lTColdCrit; break;
  case 6: v = &T.bModCold; break;
  case 7: v = &T.dPdECold; break;
  case 8: v = &T.dEdTCold; break;
  case 9: v = &T.bMod; break;
  case 10:
    if (out) {
      out[0] = T.hdev.PMin;
      out[1] = T.hdev.rhoMax;
      out[2] = T.hdev.TMax;
    }
    return 3;
  default: return SGM_E_ARG;
  }
  if (out) std::memcpy(out, v->data(), v->size() * sizeof(Real));
  return (long)v->size();
}

int sgm_vec_eval(sgm_ctx *ctx, int m, int func, const double *in1, const double *in2,
                 double *out, long n, int lstride, double *lambda, sgm_stream_t stream) {
  const int r = check_launch(ctx, m);
  if (r != SGM_OK) return r;
  if (n < 0 || func < 0 || func > SGM_F_MIN_SIE_RHO) return SGM_E_ARG;
  if (n > 0 && (!in1 || !out || (func != SGM_F_MIN_SIE_RHO && !in2))) return SGM_E_ARG;
  if (lambda && lstride < 2) return SGM_E_ARG;
  if (sync_registry(ctx) != SGM_OK) return SGM_E_HIP;
  return launch_vec(ctx->host_eos[m], ctx->d_eos, m, func, in1, in2, out, n, lstride, lambda,
                    (hipStream_t)stream) == hipSuccess
             ? SGM_OK
             : SGM_E_HIP;
}

int sgm_vec_P_from_rho_sie(sgm_ctx *ctx, int m, const double *rho, const double *sie, double *P,
                           long n, int lstride, double *lambda, sgm_stream_t stream) {
  return sgm_vec_eval(ctx, m, SGM_F_P_RHO_SIE, rho, sie, P, n, lstride, lambda, stream);
}
int sgm_vec_T_from_rho_sie(sgm_ctx *ctx, int m, const double *rho, const double *sie, double *T,
                           long n, int lstride, double *lambda, sgm_stream_t stream) {
  return sgm_vec_eval(ctx, m, SGM_F_T_RHO_SIE, rho, sie, T, n, lstride, lambda, stream);
}

int sgm_vec_fill_eos(sgm_ctx *ctx, int m, double *rho, double *temp, double *sie, double *press,
                     double *cv, double *bmod, long n, unsigned long output, int lstride,
                     double *lambda, int *errs, sgm_stream_t stream) {
  const int r = check_launch(ctx, m);
  if (r != SGM_OK) return r;
  if (n < 0 || (lambda && lstride < 2)) return SGM_E_ARG;
  // density is always read or written; every requested output needs its
  // array (the kernel reads only non-null inputs and writes only non-null outputs)
  if (n > 0) {
    if (!rho) return SGM_E_ARG;
    if ((output & SGM_Q_TEMPERATURE) && !temp) return SGM_E_ARG;
    if ((output & SGM_Q_SIE) && !sie) return SGM_E_ARG;
    if ((output & SGM_Q_PRESSURE) && !press) return SGM_E_ARG;
    if ((output & SGM_Q_CV) && !cv) return SGM_E_ARG;
    if ((output & SGM_Q_BMOD) && !bmod) return SGM_E_ARG;
  }
  if (sync_registry(ctx) != SGM_OK) return SGM_E_HIP;
  return launch_fill(ctx->host_eos[m], ctx->d_eos, m, rho, temp, sie, press, cv, bmod, n, output,
                     lstride, lambda, errs, (hipStream_t)stream) == hipSuccess
             ? SGM_OK
             : SGM_E_HIP;
}

int sgm_vec_rhoe_from_PT(sgm_ctx *ctx, int m, const double *P, const double *T, double *rho,
                         double *sie, long n, int lstride, double *lambda, int *errs,
                         sgm_stream_t stream) {
  const int r = check_launch(ctx, m);
  if (r != SGM_OK) return r;
  if (n < 0 || (n > 0 && (!P || !T || !rho || !sie)) || (lambda && lstride < 2)) return SGM_E_ARG;
  if (sync_registry(ctx) != SGM_OK) return SGM_E_HIP;
  return launch_rhoe_PT(ctx->host_eos[m], ctx->d_eos, m, P, T, rho, sie, n, lstride, lambda,
                        errs, (hipStream_t)stream) == hipSuccess
             ? SGM_OK
             : SGM_E_HIP;
}

static int pte_common(int method, sgm_ctx *ctx, long ncell, long nnz, const int *offsets,
                      const uint8_t *mat_ids, double *rho, double *vfrac, double *sie,
                      double *temp, double *press, const double *vfrac_tot,
                      const double *sie_tot, const double *Tnorm, double *lambda,
                      const sgm_mix_params *params, sgm_pte_status st, sgm_stream_t stream) {
  if (!ctx || ncell < 0 || nnz < 0) return SGM_E_ARG;
  if (ctx->host_only) return SGM_E_NOGPU;
  if (ncell == 0) return SGM_OK;
  if (!offsets || !mat_ids || !rho || !vfrac || !sie || !temp || !press || !vfrac_tot ||
      !sie_tot)
    return SGM_E_ARG;
  if (nnz > 2147483647L) return SGM_E_ARG;
  // cells naming an index >= nmat are rejected on the device; unset slots hold
  // a zeroed descriptor (no pointers), so naming one cannot fault
  if (sync_registry(ctx) != SGM_OK) return SGM_E_HIP;
  PteArgs a;
  a.d_eos = ctx->d_eos;
  a.nreg = ctx->nmat;
  a.ncell = ncell;
  a.nnz = nnz;
  a.offsets = offsets;
  a.mat_ids = mat_ids;
  a.rho = rho;
  a.vfrac = vfrac;
  a.sie = sie;
  a.temp = temp;
  a.press = press;
  a.vfrac_tot = vfrac_tot;
  a.sie_tot = sie_tot;
  a.Tnorm = Tnorm;
  a.lambda = lambda;
  a.conv = st.converged;
  a.niter = st.niter;
  a.err = st.err;
  a.resid = st.residual;
  const MixParamsDev P = to_dev(params);
  const hipError_t e = launch_pte_plain(method, a, P, (hipStream_t)stream);
  return e == hipSuccess ? SGM_OK : SGM_E_HIP;
}

int sgm_pte_rhoT(sgm_ctx *ctx, long ncell, long nnz, const int *offsets, const uint8_t *mat_ids,
                 double *rho, double *vfrac, double *sie, double *temp, double *press,
                 const double *vfrac_tot, const double *sie_tot, const double *Tnorm,
                 double *lambda, const sgm_mix_params *params, sgm_pte_status status,
                 sgm_stream_t stream) {
  return pte_common(0, ctx, ncell, nnz, offsets, mat_ids, rho, vfrac, sie, temp, press,
                    vfrac_tot, sie_tot, Tnorm, lambda, params, status, stream);
}

int sgm_pte_PT(sgm_ctx *ctx, long ncell, long nnz, const int *offsets, const uint8_t *mat_ids,
               double *rho, double *vfrac, double *sie, double *temp, double *press,
               const double *vfrac_tot, const double *sie_tot, const double *Tnorm,
               double *lambda, const sgm_mix_params *params, sgm_pte_status status,
               sgm_stream_t stream) {
  return pte_common(1, ctx, ncell, nnz, offsets, mat_ids, rho, vfrac, sie, temp, press,
                    vfrac_tot, sie_tot, Tnorm, lambda, params, status, stream);
}

int sgm_pte_fixedT(sgm_ctx *ctx, long ncell, long nnz, const int *offsets,
                   const uint8_t *mat_ids, double *rho, double *vfrac, double *sie, double *temp,
                   double *press, const double *vfrac_tot, const double *T_true, double *lambda,
                   const sgm_mix_params *params, sgm_pte_status status, sgm_stream_t stream) {
  return pte_common(PTE_FIXED_T, ctx, ncell, nnz, offsets, mat_ids, rho, vfrac, sie, temp,
                    press, vfrac_tot, T_true, nullptr, lambda, params, status, stream);
}

int sgm_pte_fixedP(sgm_ctx *ctx, long ncell, long nnz, const int *offsets,
                   const uint8_t *mat_ids, double *rho, double *vfrac, double *sie, double *temp,
                   double *press, const double *vfrac_tot, const double *P, double *lambda,
                   const sgm_mix_params *params, sgm_pte_status status, sgm_stream_t stream) {
  return pte_common(PTE_FIXED_P, ctx, ncell, nnz, offsets, mat_ids, rho, vfrac, sie, temp,
                    press, vfrac_tot, P, nullptr, lambda, params, status, stream);
}

int sgm_pte_rhoU(sgm_ctx *ctx, long ncell, long nnz, const int *offsets, const uint8_t *mat_ids,
                 double *rho, double *vfrac, double *sie, double *temp, double *press,
                 const double *vfrac_tot, const double *sie_tot, const double *Tnorm,
                 double *lambda, const sgm_mix_params *params, sgm_pte_status status,
                 sgm_stream_t stream) {
  return pte_common(PTE_RHOU, ctx, ncell, nnz, offsets, mat_ids, rho, vfrac, sie, temp, press,
                    vfrac_tot, sie_tot, Tnorm, lambda, params, status, stream);
}

int sgm_device_registry(sgm_ctx *ctx, const void **d_registry, int *n) {
  if (!ctx || !d_registry || !n) return SGM_E_ARG;
  if (ctx->host_only) return SGM_E_NOGPU;
  std::lock_guard<std::mutex> g(ctx->mu);
  if (sync_registry(ctx) != SGM_OK) return SGM_E_HIP;
  *d_registry = ctx->d_eos;
  *n = ctx->nmat;
  return SGM_OK;
}

int sgm_approx_temperature(sgm_ctx *ctx, long ncell, long nnz, const int *offsets,
                           const uint8_t *mat_ids, const double *rho, const double *vfrac,
                           const double *u_tot, const double *Tguess, double *lambda,
                           double *T_out, sgm_stream_t stream) {
  if (!ctx || ncell < 0 || nnz < 0) return SGM_E_ARG;
  if (ctx->host_only) return SGM_E_NOGPU;
  if (ncell == 0) return SGM_OK;
  if (!offsets || !mat_ids || !rho || !vfrac || !u_tot || !T_out) return SGM_E_ARG;
  if (nnz > 2147483647L) return SGM_E_ARG;
  std::lock_guard<std::mutex> g(ctx->mu);
  if (sync_registry(ctx) != SGM_OK) return SGM_E_HIP;
  return launch_approx_temperature(ctx->d_eos, ctx->nmat, ncell, nnz, offsets, mat_ids, rho,
                                   vfrac, u_tot, Tguess, lambda, T_out,
                                   (hipStream_t)stream) == hipSuccess
             ? SGM_OK
             : SGM_E_HIP;
}

} // extern "C"
