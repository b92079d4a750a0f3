// Synthetic inputs for the closure and table workloads (host memory).
//
// Random streams: std::mt19937_64 (sequence fixed by the C++ standard).
//   U(a,b)       = a + (b-a) * u,  u = (x >> 11) * 2^-53
//   N(0,1)       = sqrt(-2 ln(1-u1)) * cos(2 pi u2)         (Box-Muller, one value per pair)
//   LogNormal    = exp(mu + sigma * N(0,1))
//   Exp(1)       = -ln(1-u)
//   discrete pmf = inverse CDF on one u
// Draws happen per cell in the order documented at each generator.
#include "sgm_internal.hpp"

#include <cmath>
#include <random>

using namespace sgm;

namespace {
struct Rng {
  std::mt19937_64 g;
  explicit Rng(uint64_t seed) : g(seed) {}
  double u() { return (double)(g() >> 11) * (1.0 / 9007199254740992.0); }
  double uni(double a, double b) { return a + (b - a) * u(); }
  double normal() {
    const double u1 = u(), u2 = u();
    return std::sqrt(-2.0 * std::log(1.0 - u1)) * std::cos(6.283185307179586 * u2);
  }
  double lognormal(double mu, double sigma) { return std::exp(mu + sigma * normal()); }
  double expo() { return -std::log(1.0 - u()); }
};
} // namespace

extern "C" {

int sgm_build_gruneisen_table(double C0, double s1, double s2, double s3, double G0, double b,
                              double rho0, double T0, double P0, double Cv, double lRhoMin,
                              double lRhoMax, int nRho, double lTMin, double lTMax, int nT,
                              double *P, double *sie, double *bMod, double *dPdRho, double *dPdE,
                              double *dEdRho, double *dEdT, double *PCold, double *sieCold,
                              double *bModCold) {
  if (nRho < 2 || nT < 2 || !P || !sie || !bMod || !dPdRho || !dPdE || !dEdRho || !dEdT ||
      !PCold || !sieCold || !bModCold)
    return SGM_E_ARG;
  // Gruneisen parameters with the automatic rho_max (0.99 * ComputeRhoMax)
  const Real p[11] = {C0, s1, s2, s3, G0, b, rho0, T0, P0, Cv,
                      0.99 * gruneisen_compute_rho_max(s1, s2, s3, rho0)};
  const Real dlR = (lRhoMax - lRhoMin) / ((Real)(nRho - 1));
  const Real dlT = (lTMax - lTMin) / ((Real)(nT - 1));
  for (int j = 0; j < nRho; ++j) {
    const Real rho = from_log(lRhoMin + j * dlR, 0.0);
    for (int i = 0; i < nT; ++i) {
      const Real T = from_log(lTMin + i * dlT, 0.0);
      const size_t k = (size_t)j * nT + i;
      const Real e = grun::e_rT(p, rho, T);
      P[k] = grun::P_rT(p, rho, T);
      sie[k] = e;
      bMod[k] = grun::bmod_rT(p, rho, T);
      dPdRho[k] = grun::dPres_drho_e(p, rho, e);
      dPdE[k] = grun::Gamma(p, rho) * smin(rho, p[10]);
      dEdRho[k] = 0.0;
      dEdT[k] = Cv;
    }
    PCold[j] = P[(size_t)j * nT];
    sieCold[j] = sie[(size_t)j * nT];
    bModCold[j] = bMod[(size_t)j * nT];
  }
  return SGM_OK;
}

// Two-material closure: per cell draw alpha_A ~ U(0.05,0.95), rho_A ~
// LogNormal(ln 1e-3, 0.3), rho_B ~ U(0.9,1.1), T_A ~ U(300,3000), T_B ~
// U(300,3000); A = IdealGas(0.4, 7.176e6), B = StiffGas(3.4, 3.6e7, 6e9, 0);
// sie and press from each EOS at its own (rho, T).
int sgm_workload_pte2(uint64_t seed, long n, double *rho, double *vfrac, double *sie,
                      double *temp, double *press, double *sie_tot) {
  if (n < 0 || !rho || !vfrac || !sie || !temp || !press || !sie_tot) return SGM_E_ARG;
  const Real pa[2] = {7.176e6, 0.4};
  const Real pb[4] = {3.6e7, 3.4, 6e9, 0.0};
  Rng r(seed);
  for (long c = 0; c < n; ++c) {
    const double aA = r.uni(0.05, 0.95);
    const double rA = r.lognormal(std::log(1e-3), 0.3);
    const double rB = r.uni(0.9, 1.1);
    const double TA = r.uni(300.0, 3000.0);
    const double TB = r.uni(300.0, 3000.0);
    const long k = 2 * c;
    rho[k] = rA;
    rho[k + 1] = rB;
    vfrac[k] = aA;
    vfrac[k + 1] = 1.0 - aA;
    temp[k] = TA;
    temp[k + 1] = TB;
    sie[k] = ideal::e_rT(pa, rA, TA);
    sie[k + 1] = stiff::e_rT(pb, rB, TB);
    press[k] = ideal::P_rT(pa, rA, TA);
    press[k + 1] = stiff::P_rT(pb, rB, TB);
    const double rbA = rA * aA, rbB = rB * (1.0 - aA);
    sie_tot[c] = (rbA * sie[k] + rbB * sie[k + 1]) / (rbA + rbB);
  }
  return SGM_OK;
}

// Uneven closure: per cell draw nmat from {2:.40, 3:.25, 4:.15, 5:.12, 6:.08};
// materials without replacement from the pool of 6 (partial Fisher-Yates);
// alphas = Exp(1) + EPS normalised and clamped to [EPS, 1]; per material rho
// then T. Gases (IdealGas slots) rho ~ LogNormal(ln 1e-3, 0.3), condensed
// rho ~ rho_ref * U(0.9, 1.1) with rho_ref from the EOS (StiffGas 1.0,
// Gruneisen rho0, Vinet rho0, table 2.7); T ~ U(300, 3000).
long sgm_workload_uneven(sgm_ctx *ctx, uint64_t seed, long n, const int *pool, int *offsets,
                         uint8_t *mat_ids, double *rho, double *vfrac, double *sie,
                         double *temp, double *press, double *sie_tot) {
  if (!ctx || !pool || n < 0 || !offsets || !mat_ids || !rho || !vfrac || !sie || !temp ||
      !press || !sie_tot)
    return SGM_E_ARG;
  for (int i = 0; i < 6; ++i)
    if (pool[i] < 0 || pool[i] >= ctx->nmat || !ctx->is_set[pool[i]]) return SGM_E_MAT;
  const double cdf[5] = {0.40, 0.65, 0.80, 0.92, 1.0};
  Rng r(seed);
  long nnz = 0;
  offsets[0] = 0;
  for (long c = 0; c < n; ++c) {
    const double un = r.u();
    int nm = 6;
    for (int k = 0; k < 5; ++k)
      if (un < cdf[k]) {
        nm = k + 2;
        break;
      }
    int idx[6] = {0, 1, 2, 3, 4, 5};
    for (int k = 0; k < nm; ++k) {
      const int j = k + (int)(r.u() * (6 - k));
      const int t = idx[k];
      idx[k] = idx[j < 6 ? j : 5];
      idx[j < 6 ? j : 5] = t;
    }
    double a[6], asum = 0;
    for (int k = 0; k < nm; ++k) {
      a[k] = r.expo() + robust::EPS();
      asum += a[k];
    }
    double rbsum = 0, usum = 0;
    for (int k = 0; k < nm; ++k) {
      const long q = nnz + k;
      const int mi = pool[idx[k]];
      const EosDev &e = ctx->host_eval[mi];
      double v = a[k] * (1.0 / asum);
      v = smax(robust::EPS(), smin(v, 1.0));
      double rr;
      if (e.kind == KIND_IDEAL) {
        rr = r.lognormal(std::log(1e-3), 0.3);
      } else {
        const double ref = e.kind == KIND_STIFF   ? 1.0
                           : e.kind == KIND_GRUN  ? e.m.p[6]
                           : e.kind == KIND_VINET ? e.m.p[0]
                                                  : 2.7;
        rr = ref * r.uni(0.9, 1.1);
      }
      const double T = r.uni(300.0, 3000.0);
      Lam none;
      mat_ids[q] = (uint8_t)mi;
      rho[q] = rr;
      vfrac[q] = v;
      temp[q] = T;
      sie[q] = eos_e_rT(e, rr, T, none);
      press[q] = eos_P_rT(e, rr, T, none);
      const double rb = rr * v;
      rbsum += rb;
      usum += rb * sie[q];
    }
    sie_tot[c] = usum / rbsum;
    nnz += nm;
    offsets[c + 1] = (int)nnz;
  }
  return nnz;
}

} // extern "C"
