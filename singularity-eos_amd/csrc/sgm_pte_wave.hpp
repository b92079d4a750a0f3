// The closure for cells of more materials than the register-resident solvers
// hold (9..64): one wave per cell, lane m carrying material m's state in
// VGPRs, the Newton system (Jacobian, residual, step) in the wave's LDS. Every
// sum over materials runs in material order on values read from the lanes
// (v_readlane), so each lane holds the same cell scalars and each value is
// formed in the order of the register-resident solvers in sgm_pte.hpp (and of
// the reference's loops); the dense solve spreads the columns of the
// pivoted QR over the lanes, each element again in the serial order.
//
// Restates singularity-eos/closure/mixed_cell_models.hpp: solve_Ax_b_wscr
// :176-278, PTESolverBase :329-786 (GetIdealPTE / TryIdealPTE :672-763),
// PTESolverRhoT :873-1193, PTESolverPT :1198-1446, PTESolverFixedT
// :1451-1669, PTESolverFixedP :1674-1921, PTESolverRhoU :1924-2183, with the
// driver loop of PTESolver :2188-2291 (pte_step).
//
// Lanes at or past the material count mirror material 0 so that their EOS
// calls stay on valid inputs; nothing they compute is read.
#pragma once
#include "sgm_pte.hpp"

namespace sgm {
namespace wave {

constexpr int MAX_MATS = 64;              // one lane per material
constexpr int MAX_NEQ = MAX_MATS + 1;     // FixedP: the materials and T
// LDS words of one wave: the Jacobian, then res, dx, x, z, cu, cd, perm
constexpr int WORK_WORDS = MAX_NEQ * MAX_NEQ + 7 * MAX_NEQ;

__device__ __forceinline__ int uni(int x) { return __builtin_amdgcn_readfirstlane(x); }
__device__ __forceinline__ Real uni(Real x) {
  const unsigned long long b = __builtin_bit_cast(unsigned long long, x);
  const unsigned lo = (unsigned)__builtin_amdgcn_readfirstlane((int)(unsigned)b);
  const unsigned hi = (unsigned)__builtin_amdgcn_readfirstlane((int)(unsigned)(b >> 32));
  return __builtin_bit_cast(Real, ((unsigned long long)hi << 32) | lo);
}
__device__ __forceinline__ bool uni(bool x) { return __builtin_amdgcn_readfirstlane((int)x) != 0; }
// lane k's value (k wave-uniform)
__device__ __forceinline__ Real at(Real x, int k) { return lane_value(x, k); }
__device__ __forceinline__ int at(int x, int k) { return __builtin_amdgcn_readlane(x, k); }
__device__ __forceinline__ bool any_lane(bool x) { return __ballot(x) != 0; }
__device__ __forceinline__ bool all_lanes(bool x) { return __ballot(!x) == 0; }
// the workgroup is this one wave: orders the LDS accesses of its lanes
__device__ __forceinline__ void wsync() { __syncthreads(); }

struct Work {
  Real *jac, *res, *dx, *x, *z, *cu, *cd;
  int *perm;
  __device__ void carve(Real *w) {
    jac = w;
    res = jac + MAX_NEQ * MAX_NEQ;
    dx = res + MAX_NEQ;
    x = dx + MAX_NEQ;
    z = x + MAX_NEQ;
    cu = z + MAX_NEQ;
    cd = cu + MAX_NEQ;
    perm = reinterpret_cast<int *>(cd + MAX_NEQ);
  }
};

// qr_colpiv_solve of sgm_pte.hpp on an N x N row-major system in LDS: lane j
// owns columns j and j + 64 (the pivot norms, the reflections of its columns,
// the norm downdates); the pivot search, the Householder vector's norm and
// the reflection of b are formed by every lane alike, in index order.
__device__ __noinline__ void qr_colpiv_solve(Real *a, Real *b, Real *x, int N, const Work &w,
                                             int lane) {
  for (int j = lane; j < N; j += 64) {
    w.perm[j] = j;
    Real s = 0;
    for (int i = 0; i < N; ++i) s += a[i * N + j] * a[i * N + j];
    w.cd[j] = sqrt(s);
    w.cu[j] = w.cd[j];
  }
  wsync();
  const Real eps = 2.220446049250313e-16;
  const Real norm_downdate_threshold = sqrt(eps);
  int nzp = N;
  Real th = 0;
  for (int k = 0; k < N; ++k) {
    int big = k;
    Real bign = w.cu[k];
    for (int j = k + 1; j < N; ++j) {
      const Real c = w.cu[j];
      if (c > bign) {
        bign = c;
        big = j;
      }
    }
    big = uni(big);
    bign = uni(bign);
    const Real big_sq = bign * bign;
    if (k == 0) th = big_sq * eps * eps / N;
    if (nzp == N && big_sq < th * (N - k)) nzp = k;
    if (big != k) {
      for (int i = lane; i < N; i += 64) {
        const Real t = a[i * N + k];
        a[i * N + k] = a[i * N + big];
        a[i * N + big] = t;
      }
      if (lane == 0) {
        const int tp = w.perm[k];
        w.perm[k] = w.perm[big];
        w.perm[big] = tp;
        Real t = w.cu[k];
        w.cu[k] = w.cu[big];
        w.cu[big] = t;
        t = w.cd[k];
        w.cd[k] = w.cd[big];
        w.cd[big] = t;
      }
      wsync();
    }
    Real tailsq = 0;
    for (int i = k + 1; i < N; ++i) tailsq += a[i * N + k] * a[i * N + k];
    tailsq = uni(tailsq);
    const Real c0 = uni(a[k * N + k]);
    wsync();
    Real tau, beta;
    if (tailsq <= 2.2250738585072014e-308) {
      tau = 0;
      beta = c0;
      for (int i = k + 1 + lane; i < N; i += 64) a[i * N + k] = 0;
    } else {
      beta = sqrt(c0 * c0 + tailsq);
      if (c0 >= 0) beta = -beta;
      const Real den = c0 - beta;
      for (int i = k + 1 + lane; i < N; i += 64) a[i * N + k] = a[i * N + k] / den;
      tau = (beta - c0) / beta;
    }
    if (lane == 0) a[k * N + k] = beta;
    wsync();
    for (int j = k + 1 + lane; j < N; j += 64) {
      Real s = 0;
      for (int i = k + 1; i < N; ++i) s += a[i * N + k] * a[i * N + j];
      s += a[k * N + j];
      a[k * N + j] -= tau * s;
      for (int i = k + 1; i < N; ++i) a[i * N + j] -= (tau * a[i * N + k]) * s;
    }
    {
      Real s = 0;
      for (int i = k + 1; i < N; ++i) s += a[i * N + k] * b[i];
      s += b[k];
      s = uni(s);
      wsync();
      if (lane == 0) b[k] -= tau * s;
      for (int i = k + 1 + lane; i < N; i += 64) b[i] -= (tau * a[i * N + k]) * s;
    }
    wsync();
    for (int j = k + 1 + lane; j < N; j += 64) {
      if (w.cu[j] != 0) {
        Real t = fabs(a[k * N + j]) / w.cu[j];
        t = (1 + t) * (1 - t);
        if (t < 0) t = 0;
        const Real r = w.cu[j] / w.cd[j];
        const Real t2 = t * r * r;
        if (t2 <= norm_downdate_threshold) {
          Real s = 0;
          for (int i = k + 1; i < N; ++i) s += a[i * N + j] * a[i * N + j];
          w.cd[j] = sqrt(s);
          w.cu[j] = w.cd[j];
        } else {
          w.cu[j] *= sqrt(t);
        }
      }
    }
    wsync();
  }
  for (int i = lane; i < N; i += 64) w.z[i] = b[i];
  wsync();
  for (int i = N - 1; i >= 0; --i) {
    if (i < nzp) {
      const Real zi = uni(w.z[i] / a[i * N + i]);
      wsync();
      if (lane == 0) w.z[i] = zi;
      for (int r = lane; r < i; r += 64) w.z[r] -= zi * a[r * N + i];
      wsync();
    }
  }
  for (int i = lane; i < N; i += 64) x[w.perm[i]] = (i < nzp) ? w.z[i] : 0.0;
  wsync();
}

// solve_small of sgm_pte.hpp: Jacobi row scaling (lane per row), then
// direct (N = 1), Cramer (N = 2) or pivoted QR; b is overwritten with x.
// false on a zero row or a non-finite x.
__device__ __forceinline__ bool solve_small(Real *a, Real *b, int N, const Work &w, int lane) {
  bool zero = false;
  for (int row = lane; row < N; row += 64) {
    Real maxabs = 0;
    for (int c = 0; c < N; ++c) maxabs = smax(fabs(a[row * N + c]), maxabs);
    if (maxabs == 0) {
      zero = true;
    } else {
      for (int c = 0; c < N; ++c) a[row * N + c] /= maxabs;
      b[row] /= maxabs;
    }
  }
  if (any_lane(zero)) return false;
  wsync();
  if (N == 1) {
    const Real x0 = robust::ratio(b[0], a[0]);
    wsync();
    if (lane == 0) b[0] = x0;
  } else if (N == 2) {
    const Real det = a[0] * a[3] - a[1] * a[2];
    const Real x0 = robust::ratio(a[3] * b[0] - a[1] * b[1], det);
    const Real x1 = robust::ratio(a[0] * b[1] - a[2] * b[0], det);
    wsync();
    if (lane == 0) {
      b[0] = x0;
      b[1] = x1;
    }
  } else {
    qr_colpiv_solve(a, b, w.x, N, w, lane);
    for (int i = lane; i < N; i += 64) b[i] = w.x[i];
  }
  wsync();
  bool ok = true;
  for (int i = lane; i < N; i += 64) ok = ok && isfinite_ref(b[i]);
  return all_lanes(ok);
}

// ---------------------------------------------------------------------------
// Shared state (PTESolverBase): this lane's material, the cell's scalars
// ---------------------------------------------------------------------------
template <bool MODS>
struct Base {
  using Ev = sgm::Ev<MODS>;
  int n = 0;      // materials (wave-uniform)
  int neq = 0;    // equations (wave-uniform)
  int lane = 0;
  bool act = false;  // lane < n
  Work w;
  Real *jac = nullptr, *res = nullptr, *dx = nullptr;
  unsigned char eidx = 0;
  Lam lam;
  Real rho = 0, vfrac = 0, sie = 0, press = 0, u = 0, rhobar = 0;
  Real vfrac_total = 0, sie_total = 0, rho_total = 0, uscale = 0, utotal_scale = 0, Tnorm = 0;
  int err = 0;
  MixParamsDev P;

  __device__ void Setup(int nm, int nq, int ln, Real *work) {
    n = nm;
    neq = nq;
    lane = ln;
    act = ln < nm;
    w.carve(work);
    jac = w.jac;
    res = w.res;
    dx = w.dx;
  }
  __device__ const EosDev &E() const { return registry_entry(eidx); }
  // sums over the materials in material order
  __device__ Real Sum(Real x) const {
    Real s = 0;
    for (int k = 0; k < n; ++k) s += at(x, k);
    return uni(s);
  }
  __device__ Real NSum(Real x) const {
    NeumaierSum s;
    for (int k = 0; k < n; ++k) s.add(at(x, k));
    return uni(s.result());
  }
  // the error bits of every material's EOS calls
  __device__ int ErrAll() const {
    int e = 0;
    for (int k = 0; k < n; ++k) e |= at(err, k);
    return uni(e);
  }

  // :388-399
  __device__ void InitRhoBarandRho() {
    rhobar = rho * vfrac;
    rho_total = NSum(rhobar);
  }
  __device__ void NormalizeVfrac() {
    const Real f = robust::ratio(vfrac_total, NSum(vfrac));
    vfrac *= f;
  }
  // :423-498; true on FAIL
  __device__ bool SetVfracFromT(Real T) {
    constexpr Real min_dv = 1e-12;
    constexpr Real safety = 0.9;
    const Real vmax = smin(safety * robust::ratio(rhobar, Ev::rho_pmin(E(), T)), 1.0);
    if (vfrac <= 0) vfrac = vmax;
    if (Sum(vmax) < vfrac_total) return true;
    NormalizeVfrac();
    for (int it = 0; it < n; ++it) {
      const Real vnew = smin(vmax, vfrac);
      const Real dv = vfrac - vnew;
      vfrac -= dv;
      Real remaining = 0;
      int n_uncapped = n;
      for (int k = 0; k < n; ++k) {
        const Real d = at(dv, k);
        remaining += d;
        n_uncapped -= (d > min_dv);
      }
      remaining = uni(remaining);
      if (fabs(remaining) < min_dv) break;
      if ((vmax - vfrac) > min_dv) vfrac += remaining / (Real)n_uncapped;
    }
    NormalizeVfrac();
    return false;
  }
  // :500-581; Tin: this lane's input temperature
  __device__ Real GetTguess(Real Tin) {
    Real Tguess;
    if (Tnorm > 0.0) {
      Tguess = Tnorm;
    } else {
      Tguess = P.default_tguess;
      for (int k = 0; k < n; ++k) Tguess = smax(Tguess, at(Tin, k));
    }
    Tguess = uni(Tguess);
    bool bad = false;
    for (int i = 0; i < 3; i++) {
      if (P.iterate_t_guess) {
        const Real rho_max = Ev::max_density(E());
        Real s, cv;
        sie_cv_from_T_preferred<MODS>(E(), lam, smin(rho, rho_max), press, Tguess, s, cv);
        const Real us = rhobar * s, uc = rhobar * cv;
        Real dudt = 0, usum = 0;
        for (int k = 0; k < n; ++k) {
          usum += at(us, k);
          dudt += at(uc, k);
        }
        const Real Tn = Tguess - robust::ratio((usum - utotal_scale * uscale), dudt);
        Tguess = uni(smin(P.temperature_limit, smax(Tguess, Tn)));
      }
      const Real tmin = Ev::min_temperature(E());
      for (int k = 0; k < n; ++k) Tguess = smax(at(tmin, k), Tguess);
      Tguess = uni(Tguess);
      bad = SetVfracFromT(Tguess);
      if (bad) {
        Tguess *= 10.0;
        continue;
      }
      rho = robust::ratio(rhobar, vfrac);
      bad = any_lane(act && rho < Ev::rho_pmin(E(), Tguess));
      if (bad) Tguess *= 10.0;
      if (!bad) break;
    }
    return Tguess;
  }
  // :583-597, the P(rho, T) beside e already evaluated
  __device__ Real PressureFromPreferredP(Real r, Real s, Real P_T) {
    if (Ev::preferred_input(E()) & tq::specific_internal_energy) return Ev::P_re(E(), r, s, lam);
    return P_T;
  }
  // :627-662
  __device__ void InitBase(Real Tin) {
    InitRhoBarandRho();
    const Real utotal = rho_total * sie_total;
    uscale = smax(fabs(utotal), 1.0e-14);
    utotal_scale = robust::ratio(utotal, uscale);
    const Real Tguess = GetTguess(Tin);
    Tnorm = Tguess;
    Real P_T;
    sie = Ev::e_rT_with_P(E(), rho, Tguess, lam, P_T);
    if ((Ev::preferred_input(E()) & tq::pressure) && !bad_value(press)) {
      press = robust::ratio(press, uscale);
    } else {
      press = robust::ratio(PressureFromPreferredP(rho, sie, P_T), uscale);
    }
    u = sie * robust::ratio(rhobar, uscale);
  }
  __device__ Real ResidualNorm() const {
    NeumaierSum s;
    for (int i = 0; i < neq; ++i) s.add(res[i] * res[i]);
    return uni(0.5 * s.result());
  }
  __device__ bool Solve() {
    for (int i = lane; i < neq; i += 64) dx[i] = res[i];
    wsync();
    return solve_small(jac, dx, neq, w, lane);
  }
  __device__ void ZeroJac() {
    for (int i = lane; i < neq * neq; i += 64) jac[i] = 0.0;
    wsync();
  }
};

// ---------------------------------------------------------------------------
// PTESolverRhoT (:873-1193): unknowns T and alpha_m for m != ms; material m
// (m != ms) is row and column r(m) = m + 1 below ms, m above.
// ---------------------------------------------------------------------------
template <bool MODS>
struct RhoT : Base<MODS> {
  using Ev = sgm::Ev<MODS>;
  using B = Base<MODS>;
  using B::act;
  using B::dx;
  using B::jac;
  using B::lam;
  using B::lane;
  using B::n;
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
  static __device__ int Equations(int nm) { return nm; }
  Real vtemp = 0;
  Real Tequil = 0, Ttemp = 0;
  int ms = 0;

  __device__ void SetCellInputs(Real vfrac_tot, Real sie_tot, Real Tnorm_in) {
    B::vfrac_total = vfrac_tot;
    B::sie_total = sie_tot;
    Tnorm = Tnorm_in;
  }
  __device__ int Row() const { return lane < ms ? lane + 1 : lane; }
  __device__ void Residual() {
    res[0] = utotal_scale - B::NSum(u);
    const Real pms = at(press, ms);
    if (act && lane != ms) res[Row()] = pms - press;
    wsync();
  }
  // :947-984
  __device__ Real Init(Real Tin) {
    B::InitBase(Tin);
    if (P.pte_reduced_system_exclude_idx >= 0) {
      ms = P.pte_reduced_system_exclude_idx;
    } else {
      const Real x = robust::ratio(rhobar, B::rho_total);
      ms = 0;
      Real xmax = 0;
      for (int k = 0; k < n; ++k) {
        const Real xk = at(x, k);
        if (xk > xmax) {
          xmax = xk;
          ms = k;
        }
      }
      ms = uni(ms);
    }
    Residual();
    Tequil = 1.0;
    return B::ResidualNorm();
  }
  // :986-1008
  __device__ void Fixup() {
    if (P.pte_reduced_system_exclude_idx < 0) {
      Real amax = 0;
      ms = 0;
      for (int k = 0; k < n; ++k) {
        const Real v = at(vfrac, k);
        if (v > amax) {
          amax = v;
          ms = k;
        }
      }
      ms = uni(ms);
      Residual();
    }
  }
  // :1021-1040
  __device__ void CheckPTE(bool &conv, bool &ok) const {
    const Real mean_p = B::Sum(vfrac * press);
    NeumaierSum s;
    for (int k = 1; k < n; ++k) s.add(res[k] * res[k]);
    const Real error_p = sqrt(s.result());
    const Real error_u = fabs(res[0]);
    const Real amp = fabs(mean_p);
    const bool cp = (error_p < P.pte_rel_tolerance_p * amp || error_p < P.pte_abs_tolerance_p);
    const bool cu = (error_u < P.pte_rel_tolerance_e);
    const bool ap = (error_p < P.pte_rel_tolerance_p_sufficient * amp ||
                     error_p < P.pte_abs_tolerance_p_sufficient);
    const bool au = (error_u < P.pte_rel_tolerance_e_sufficient);
    conv = uni(cp && cu);
    ok = uni(ap && au);
  }
  static __device__ Real PressureFromPreferredE(const EosDev &e, Lam &l, Real r, Real s,
                                                Real P_T) {
    if (Ev::preferred_input(e) & tq::specific_internal_energy) return Ev::P_re(e, r, s, l);
    return P_T;
  }
  // :1042-1105: this lane's material derivatives (MaterialDerivs of
  // sgm_pte.hpp), then the rows of the reduced system
  __device__ void Jacobian() {
    const Real Tcur = Tnorm * Tequil;
    const Real dT = Tequil * P.derivative_eps;
    const Real Tpert = Tnorm * (Tequil + dT);
    const EosDev &e = B::E();
    const Real dv = (vfrac < 0.5 ? 1.0 : -1.0) * vfrac * P.derivative_eps;
    const Real vf_pert = vfrac + dv;
    const Real rho_pert = robust::ratio(rhobar, vf_pert);
    Real P_T;
    Real e_pert = Ev::e_rT_with_P(e, rho_pert, Tcur, lam, P_T);
    Real p_pert = robust::ratio(PressureFromPreferredE(e, lam, rho_pert, e_pert, P_T), uscale);
    const Real dpdv = smin(robust::ratio((p_pert - press), dv), -fabs(P.pte_max_dpdv));
    const Real dedv = robust::ratio(rhobar * robust::ratio(e_pert, uscale) - u, dv);
    e_pert = Ev::e_rT_with_P(e, rho, Tpert, lam, P_T);
    p_pert = robust::ratio(PressureFromPreferredE(e, lam, rho, e_pert, P_T), uscale);
    const Real dpdT = robust::ratio((p_pert - press), dT);
    const Real dedT = robust::ratio(rhobar * robust::ratio(e_pert, uscale) - u, dT);
    const Real dedT_sum = B::Sum(dedT);
    const Real dedv_s = at(dedv, ms), dpdT_s = at(dpdT, ms), dpdv_s = at(dpdv, ms);
    const int N = n;
    if (lane == 0) jac[0] = dedT_sum;
    if (act && lane != ms) {
      const int r = Row();
      jac[r] = dedv - dedv_s;
      jac[r * N + 0] = dpdT - dpdT_s;
      for (int c = 1; c < N; ++c) jac[r * N + c] = (c == r) ? dpdv + dpdv_s : dpdv_s;
    }
    wsync();
  }
  // :1107-1151 (with the reference's idx bookkeeping, which does not advance
  // for materials skipped in the second pass)
  __device__ Real ScaleDx() {
    Real scale = 1.0;
    NeumaierSum s;
    for (int i = 0; i < n - 1; ++i) s.add(dx[1 + i]);
    const Real dalphaskip = uni(-s.result());
    const Real mydx = (lane == ms) ? dalphaskip : dx[act ? Row() : 0];
    for (int k = 0; k < n; ++k) {
      const Real v = at(vfrac, k), d = at(mydx, k);
      if (scale * d < -P.vfrac_safety_fac * v)
        scale = smin(scale, fabs(P.vfrac_safety_fac * robust::ratio(v, d)));
    }
    scale = uni(scale);
    const Real Tnew = Tequil + scale * dx[0];
    const Real rho_min = smax(Ev::rho_pmin(B::E(), Tnorm * Tequil), Ev::rho_pmin(B::E(), Tnorm * Tnew));
    const Real alpha_max = smin(1.0, robust::ratio(rhobar, rho_min));
    int idx = 1;
    for (int k = 0; k < n; ++k) {
      const Real am = at(alpha_max, k), v = at(vfrac, k);
      if (!(am < v)) {
        Real d;
        if (k == ms) {
          d = dalphaskip;
        } else {
          d = dx[idx];
          idx++;
        }
        if (scale * d > 0.5 * (am - v)) scale = smin(robust::ratio(0.5 * (am - v), d), scale);
      }
    }
    scale = uni(scale);
    if (scale * dx[0] < -0.95 * Tequil) scale = smin(scale, robust::ratio(-0.95 * Tequil, dx[0]));
    scale = uni(scale);
    wsync();
    for (int i = lane; i < n; i += 64) dx[i] *= scale;
    wsync();
    return scale;
  }
  // :1155-1187
  __device__ Real TestUpdate(Real scale, bool cache_state) {
    if (cache_state) {
      Ttemp = Tequil;
      vtemp = vfrac;
    }
    Tequil = uni(Ttemp + scale * dx[0]);
    if (act && lane != ms) vfrac = vtemp + scale * dx[Row()];
    NeumaierSum s;
    for (int k = 0; k < n; ++k)
      if (k != ms) s.add(at(vfrac, k));
    const Real vs = uni(1 - s.result());
    if (lane == ms) vfrac = vs;
    const Real Tc = Tnorm * Tequil;
    rho = robust::ratio(rhobar, vfrac);
    Real P_T;
    const Real um = rhobar * Ev::e_rT_with_P(B::E(), rho, Tc, lam, P_T);
    sie = robust::ratio(um, rhobar);
    u = robust::ratio(um, uscale);
    press = robust::ratio(B::PressureFromPreferredP(rho, sie, P_T), uscale);
    wsync();
    Residual();
    return B::ResidualNorm();
  }
  __device__ Real Tout() const { return Tequil * Tnorm; }
  __device__ Real Pout() const { return press * uscale; }
};

// ---------------------------------------------------------------------------
// PTESolverPT (:1198-1446): unknowns T and P, a 2 x 2 system
// ---------------------------------------------------------------------------
template <bool MODS>
struct PT : Base<MODS> {
  using Ev = sgm::Ev<MODS>;
  using B = Base<MODS>;
  using B::act;
  using B::dx;
  using B::err;
  using B::jac;
  using B::lam;
  using B::lane;
  using B::n;
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
  static __device__ int Equations(int) { return 2; }
  Real Pequil = 0, Tequil = 0, Ptemp = 0, Ttemp = 0;

  __device__ void SetCellInputs(Real vfrac_tot, Real sie_tot, Real Tnorm_in) {
    B::vfrac_total = vfrac_tot;
    B::sie_total = sie_tot;
    Tnorm = Tnorm_in;
  }
  __device__ void Residual() {
    const Real sv = B::NSum(vfrac), se = B::NSum(u);
    wsync();
    if (lane == 0) {
      res[0] = B::vfrac_total - sv;
      res[1] = utotal_scale - se;
    }
    wsync();
  }
  // :1258-1297
  __device__ Real Init(Real Tin) {
    B::InitBase(Tin);
    Residual();
    Pequil = B::Sum(fabs(press) * vfrac);
    Pequil = uni(Pequil / B::NSum(vfrac));
    Tequil = 1;
    Ev::rhoe_PT(B::E(), Pequil * uscale, Tequil * Tnorm, rho, sie, lam, err);
    vfrac = robust::ratio(rhobar, rho);
    u = robust::ratio(sie * rhobar, uscale);
    Residual();
    return B::ResidualNorm();
  }
  __device__ void Fixup() {}
  __device__ void CheckPTE(bool &conv, bool &ok) const {
    const Real ev = fabs(res[0]), eu = fabs(res[1]);
    conv = uni((ev < P.pte_rel_tolerance_v) && (eu < P.pte_rel_tolerance_e));
    ok = uni((eu < P.pte_rel_tolerance_e_sufficient) && (ev < P.pte_rel_tolerance_v_sufficient));
  }
  // :1336-1385
  __device__ void Jacobian() {
    const Real dp = -Pequil * P.derivative_eps;
    const Real dT = Tequil * P.derivative_eps;
    Real r_pert = rho;
    Real e_pert = sie;
    Ev::rhoe_PT(B::E(), uscale * (Pequil + dp), Tnorm * Tequil, r_pert, e_pert, lam, err);
    const Real drdp = robust::ratio(r_pert - rho, dp);
    const Real dudp = robust::ratio(robust::ratio(rhobar * e_pert, uscale) - u, dp);
    const Real rbp = robust::ratio(rhobar, rho * rho) * drdp;
    Ev::rhoe_PT(B::E(), uscale * Pequil, Tnorm * (Tequil + dT), r_pert, e_pert, lam, err);
    const Real drdT = robust::ratio(r_pert - rho, dT);
    const Real dudTm = robust::ratio(robust::ratio(rhobar * e_pert, uscale) - u, dT);
    const Real rbt = robust::ratio(rhobar, rho * rho) * drdT;
    // the reference's four sums interleave per material; each is its own chain
    const Real rbP = B::Sum(rbp), dudP = B::Sum(dudp), rbT = B::Sum(rbt), dudT = B::Sum(dudTm);
    if (lane == 0) {
      jac[0] = -rbT;
      jac[1] = -rbP;
      jac[2] = dudT;
      jac[3] = dudP;
    }
    wsync();
  }
  // :1387-1409: the limits chain through the materials in order, the
  // maximum pressure at the temperature the chain has reached
  __device__ Real ScaleDx() {
    Real scale = 1.0;
    const Real dx0 = dx[0], dx1 = dx[1];
    if (scale * dx0 < -0.95 * Tequil) scale = robust::ratio(-0.95 * Tequil, dx0);
    const Real Pmin = Ev::min_pressure(B::E());
    for (int k = 0; k < n; ++k)
      scale = smin(fabs(scale),
                   fabs(0.95 * robust::ratio(robust::ratio(at(Pmin, k), uscale) - Pequil, dx1)));
    scale = uni(scale);
    for (int k = 0; k < n; ++k) {
      const Real Ttest = (Tequil + scale * dx0) * Tnorm;
      int e2 = 0;
      const Real Pmax = Ev::max_pressure_at_T(B::E(), Ttest, e2);
      if (lane == k) err |= e2;
      scale = uni(smin(fabs(scale),
                       fabs(0.95 * robust::ratio(robust::ratio(at(Pmax, k), uscale) - Pequil, dx0))));
    }
    wsync();
    if (lane == 0) {
      dx[0] = dx0 * scale;
      dx[1] = dx1 * scale;
    }
    wsync();
    return scale;
  }
  // :1413-1435
  __device__ Real TestUpdate(Real scale, bool cache_state) {
    if (cache_state) {
      Ttemp = Tequil;
      Ptemp = Pequil;
    }
    Tequil = uni(Ttemp + scale * dx[0]);
    Pequil = uni(Ptemp + scale * dx[1]);
    Ev::rhoe_PT(B::E(), Pequil * uscale, Tequil * Tnorm, rho, sie, lam, err);
    vfrac = robust::ratio(rhobar, rho);
    u = robust::ratio(sie * rhobar, uscale);
    Residual();
    return B::ResidualNorm();
  }
  __device__ Real Tout() const { return Tequil * Tnorm; }
  __device__ Real Pout() const { return Pequil * uscale; }
};

// ---------------------------------------------------------------------------
// PTESolverFixedT (:1451-1669): temperature given; unknowns the volume
// fractions; rows: volume closure, neighbouring pressure differences
// ---------------------------------------------------------------------------
template <bool MODS>
struct FixedT : Base<MODS> {
  using Ev = sgm::Ev<MODS>;
  using B = Base<MODS>;
  using B::act;
  using B::dx;
  using B::jac;
  using B::lam;
  using B::lane;
  using B::n;
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
  static constexpr bool kStoresTemp = false;
  static __device__ int Equations(int nm) { return nm; }
  Real vtemp = 0;
  Real Tequil = 0;

  // constructor (:1496-1509)
  __device__ void SetCellInputs(Real vfrac_tot, Real T_true, Real) {
    B::vfrac_total = vfrac_tot;
    B::sie_total = 1.0;
    Tequil = T_true;
    Tnorm = 1.0;
  }
  __device__ void Residual() {
    const Real sv = B::NSum(vfrac);
    const Real pnext = __shfl(press, lane + 1 < 64 ? lane + 1 : lane, 64);
    wsync();
    if (lane == 0) res[0] = B::vfrac_total - sv;
    if (lane < n - 1) res[1 + lane] = press - pnext;
    wsync();
  }
  // :1523-1551
  __device__ Real Init(Real) {
    Tnorm = 1.0;
    B::InitRhoBarandRho();
    B::SetVfracFromT(Tequil);
    rho = robust::ratio(rhobar, vfrac);
    Real P_T;
    sie = Ev::e_and_P_rT(B::E(), rho, Tequil, lam, P_T);
    uscale = B::Sum(fabs(sie) * rho);
    press = P_T;
    uscale = smax(uscale, 1.0e-14);
    press = robust::ratio(press, uscale);
    u = sie * robust::ratio(rhobar, uscale);
    Residual();
    return B::ResidualNorm();
  }
  __device__ void Fixup() {}
  // :1562-1581
  __device__ void CheckPTE(bool &conv, bool &ok) const {
    const Real vp = vfrac * press;
    Real mean_p = at(vp, 0);
    Real error_p = 0;
    for (int k = 1; k < n; ++k) {
      mean_p += at(vp, k);
      error_p += res[k] * res[k];
    }
    error_p = sqrt(error_p);
    const Real error_v = fabs(res[0]);
    const Real amp = fabs(mean_p);
    conv = uni((error_p < P.pte_rel_tolerance_p * amp || error_p < P.pte_abs_tolerance_p) &&
               (error_v < P.pte_rel_tolerance_v));
    ok = uni((error_p < P.pte_rel_tolerance_p_sufficient * amp ||
              error_p < P.pte_abs_tolerance_p_sufficient) &&
             (error_v < P.pte_rel_tolerance_v_sufficient));
  }
  // :1583-1609
  __device__ void Jacobian() {
    const Real dv = (vfrac < 0.5 ? 1.0 : -1.0) * vfrac * P.derivative_eps;
    const Real vf_pert = vfrac + dv;
    const Real rho_pert = robust::ratio(rhobar, vf_pert);
    const Real p_pert = robust::ratio(Ev::P_rT(B::E(), rho_pert, Tequil, lam), uscale);
    const Real dpdv = smin(robust::ratio((p_pert - press), dv), -fabs(P.pte_max_dpdv));
    const Real dnext = __shfl(dpdv, lane + 1 < 64 ? lane + 1 : lane, 64);
    const int N = n;
    B::ZeroJac();
    if (act) jac[lane] = 1.0;
    if (lane < n - 1) {
      jac[(lane + 1) * N + lane] = -dpdv;
      jac[(lane + 1) * N + lane + 1] = dnext;
    }
    wsync();
  }
  // :1611-1641
  __device__ Real ScaleDx() {
    Real scale = 1.0;
    const Real mydx = dx[act ? lane : 0];
    for (int k = 0; k < n; ++k) {
      const Real v = at(vfrac, k), d = at(mydx, k);
      if (scale * d < -P.vfrac_safety_fac * v) scale = robust::ratio(-P.vfrac_safety_fac * v, d);
    }
    scale = uni(scale);
    const Real alpha_max = robust::ratio(rhobar, Ev::rho_pmin(B::E(), Tequil));
    for (int k = 0; k < n; ++k) {
      const Real am = at(alpha_max, k), v = at(vfrac, k), d = at(mydx, k);
      if (!(am < v) && scale * d > 0.5 * (am - v)) scale = robust::ratio(0.5 * (am - v), d);
    }
    scale = uni(scale);
    wsync();
    if (act) dx[lane] = mydx * scale;
    wsync();
    return scale;
  }
  // :1643-1664
  __device__ Real TestUpdate(Real scale, bool cache_state) {
    if (cache_state) vtemp = vfrac;
    vfrac = vtemp + scale * dx[act ? lane : 0];
    rho = robust::ratio(rhobar, vfrac);
    Real P_T;
    const Real um = rhobar * Ev::e_and_P_rT(B::E(), rho, Tequil, lam, P_T);
    sie = robust::ratio(um, rhobar);
    u = robust::ratio(um, uscale);
    press = robust::ratio(P_T, uscale);
    Residual();
    return B::ResidualNorm();
  }
  __device__ Real Tout() const { return Tequil; }
  __device__ Real Pout() const { return press * uscale; }
};

// ---------------------------------------------------------------------------
// PTESolverFixedP (:1674-1921): pressure given; unknowns the volume
// fractions (rows 0..n-1: each material's pressure against P) and the
// temperature (row n: the volume closure)
// ---------------------------------------------------------------------------
template <bool MODS>
struct FixedP : Base<MODS> {
  using Ev = sgm::Ev<MODS>;
  using B = Base<MODS>;
  using B::act;
  using B::dx;
  using B::jac;
  using B::lam;
  using B::lane;
  using B::n;
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
  static __device__ int Equations(int nm) { return nm + 1; }
  Real vtemp = 0;
  Real Tequil = 0, Ttemp = 0, Pequil = 0;

  // constructor (:1721-1733), with the base's uscale and utotal_scale zero
  // when GetTguess reads them (as the register-resident solver)
  __device__ void SetCellInputs(Real vfrac_tot, Real P_in, Real) {
    B::vfrac_total = vfrac_tot;
    B::sie_total = 1.0;
    Pequil = P_in;
    Tnorm = 0.0;
    uscale = 0.0;
    utotal_scale = 0.0;
  }
  // GetPressureFromPreferred(..., do_e_lookup = true) (:583-597)
  __device__ Real PressureLookup(Real r, Real T) {
    const EosDev &e = B::E();
    if (Ev::preferred_input(e) & tq::specific_internal_energy)
      return Ev::P_re(e, r, Ev::e_rT(e, r, T, lam), lam);
    return Ev::P_rT(e, r, T, lam);
  }
  __device__ void Residual() {
    const Real sv = B::NSum(vfrac);
    const Real pt = robust::ratio(Pequil, uscale);
    wsync();
    if (act) res[lane] = pt - press;
    if (lane == 0) res[n] = B::vfrac_total - sv;
    wsync();
  }
  // :1747-1779
  __device__ Real Init(Real Tin) {
    B::InitRhoBarandRho();
    const Real Tguess = B::GetTguess(Tin);
    Tnorm = Tguess;
    uscale = smax(fabs(Pequil), 1.0e-14);
    Real P_T;
    sie = Ev::e_and_P_rT(B::E(), rho, Tguess, lam, P_T);
    u = robust::ratio(sie * rhobar, uscale);
    press = robust::ratio(P_T, uscale);
    Residual();
    Tequil = 1.0;
    Ttemp = Tequil;
    return B::ResidualNorm();
  }
  __device__ void Fixup() {}
  // :1790-1808
  __device__ void CheckPTE(bool &conv, bool &ok) const {
    Real error_p = 0;
    for (int k = 0; k < n; ++k) error_p += res[k] * res[k];
    error_p = sqrt(error_p);
    error_p *= uscale;
    const Real error_v = fabs(res[n]);
    conv = uni((error_p < P.pte_rel_tolerance_p || error_p < P.pte_abs_tolerance_p) &&
               (error_v < P.pte_rel_tolerance_v));
    ok = uni((error_p < P.pte_rel_tolerance_p_sufficient ||
              error_p < P.pte_abs_tolerance_p_sufficient) &&
             (error_v < P.pte_rel_tolerance_v_sufficient));
  }
  // :1809-1850
  __device__ void Jacobian() {
    const Real dT = Tequil * P.derivative_eps;
    const Real dv = (vfrac < 0.5 ? 1.0 : -1.0) * vfrac * P.derivative_eps;
    const Real vf_pert = vfrac + dv;
    const Real rho_pert = robust::ratio(rhobar, vf_pert);
    Real p_pert = robust::ratio(PressureLookup(rho_pert, Tnorm * Tequil), uscale);
    const Real dpdv = smin(robust::ratio((p_pert - press), dv), -fabs(P.pte_max_dpdv));
    p_pert = robust::ratio(PressureLookup(rho, Tnorm * (Tequil + dT)), uscale);
    const Real dpdT = robust::ratio((p_pert - press), dT);
    const int NEQ = n + 1;
    B::ZeroJac();
    if (act) {
      jac[lane * NEQ + lane] = dpdv;
      jac[lane * NEQ + n] = dpdT;
      jac[n * NEQ + lane] = 1.0;
    }
    wsync();
  }
  // :1851-1887
  __device__ Real ScaleDx() {
    Real scale = 1.0;
    const Real mydx = dx[act ? lane : 0];
    const Real dT = dx[n];
    for (int k = 0; k < n; ++k) {
      const Real v = at(vfrac, k), d = at(mydx, k);
      if (scale * d < -P.vfrac_safety_fac * v) scale = -P.vfrac_safety_fac * robust::ratio(v, d);
    }
    scale = uni(scale);
    const Real Tnew = Tequil + scale * dT;
    const Real rho_min =
        smax(Ev::rho_pmin(B::E(), Tnorm * Tequil), Ev::rho_pmin(B::E(), Tnorm * Tnew));
    const Real alpha_max = robust::ratio(rhobar, rho_min);
    for (int k = 0; k < n; ++k) {
      const Real am = at(alpha_max, k), v = at(vfrac, k), d = at(mydx, k);
      if (!(am < v) && scale * d > 0.5 * (am - v)) scale = 0.5 * robust::ratio(am - v, d);
    }
    scale = uni(scale);
    if (scale * dT < -0.95 * Tequil) scale = -0.95 * robust::ratio(Tequil, dT);
    scale = uni(scale);
    wsync();
    for (int i = lane; i < n + 1; i += 64) dx[i] *= scale;
    wsync();
    return scale;
  }
  // :1889-1917
  __device__ Real TestUpdate(Real scale, bool cache_state) {
    if (cache_state) {
      Ttemp = Tequil;
      vtemp = vfrac;
    }
    Tequil = uni(Ttemp + scale * dx[n]);
    const Real Tc = Tnorm * Tequil;
    vfrac = vtemp + scale * dx[act ? lane : 0];
    rho = robust::ratio(rhobar, vfrac);
    Real P_T;
    const Real um = rhobar * Ev::e_and_P_rT(B::E(), rho, Tc, lam, P_T);
    press = robust::ratio(P_T, uscale);
    sie = robust::ratio(um, rhobar);
    u = robust::ratio(um, uscale);
    Residual();
    return B::ResidualNorm();
  }
  __device__ Real Tout() const { return Tequil * Tnorm; }
  __device__ Real Pout() const { return press * uscale; }
};

// ---------------------------------------------------------------------------
// PTESolverRhoU (:1924-2183, with GetIdealPTE / TryIdealPTE :672-763):
// unknowns the volume fractions (columns 0..n-1) and scaled energy densities
// (n..2n-1) of every material; rows: volume and energy closure, neighbouring
// pressure and temperature differences. 2n equations, so at most 32
// materials fit the wave's workspace.
// ---------------------------------------------------------------------------
template <bool MODS>
struct RhoU : Base<MODS> {
  using Ev = sgm::Ev<MODS>;
  using B = Base<MODS>;
  using B::act;
  using B::dx;
  using B::jac;
  using B::lam;
  using B::lane;
  using B::n;
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
  static __device__ int Equations(int nm) { return 2 * nm; }
  Real temp = 0, vtemp = 0, utemp = 0, dtdv = 0, dtde = 0;

  __device__ void SetCellInputs(Real vfrac_tot, Real sie_tot, Real Tnorm_in) {
    B::vfrac_total = vfrac_tot;
    B::sie_total = sie_tot;
    Tnorm = Tnorm_in;
  }
  // GetPressureFromPreferred (:583-597)
  __device__ Real PressureFromPreferred(Real r, Real T, Real s) {
    const EosDev &e = B::E();
    if (Ev::preferred_input(e) & tq::specific_internal_energy) return Ev::P_re(e, r, s, lam);
    return Ev::P_rT(e, r, T, lam);
  }
  // :2014-2026
  __device__ void Residual() {
    const Real vs = B::NSum(vfrac), es = B::NSum(u);
    const int nx = lane + 1 < 64 ? lane + 1 : lane;
    const Real pnext = __shfl(press, nx, 64), tnext = __shfl(temp, nx, 64);
    wsync();
    if (lane == 0) {
      res[0] = B::vfrac_total - vs;
      res[1] = utotal_scale - es;
    }
    if (lane < n - 1) {
      res[2 + lane] = pnext - press;
      res[n + 1 + lane] = tnext - temp;
    }
    wsync();
  }
  // :672-687
  __device__ void GetIdealPTE(Real &Pideal, Real &Tideal) const {
    const Real a = vfrac * robust::ratio(press, temp);
    const Real rb = rho * vfrac * robust::ratio(sie, temp);
    Real Asum = 0.0, rhoBsum = 0.0;
    for (int k = 0; k < n; ++k) {
      Asum += at(a, k);
      rhoBsum += at(rb, k);
    }
    Asum *= uscale / Tnorm;
    rhoBsum /= Tnorm;
    Tideal = uni(uscale / rhoBsum / Tnorm);
    Pideal = uni(robust::ratio(Tnorm * Tideal * Asum, uscale * B::vfrac_total));
  }
  __device__ Real SumSquaresRes() const {
    Real r = 0.0;
    for (int i = 0; i < 2 * n; ++i) r += res[i] * res[i];
    return uni(r);
  }
  // :691-763. A material whose scaled volume fraction falls below its
  // minimum-pressure density stops the reference's loop and restores the
  // fractions it changed: none changes then.
  __device__ void TryIdealPTE() {
    Real Pideal, Tideal;
    GetIdealPTE(Pideal, Tideal);
    const Real etemp = u, ptemp = press, vt = vfrac, rtemp = rho;
    Real *rs = B::w.x;  // the residual before the ideal-gas guess
    for (int i = lane; i < 2 * n; i += 64) rs[i] = res[i];
    const Real res_norm_old = SumSquaresRes();
    const Real alpha = robust::ratio(Pideal, Tideal);
    const Real vnew = vfrac * robust::ratio(press, (temp * alpha));
    const bool bad = Tnorm * Tideal < 0 ||
                     robust::ratio(rhobar, vnew) < Ev::rho_pmin(B::E(), Tnorm * Tideal);
    if (any_lane(act && bad)) return;
    vfrac = vnew;
    rho = robust::ratio(rhobar, vfrac);
    Real P_T;
    const Real sie_m = Ev::e_rT_with_P(B::E(), rho, Tnorm * Tideal, lam, P_T);
    u = rhobar * robust::ratio(sie_m, uscale);
    press = robust::ratio(B::PressureFromPreferredP(rho, sie_m, P_T), uscale);
    Residual();
    const Real res_norm_new = SumSquaresRes();
    if (res_norm_new > res_norm_old) {
      vfrac = vt;
      rho = rtemp;
      u = etemp;
      press = ptemp;
      for (int i = lane; i < 2 * n; i += 64) res[i] = rs[i];
      wsync();
    } else {
      temp = Tideal;
      sie = uscale * robust::ratio(u, rhobar);
    }
  }
  // :2005-2011 (InitBase sets every scaled temperature to 1, :641)
  __device__ Real Init(Real Tin) {
    B::InitBase(Tin);
    temp = 1.0;
    Residual();
    TryIdealPTE();
    return B::ResidualNorm();
  }
  __device__ void Fixup() {}
  // :2029-2055
  __device__ void CheckPTE(bool &conv, bool &ok) const {
    const Real vp = vfrac * press, rt = rhobar * temp;
    Real mean_p = at(vp, 0), mean_t = at(rt, 0);
    Real error_p = 0.0, error_t = 0.0;
    for (int k = 1; k < n; ++k) {
      mean_p += at(vp, k);
      mean_t += at(rt, k);
      error_p += res[k + 1] * res[k + 1];
      error_t += res[k + n] * res[k + n];
    }
    mean_t = robust::ratio(mean_t, B::rho_total);
    error_p = sqrt(error_p);
    error_t = sqrt(error_t);
    const bool cp = (error_p < P.pte_rel_tolerance_p * fabs(mean_p) ||
                     error_p < P.pte_abs_tolerance_p);
    const bool ct = (error_t < P.pte_rel_tolerance_t * mean_t || error_t < P.pte_abs_tolerance_t);
    const bool ap = (error_p < P.pte_rel_tolerance_p_sufficient * fabs(mean_p) ||
                     error_p < P.pte_abs_tolerance_p_sufficient);
    const bool at_ = (error_t < P.pte_rel_tolerance_t_sufficient * mean_t ||
                      error_t < P.pte_abs_tolerance_t_sufficient);
    conv = uni(cp && ct);
    ok = uni(ap && at_);
  }
  // :2058-2121: this lane's derivatives, then its columns m and n + m
  __device__ void Jacobian() {
    const EosDev &e = B::E();
    const Real dv = (vfrac < 0.5 ? 1.0 : -1.0) * vfrac * P.derivative_eps;
    const Real vf_pert = vfrac + dv;
    const Real rho_pert = robust::ratio(rhobar, vf_pert);
    Real t_pert = robust::ratio(Ev::T_re(e, rho_pert, sie, lam), Tnorm);
    Real p_pert = robust::ratio(PressureFromPreferred(rho_pert, Tnorm * t_pert, sie), uscale);
    const Real dpdv = smin(robust::ratio(p_pert - press, dv), -fabs(P.pte_max_dpdv));
    dtdv = robust::ratio(t_pert - temp, dv);
    const Real de = fabs(u) * P.derivative_eps;
    const Real e_pert = robust::ratio(u + de, rhobar);
    t_pert = robust::ratio(Ev::T_re(e, rho, uscale * e_pert, lam), Tnorm);
    p_pert = robust::ratio(PressureFromPreferred(rho, Tnorm * t_pert, uscale * e_pert), uscale);
    const Real dpde = robust::ratio(p_pert - press, de);
    dtde = robust::ratio(t_pert - temp, de);
    if (fabs(dtde) < P.min_dtde) dtde = P.derivative_eps;
    const int N = 2 * n, m = lane;
    B::ZeroJac();
    if (act) {
      jac[m] = 1.0;
      jac[2 * n + n + m] = 1.0;
      if (m == 0) {
        jac[2 * N] = dpdv;
        jac[(n + 1) * N] = dtdv;
        jac[2 * N + n] = dpde;
        jac[(n + 1) * N + n] = dtde;
      } else if (m == n - 1) {
        jac[n * N + n - 1] = -dpdv;
        jac[(2 * n - 1) * N + n - 1] = -dtdv;
        jac[n * N + 2 * n - 1] = -dpde;
        jac[(2 * n - 1) * N + 2 * n - 1] = -dtde;
      } else {
        jac[(1 + m) * N + m] = -dpdv;
        jac[(2 + m) * N + m] = dpdv;
        jac[(n + m) * N + m] = -dtdv;
        jac[(n + m + 1) * N + m] = dtdv;
        jac[(1 + m) * N + n + m] = -dpde;
        jac[(2 + m) * N + n + m] = dpde;
        jac[(n + m) * N + n + m] = -dtde;
        jac[(n + m + 1) * N + n + m] = dtde;
      }
    }
    wsync();
  }
  // :2124-2158: the limits chain through the materials in order, each
  // material's minimum-pressure density at the temperature the chain has
  // reached (every lane evaluates its material there; lane k's is used)
  __device__ Real ScaleDx() {
    Real scale = 1.0;
    const Real rp0 = Ev::rho_pmin(B::E(), Tnorm * temp);
    for (int k = 0; k < n; ++k) {
      const Real v = at(vfrac, k), tk = at(temp, k), d = dx[k], dE = dx[k + n];
      if (scale * d < -0.1 * v) scale = -0.1 * robust::ratio(v, d);
      const Real dt = (at(dtdv, k) * d + at(dtde, k) * dE);
      if (scale * dt < -0.1 * tk) scale = -0.1 * robust::ratio(tk, dt);
      const Real tt = uni(tk + scale * dt);
      const Real rho_min = smax(at(rp0, k), at(Ev::rho_pmin(B::E(), Tnorm * tt), k));
      const Real alpha_max = robust::ratio(at(rhobar, k), rho_min);
      if (!(alpha_max < v)) {
        if (scale * d > 0.5 * (alpha_max - v)) scale = 0.5 * robust::ratio(alpha_max - v, d);
      }
      scale = uni(scale);
    }
    wsync();
    for (int i = lane; i < 2 * n; i += 64) dx[i] *= scale;
    wsync();
    return scale;
  }
  // :2162-2181
  __device__ Real TestUpdate(Real scale, bool cache_state) {
    if (cache_state) {
      vtemp = vfrac;
      utemp = u;
    }
    const int m = act ? lane : 0;
    vfrac = vtemp + scale * dx[m];
    rho = robust::ratio(rhobar, vfrac);
    u = utemp + scale * dx[n + m];
    sie = uscale * robust::ratio(u, rhobar);
    temp = robust::ratio(Ev::T_re(B::E(), rho, sie, lam), Tnorm);
    press = robust::ratio(PressureFromPreferred(rho, Tnorm * temp, sie), uscale);
    Residual();
    return B::ResidualNorm();
  }
  __device__ Real Tout() const { return temp * Tnorm; }
  __device__ Real Pout() const { return press * uscale; }
};

// formulation by method id (0 RhoT, 1 PT, 2 FixedT, 3 FixedP, 4 RhoU)
template <int METHOD, bool MODS = true>
using SysOf = typename std::conditional<
    METHOD == 0, RhoT<MODS>,
    typename std::conditional<
        METHOD == 1, PT<MODS>,
        typename std::conditional<
            METHOD == 2, FixedT<MODS>,
            typename std::conditional<METHOD == 3, FixedP<MODS>, RhoU<MODS>>::type>::type>::
            type>::type;
// the largest material count of the one-wave form by method: RhoU's 2 nm
// equations fit the workspace up to 32 materials
constexpr int WaveMaxNM(int method) { return method == 4 ? 32 : MAX_MATS; }

} // namespace wave
} // namespace sgm
