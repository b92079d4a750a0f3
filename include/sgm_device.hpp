/*
 * sgm_device.hpp — the mixed-cell closure as device code, for HIP kernels
 * written by the caller (one cell per thread, as a hydro code's own kernels
 * call singularity-eos's closure objects).
 *
 * Replaces the device-callable templates of
 * singularity-eos/closure/mixed_cell_models.hpp:
 *   MixParams, SolverStatus (:43-87);
 *   ApproxTemperatureFromRhoMatU(nmat, eos, u_tot, rho, vfrac, Tguess, lambda) (:790-847);
 *   PTESolverRhoT (:873-1193), PTESolverPT (:1198-1446), PTESolverFixedT (:1451-1669),
 *   PTESolverFixedP (:1674-1921), PTESolverRhoU (:1924-2183) and their
 *   PTESolver*RequiredScratch[InBytes](nmat);
 *   SolverStatus PTESolver(System &) (:2188-2291).
 * Same constructor arguments, same meaning, same results as the library's
 * sgm_pte_* kernels bit for bit (the library's own register-resident solver
 * runs inside the caller's kernel; compile with hipcc --offload-arch=gfx950
 * -ffp-contract=off -Xclang -target-feature -Xclang -packed-fp32-ops).
 *
 * Differences a caller sees:
 *   * EOS objects are the context's materials. An EOSIndexer yields, for
 *     material m of the cell, its context material index (eos[m], an
 *     integer, what mat_ids holds for the batch calls). A kernel that builds
 *     closures first stages the context's registry into LDS with
 *     sgm::device::stage_registry (every thread of the block, before any
 *     solver), and is launched with at least registry_lds_bytes(n) bytes of
 *     dynamic LDS, which the registry occupies from its start. The registry
 *     (device pointer and count) comes from sgm_device_registry().
 *   * The solver keeps its whole state in registers, templated on the cell's
 *     material count: PTESolver dispatches on nmat over 1..SGM_DEVICE_MAX_MATS
 *     (define it before including to bound the code size; default 8). A cell
 *     with more materials is not solved (converged = false, max_niter = 0).
 *   * `scratch` is accepted and not used; PTESolver*RequiredScratch return the
 *     reference's sizes, so callers that size scratch keep working.
 *   * A LambdaIndexer yields per material a Real* (or nullptr): [0] log10
 *     rho and [1] log10 T of the table EOS's root-find cache, read and
 *     updated as the reference's lambda.
 *   * Indexers are held by value: pointers, or views whose operator[] returns
 *     a reference into the caller's storage.
 */
#ifndef SGM_DEVICE_HPP_
#define SGM_DEVICE_HPP_

#include "sgm.h"
#include "../singularity-eos_amd/csrc/sgm_approx.hpp"

#include <cstddef>
#include <type_traits>

#ifndef SGM_DEVICE_MAX_MATS
#define SGM_DEVICE_MAX_MATS 8
#endif

namespace sgm {
namespace device {

using Real = double;

/* MixParams (mixed_cell_models.hpp:43-78), the reference's defaults */
struct MixParams : MixParamsDev {
  __host__ __device__ MixParams() {
    derivative_eps = 3.0e-6;
    pte_rel_tolerance_p = 1.0e-6;
    pte_rel_tolerance_t = 1.0e-4;
    pte_rel_tolerance_e = 1.0e-6;
    pte_rel_tolerance_v = 1.0e-6;
    pte_abs_tolerance_p = 1.0e-6;
    pte_abs_tolerance_t = 1.0e-4;
    pte_slow_convergence_thresh = 0.99;
    pte_rel_tolerance_p_sufficient = 1.e2 * pte_rel_tolerance_p;
    pte_rel_tolerance_t_sufficient = 10 * pte_rel_tolerance_t;
    pte_rel_tolerance_e_sufficient = 1.e2 * pte_rel_tolerance_e;
    pte_rel_tolerance_v_sufficient = 1.e2 * pte_rel_tolerance_v;
    pte_abs_tolerance_p_sufficient = 1.e2 * pte_abs_tolerance_p;
    pte_abs_tolerance_t_sufficient = 10 * pte_abs_tolerance_t;
    line_search_alpha = 1.e-2;
    line_search_fac = 0.5;
    vfrac_safety_fac = 0.95;
    temperature_limit = 1.0e15;
    default_tguess = 300.;
    min_dtde = 1.0e-16;
    pte_small_step_thresh = 1e-16;
    pte_max_dpdv = -1e-8;
    pte_max_iter_per_mat = 128;
    line_search_max_iter = 6;
    pte_small_step_tries = 2;
    pte_reduced_system_exclude_idx = -1;
    iterate_t_guess = 1;
    pad = 0;
  }
  /* the C ABI's parameter block (sgm_mix_params_default / _tight fill it) */
  __host__ __device__ explicit MixParams(const sgm_mix_params &p) {
    static_assert(sizeof(MixParamsDev) == sizeof(sgm_mix_params), "MixParams layout");
    __builtin_memcpy(static_cast<MixParamsDev *>(this), &p, sizeof(p));
  }
};

/* SolverStatus (mixed_cell_models.hpp:80-87); err: SGM_ERRB_* bits */
struct SolverStatus {
  bool converged = false;
  bool slow_convergence_detected = false;
  std::size_t max_niter = 0;
  std::size_t max_line_niter = 0;
  std::size_t small_step_iters = 0;
  Real residual = 0;
  int err = 0;
};

/* the context's material registry (sgm_device_registry) */
struct Registry {
  const void *d_eos;
  int n;
};
__host__ __device__ inline std::size_t registry_lds_bytes(int n) {
  return (std::size_t)n * sizeof(EosDev);
}
/* all threads of the block, before any closure object is used */
__device__ inline void stage_registry(const Registry &r) {
  extern __shared__ EosDev sgm_lds_registry[];
  const double *s = reinterpret_cast<const double *>(r.d_eos);
  double *d = reinterpret_cast<double *>(sgm_lds_registry);
  const int nd = r.n * (int)(sizeof(EosDev) / 8);
  for (int i = threadIdx.x; i < nd; i += blockDim.x) d[i] = s[i];
  __syncthreads();
}

/* a LambdaIndexer with no caches (the reference's NullIndexer) */
struct NullIndexer {
  __host__ __device__ Real *operator[](std::size_t) const { return nullptr; }
};

/* ApproxTemperatureFromRhoMatU (mixed_cell_models.hpp:790-847) */
template <typename EOSIndexer, typename RealIndexer, typename LambdaIndexer = NullIndexer>
__device__ Real ApproxTemperatureFromRhoMatU(const std::size_t nmat, const EOSIndexer &eos,
                                             const Real u_tot, const RealIndexer &rho,
                                             const RealIndexer &vfrac, const Real Tguess = 0.0,
                                             const LambdaIndexer &lambda = LambdaIndexer()) {
  unsigned char ids[SGM_DEVICE_MAX_MATS];
  Real r[SGM_DEVICE_MAX_MATS], v[SGM_DEVICE_MAX_MATS];
  Lam l[SGM_DEVICE_MAX_MATS];
  const int n = nmat < SGM_DEVICE_MAX_MATS ? (int)nmat : SGM_DEVICE_MAX_MATS;
  bool any_lambda = false;
  for (int m = 0; m < n; ++m) {
    ids[m] = (unsigned char)eos[m];
    r[m] = rho[m];
    v[m] = vfrac[m];
    Real *lm = lambda[m];
    if (lm) {
      l[m].present = true;
      l[m].lRho = lm[0];
      l[m].lT = lm[1];
      any_lambda = true;
    }
  }
  const Real T = approx_temperature_from_rho_mat_u(n, ids, u_tot, r, v, Tguess,
                                                   any_lambda ? l : nullptr);
  for (int m = 0; m < n && any_lambda; ++m) {
    Real *lm = lambda[m];
    if (lm) {
      lm[0] = l[m].lRho;
      lm[1] = l[m].lT;
    }
  }
  return T;
}

namespace detail {
// closure formulation by method: 0 RhoT, 1 PT, 2 FixedT, 3 FixedP, 4 RhoU
template <int NM, int METHOD>
using System = typename std::conditional<
    METHOD == 0, PteRhoT<NM>,
    typename std::conditional<
        METHOD == 1, PtePT<NM>,
        typename std::conditional<
            METHOD == 2, PteFixedT<NM>,
            typename std::conditional<METHOD == 3, PteFixedP<NM>, PteRhoU<NM>>::type>::type>::
        type>::type;
template <class S, class = void>
struct PerMaterialT : std::false_type {};
template <class S>
struct PerMaterialT<S, std::void_t<decltype(S::kPerMaterialT)>>
    : std::integral_constant<bool, S::kPerMaterialT> {};

// A cell's closure: the caller's indexers and cell inputs. `cell_in` is
// sie_tot (RhoT, PT, RhoU), the temperature (FixedT) or the pressure (FixedP).
template <int METHOD, typename EOSIndexer, typename RealIndexer, typename LambdaIndexer>
class Closure {
 public:
  __device__ Closure(std::size_t nmat, const EOSIndexer &eos, Real vfrac_tot, Real cell_in,
                     const RealIndexer &rho, const RealIndexer &vfrac, const RealIndexer &sie,
                     const RealIndexer &temp, const RealIndexer &press,
                     const LambdaIndexer &lambda, Real Tnorm, const MixParams &params)
      : nmat_(nmat), eos_(eos), rho_(rho), vfrac_(vfrac), sie_(sie), temp_(temp),
        press_(press), lambda_(lambda), vfrac_tot_(vfrac_tot), cell_in_(cell_in),
        Tnorm_(Tnorm), params_(params) {}
  __device__ std::size_t Nmat() const { return nmat_; }
  __device__ std::size_t Niter() const { return niter_; }
  __device__ const MixParams &GetParams() const { return params_; }

  template <int NM>
  __device__ SolverStatus Run() {
    System<NM, METHOD> s;
    Real Tin[NM];
#pragma unroll
    for (int m = 0; m < NM; ++m) {
      s.eidx[m] = (unsigned char)eos_[m];
      s.rho[m] = rho_[m];
      s.vfrac[m] = vfrac_[m];
      s.press[m] = press_[m];
      Tin[m] = temp_[m];
      Real *lm = lambda_[m];
      s.lam[m].present = lm != nullptr;
      s.lam[m].lRho = lm ? lm[0] : 0.0;
      s.lam[m].lT = lm ? lm[1] : 0.0;
    }
    s.SetCellInputs(vfrac_tot_, cell_in_, Tnorm_);
    s.err = 0;
    s.P = params_;
    CellStatus cs;
    DriverState d;
    pte_begin<NM>(s, Tin, d, cs);
    while (!pte_iterate<NM>(s, d, cs)) {
    }
    pte_end<NM>(s, d, cs);
    const Real Tout = s.Tout();
#pragma unroll
    for (int m = 0; m < NM; ++m) {
      rho_[m] = s.rho[m];
      vfrac_[m] = s.vfrac[m];
      sie_[m] = s.sie[m];
      if constexpr (PerMaterialT<System<NM, METHOD>>::value)
        temp_[m] = s.TempOut(m);
      else if (System<NM, METHOD>::kStoresTemp)
        temp_[m] = Tout;
      press_[m] = s.Pout(m);
      Real *lm = lambda_[m];
      if (lm) {
        lm[0] = s.lam[m].lRho;
        lm[1] = s.lam[m].lT;
      }
    }
    niter_ = (std::size_t)cs.niter;
    SolverStatus st;
    st.converged = cs.converged != 0;
    st.slow_convergence_detected = d.slow;
    st.max_niter = (std::size_t)cs.niter;
    st.max_line_niter = (std::size_t)d.max_line_niter;
    st.small_step_iters = (std::size_t)d.small_step_iters;
    st.residual = cs.residual;
    st.err = cs.err;
    return st;
  }

 private:
  std::size_t nmat_;
  EOSIndexer eos_;
  RealIndexer rho_, vfrac_, sie_, temp_, press_;
  LambdaIndexer lambda_;
  Real vfrac_tot_, cell_in_, Tnorm_;
  MixParams params_;
  std::size_t niter_ = 0;
};
} // namespace detail

/* the reference's scratch sizes (mixed_cell_models.hpp:873-881, 1198-1206,
 * 1451-1460, 1674-1683, 1924-1932); the device solver needs none */
__host__ __device__ constexpr int PTESolverRhoTRequiredScratch(std::size_t nmat) {
  return (int)(nmat * nmat + 4 * nmat + 6 * nmat);
}
__host__ __device__ constexpr std::size_t PTESolverRhoTRequiredScratchInBytes(std::size_t nmat) {
  return PTESolverRhoTRequiredScratch(nmat) * sizeof(Real);
}
__host__ __device__ constexpr int PTESolverPTRequiredScratch(std::size_t nmat) {
  return (int)(2 * 2 + 4 * 2 + 2 * nmat);
}
__host__ __device__ constexpr std::size_t PTESolverPTRequiredScratchInBytes(std::size_t nmat) {
  return PTESolverPTRequiredScratch(nmat) * sizeof(Real);
}
__host__ __device__ constexpr std::size_t PTESolverFixedTRequiredScratch(std::size_t nmat) {
  return nmat * nmat + 4 * nmat + 2 * nmat + 2 * nmat;
}
__host__ __device__ constexpr std::size_t PTESolverFixedTRequiredScratchInBytes(std::size_t nmat) {
  return PTESolverFixedTRequiredScratch(nmat) * sizeof(Real);
}
__host__ __device__ constexpr std::size_t PTESolverFixedPRequiredScratch(std::size_t nmat) {
  return (nmat + 1) * (nmat + 1) + 4 * (nmat + 1) + 2 * nmat + 3 * nmat;
}
__host__ __device__ constexpr std::size_t PTESolverFixedPRequiredScratchInBytes(std::size_t nmat) {
  return PTESolverFixedPRequiredScratch(nmat) * sizeof(Real);
}
__host__ __device__ constexpr std::size_t PTESolverRhoURequiredScratch(std::size_t nmat) {
  return (2 * nmat) * (2 * nmat) + 4 * (2 * nmat) + 8 * nmat;
}
__host__ __device__ constexpr std::size_t PTESolverRhoURequiredScratchInBytes(std::size_t nmat) {
  return PTESolverRhoURequiredScratch(nmat) * sizeof(Real);
}

/* PTESolverRhoT(nmat, eos, vfrac_tot, sie_tot, rho, vfrac, sie, temp, press,
 * lambda, scratch, Tnorm, params) (mixed_cell_models.hpp:922-935) */
template <typename EOSIndexer, typename RealIndexer, typename LambdaIndexer>
class PTESolverRhoT : public detail::Closure<0, EOSIndexer, RealIndexer, LambdaIndexer> {
 public:
  __device__ PTESolverRhoT(std::size_t nmat, const EOSIndexer &eos, Real vfrac_tot,
                           Real sie_tot, const RealIndexer &rho, const RealIndexer &vfrac,
                           const RealIndexer &sie, const RealIndexer &temp,
                           const RealIndexer &press, const LambdaIndexer &lambda, Real *scratch,
                           Real Tnorm = 0.0, const MixParams &params = MixParams())
      : detail::Closure<0, EOSIndexer, RealIndexer, LambdaIndexer>(
            nmat, eos, vfrac_tot, sie_tot, rho, vfrac, sie, temp, press, lambda, Tnorm, params) {}
  static constexpr int RequiredScratch(std::size_t nmat) {
    return PTESolverRhoTRequiredScratch(nmat);
  }
};

/* PTESolverPT (mixed_cell_models.hpp:1245-1253) */
template <typename EOSIndexer, typename RealIndexer, typename LambdaIndexer>
class PTESolverPT : public detail::Closure<1, EOSIndexer, RealIndexer, LambdaIndexer> {
 public:
  __device__ PTESolverPT(std::size_t nmat, const EOSIndexer &eos, Real vfrac_tot, Real sie_tot,
                         const RealIndexer &rho, const RealIndexer &vfrac, const RealIndexer &sie,
                         const RealIndexer &temp, const RealIndexer &press,
                         const LambdaIndexer &lambda, Real *scratch, Real Tnorm = 0.0,
                         const MixParams &params = MixParams())
      : detail::Closure<1, EOSIndexer, RealIndexer, LambdaIndexer>(
            nmat, eos, vfrac_tot, sie_tot, rho, vfrac, sie, temp, press, lambda, Tnorm, params) {}
  static constexpr int RequiredScratch(std::size_t nmat) {
    return PTESolverPTRequiredScratch(nmat);
  }
};

/* PTESolverFixedT(nmat, eos, vfrac_tot, T_true, rho, vfrac, sie, temp, press,
 * lambda, scratch, params) (mixed_cell_models.hpp:1497-1510); temp is not
 * written, as in the reference */
template <typename EOSIndexer, typename RealIndexer, typename LambdaIndexer>
class PTESolverFixedT : public detail::Closure<2, EOSIndexer, RealIndexer, LambdaIndexer> {
 public:
  __device__ PTESolverFixedT(std::size_t nmat, const EOSIndexer &eos, Real vfrac_tot,
                             Real T_true, const RealIndexer &rho, const RealIndexer &vfrac,
                             const RealIndexer &sie, const RealIndexer &temp,
                             const RealIndexer &press, const LambdaIndexer &lambda, Real *scratch,
                             const MixParams &params = MixParams())
      : detail::Closure<2, EOSIndexer, RealIndexer, LambdaIndexer>(
            nmat, eos, vfrac_tot, T_true, rho, vfrac, sie, temp, press, lambda, 0.0, params) {}
  static constexpr std::size_t RequiredScratch(std::size_t nmat) {
    return PTESolverFixedTRequiredScratch(nmat);
  }
};

/* PTESolverFixedP(nmat, eos, vfrac_tot, P, rho, vfrac, sie, temp, press,
 * lambda, scratch, params) (mixed_cell_models.hpp:1722-1735) */
template <typename EOSIndexer, typename RealIndexer, typename LambdaIndexer>
class PTESolverFixedP : public detail::Closure<3, EOSIndexer, RealIndexer, LambdaIndexer> {
 public:
  __device__ PTESolverFixedP(std::size_t nmat, const EOSIndexer &eos, Real vfrac_tot, Real P,
                             const RealIndexer &rho, const RealIndexer &vfrac,
                             const RealIndexer &sie, const RealIndexer &temp,
                             const RealIndexer &press, const LambdaIndexer &lambda, Real *scratch,
                             const MixParams &params = MixParams())
      : detail::Closure<3, EOSIndexer, RealIndexer, LambdaIndexer>(
            nmat, eos, vfrac_tot, P, rho, vfrac, sie, temp, press, lambda, 0.0, params) {}
  static constexpr std::size_t RequiredScratch(std::size_t nmat) {
    return PTESolverFixedPRequiredScratch(nmat);
  }
};

/* PTESolverRhoU (mixed_cell_models.hpp:1973-1986) */
template <typename EOSIndexer, typename RealIndexer, typename LambdaIndexer>
class PTESolverRhoU : public detail::Closure<4, EOSIndexer, RealIndexer, LambdaIndexer> {
 public:
  __device__ PTESolverRhoU(std::size_t nmat, const EOSIndexer &eos, Real vfrac_tot, Real sie_tot,
                           const RealIndexer &rho, const RealIndexer &vfrac,
                           const RealIndexer &sie, const RealIndexer &temp,
                           const RealIndexer &press, const LambdaIndexer &lambda, Real *scratch,
                           Real Tnorm = 0.0, const MixParams &params = MixParams())
      : detail::Closure<4, EOSIndexer, RealIndexer, LambdaIndexer>(
            nmat, eos, vfrac_tot, sie_tot, rho, vfrac, sie, temp, press, lambda, Tnorm, params) {}
  static constexpr std::size_t RequiredScratch(std::size_t nmat) {
    return PTESolverRhoURequiredScratch(nmat);
  }
};

/* SolverStatus PTESolver(System &) (mixed_cell_models.hpp:2188-2291): runs
 * the closure to the end of its Newton loop and writes the results through
 * the indexers given to the constructor */
template <class System>
__device__ SolverStatus PTESolver(System &s) {
  switch (s.Nmat()) {
#define SGM_DEVICE_NM(K)                                                                       \
  case K:                                                                                      \
    if constexpr (K <= SGM_DEVICE_MAX_MATS) return s.template Run<K>();                        \
    break;
    SGM_DEVICE_NM(1)
    SGM_DEVICE_NM(2)
    SGM_DEVICE_NM(3)
    SGM_DEVICE_NM(4)
    SGM_DEVICE_NM(5)
    SGM_DEVICE_NM(6)
    SGM_DEVICE_NM(7)
    SGM_DEVICE_NM(8)
#undef SGM_DEVICE_NM
  default: break;
  }
  SolverStatus st;
  st.err = SGM_ERRB_INPUT;
  return st;
}

} // namespace device
} // namespace sgm

#endif /* SGM_DEVICE_HPP_ */
