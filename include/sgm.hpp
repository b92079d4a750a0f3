// C++ adapter over the C ABI of sgm.h, mirroring the singularity-eos host
// API for the vector EOS calls and the mixed-cell closure so that a host code
// changes a namespace and a handle type, not its call sites:
//
//   singularity::EOS eos = singularity::IdealGas(gm1, Cv);        (reference)
//   eos.PressureFromDensityInternalEnergy(rhos, sies, pressures, num, lambdas);
//
//   sgm::Registry reg(nmat);                                    (this adapter)
//   sgm::EOS eos = reg.IdealGas(m, gm1, Cv);
//   eos.PressureFromDensityInternalEnergy(rhos, sies, pressures, num, lambdas);
//
// Arrays are device pointers (HIP memory); calls are asynchronous on the
// handle's stream (Exec::Device of the reference). Errors throw sgm::Error
// (the reference aborts or throws via EOS_ERROR, base/eos_error.hpp:19-20).
// The closure is the batched form of PTESolverRhoT / PTESolverPT +
// PTESolver (mixed_cell_models.hpp:922-935, 1245-1253, 2188-2291): one
// object covers every cell of a CSR list instead of one object per cell.
#pragma once

#include "sgm.h"

#include <stdexcept>
#include <string>

namespace sgm {

struct Error : std::runtime_error {
  int code;
  Error(const std::string &what, int c)
      : std::runtime_error(what + " failed with code " + std::to_string(c)), code(c) {}
};

inline void check(int rc, const char *what) {
  if (rc != SGM_OK) throw Error(what, rc);
}

// thermalqs (base/constants.hpp:22-35)
namespace thermalqs {
constexpr unsigned long none = 0;
constexpr unsigned long density = SGM_Q_DENSITY;
constexpr unsigned long specific_internal_energy = SGM_Q_SIE;
constexpr unsigned long pressure = SGM_Q_PRESSURE;
constexpr unsigned long temperature = SGM_Q_TEMPERATURE;
constexpr unsigned long specific_heat = SGM_Q_CV;
constexpr unsigned long bulk_modulus = SGM_Q_BMOD;
} // namespace thermalqs

// MixParams (mixed_cell_models.hpp:43-78) with the reference's defaults
struct MixParams : sgm_mix_params {
  MixParams() { sgm_mix_params_default(this); }
};

// the lambda side arrays of the vector calls: lambda[i*stride + k]
// (get_sg_PressureFromDensityInternalEnergy's lambda_data/stride,
// singularity_eos.cpp:39-49, lambdaIndexer2D)
struct Lambdas {
  double *data = nullptr;
  int stride = 0;
};

class EOS {
 public:
  EOS() = default;
  EOS(sgm_ctx *ctx, int m, sgm_stream_t s = nullptr) : ctx_(ctx), m_(m), s_(s) {}
  int matindex() const { return m_; }
  void SetStream(sgm_stream_t s) { s_ = s; }

  // SG_VARIANT_VEC_2IN_1OUT calls (eos_variant.hpp:56-149, 543-560)
#define SGM_VEC(NAME, SEL)                                                                   \
  void NAME(const double *in1, const double *in2, double *out, int num) const {              \
    check(sgm_vec_eval(ctx_, m_, SEL, in1, in2, out, num, 0, nullptr, s_), #NAME);           \
  }                                                                                          \
  void NAME(const double *in1, const double *in2, double *out, int num, Lambdas lam) const { \
    check(sgm_vec_eval(ctx_, m_, SEL, in1, in2, out, num, lam.stride, lam.data, s_), #NAME); \
  }
  SGM_VEC(PressureFromDensityInternalEnergy, SGM_F_P_RHO_SIE)
  SGM_VEC(TemperatureFromDensityInternalEnergy, SGM_F_T_RHO_SIE)
  SGM_VEC(InternalEnergyFromDensityTemperature, SGM_F_SIE_RHO_T)
  SGM_VEC(PressureFromDensityTemperature, SGM_F_P_RHO_T)
  SGM_VEC(SpecificHeatFromDensityInternalEnergy, SGM_F_CV_RHO_SIE)
  SGM_VEC(SpecificHeatFromDensityTemperature, SGM_F_CV_RHO_T)
  SGM_VEC(BulkModulusFromDensityInternalEnergy, SGM_F_BMOD_RHO_SIE)
  SGM_VEC(BulkModulusFromDensityTemperature, SGM_F_BMOD_RHO_T)
#undef SGM_VEC

  // FillEos (eos_variant.hpp:652-708)
  void FillEos(double *rhos, double *temps, double *energies, double *presses, double *cvs,
               double *bmods, int num, unsigned long output, Lambdas lam = {},
               int *errs = nullptr) const {
    check(sgm_vec_fill_eos(ctx_, m_, rhos, temps, energies, presses, cvs, bmods, num, output,
                           lam.stride, lam.data, errs, s_),
          "FillEos");
  }
  // DensityEnergyFromPressureTemperature over arrays (eos_base.hpp:630-671)
  void DensityEnergyFromPressureTemperature(const double *presses, const double *temps,
                                            double *rhos, double *sies, int num,
                                            Lambdas lam = {}, int *errs = nullptr) const {
    check(sgm_vec_rhoe_from_PT(ctx_, m_, presses, temps, rhos, sies, num, lam.stride, lam.data,
                               errs, s_),
          "DensityEnergyFromPressureTemperature");
  }

 private:
  sgm_ctx *ctx_ = nullptr;
  int m_ = -1;
  sgm_stream_t s_ = nullptr;
};

// Material registry on the current GPU: init_sg_eos / init_sg_* /
// finalize_sg_eos (singularity_eos.hpp:28-134) with owned lifetime.
class Registry {
 public:
  explicit Registry(int nmat, bool host_only = false) {
    check(host_only ? sgm_init_host(nmat, &ctx_) : sgm_init(nmat, &ctx_), "init");
  }
  ~Registry() {
    if (ctx_) sgm_finalize(ctx_);
  }
  Registry(const Registry &) = delete;
  Registry &operator=(const Registry &) = delete;
  sgm_ctx *get() const { return ctx_; }

  EOS IdealGas(int m, double gm1, double Cv) {
    check(sgm_set_IdealGas(ctx_, m, gm1, Cv), "IdealGas");
    return EOS(ctx_, m);
  }
  EOS StiffGas(int m, double gm1, double Cv, double Pinf, double qq) {
    check(sgm_set_StiffGas(ctx_, m, gm1, Cv, Pinf, qq), "StiffGas");
    return EOS(ctx_, m);
  }
  EOS Gruneisen(int m, double C0, double s1, double s2, double s3, double G0, double b,
                double rho0, double T0, double P0, double Cv, double rho_max = 0) {
    check(sgm_set_Gruneisen(ctx_, m, C0, s1, s2, s3, G0, b, rho0, T0, P0, Cv, rho_max),
          "Gruneisen");
    return EOS(ctx_, m);
  }
  EOS Vinet(int m, double rho0, double T0, double B0, double BP0, double A0, double Cv0,
            double E0, double S0, const double *d2to40) {
    check(sgm_set_Vinet(ctx_, m, rho0, T0, B0, BP0, A0, Cv0, E0, S0, d2to40), "Vinet");
    return EOS(ctx_, m);
  }
  EOS SpinerEOSDependsRhoT(int m, const sgm_table_desc &t) {
    check(sgm_set_TableRhoT(ctx_, m, &t), "SpinerEOSDependsRhoT");
    return EOS(ctx_, m);
  }

 private:
  sgm_ctx *ctx_ = nullptr;
};

// The closure over a CSR list of cells (one PTESolverRhoT / PTESolverPT per
// cell in the reference). Arrays as in the reference's constructor, per
// (cell, material) entry; vfrac_tot, sie_tot, Tnorm per cell.
enum class Method { RhoT = 0, PT = 1 };

struct CellArrays {
  long ncell = 0, nnz = 0;
  const int *offsets = nullptr;
  const uint8_t *mat_ids = nullptr;
  double *rho = nullptr, *vfrac = nullptr, *sie = nullptr, *temp = nullptr, *press = nullptr;
  const double *vfrac_tot = nullptr, *sie_tot = nullptr, *Tnorm = nullptr;
  double *lambda = nullptr;
};

// PTESolver(system) over every cell: the per-cell call, batched into one launch
inline void PTESolver(const Registry &reg, Method method, const CellArrays &c,
                      const MixParams &params = MixParams(), sgm_pte_status status = {},
                      sgm_stream_t s = nullptr) {
  const auto f = method == Method::RhoT ? sgm_pte_rhoT : sgm_pte_PT;
  check(f(reg.get(), c.ncell, c.nnz, c.offsets, c.mat_ids, c.rho, c.vfrac, c.sie, c.temp,
          c.press, c.vfrac_tot, c.sie_tot, c.Tnorm, c.lambda, &params, status, s),
        "PTESolver");
}

// The batched, graph-captured form for a fixed cell topology (no reference
// counterpart): build once, Solve every step.
class PTEPlan {
 public:
  PTEPlan(const Registry &reg, Method method, long ncell, long nnz, const int *offsets,
          const uint8_t *mat_ids, sgm_stream_t s = nullptr) {
    check(sgm_plan_create(reg.get(), ncell, nnz, offsets, mat_ids, (int)method, &p_, s),
          "plan_create");
  }
  ~PTEPlan() {
    if (p_) sgm_plan_destroy(p_);
  }
  PTEPlan(const PTEPlan &) = delete;
  PTEPlan &operator=(const PTEPlan &) = delete;
  void Solve(const CellArrays &c, const MixParams &params = MixParams(),
             sgm_pte_status status = {}, sgm_stream_t s = nullptr) {
    check(sgm_plan_execute(p_, c.rho, c.vfrac, c.sie, c.temp, c.press, c.vfrac_tot, c.sie_tot,
                           c.Tnorm, c.lambda, &params, status, s),
          "plan_execute");
  }

 private:
  sgm_plan *p_ = nullptr;
};

} // namespace sgm
