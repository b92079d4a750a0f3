// Library-internal declarations shared by the C-ABI translation unit and the
// kernel translation units.
#pragma once

#include "../../include/sgm.h"
#include "sgm_kernels.hpp"

#include <hip/hip_runtime.h>

#include <memory>
#include <mutex>
#include <vector>

namespace sgm {

// Host-side ownership of one table: host arrays (for host-side evaluation and
// queries) and the device copies the kernels read.
struct TableHost {
  std::vector<Real> P, sie, bMod, dEdT, dPdE;
  std::vector<Real> PCold, sieCold, bModCold, dPdECold, dEdTCold, lTColdCrit;
  std::vector<Real> PMax, sielTMax, dEdTMax, gm1Max, rho_at_pmin;
  TableDev hdev{};       // pointers into the host vectors above
  TableDev ddev{};       // the same record with pointers to the device arrays
  TableDev *d_tab = nullptr;  // device copy of ddev
  std::vector<void *> d_arrays;
};

struct VinetHost {
  std::vector<Real> coef; // d2..d40 (39) then VIP[0..42] (43)
  Real *d_coef = nullptr;
};

} // namespace sgm

struct sgm_ctx {
  int device = 0;
  bool host_only = false; // sgm_init_host: no device copies, launches refused
  int nmat = 0;
  std::vector<sgm::EosDev> host_eos; // device-pointer view (what kernels get)
  std::vector<sgm::EosDev> host_eval; // host-pointer view (for host-side evaluation)
  std::vector<int> is_set;
  std::vector<std::unique_ptr<sgm::TableHost>> tables;
  std::vector<std::unique_ptr<sgm::VinetHost>> vinets;
  sgm::EosDev *d_eos = nullptr;
  bool dirty = true;
  std::mutex mu;
  // get_sg_eos keeps the batched plan of its last cell topology (CSR offsets
  // and material ids of the participating entries) and reuses it while the
  // topology repeats; drv_mu serializes the driver calls that use it
  sgm_plan *drv_plan = nullptr;
  void *drv_cells = nullptr, *drv_entries = nullptr;  // its grow-only work buffers
  size_t drv_cells_bytes = 0, drv_entries_bytes = 0;
  long drv_plan_builds = 0, drv_plan_reuses = 0;  // calls that built / reused the plan
  std::mutex drv_mu;
};

namespace sgm {
// launchers (kernels translation units)
hipError_t launch_vec(const EosDev &e_host, const EosDev *d_eos, int mat, int func,
                      const double *in1, const double *in2, double *out, long n, int lstride,
                      double *lambda, hipStream_t s);
hipError_t launch_fill(const EosDev &e_host, const EosDev *d_eos, int mat, double *rho,
                       double *temp, double *sie, double *press, double *cv, double *bmod,
                       long n, unsigned long output, int lstride, double *lambda, int *errs,
                       hipStream_t s);
hipError_t launch_rhoe_PT(const EosDev &e_host, const EosDev *d_eos, int mat, const double *P,
                          const double *T, double *rho, double *sie, long n, int lstride,
                          double *lambda, int *errs, hipStream_t s);

// copies the host registry to the device when it changed (sgm_capi.cpp)
int sync_registry(sgm_ctx *ctx);
// bit i: registry entry i has modifiers (the batched bins leave its cells to
// the list kernels)
inline unsigned long long ctx_mod_mask(const sgm_ctx *ctx) {
  unsigned long long m = 0;
  for (int i = 0; i < ctx->nmat && i < 64; ++i)
    if (ctx->host_eos[i].mods) m |= 1ull << i;
  return m;
}
// MixParams as the kernels take them (NULL = the reference's defaults)
MixParamsDev to_dev(const sgm_mix_params *p);
// A batched plan over a CSR topology (sgm_plan_create) whose cells with fewer
// than min_nm materials are left out of every bin, and a direct (uncaptured)
// launch of all its bins on stream s. Used by the get_sg_eos driver, which
// solves only its cells of two or more materials.
int plan_create(sgm_ctx *ctx, long ncell, long nnz, const int *offsets, const uint8_t *mat_ids,
                int method, int min_nm, sgm_plan **out, hipStream_t s);
// match = the plan was made for this method and exactly this CSR topology
// (a device comparison; synchronizes s)
int plan_match(sgm_plan *p, long ncell, long nnz, const int *offsets, const uint8_t *mat_ids,
               int method, hipStream_t s, bool &match);
int plan_launch(sgm_plan *p, double *rho, double *vfrac, double *sie, double *temp,
                double *press, const double *vfrac_tot, const double *sie_tot, double *lambda,
                const sgm_mix_params *params, sgm_pte_status status, hipStream_t s);
// Gruneisen::ComputeRhoMax (eos_gruneisen.hpp:231-329), host side
Real gruneisen_compute_rho_max(Real s1, Real s2, Real s3, Real rho0);
} // namespace sgm
