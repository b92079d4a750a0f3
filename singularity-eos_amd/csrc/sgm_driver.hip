// get_sg_eos: the closure driver a hydro code calls once per step
// (singularity-eos/eos/get_sg_eos.cpp:35-280 with get_sg_eos_functors.hpp and
// the per-mode bodies get_sg_eos_{rho_e,rho_t,rho_p,p_t}.cpp). On the device
// it is four passes over the cells around the closure kernels:
//
//  1. prologue: normalise each cell's mass fractions in place, apply the
//     mass-fraction cutoff, count the participating materials (the reference's
//     init_functor);
//  2. an exclusive scan of the counts gives a CSR layout of the participating
//     (cell, material) entries;
//  3. fill: the PTE inputs of every entry (densities at the cell's specific
//     volume, the mode's temperature / pressure / energy guesses, zeroed
//     lambda caches);
//  4. the batched PTE path (RhoT, FixedT or FixedP by mode) over the cells of
//     two or more materials: a plan bins them by material count and
//     signature and launches the register-resident persistent kernels (with
//     straggler parking) on the caller's stream; the context keeps the plan
//     and reuses it while the next calls bring the same topology (an exact
//     device comparison of the CSR offsets and material ids); then the
//     epilogue:
//     the single-material paths, the P-T mode's
//     per-material inversions, and the reference's final_functor (cell
//     averages, per-material bulk modulus, heat capacity, Gruneisen
//     parameter, volume and energy fractions, pmax).
//
// Per-material arrays are column-major [cell_dim x nmat] as the reference's
// Kokkos LayoutLeft views: x(i, m) = x[i + m * cell_dim]. The lambda cache of
// each entry has storage of its own (the reference places it inside the
// solver scratch, see sgm.h).
#include "sgm_internal.hpp"

#include <hipcub/hipcub.hpp>

#include <cstring>
#include <mutex>
#include <vector>

using namespace sgm;

namespace {

constexpr Real EV2K = 1.160451930280894026e4;  // get_sg_eos.hpp:68
constexpr int MAX_DRIVER_MATS = 64;

enum { IN_RHO_T = -3, IN_RHO_P = -2, IN_P_T = -1, IN_NORM_RHO_E = 0, IN_RHO_E = 1 };

struct DrvArgs {
  int nmat = 0, ncell = 0, input = 1;
  long cd = 0;
  int eos_idx[MAX_DRIVER_MATS] = {};  // context material index per material slot
  const int *offsets = nullptr;       // 1-based cell index per loop index
  double *press, *pmax, *vol, *spvol, *sie, *temp, *bmod, *dpde, *cv;
  double *fm, *fv, *fi, *fb, *fd, *fc;
  double cutoff = 0;
  const EosDev *d_eos = nullptr;
  int nreg = 0;
  // per loop index
  int *npte = nullptr;
  double *mass_sum = nullptr;
  int *eoff = nullptr;  // [ncell + 1] entry offsets
  double *cell_in = nullptr, *vtot = nullptr;
  int *conv = nullptr;  // optional: PTE convergence per loop index
  // per entry
  int *emat = nullptr;
  uint8_t *eid = nullptr;
  double *rho = nullptr, *vfrac = nullptr, *esie = nullptr, *etemp = nullptr,
         *epress = nullptr, *lam = nullptr;
};

__device__ __forceinline__ Real mults_t(int in) { return in == IN_RHO_T ? 1.0 : 0.0; }
__device__ __forceinline__ Real mults_p(int in) { return in == IN_RHO_P ? 1.0 : 0.0; }
__device__ __forceinline__ Real mults_s(int in) {
  return (in == IN_RHO_E || in == IN_NORM_RHO_E) ? 1.0 : 0.0;
}

// init_functor, first half (get_sg_eos_functors.hpp:64-100); P-T mode:
// get_sg_eos_p_t.cpp:40-58
__global__ void prologue_kernel(DrvArgs a) {
  const int il = blockIdx.x * blockDim.x + threadIdx.x;
  if (il >= a.ncell) return;
  const long i = a.offsets[il] - 1;
  const long cd = a.cd;
  if (i < 0 || i >= cd) {  // not a cell of the arrays: no materials, left untouched
    a.npte[il] = 0;
    a.mass_sum[il] = 1.0;
    return;
  }
  Real ms = 0.0;
  int n = 0;
  if (a.input != IN_P_T) {
    const Real s_mult = mults_s(a.input);
    for (int m = 0; m < a.nmat; ++m) {
      ms += a.fm[i + m * cd];
      a.fv[i + m * cd] = a.fm[i + m * cd];
      a.fi[i + m * cd] = a.sie[i] * a.fm[i + m * cd] * s_mult;
    }
    for (int m = 0; m < a.nmat; ++m) a.fm[i + m * cd] /= ms;
    for (int m = 0; m < a.nmat; ++m) n += a.fm[i + m * cd] > a.cutoff ? 1 : 0;
  } else {
    for (int m = 0; m < a.nmat; ++m) ms += a.fm[i + m * cd];
    for (int m = 0; m < a.nmat; ++m) a.fm[i + m * cd] /= ms;
    n = a.nmat;
  }
  a.npte[il] = n;
  a.mass_sum[il] = ms;
}

// init_functor, second half (:101-113): the PTE inputs of each entry
__global__ void fill_kernel(DrvArgs a) {
  const int il = blockIdx.x * blockDim.x + threadIdx.x;
  if (il >= a.ncell) return;
  const long i = a.offsets[il] - 1;
  const long cd = a.cd;
  const int o = a.eoff[il];
  const int n = a.npte[il];
  if (a.conv) a.conv[il] = (i >= 0 && i < cd) ? 1 : 0;
  if (i < 0 || i >= cd) {
    a.vtot[il] = 0;
    a.cell_in[il] = 0;
    return;
  }
  if (a.input == IN_P_T) {
    for (int m = 0; m < a.nmat; ++m) {
      const int e = o + m;
      a.emat[e] = m;
      a.eid[e] = (uint8_t)a.eos_idx[m];
      a.etemp[e] = a.temp[i] * EV2K;
      a.epress[e] = a.press[i];
      a.rho[e] = 0.0;  // density guess of the P-T inversions
      a.esie[e] = 0.0;
      a.lam[3L * e] = a.lam[3L * e + 1] = a.lam[3L * e + 2] = 0.0;
    }
    return;
  }
  const Real t_mult = mults_t(a.input), p_mult = mults_p(a.input), s_mult = mults_s(a.input);
  Real vsum = 0.;
  int mp = 0;
  for (int m = 0; m < a.nmat; ++m) {
    const Real f = a.fm[i + m * cd];
    if (!(f > a.cutoff)) continue;
    const int e = o + mp;
    a.emat[e] = m;
    a.eid[e] = (uint8_t)a.eos_idx[m];
    a.vfrac[e] = f;
    vsum += a.vfrac[e];
    a.rho[e] = f / a.spvol[i] / a.vfrac[e];
    a.etemp[e] = a.temp[i] * EV2K * t_mult;
    a.epress[e] = a.press[i] * p_mult;
    a.esie[e] = a.sie[i] * f * s_mult;
    a.lam[3L * e] = a.lam[3L * e + 1] = a.lam[3L * e + 2] = 0.0;
    ++mp;
  }
  a.vtot[il] = vsum;
  // the PTE cell input: the cell energy (rho-e), the first entry's
  // temperature (rho-T) or pressure (rho-P), as the reference passes them
  a.cell_in[il] = a.input == IN_RHO_T ? (n > 0 ? a.etemp[o] : 0.0)
                  : a.input == IN_RHO_P ? (n > 0 ? a.epress[o] : 0.0)
                                        : a.sie[i];
}

// the single-material paths, the P-T mode's inversions and final_functor
// (get_sg_eos_functors.hpp:206-282)
__global__ void epilogue_kernel(DrvArgs a) {
  const int il = blockIdx.x * blockDim.x + threadIdx.x;
  if (il >= a.ncell) return;
  const long i = a.offsets[il] - 1;
  const long cd = a.cd;
  if (i < 0 || i >= cd) return;
  const int o = a.eoff[il];
  const int n = a.npte[il];
  const Real ms = a.mass_sum[il];
  auto lam_of = [&](int e) {
    Lam l;
    l.present = true;
    l.lRho = a.lam[3L * e];
    l.lT = a.lam[3L * e + 1];
    return l;
  };
  Real ft = 0, fs = 0, fp = 0;
  if (a.input == IN_RHO_E || a.input == IN_NORM_RHO_E) {
    if (n == 1) {  // get_sg_eos_rho_e.cpp:73-80
      const EosDev &e = a.d_eos[a.eid[o]];
      Lam l = lam_of(o);
      a.etemp[o] = eos_T_re(e, a.rho[o], a.esie[o], l);
      a.epress[o] = eos_P_rT(e, a.rho[o], a.etemp[o], l);
      a.lam[3L * o] = l.lRho;
      a.lam[3L * o + 1] = l.lT;
    }
    ft = 1.0, fs = 0.0, fp = 1.0;
  } else if (a.input == IN_RHO_T || a.input == IN_RHO_P) {
    Real sie_true = 0.0;
    if (n > 1) {  // get_sg_eos_rho_t.cpp:75-79, get_sg_eos_rho_p.cpp:73-77
      for (int mp = 0; mp < n; ++mp) sie_true += a.esie[o + mp] * a.fm[i + a.emat[o + mp] * cd];
    } else if (n == 1) {
      const EosDev &e = a.d_eos[a.eid[o]];
      Lam l = lam_of(o);
      if (a.input == IN_RHO_T) {  // get_sg_eos_rho_t.cpp:80-91
        a.esie[o] = eos_e_rT(e, a.rho[o], a.etemp[o], l);
        sie_true = a.esie[o];
        a.epress[o] = eos_P_rT(e, a.rho[o], a.etemp[o], l);
      } else {  // get_sg_eos_rho_p.cpp:78-101
        const Real r = a.rho[o];
        auto p_of_t = [&](Real t) { return eos_P_rT(e, r, t, l); };
        const Real r_temp = eos_T_reference(e);
        Real temp_i = 0;
        root::find_root(p_of_t, a.press[i], r_temp, 1.e-10 * r_temp, 1.e10 * r_temp, 1.e-12,
                        1.e-12, temp_i);
        a.esie[o] = eos_e_rT(e, r, temp_i, l);
        sie_true = a.esie[o];
        a.etemp[o] = temp_i;
      }
      a.lam[3L * o] = l.lRho;
      a.lam[3L * o + 1] = l.lT;
    }
    a.sie[i] = sie_true;
    if (a.input == IN_RHO_T)
      ft = 0.0, fs = 0.0, fp = 1.0;
    else
      ft = 1.0, fs = 0.0, fp = 0.0;
  } else {  // P-T: get_sg_eos_p_t.cpp:59-90
    Real vt = 0.0, st = 0.0;
    int err = 0;
    for (int m = 0; m < n; ++m) {
      const int e = o + m;
      Lam l = lam_of(e);
      eos_rhoe_PT(a.d_eos[a.eid[e]], a.epress[e], a.etemp[e], a.rho[e], a.esie[e], l, err);
      a.lam[3L * e] = l.lRho;
      a.lam[3L * e + 1] = l.lT;
      a.vfrac[e] = a.fm[i + m * cd] / a.rho[e] * ms;
      vt += a.vfrac[e];
      st += a.esie[e] * a.fm[i + m * cd];
    }
    a.sie[i] = st;
    a.vol[i] = vt;
    a.spvol[i] = a.vol[i] / ms;
    for (int m = 0; m < n; ++m) a.vfrac[o + m] /= vt;
  }
  // final_functor
  const bool do_t = ft == 1.0, do_s = fs == 1.0, do_p = fp == 1.0;
  if (do_t) a.temp[i] = 0.0;
  if (do_p) a.press[i] = 0.0;
  if (do_s) a.sie[i] = 0.0;
  a.bmod[i] = 0.0;
  a.cv[i] = 0.0;
  a.dpde[i] = 0.0;
  for (int mp = 0; mp < n; ++mp) {
    const int e = o + mp;
    const int m = a.emat[e];
    const EosDev &eo = a.d_eos[a.eid[e]];
    Lam l = lam_of(e);
    a.press[i] += a.epress[e] * a.vfrac[e] * fp;
    a.temp[i] += a.etemp[e] * a.vfrac[e] * ft;
    const Real ie_m = a.esie[e] * a.fm[i + m * cd] * ms;
    a.sie[i] += ie_m * fs;
    a.fi[i + m * cd] = ie_m;
    a.fv[i + m * cd] = a.vfrac[e] * a.vol[i];
    const Real bmod_m = eos_bmod_rT(eo, a.rho[e], a.etemp[e], l);
    a.bmod[i] += bmod_m * a.vfrac[e];
    const Real cv_m = EV2K * eos_cv_rT(eo, a.rho[e], a.etemp[e], l);
    a.cv[i] += cv_m * a.fm[i + m * cd];
    const Real dpde_m = eos_gruneisen_rT(eo, a.rho[e], a.etemp[e], l);
    a.dpde[i] += dpde_m * a.vfrac[e];
    if (a.fb) a.fb[i + m * cd] = bmod_m;
    if (a.fd) a.fc[i + m * cd] = cv_m;  // written when frac_dpde is given (get_sg_eos.cpp:61-62)
    if (a.fd) a.fd[i + m * cd] = dpde_m;
  }
  if (do_t) a.temp[i] /= EV2K;
  if (do_s) a.sie[i] /= ms;
  for (int m = 0; m < a.nmat; ++m) a.fm[i + m * cd] *= ms;
  a.pmax[i] = a.press[i] > a.pmax[i] ? a.press[i] : a.pmax[i];
}

// Carves typed arrays (256-byte aligned) out of a grow-only device buffer:
// take() records the layout, ok() grows the buffer when the layout does not
// fit (after the stream's earlier work, which may still use it) and assigns
// the pointers.
struct Arena {
  void *&buf;
  size_t &cap;
  size_t used = 0;
  std::vector<std::pair<void **, size_t>> parts;
  Arena(void *&b, size_t &c) : buf(b), cap(c) {}
  template <class T>
  void take(T *&p, size_t n) {
    p = nullptr;
    parts.push_back({(void **)&p, used});
    used += (sizeof(T) * (n ? n : 1) + 255) / 256 * 256;
  }
  bool ok(hipStream_t s) {
    if (used > cap) {
      if (hipStreamSynchronize(s) != hipSuccess) return false;
      if (buf) (void)hipFree(buf);
      buf = nullptr;
      cap = 0;
      const size_t want = used + used / 4;
      if (hipMalloc(&buf, want) != hipSuccess) return false;
      cap = want;
    }
    for (auto &q : parts) *q.first = static_cast<unsigned char *>(buf) + q.second;
    return true;
  }
};

// argument checks and the kernels' argument record, shared by both entry points
int driver_setup(sgm_ctx *ctx, int nmat, int ncell, int cell_dim, int input_int,
                 const int *eos_offsets, const int *offsets, double *press, double *pmax,
                 double *vol, double *spvol, double *sie, double *temp, double *bmod,
                 double *dpde, double *cv, double *frac_mass, double *frac_vol, double *frac_ie,
                 double *frac_bmod, double *frac_dpde, double *frac_cv,
                 double mass_frac_cutoff, int *pte_converged, DrvArgs &a) {
  if (!ctx || nmat < 1 || nmat > MAX_DRIVER_MATS || ncell < 0 || cell_dim < ncell ||
      input_int < IN_RHO_T || input_int > IN_RHO_E || !eos_offsets || !offsets || !press ||
      !pmax || !vol || !spvol || !sie || !temp || !bmod || !dpde || !cv || !frac_mass ||
      !frac_vol || !frac_ie || (frac_dpde && !frac_cv))
    return SGM_E_ARG;
  if (ctx->host_only) return SGM_E_NOGPU;
  a.nmat = nmat;
  a.ncell = ncell;
  a.input = input_int;
  a.cd = cell_dim;
  for (int m = 0; m < nmat; ++m) {
    a.eos_idx[m] = eos_offsets[m] - 1;
    if (a.eos_idx[m] < 0 || a.eos_idx[m] >= ctx->nmat || !ctx->is_set[a.eos_idx[m]])
      return SGM_E_MAT;
  }
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
  a.fm = frac_mass;
  a.fv = frac_vol;
  a.fi = frac_ie;
  a.fb = frac_bmod;
  a.fd = frac_dpde;
  a.fc = frac_cv;
  a.cutoff = mass_frac_cutoff;
  a.conv = pte_converged;
  {
    std::lock_guard<std::mutex> g(ctx->mu);
    if (sync_registry(ctx) != SGM_OK) return SGM_E_HIP;
  }
  a.d_eos = ctx->d_eos;
  a.nreg = ctx->nmat;
  return SGM_OK;
}

}  // namespace

extern "C" int sgm_get_sg_eos(sgm_ctx *ctx, int nmat, int ncell, int cell_dim, int input_int,
                              const int *eos_offsets, const int *offsets, double *press,
                              double *pmax, double *vol, double *spvol, double *sie,
                              double *temp, double *bmod, double *dpde, double *cv,
                              double *frac_mass, double *frac_vol, double *frac_ie,
                              double *frac_bmod, double *frac_dpde, double *frac_cv,
                              double mass_frac_cutoff, const sgm_mix_params *params,
                              int *pte_converged, sgm_stream_t stream) {
  DrvArgs a;
  const int rc0 = driver_setup(ctx, nmat, ncell, cell_dim, input_int, eos_offsets, offsets,
                               press, pmax, vol, spvol, sie, temp, bmod, dpde, cv, frac_mass,
                               frac_vol, frac_ie, frac_bmod, frac_dpde, frac_cv,
                               mass_frac_cutoff, pte_converged, a);
  if (rc0 != SGM_OK) return rc0;
  if (ncell == 0) return SGM_OK;
  hipStream_t s = (hipStream_t)stream;
  // the driver's buffers and plan live in the context (grow-only arenas,
  // reused by the next call); one call at a time per context
  std::unique_lock<std::mutex> lock(ctx->drv_mu);
  auto cleanup = [&](int rc) {
    hipStreamSynchronize(s);
    return rc;
  };
  const int nb = (ncell + 255) / 256;
  size_t scan_bytes = 0;
  if (hipcub::DeviceScan::ExclusiveSum(nullptr, scan_bytes, (int *)nullptr, (int *)nullptr,
                                       ncell + 1, s) != hipSuccess)
    return SGM_E_HIP;
  {
    Arena ar(ctx->drv_cells, ctx->drv_cells_bytes);
    ar.take(a.npte, ncell + 1);
    ar.take(a.mass_sum, ncell);
    ar.take(a.eoff, ncell + 1);
    ar.take(a.cell_in, ncell);
    ar.take(a.vtot, ncell);
    unsigned char *scan_tmp = nullptr;
    ar.take(scan_tmp, scan_bytes + 16);
    if (!ar.ok(s)) return cleanup(SGM_E_HIP);
    if (launch_zero_ints(a.npte + ncell, 1, s) != hipSuccess) return cleanup(SGM_E_HIP);
    prologue_kernel<<<nb, 256, 0, s>>>(a);
    if (hipGetLastError() != hipSuccess) return cleanup(SGM_E_HIP);
    if (hipcub::DeviceScan::ExclusiveSum(scan_tmp, scan_bytes, a.npte, a.eoff, ncell + 1, s) !=
        hipSuccess)
      return cleanup(SGM_E_HIP);
  }
  int nnz = 0;
  if (hipMemcpyAsync(&nnz, a.eoff + ncell, sizeof(int), hipMemcpyDeviceToHost, s) !=
          hipSuccess ||
      hipStreamSynchronize(s) != hipSuccess)
    return cleanup(SGM_E_HIP);
  {
    Arena ar(ctx->drv_entries, ctx->drv_entries_bytes);
    ar.take(a.emat, nnz);
    ar.take(a.eid, nnz);
    ar.take(a.rho, nnz);
    ar.take(a.vfrac, nnz);
    ar.take(a.esie, nnz);
    ar.take(a.etemp, nnz);
    ar.take(a.epress, nnz);
    ar.take(a.lam, 3L * nnz);
    if (!ar.ok(s)) return cleanup(SGM_E_HIP);
  }
  fill_kernel<<<nb, 256, 0, s>>>(a);
  if (hipGetLastError() != hipSuccess) return cleanup(SGM_E_HIP);
  // the context's plan for the last topology, reused while it repeats (a
  // hydro code's mixed cells change little from one step to the next)
  if (input_int != IN_P_T && nnz > 0) {
    const int method = input_int == IN_RHO_T   ? PTE_FIXED_T
                       : input_int == IN_RHO_P ? PTE_FIXED_P
                                               : PTE_RHOT;
    int rc = SGM_OK;
    if (ctx->drv_plan) {
      bool same = false;
      rc = plan_match(ctx->drv_plan, ncell, nnz, a.eoff, a.eid, method, s, same);
      if (rc != SGM_OK) return cleanup(rc);
      if (!same) {
        sgm_plan_destroy(ctx->drv_plan);
        ctx->drv_plan = nullptr;
      } else {
        ++ctx->drv_plan_reuses;
      }
    }
    if (!ctx->drv_plan) {
      rc = plan_create(ctx, ncell, nnz, a.eoff, a.eid, method, 2, &ctx->drv_plan, s);
      if (rc != SGM_OK) return cleanup(rc);
      ++ctx->drv_plan_builds;
    }
    sgm_pte_status st{a.conv, nullptr, nullptr, nullptr};
    rc = plan_launch(ctx->drv_plan, a.rho, a.vfrac, a.esie, a.etemp, a.epress, a.vtot,
                     a.cell_in, a.lam, params, st, s);
    if (rc != SGM_OK) return cleanup(rc);
  }
  epilogue_kernel<<<nb, 256, 0, s>>>(a);
  if (hipGetLastError() != hipSuccess) return cleanup(SGM_E_HIP);
  return cleanup(SGM_OK);
}

// ---- get_sg_eos without allocation -------------------------------------------
// sgm_get_sg_eos_noalloc: the driver's work arrays live in one static device
// buffer per GPU (`scratch`, NOALLOC_SCRATCH_BYTES, held until the process
// ends), carved per call for the worst case of ncell x nmat participating
// entries, so no size has to come back to the host; the closure runs on the
// plain list kernel over the cells of two or more materials (the same
// results as the batched plan, which would need allocations of its own).
namespace {
constexpr size_t NOALLOC_SCRATCH_BYTES = size_t(64) << 20;
constexpr int NOALLOC_MAX_DEVICES = 64;
__device__ __align__(256) unsigned char scratch[NOALLOC_SCRATCH_BYTES];
std::mutex g_scratch_mu;
void *g_scratch_ptr[NOALLOC_MAX_DEVICES] = {};

// The device address of `scratch` on the current device: looked up once per
// device (each GPU holds its own copy of the array) with that device
// current. Host code uses this address, never the array's name.
int scratch_address(void **out) {
  int dev = -1;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= NOALLOC_MAX_DEVICES)
    return SGM_E_HIP;
  std::lock_guard<std::mutex> g(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void *p = nullptr;
    if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) return SGM_E_HIP;
    g_scratch_ptr[dev] = p;
  }
  *out = g_scratch_ptr[dev];
  return SGM_OK;
}

// 256-byte aligned arrays out of a fixed buffer; ok() once every take fits
struct FixedArena {
  unsigned char *base;
  size_t cap, used = 0;
  template <class T>
  void take(T *&p, size_t n) {
    p = reinterpret_cast<T *>(base + used);
    used += (sizeof(T) * (n ? n : 1) + 255) / 256 * 256;
  }
  bool ok() const { return used <= cap; }
};

// the loop indices whose cells take part in the closure (two or more
// materials), appended in any order; range = {0, count} for the list kernel
__global__ void select_pte_cells_kernel(const int *npte, int ncell, int *cells,
                                        unsigned long long *range) {
  const int il = blockIdx.x * blockDim.x + threadIdx.x;
  if (il >= ncell || npte[il] < 2) return;
  const unsigned long long k = atomicAdd(range + 1, 1ull);
  cells[k] = il;
}
}  // namespace

extern "C" int sgm_get_sg_eos_noalloc(sgm_ctx *ctx, int nmat, int ncell, int cell_dim,
                                      int input_int, const int *eos_offsets, const int *offsets,
                                      double *press, double *pmax, double *vol, double *spvol,
                                      double *sie, double *temp, double *bmod, double *dpde,
                                      double *cv, double *frac_mass, double *frac_vol,
                                      double *frac_ie, double *frac_bmod, double *frac_dpde,
                                      double *frac_cv, double mass_frac_cutoff,
                                      const sgm_mix_params *params, int *pte_converged,
                                      sgm_stream_t stream) {
  DrvArgs a;
  const int rc0 = driver_setup(ctx, nmat, ncell, cell_dim, input_int, eos_offsets, offsets,
                               press, pmax, vol, spvol, sie, temp, bmod, dpde, cv, frac_mass,
                               frac_vol, frac_ie, frac_bmod, frac_dpde, frac_cv,
                               mass_frac_cutoff, pte_converged, a);
  if (rc0 != SGM_OK) return rc0;
  if (ncell == 0) return SGM_OK;
  const long nnz_max = (long)ncell * nmat;
  if (nnz_max > 2147483647L) return SGM_E_BUDGET;
  hipStream_t s = (hipStream_t)stream;
  size_t scan_bytes = 0;
  if (hipcub::DeviceScan::ExclusiveSum(nullptr, scan_bytes, (int *)nullptr, (int *)nullptr,
                                       ncell + 1, s) != hipSuccess)
    return SGM_E_HIP;
  void *base = nullptr;
  if (scratch_address(&base) != SGM_OK) return SGM_E_HIP;
  FixedArena ar{static_cast<unsigned char *>(base), NOALLOC_SCRATCH_BYTES};
  ar.take(a.npte, ncell + 1);
  ar.take(a.mass_sum, ncell);
  ar.take(a.eoff, ncell + 1);
  ar.take(a.cell_in, ncell);
  ar.take(a.vtot, ncell);
  unsigned char *scan_tmp = nullptr;
  ar.take(scan_tmp, scan_bytes + 16);
  int *cells = nullptr;
  unsigned long long *range = nullptr;
  ar.take(cells, ncell);
  ar.take(range, 2);
  ar.take(a.emat, nnz_max);
  ar.take(a.eid, nnz_max);
  ar.take(a.rho, nnz_max);
  ar.take(a.vfrac, nnz_max);
  ar.take(a.esie, nnz_max);
  ar.take(a.etemp, nnz_max);
  ar.take(a.epress, nnz_max);
  ar.take(a.lam, 3L * nnz_max);
  if (!ar.ok()) return SGM_E_BUDGET;  // nothing launched
  const int nb = (ncell + 255) / 256;
  // what must start at zero: the scan's end sentinel and the cell counter
  if (launch_zero_ints(a.npte + ncell, 1, s) != hipSuccess) return SGM_E_HIP;
  if (launch_zero_ints(reinterpret_cast<int *>(range), 4, s) != hipSuccess) return SGM_E_HIP;
  prologue_kernel<<<nb, 256, 0, s>>>(a);
  if (hipGetLastError() != hipSuccess) return SGM_E_HIP;
  if (hipcub::DeviceScan::ExclusiveSum(scan_tmp, scan_bytes, a.npte, a.eoff, ncell + 1, s) !=
      hipSuccess)
    return SGM_E_HIP;
  fill_kernel<<<nb, 256, 0, s>>>(a);
  if (hipGetLastError() != hipSuccess) return SGM_E_HIP;
  if (input_int != IN_P_T) {
    const int method = input_int == IN_RHO_T   ? PTE_FIXED_T
                       : input_int == IN_RHO_P ? PTE_FIXED_P
                                               : PTE_RHOT;
    select_pte_cells_kernel<<<nb, 256, 0, s>>>(a.npte, ncell, cells, range);
    if (hipGetLastError() != hipSuccess) return SGM_E_HIP;
    PteArgs pa;
    pa.d_eos = a.d_eos;
    pa.nreg = a.nreg;
    pa.ncell = ncell;
    pa.nnz = nnz_max;
    pa.offsets = a.eoff;
    pa.mat_ids = a.eid;
    pa.rho = a.rho;
    pa.vfrac = a.vfrac;
    pa.sie = a.esie;
    pa.temp = a.etemp;
    pa.press = a.epress;
    pa.vfrac_tot = a.vtot;
    pa.sie_tot = a.cell_in;
    pa.lambda = a.lam;
    pa.conv = a.conv;
    if (launch_pte_list(method, pa, to_dev(params), cells, ncell, s,
                        reinterpret_cast<const long *>(range)) != hipSuccess)
      return SGM_E_HIP;
  }
  epilogue_kernel<<<nb, 256, 0, s>>>(a);
  return hipGetLastError() == hipSuccess ? SGM_OK : SGM_E_HIP;
}
