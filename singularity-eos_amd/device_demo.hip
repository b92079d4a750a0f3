// A caller's own HIP kernels written against include/sgm_device.hpp, the way
// a hydro code's kernels construct singularity-eos closure objects per cell
// (singularity-eos/eos/get_sg_eos_rho_e.cpp:59-72, test/test_pte_2phaseVinetSn.cpp:164-201):
// one thread per cell, pointer indexers into the cell's slice of the
// per-material arrays, ApproxTemperatureFromRhoMatU for the temperature
// guess, then PTESolverRhoT or PTESolverPT and PTESolver. Built into
// libsgm_device_demo.so, which the tests call to check that the device-side
// interface gives the library's results bit for bit.
#define SGM_DEVICE_MAX_MATS 6
#include "../include/sgm_device.hpp"

#include <hip/hip_runtime.h>

namespace {

using sgm::device::Real;

template <int METHOD>
__global__ void __launch_bounds__(256)
    demo_closure_kernel(sgm::device::Registry reg, long ncell, const int *offsets,
                        const uint8_t *mat_ids, double *rho, double *vfrac, double *sie,
                        double *temp, double *press, const double *vfrac_tot,
                        const double *sie_tot, const double *Tnorm, int approx_tnorm,
                        sgm::device::MixParams params, int *conv, int *niter, double *resid,
                        int *slow, int *max_line) {
  sgm::device::stage_registry(reg);
  const long c = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= ncell) return;
  const int o = offsets[c];
  const std::size_t nmat = (std::size_t)(offsets[c + 1] - o);
  const uint8_t *eos = mat_ids + o;
  double *r = rho + o, *v = vfrac + o, *e = sie + o, *t = temp + o, *p = press + o;
  sgm::device::NullIndexer lambda;
  Real scratch[1];
  Real Tn = Tnorm ? Tnorm[c] : 0.0;
  if (approx_tnorm) {
    Real rho_tot = 0, u_tot = 0;
    for (std::size_t m = 0; m < nmat; ++m) rho_tot += r[m] * v[m];
    u_tot = rho_tot * sie_tot[c];
    Tn = sgm::device::ApproxTemperatureFromRhoMatU(nmat, eos, u_tot, r, v);
  }
  sgm::device::SolverStatus st;
  if (METHOD == 0) {
    sgm::device::PTESolverRhoT<const uint8_t *, double *, sgm::device::NullIndexer> method(
        nmat, eos, vfrac_tot[c], sie_tot[c], r, v, e, t, p, lambda, scratch, Tn, params);
    st = sgm::device::PTESolver(method);
  } else {
    sgm::device::PTESolverPT<const uint8_t *, double *, sgm::device::NullIndexer> method(
        nmat, eos, vfrac_tot[c], sie_tot[c], r, v, e, t, p, lambda, scratch, Tn, params);
    st = sgm::device::PTESolver(method);
  }
  conv[c] = st.converged ? 1 : 0;
  niter[c] = (int)st.max_niter;
  resid[c] = st.residual;
  if (slow) slow[c] = st.slow_convergence_detected ? 1 : 0;
  if (max_line) max_line[c] = (int)st.max_line_niter;
}

} // namespace

extern "C" {

// One closure per CSR cell (2..6 materials) through the device interface.
// method 0 PTESolverRhoT, 1 PTESolverPT; approx_tnorm: the temperature guess
// from ApproxTemperatureFromRhoMatU(nmat, eos, rho_tot * sie_tot, rho,
// vfrac) instead of Tnorm. params: HOST (NULL = defaults). Every array is
// device memory; slow / max_line may be NULL.
int sgm_demo_pte_per_cell(sgm_ctx *ctx, int method, long ncell, const int *offsets,
                          const uint8_t *mat_ids, double *rho, double *vfrac, double *sie,
                          double *temp, double *press, const double *vfrac_tot,
                          const double *sie_tot, const double *Tnorm, int approx_tnorm,
                          const sgm_mix_params *params, int *conv, int *niter, double *resid,
                          int *slow, int *max_line, sgm_stream_t stream) {
  sgm::device::Registry reg{};
  if (sgm_device_registry(ctx, &reg.d_eos, &reg.n) != SGM_OK) return SGM_E_ARG;
  if (ncell <= 0) return SGM_OK;
  sgm::device::MixParams P;
  if (params) P = sgm::device::MixParams(*params);
  const int blocks = (int)((ncell + 255) / 256);
  const size_t lds = sgm::device::registry_lds_bytes(reg.n);
  hipStream_t s = (hipStream_t)stream;
  if (method == 0)
    demo_closure_kernel<0><<<blocks, 256, lds, s>>>(reg, ncell, offsets, mat_ids, rho, vfrac,
                                                    sie, temp, press, vfrac_tot, sie_tot, Tnorm,
                                                    approx_tnorm, P, conv, niter, resid, slow,
                                                    max_line);
  else
    demo_closure_kernel<1><<<blocks, 256, lds, s>>>(reg, ncell, offsets, mat_ids, rho, vfrac,
                                                    sie, temp, press, vfrac_tot, sie_tot, Tnorm,
                                                    approx_tnorm, P, conv, niter, resid, slow,
                                                    max_line);
  return hipGetLastError() == hipSuccess ? SGM_OK : SGM_E_HIP;
}

} // extern "C"
