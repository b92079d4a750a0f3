// A caller's own HIP kernels written against include/sgm_device.hpp, the way
// a hydro code's kernels construct singularity-eos closure objects per cell
// (singularity-eos/eos/get_sg_eos_rho_e.cpp:59-72, test/test_pte_2phaseVinetSn.cpp:164-201):
// one thread per cell, pointer indexers into the cell's slice of the
// per-material arrays, ApproxTemperatureFromRhoMatU for the temperature
// guess, then PTESolverRhoT or PTESolverPT and PTESolver. Built into
// libsgm_device_demo.so, which the tests call to check that the device-side
// interface gives the library's results bit for bit (and that the library's
// two forms of its dense linear solve agree bit for bit).
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

// The library's two forms of the dense solve on one system per wave: every
// lane runs solve_small<N> (the register-resident bins' form) and the wave
// runs solve_small_wave<N> (the whole-wave tails' form); lane 0 stores both
// solutions, and the kernel counts the lanes whose wave-form result is not
// lane 0's bit for bit.
template <int N>
__global__ void __launch_bounds__(64)
    demo_solve_kernel(const double *A, const double *b, double *x_serial, double *x_wave,
                      int *ok, int *lanes_differ) {
  const long sys = blockIdx.x;  // the grid holds one block per system
  const int lane = threadIdx.x;
  Real a[N * N], bs[N], bw[N];
#pragma unroll
  for (int k = 0; k < N * N; ++k) a[k] = A[sys * N * N + k];
#pragma unroll
  for (int i = 0; i < N; ++i) bs[i] = bw[i] = b[sys * N + i];
  const bool okw = sgm::solve_small_wave<N>(a, bw, lane);
  const bool oks = sgm::solve_small<N>(a, bs);
  bool differ = false;
#pragma unroll
  for (int i = 0; i < N; ++i)
    differ = differ || __builtin_bit_cast(unsigned long long, bw[i]) !=
                           __builtin_bit_cast(unsigned long long, sgm::lane_value(bw[i], 0));
  const unsigned long long mask = __ballot(differ);
  if (lane == 0) {
#pragma unroll
    for (int i = 0; i < N; ++i) {
      x_serial[sys * N + i] = bs[i];
      x_wave[sys * N + i] = bw[i];
    }
    ok[sys] = (oks ? 1 : 0) | (okw ? 2 : 0);
    lanes_differ[sys] = __popcll(mask);
  }
}

} // namespace

extern "C" {

// Both dense solves (above) on `count` systems of n = 3..8 equations, A
// row-major n x n per system, b n per system, every array device memory.
// ok[s]: bit 0 the serial form succeeded, bit 1 the wave form did.
int sgm_demo_solve_forms(int n, long count, const double *A, const double *b,
                         double *x_serial, double *x_wave, int *ok, int *lanes_differ,
                         sgm_stream_t stream) {
  if (n < 3 || n > 8 || count < 0 || count > 2147483647L) return SGM_E_ARG;
  if (count == 0) return SGM_OK;
  if (!A || !b || !x_serial || !x_wave || !ok || !lanes_differ) return SGM_E_ARG;
  hipStream_t s = (hipStream_t)stream;
  const dim3 grid((unsigned)count);
  switch (n) {
  case 3: demo_solve_kernel<3><<<grid, 64, 0, s>>>(A, b, x_serial, x_wave, ok, lanes_differ); break;
  case 4: demo_solve_kernel<4><<<grid, 64, 0, s>>>(A, b, x_serial, x_wave, ok, lanes_differ); break;
  case 5: demo_solve_kernel<5><<<grid, 64, 0, s>>>(A, b, x_serial, x_wave, ok, lanes_differ); break;
  case 6: demo_solve_kernel<6><<<grid, 64, 0, s>>>(A, b, x_serial, x_wave, ok, lanes_differ); break;
  case 7: demo_solve_kernel<7><<<grid, 64, 0, s>>>(A, b, x_serial, x_wave, ok, lanes_differ); break;
  default: demo_solve_kernel<8><<<grid, 64, 0, s>>>(A, b, x_serial, x_wave, ok, lanes_differ); break;
  }
  return hipGetLastError() == hipSuccess ? SGM_OK : SGM_E_HIP;
}

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
