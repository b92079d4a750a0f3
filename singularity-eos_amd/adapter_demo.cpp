// Host program written against include/sgm.hpp the way a singularity-eos
// host code is written against singularity::EOS and PTESolverRhoT: set up
// materials, call the vector EOS functions and the closure on device arrays.
//
//   sgm_adapter_demo --host        registry checks on a host-only registry
//   sgm_adapter_demo OUT.bin       two-material closure on 65,536 cells plus
//                                  P(rho, e) and T(rho, e) of material 1 on
//                                  the same entries; raw doubles to OUT.bin
#include "../include/sgm.hpp"

#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstring>
#include <vector>

namespace {
template <class T>
T *dev_copy(const std::vector<T> &h) {
  T *d = nullptr;
  if (hipMalloc(&d, h.size() * sizeof(T)) != hipSuccess ||
      hipMemcpy(d, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice) != hipSuccess)
    throw sgm::Error("device copy", SGM_E_HIP);
  return d;
}
template <class T>
std::vector<T> host_copy(const T *d, size_t n) {
  std::vector<T> h(n);
  if (hipMemcpy(h.data(), d, n * sizeof(T), hipMemcpyDeviceToHost) != hipSuccess)
    throw sgm::Error("host copy", SGM_E_HIP);
  return h;
}
} // namespace

int main(int argc, char **argv) {
  try {
    if (argc > 1 && std::strcmp(argv[1], "--host") == 0) {
      sgm::Registry reg(2, true);
      sgm::EOS a = reg.IdealGas(0, 0.4, 7.176e6);
      reg.StiffGas(1, 3.4, 3.6e7, 6e9, 0.0);
      if (sgm_get_kind(reg.get(), 0) != SGM_IDEAL || sgm_get_kind(reg.get(), 1) != SGM_STIFF) {
        std::printf("kind mismatch\n");
        return 1;
      }
      try {
        a.PressureFromDensityInternalEnergy(nullptr, nullptr, nullptr, 1);
        std::printf("a launch on a host-only registry did not fail\n");
        return 1;
      } catch (const sgm::Error &e) {
        if (e.code != SGM_E_NOGPU) {
          std::printf("unexpected code %d\n", e.code);
          return 1;
        }
      }
      std::printf("adapter host OK\n");
      return 0;
    }
    if (argc < 2) {
      std::printf("usage: %s --host | OUT.bin\n", argv[0]);
      return 2;
    }
    const long n = 65536;
    std::vector<double> rho(2 * n), vfrac(2 * n), sie(2 * n), temp(2 * n), press(2 * n),
        sie_tot(n), vfrac_tot(n, 1.0);
    sgm::check(sgm_workload_pte2(0x5EED0000ull, n, rho.data(), vfrac.data(), sie.data(),
                                 temp.data(), press.data(), sie_tot.data()),
               "workload");
    std::vector<int> off(n + 1);
    std::vector<uint8_t> mat(2 * n);
    for (long c = 0; c <= n; ++c) off[c] = (int)(2 * c);
    for (long k = 0; k < 2 * n; ++k) mat[k] = (uint8_t)(k & 1);

    sgm::Registry reg(2);
    sgm::EOS gas = reg.IdealGas(0, 0.4, 7.176e6);
    sgm::EOS water = reg.StiffGas(1, 3.4, 3.6e7, 6e9, 0.0);
    (void)gas;
    double *d_rho = dev_copy(rho), *d_sie = dev_copy(sie);
    double *d_P = dev_copy(std::vector<double>(2 * n)), *d_T = dev_copy(std::vector<double>(2 * n));
    // the vector calls, as Variant::PressureFromDensityInternalEnergy(rhos, sies, pressures, num)
    water.PressureFromDensityInternalEnergy(d_rho, d_sie, d_P, (int)(2 * n));
    water.TemperatureFromDensityInternalEnergy(d_rho, d_sie, d_T, (int)(2 * n));
    // the closure, as one PTESolverRhoT + PTESolver per cell
    sgm::CellArrays c;
    c.ncell = n;
    c.nnz = 2 * n;
    c.offsets = dev_copy(off);
    c.mat_ids = dev_copy(mat);
    c.rho = dev_copy(rho);
    c.vfrac = dev_copy(vfrac);
    c.sie = dev_copy(sie);
    c.temp = dev_copy(temp);
    c.press = dev_copy(press);
    c.vfrac_tot = dev_copy(vfrac_tot);
    c.sie_tot = dev_copy(sie_tot);
    sgm::MixParams params;
    sgm_mix_params_tight(&params);
    sgm::PTESolver(reg, sgm::Method::RhoT, c, params);
    if (hipDeviceSynchronize() != hipSuccess) throw sgm::Error("sync", SGM_E_HIP);
    FILE *f = std::fopen(argv[1], "wb");
    if (!f) return 3;
    for (double *d : {d_P, d_T, c.rho, c.vfrac, c.sie, c.temp, c.press}) {
      const auto h = host_copy(d, 2 * n);
      std::fwrite(h.data(), sizeof(double), h.size(), f);
    }
    std::fclose(f);
    for (const void *d : {(const void *)d_rho, (const void *)d_sie, (const void *)d_P,
                          (const void *)d_T, (const void *)c.offsets, (const void *)c.mat_ids,
                          (const void *)c.rho, (const void *)c.vfrac, (const void *)c.sie,
                          (const void *)c.temp, (const void *)c.press, (const void *)c.vfrac_tot,
                          (const void *)c.sie_tot})
      (void)hipFree(const_cast<void *>(d));
    std::printf("adapter device OK\n");
    return 0;
  } catch (const sgm::Error &e) {
    std::printf("error: %s\n", e.what());
    return 1;
  }
}
