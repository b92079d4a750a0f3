// ApproxTemperatureFromRhoMatU over a CSR list of cells (1..64 materials):
// one lane per cell,
// the material registry staged once per block into LDS (every EOS parameter
// read of the bisection is an LDS access). Per cell it reads the entries'
// densities and volume fractions and the cell's energy density, and writes
// one temperature.
#include "sgm_internal.hpp"
#include "sgm_approx.hpp"

namespace sgm {
namespace {

__global__ void __launch_bounds__(256)
    approx_temperature_kernel(const EosDev *__restrict__ reg, int nreg, long ncell, long nnz,
                              const int *__restrict__ offsets, const uint8_t *__restrict__ mat,
                              const double *__restrict__ rho, const double *__restrict__ vfrac,
                              const double *__restrict__ u_tot, const double *__restrict__ Tguess,
                              double *lambda, double *__restrict__ T_out) {
  extern __shared__ EosDev sgm_lds_registry[];
  {
    const double *s = reinterpret_cast<const double *>(reg);
    double *d = reinterpret_cast<double *>(sgm_lds_registry);
    const int nd = nreg * (int)(sizeof(EosDev) / 8);
    for (int i = threadIdx.x; i < nd; i += blockDim.x) d[i] = s[i];
    __syncthreads();
  }
  const long nthr = (long)gridDim.x * blockDim.x;
  for (long c = (long)blockIdx.x * blockDim.x + threadIdx.x; c < ncell; c += nthr) {
    const long o0 = offsets[c], o1 = offsets[c + 1];
    const long nm = o1 - o0;
    if (o0 >= 0 && o1 >= o0 && o1 <= nnz && nm > MAX_MATS_PER_CELL && nm <= MAX_MATS_WAVE) {
      // a wider cell: its entries are read where they lie, the lambda caches
      // read and written back around each evaluation (the same values as the
      // register copies of a narrower cell)
      bool ok = true;
      for (long k = o0; k < o1; ++k) ok = ok && mat[k] < nreg;
      if (!ok) {
        T_out[c] = __builtin_nan("");
        continue;
      }
      auto ufunc = [&](Real T) {
        Real usum = 0.0;
        for (long k = o0; k < o1; ++k) {
          Lam l;
          if (lambda) {
            l.present = true;
            l.lRho = lambda[3 * k + 0];
            l.lT = lambda[3 * k + 1];
          }
          usum += rho[k] * vfrac[k] * eos_e_rT(registry_entry(mat[k]), rho[k], T, l);
          if (lambda) {
            lambda[3 * k + 0] = l.lRho;
            lambda[3 * k + 1] = l.lT;
          }
        }
        return usum;
      };
      T_out[c] = approx_temperature_core(u_tot[c], Tguess ? Tguess[c] : 0.0, ufunc);
      continue;
    }
    bool ok = o0 >= 0 && o1 >= o0 && o1 <= nnz && nm >= 1 && nm <= MAX_MATS_PER_CELL;
    unsigned char ids[MAX_MATS_PER_CELL];
    Real r[MAX_MATS_PER_CELL], v[MAX_MATS_PER_CELL];
    Lam l[MAX_MATS_PER_CELL];
    for (int m = 0; ok && m < nm; ++m) {
      const long k = o0 + m;
      ids[m] = mat[k];
      ok = ids[m] < nreg;
      r[m] = rho[k];
      v[m] = vfrac[k];
      if (lambda) {
        l[m].present = true;
        l[m].lRho = lambda[3 * k + 0];
        l[m].lT = lambda[3 * k + 1];
      }
    }
    if (!ok) {
      T_out[c] = __builtin_nan("");
      continue;
    }
    T_out[c] = approx_temperature_from_rho_mat_u((int)nm, ids, u_tot[c], r, v,
                                                 Tguess ? Tguess[c] : 0.0, lambda ? l : nullptr);
    if (lambda) {
      for (int m = 0; m < nm; ++m) {
        lambda[3 * (o0 + m) + 0] = l[m].lRho;
        lambda[3 * (o0 + m) + 1] = l[m].lT;
      }
    }
  }
}

} // namespace

hipError_t launch_approx_temperature(const EosDev *d_eos, int nreg, long ncell, long nnz,
                                     const int *offsets, const uint8_t *mat, const double *rho,
                                     const double *vfrac, const double *u_tot,
                                     const double *Tguess, double *lambda, double *T_out,
                                     hipStream_t s) {
  if (ncell <= 0) return hipSuccess;
  const long blocks = (ncell + 255) / 256;
  const long cap = (long)device_cu_count() * 16;
  const int g = (int)(blocks < cap ? blocks : cap);
  approx_temperature_kernel<<<g, 256, (size_t)nreg * sizeof(EosDev), s>>>(
      d_eos, nreg, ncell, nnz, offsets, mat, rho, vfrac, u_tot, Tguess, lambda, T_out);
  return hipGetLastError();
}

} // namespace sgm
