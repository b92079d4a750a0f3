// ApproxTemperatureFromRhoMatU as a device function: shared by the per-cell
// kernel of the C ABI and the device-side closure interface.
#pragma once

#include "sgm_pte.hpp"

namespace sgm {

// ApproxTemperatureFromRhoMatU (mixed_cell_models.hpp:790-847): a rough
// temperature for a cell from its materials' densities and volume fractions
// and the cell's internal energy density u_tot = rho e (erg/cm^3). The
// energy sum u(T) = sum_m rho_m alpha_m e_m(rho_m, T) is bracketed on
// [1e-9, 1e9] K (beyond either end the end is returned), a guess inside the
// range first narrows the bracket, then at most 10 bisection steps in NQT
// log2 T (until the bracket is narrower than 0.01) and a linear
// interpolation of log2 T by u inside the bracket. eidx[m] is material m's
// registry index; lam may be NULL (no table caches).
// ufunc(T): the cell's energy density sum at temperature T
template <class U>
SGM_HD Real approx_temperature_core(Real u_tot, Real Tguess, U &&ufunc) {
  constexpr Real minimum_temperature = 1.e-9;
  constexpr Real maximum_temperature = 1.e9;
  Real ulo = ufunc(minimum_temperature);
  if (u_tot < ulo) return minimum_temperature;
  Real uhi = ufunc(maximum_temperature);
  if (u_tot > uhi) return maximum_temperature;
  Real lTlo = nqt::lg(minimum_temperature);
  Real lThi = nqt::lg(maximum_temperature);
  if (Tguess > minimum_temperature && Tguess < maximum_temperature) {
    const Real ug = ufunc(Tguess);
    if (ug < u_tot) {
      lTlo = nqt::lg(Tguess);
      ulo = ug;
    } else {
      lThi = nqt::lg(Tguess);
      uhi = ug;
    }
  }
  for (int iter = 0; lThi - lTlo > 0.01 && iter < 10; ++iter) {
    const Real lT = 0.5 * (lTlo + lThi);
    const Real uT = ufunc(nqt::pow2(lT));
    if (uT < u_tot) {
      lTlo = lT;
      ulo = uT;
    } else {
      lThi = lT;
      uhi = uT;
    }
  }
  const Real alpha = robust::ratio((u_tot - ulo), (uhi - ulo));
  return nqt::pow2((1.0 - alpha) * lTlo + alpha * lThi);
}
SGM_HD Real approx_temperature_from_rho_mat_u(int nmat, const unsigned char *eidx, Real u_tot,
                                              const Real *rho, const Real *vfrac, Real Tguess,
                                              Lam *lam) {
  auto ufunc = [&](Real T) {
    Real usum = 0.0;
    for (int m = 0; m < nmat; ++m) {
      Lam none;
      Lam &l = lam ? lam[m] : none;
      usum += rho[m] * vfrac[m] * eos_e_rT(registry_entry(eidx[m]), rho[m], T, l);
    }
    return usum;
  };
  return approx_temperature_core(u_tot, Tguess, ufunc);
}

// the per-cell kernel of sgm_approx_temperature (sgm_approx_temp.hip)
hipError_t launch_approx_temperature(const EosDev *d_eos, int nreg, long ncell, long nnz,
                                     const int *offsets, const uint8_t *mat, const double *rho,
                                     const double *vfrac, const double *u_tot,
                                     const double *Tguess, double *lambda, double *T_out,
                                     hipStream_t s);

} // namespace sgm
