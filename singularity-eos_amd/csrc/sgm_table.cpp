// Host-side construction of a Spiner-style rho-T table: load-time derived
// curves of SpinerEOSDependsRhoT::loadDataboxes_
// (singularity-eos/eos/eos_spiner_rho_temp.hpp:515-651), fixBulkModulus_
// (:653-679), setlTColdCrit_ (:681-740) and spiner_common::SetRhoPMin
// (singularity-eos/eos/eos_spiner_common.hpp:59-109), then the device copy.
#include "sgm_internal.hpp"

#include <cfloat>
#include <cmath>
#include <cstring>

namespace sgm {

namespace {
Real node_x(Real mn, Real dx, int i) { return mn + i * dx; }

void fill_view(TableHost &T) {
  TableDev &d = T.hdev;
  d.P = T.P.data();
  d.sie = T.sie.data();
  d.bMod = T.bMod.data();
  d.dEdT = T.dEdT.data();
  d.dPdE = T.dPdE.data();
  d.PCold = T.PCold.data();
  d.sieCold = T.sieCold.data();
  d.bModCold = T.bModCold.data();
  d.dPdECold = T.dPdECold.data();
  d.dEdTCold = T.dEdTCold.data();
  d.PMax = T.PMax.data();
  d.sielTMax = T.sielTMax.data();
  d.dEdTMax = T.dEdTMax.data();
  d.gm1Max = T.gm1Max.data();
  d.rho_at_pmin = T.rho_at_pmin.data();
}
} // namespace

// Builds every table field on the host. Returns false on invalid sizes.
bool build_table_host(const sgm_table_desc &in, TableHost &T) {
  const int nR = in.nRho, nT = in.nT;
  if (nR < 2 || nT < 2 || !(in.lRhoMax > in.lRhoMin) || !(in.lTMax > in.lTMin)) return false;
  if (!in.P || !in.sie || !in.bMod || !in.dPdRho || !in.dPdE || !in.dEdRho || !in.dEdT ||
      !in.PCold || !in.sieCold || !in.bModCold)
    return false;
  const size_t n2 = (size_t)nR * nT;
  TableDev &d = T.hdev;
  d.nRho = nR;
  d.nT = nT;
  d.lRhoMin = in.lRhoMin;
  d.lRhoMax = in.lRhoMax;
  d.dlRho = (in.lRhoMax - in.lRhoMin) / ((Real)(nR - 1));
  d.inv_dlRho = 1 / d.dlRho;
  d.lTMin = in.lTMin;
  d.lTMax = in.lTMax;
  d.dlT = (in.lTMax - in.lTMin) / ((Real)(nT - 1));
  d.inv_dlT = 1 / d.dlT;
  d.lRhoOffset = std::fabs(in.rho_offset);
  d.lTOffset = std::fabs(in.T_offset);
  d.reproducible = in.reproducible ? 1 : 0;
  d.rhoMax = from_log(d.lRhoMax, d.lRhoOffset);
  d.TMax = from_log(d.lTMax, d.lTOffset);

  T.P.assign(in.P, in.P + n2);
  T.sie.assign(in.sie, in.sie + n2);
  T.bMod.assign(in.bMod, in.bMod + n2);
  T.dPdE.assign(in.dPdE, in.dPdE + n2);
  T.dEdT.assign(in.dEdT, in.dEdT + n2);
  std::vector<Real> dPdRho(in.dPdRho, in.dPdRho + n2), dEdRho(in.dEdRho, in.dEdRho + n2);
  T.PCold.assign(in.PCold, in.PCold + nR);
  T.sieCold.assign(in.sieCold, in.sieCold + nR);
  T.bModCold.assign(in.bModCold, in.bModCold + nR);
  T.dPdECold.assign(nR, 0.0);
  T.dEdTCold.assign(nR, 0.0);
  T.lTColdCrit.assign(nR, 0.0);
  T.PMax.assign(nR, 0.0);
  T.sielTMax.assign(nR, 0.0);
  T.dEdTMax.assign(nR, 0.0);
  T.gm1Max.assign(nR, 0.0);
  T.rho_at_pmin.assign(nT, 0.0);
  fill_view(T);

  // fixBulkModulus_: bulk modulus rebuilt from the derivative tables (uses
  // the original bMod_ nowhere; reads interpolated P, dP/drho|e, dP/de|rho,
  // de/drho|T at the nodes)
  {
    TableDev tmp = d;
    for (int j = 0; j < nR; j++) {
      const Real lRho = node_x(d.lRhoMin, d.dlRho, j);
      const Real rho = table::rho_(d, lRho);
      for (int i = 0; i < nT; i++) {
        const Real lT = node_x(d.lTMin, d.dlT, i);
        const Real press = bilerp(tmp, T.P.data(), lRho, lT);
        const Real DPDR_E = bilerp(tmp, dPdRho.data(), lRho, lT);
        const Real DPDE_R = bilerp(tmp, T.dPdE.data(), lRho, lT);
        const Real DEDR_T = bilerp(tmp, dEdRho.data(), lRho, lT);
        const Real DPDR_T = DPDR_E + DPDE_R * DEDR_T;
        Real bm;
        if (DPDE_R > 0.0 && rho > 0.0)
          bm = rho * DPDR_E + DPDE_R * (press / rho);
        else if (rho > 0.0)
          bm = smax(rho * DPDR_T, 0.0);
        else
          bm = 0.0;
        T.bMod[(size_t)j * nT + i] = smax(bm, std::fabs(robust::EPS()));
      }
    }
  }
  // setlTColdCrit_: highest upward crossing of sie(lRho_j, .) through the
  // cold-curve energy, found by regula falsi; the lowest isotherm otherwise
  for (int j = 0; j < nR; j++) {
    const Real lRho = node_x(d.lRhoMin, d.dlRho, j);
    const Real sieCold = lerp_rho(d, T.sieCold.data(), lRho);
    int ncross = 0, idx_last_pos = -1, last_i = -1;
    for (int i = 0; i < nT - 1; i++) {
      const Real sc = T.sie[(size_t)j * nT + i] - sieCold;
      const Real sn = T.sie[(size_t)j * nT + i + 1] - sieCold;
      if (sc < 0 && sn > 0) {
        idx_last_pos = ncross;
        last_i = i;
        ncross++;
      } else if (sc > 0 && sn < 0) {
        ncross++;
      }
    }
    Real lTc;
    if (idx_last_pos <= 0) {
      lTc = d.lTMin;
    } else {
      const W1 wr = grid_weights(lRho, d.lRhoMin, d.dlRho, d.inv_dlRho, d.nRho);
      auto f = [&](Real x) { return bilerp_w(d, T.sie.data(), wr, x); };
      const Real lo = node_x(d.lTMin, d.dlT, last_i) - 1.0e-14;
      const Real hi = node_x(d.lTMin, d.dlT, last_i + 1) + 1.0e-14;
      const Real g = 0.5 * (lo + hi);
      if (!root::regula_falsi(f, sieCold, g, lo, hi, table::ROOT_THRESH, table::ROOT_THRESH,
                              lTc))
        lTc = g;
    }
    T.lTColdCrit[j] = smax(d.lTMin, lTc);
  }
  // SetRhoPMin (pmin_vapor_dome = false unless requested)
  {
    Real PMin = DBL_MAX;
    const bool vd = in.pmin_vapor_dome != 0;
    for (int i = 0; i < nT; ++i) {
      Real PMin_at_T = DBL_MAX;
      int jmax = 0;
      for (int j = 0; j < nR; ++j) {
        const Real Pji = T.P[(size_t)j * nT + i];
        if (Pji < PMin_at_T) {
          PMin_at_T = Pji;
          jmax = j;
        }
        if ((j > 1) && vd) {
          const Real dP = Pji - T.P[(size_t)(j - 2) * nT + i];
          const Real dr = from_log(node_x(d.lRhoMin, d.dlRho, j), d.lRhoOffset) -
                          from_log(node_x(d.lRhoMin, d.dlRho, j - 2), d.lRhoOffset);
          if (robust::ratio(dP, dr) < 1e-8) jmax = j;
        }
      }
      if ((PMin_at_T > 0) && !vd) {
        T.rho_at_pmin[i] = 0;
        PMin_at_T = 0;
      } else {
        T.rho_at_pmin[i] = from_log(node_x(d.lRhoMin, d.dlRho, jmax), d.lRhoOffset);
      }
      PMin = smin(PMin_at_T, PMin);
    }
    if (vd)
      for (int i = nT - 2; i >= 0; i--)
        if (T.rho_at_pmin[i] < T.rho_at_pmin[i + 1]) T.rho_at_pmin[i] = T.rho_at_pmin[i + 1];
    d.PMin = PMin;
  }
  // cold-curve bulk modulus / dP/de / Cv at lTColdCrit, then the top-isotherm curves
  for (int j = 0; j < nR; j++) {
    const Real lRho = node_x(d.lRhoMin, d.dlRho, j);
    const Real lT = T.lTColdCrit[j];
    T.bModCold[j] = bilerp(d, T.bMod.data(), lRho, lT);
    T.dPdECold[j] = bilerp(d, T.dPdE.data(), lRho, lT);
    T.dEdTCold[j] = bilerp(d, T.dEdT.data(), lRho, lT);
  }
  for (int j = 0; j < nR; j++) {
    const Real lRho = node_x(d.lRhoMin, d.dlRho, j);
    const Real rho = table::rho_(d, lRho);
    const size_t k = (size_t)j * nT + (nT - 1);
    T.PMax[j] = T.P[k];
    T.dEdTMax[j] = T.dEdT[k];
    T.gm1Max[j] = robust::ratio(T.dPdE[k], rho);
    T.sielTMax[j] = T.sie[k];
  }
  fill_view(T);
  return true;
}

// Copies every field to device memory and builds the device TableDev.
hipError_t upload_table(TableHost &T) {
  TableDev dd = T.hdev;
  auto up = [&](const std::vector<Real> &v, const Real *&dst) -> hipError_t {
    void *p = nullptr;
    hipError_t e = hipMalloc(&p, v.size() * sizeof(Real));
    if (e != hipSuccess) return e;
    T.d_arrays.push_back(p);
    e = hipMemcpy(p, v.data(), v.size() * sizeof(Real), hipMemcpyHostToDevice);
    dst = static_cast<const Real *>(p);
    return e;
  };
  hipError_t e;
#define SGM_UP(field)                                                                          \
  if ((e = up(T.field, dd.field)) != hipSuccess) return e;
  SGM_UP(P)
  SGM_UP(sie)
  SGM_UP(bMod)
  SGM_UP(dEdT)
  SGM_UP(dPdE)
  SGM_UP(PCold)
  SGM_UP(sieCold)
  SGM_UP(bModCold)
  SGM_UP(dPdECold)
  SGM_UP(dEdTCold)
  SGM_UP(PMax)
  SGM_UP(sielTMax)
  SGM_UP(dEdTMax)
  SGM_UP(gm1Max)
  SGM_UP(rho_at_pmin)
#undef SGM_UP
  void *p = nullptr;
  if ((e = hipMalloc(&p, sizeof(TableDev))) != hipSuccess) return e;
  T.d_arrays.push_back(p);
  if ((e = hipMemcpy(p, &dd, sizeof(TableDev), hipMemcpyHostToDevice)) != hipSuccess) return e;
  T.d_tab = static_cast<TableDev *>(p);
  T.ddev = dd;
  return hipSuccess;
}

} // namespace sgm
