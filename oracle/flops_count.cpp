// ORACLE — test infrastructure only. Algorithmic FP64 operation count of the
// PTE closure: the restatement compiled with -DSGO_COUNT_FLOPS (every
// floating-point +,-,*,/ and math call counted once) run over a CSR cell
// list. bench.py divides this by the kernel time for the FP64 roofline.
#include "sgo_pte.hpp"

#include <vector>

using namespace sgo;

namespace {
MixParams params_from_array(const double *a) {
  MixParams p;
  p.derivative_eps = a[0];
  p.pte_rel_tolerance_p = a[1];
  p.pte_rel_tolerance_t = a[2];
  p.pte_rel_tolerance_e = a[3];
  p.pte_rel_tolerance_v = a[4];
  p.pte_abs_tolerance_p = a[5];
  p.pte_abs_tolerance_t = a[6];
  p.pte_slow_convergence_thresh = a[7];
  p.pte_rel_tolerance_p_sufficient = a[8];
  p.pte_rel_tolerance_t_sufficient = a[9];
  p.pte_rel_tolerance_e_sufficient = a[10];
  p.pte_rel_tolerance_v_sufficient = a[11];
  p.pte_abs_tolerance_p_sufficient = a[12];
  p.pte_abs_tolerance_t_sufficient = a[13];
  p.pte_max_iter_per_mat = (long)a[14];
  p.line_search_alpha = a[15];
  p.line_search_max_iter = (long)a[16];
  p.line_search_fac = a[17];
  p.vfrac_safety_fac = a[18];
  p.temperature_limit = a[19];
  p.default_tguess = a[20];
  p.min_dtde = a[21];
  p.pte_small_step_tries = (long)a[22];
  p.pte_small_step_thresh = a[23];
  p.pte_max_dpdv = a[24];
  p.pte_reduced_system_exclude_idx = (long)a[25];
  p.iterate_t_guess = a[26] != 0.0;
  return p;
}
} // namespace

extern "C" {
// eos_params[k*16 ..]: kind, then model parameters (IdealGas gm1,Cv; StiffGas
// gm1,Cv,Pinf,qq; Gruneisen C0,s1,s2,s3,G0,b,rho0,T0,P0,Cv; Vinet
// rho0,T0,B0,BP0,A0,Cv0,E0,S0 with zero expansion coefficients; kind 4 = the
// table given by tgrid {lRhoMin,lRhoMax,nRho,lTMin,lTMax,nT} and tfields
// {P,sie,bMod,dPdRho,dPdE,dEdRho,dEdT,PCold,sieCold,bModCold}).
// flops_out[c] = operation count of cell c.
void sgo_count_pte_flops(int method, long ncell, const int *offsets, const int *mat_eos,
                         const double *eos_params, int neos, const double *rho,
                         const double *vfrac, const double *temp, const double *press,
                         const double *vfrac_tot, const double *sie_tot,
                         const double *params_arr, const double *tgrid,
                         const double *const *tfields, double *flops_out) {
  const MixParams params = params_from_array(params_arr);
  std::vector<EOS> pool(neos);
  const Real zeros[39] = {};
  SpinerRhoT *table = nullptr;
  if (tgrid && tfields) {
    const int nR = (int)tgrid[2], nT = (int)tgrid[5];
    std::vector<std::vector<Real>> f(10);
    for (int k = 0; k < 10; ++k) {
      const long len = k < 7 ? (long)nR * nT : nR;
      f[k].assign(tfields[k], tfields[k] + len);
    }
    table = new SpinerRhoT();
    table->init(Grid1D::make(tgrid[0], tgrid[1], nR), Grid1D::make(tgrid[3], tgrid[4], nT),
                f[0].data(), f[1].data(), f[2].data(), f[3].data(), f[4].data(), f[5].data(),
                f[6].data(), f[7].data(), f[8].data(), f[9].data(), 0.0, 0.0);
  }
  for (int i = 0; i < neos; ++i) {
    const double *q = eos_params + 16 * i;
    EOS &e = pool[i];
    e.kind = (int)q[0];
    switch (e.kind) {
    case K_IDEAL: e.ig = IdealGas{q[2], q[1]}; break;
    case K_STIFF: e.sg = StiffGas{q[2], q[1], q[3], q[4]}; break;
    case K_GRUN: e.gr = Gruneisen::make(q[1], q[2], q[3], q[4], q[5], q[6], q[7], q[8], q[9], q[10]); break;
    case K_VINET: e.vn = Vinet::make(q[1], q[2], q[3], q[4], q[5], q[6], q[7], q[8], zeros); break;
    default: e.tb = table; break;
    }
  }
#pragma omp parallel
  {
    std::vector<Real> scr(256), r(8), v(8), s(8), t(8), p(8);
    EOS ec[8];
#pragma omp for schedule(dynamic, 64)
    for (long c = 0; c < ncell; ++c) {
      const int off = offsets[c], nm = offsets[c + 1] - off;
      for (int m = 0; m < nm; ++m) {
        ec[m] = pool[mat_eos[off + m]];
        r[m] = rho[off + m];
        v[m] = vfrac[off + m];
        t[m] = temp[off + m];
        p[m] = press[off + m];
        s[m] = 0.0;
      }
      g_flops = 0;
      if (method == 0) {
        PTERhoT sys(nm, ec, vfrac_tot[c], sie_tot[c], r.data(), v.data(), s.data(), t.data(),
                    p.data(), nullptr, scr.data(), 0.0, params);
        PTESolver(sys);
      } else {
        PTEPT sys(nm, ec, vfrac_tot[c], sie_tot[c], r.data(), v.data(), s.data(), t.data(),
                  p.data(), nullptr, scr.data(), 0.0, params);
        PTESolver(sys);
      }
      flops_out[c] = (double)g_flops;
    }
  }
  delete table;
}
}
