// Vector EOS kernels: the SG_EOS_VEC_2IN_1OUT loops of
// singularity-eos/eos/eos_base.hpp:169-230, the vector FillEos
// (eos_base.hpp:503-531) and DensityEnergyFromPressureTemperature over arrays.
//
// One kernel instantiation per (function, EOS kind): the material descriptor
// is staged once per block in LDS, inputs stream as structure-of-arrays with
// 16-byte vector loads where alignment allows (two points per lane), and the
// grid is sized to keep every CU of the 8 XCDs busy with a grid-stride loop.
// the table inversions inline here (see sgm_eos.hpp)
#define SGM_TABLE_INVERSE_INLINE
#include "sgm_internal.hpp"

namespace sgm {

namespace {

template <int FUNC, int KIND>
__device__ __forceinline__ Real vec_point(const EosDev &ek, Real a, Real b, Lam &lam) {
  __builtin_assume(ek.kind == KIND);  // lets the compiler fold the dispatch switch
  switch (FUNC) {
  case SGM_F_P_RHO_SIE: return eos_P_re(ek, a, b, lam);
  case SGM_F_T_RHO_SIE: return eos_T_re(ek, a, b, lam);
  case SGM_F_SIE_RHO_T: return eos_e_rT(ek, a, b, lam);
  case SGM_F_P_RHO_T: return eos_P_rT(ek, a, b, lam);
  case SGM_F_CV_RHO_SIE: return eos_cv_re(ek, a, b, lam);
  case SGM_F_CV_RHO_T: return eos_cv_rT(ek, a, b, lam);
  case SGM_F_BMOD_RHO_SIE: return eos_bmod_re(ek, a, b, lam);
  case SGM_F_BMOD_RHO_T: return eos_bmod_rT(ek, a, b, lam);
  default: return eos_min_e(ek, a);
  }
}

__device__ __forceinline__ void stage_eos(EosDev &dst, const EosDev *src) {
  const double *s = reinterpret_cast<const double *>(src);
  double *d = reinterpret_cast<double *>(&dst);
  if (threadIdx.x < sizeof(EosDev) / 8) d[threadIdx.x] = s[threadIdx.x];
  __syncthreads();
}

template <int FUNC, int KIND>
__global__ void __launch_bounds__(256) vec_kernel(const EosDev *__restrict__ reg, int mat,
                                                  const double *__restrict__ in1,
                                                  const double *__restrict__ in2,
                                                  double *__restrict__ out, long n) {
  __shared__ EosDev es;
  stage_eos(es, reg + mat);
  const EosDev &e = es;
  Lam none;
  const long nthr = (long)gridDim.x * blockDim.x;
  const long tid = (long)blockIdx.x * blockDim.x + threadIdx.x;
  const bool vec_ok = ((reinterpret_cast<uintptr_t>(in1) | reinterpret_cast<uintptr_t>(in2) |
                        reinterpret_cast<uintptr_t>(out)) &
                       15) == 0;
  if (vec_ok) {
    const long npair = n >> 1;
    const double2 *a2 = reinterpret_cast<const double2 *>(in1);
    const double2 *b2 = reinterpret_cast<const double2 *>(in2);
    double2 *o2 = reinterpret_cast<double2 *>(out);
    for (long i = tid; i < npair; i += nthr) {
      const double2 a = a2[i];
      const double2 b = (FUNC == SGM_F_MIN_SIE_RHO) ? a : b2[i];
      double2 r;
      r.x = vec_point<FUNC, KIND>(e, a.x, b.x, none);
      r.y = vec_point<FUNC, KIND>(e, a.y, b.y, none);
      __builtin_nontemporal_store(r.x, &o2[i].x);
      __builtin_nontemporal_store(r.y, &o2[i].y);
    }
    if ((n & 1) && tid == 0) {
      const long i = n - 1;
      out[i] = vec_point<FUNC, KIND>(e, in1[i], (FUNC == SGM_F_MIN_SIE_RHO) ? 0.0 : in2[i], none);
    }
  } else {
    for (long i = tid; i < n; i += nthr)
      out[i] = vec_point<FUNC, KIND>(e, in1[i], (FUNC == SGM_F_MIN_SIE_RHO) ? 0.0 : in2[i], none);
  }
}

// variant with a lambda cache (table EOS): strided (lRho, lT) per point
template <int FUNC, int KIND>
__global__ void __launch_bounds__(256)
    vec_kernel_lambda(const EosDev *__restrict__ reg, int mat, const double *__restrict__ in1,
                      const double *__restrict__ in2, double *__restrict__ out, long n,
                      int lstride, double *__restrict__ lambda) {
  __shared__ EosDev es;
  stage_eos(es, reg + mat);
  const long nthr = (long)gridDim.x * blockDim.x;
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += nthr) {
    Lam lam;
    lam.present = true;
    double *lp = lambda + (long)lstride * i;
    lam.lRho = lp[0];
    lam.lT = lp[1];
    out[i] = vec_point<FUNC, KIND>(es, in1[i], (FUNC == SGM_F_MIN_SIE_RHO) ? 0.0 : in2[i], lam);
    lp[0] = lam.lRho;
    lp[1] = lam.lT;
  }
}

template <int KIND>
__global__ void __launch_bounds__(256)
    fill_kernel(const EosDev *__restrict__ reg, int mat, double *rho, double *temp, double *sie,
                double *press, double *cv, double *bmod, long n, unsigned long output,
                int lstride, double *lambda, int *errs) {
  __shared__ EosDev es;
  stage_eos(es, reg + mat);
  const EosDev &ek = es;  // read in place (a private copy would sit in scratch memory)
  __builtin_assume(ek.kind == KIND);
  // a modified material's FillEos may leave a requested output unwritten (a
  // ScaledEOS over a (rho, sie) model fills a copy of the energy,
  // scaled_eos.hpp:155-160): its arrays are read in full so that such an
  // output keeps the caller's value, as the reference's by-reference
  // arguments do; the models themselves write every output they are asked for
  const unsigned long in = ek.mods ? ~0ul : ~output;
  const long nthr = (long)gridDim.x * blockDim.x;
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += nthr) {
    Real r = (rho && (in & tq::density)) ? rho[i] : 0.0;
    Real t = (temp && (in & tq::temperature)) ? temp[i] : 0.0;
    Real s = (sie && (in & tq::specific_internal_energy)) ? sie[i] : 0.0;
    Real p = (press && (in & tq::pressure)) ? press[i] : 0.0;
    Real c = 0.0, b = 0.0;
    Lam lam;
    double *lp = nullptr;
    if (lambda) {
      lp = lambda + (long)lstride * i;
      lam.present = true;
      lam.lRho = lp[0];
      lam.lT = lp[1];
    }
    int err = 0;
    eos_fill_k<KIND>(ek, r, t, s, p, c, b, output, lam, err);
    // an output set the model cannot fill (the reference's EOS_ERROR) writes
    // nothing: the point's arrays keep their values
    // (a modified material's FillEos may also change an input in place, as
    // ScaledEOS scales the energy of a (rho, T) model, scaled_eos.hpp:150-153:
    // every array it was given is written back)
    const bool filled = !(err & ERRB_FILLEOS);
    const unsigned long wr = ek.mods ? ~0ul : output;
    if (filled && rho && (wr & tq::density)) rho[i] = r;
    if (filled && temp && (wr & tq::temperature)) temp[i] = t;
    if (filled && sie && (wr & tq::specific_internal_energy)) sie[i] = s;
    if (filled && press && (wr & tq::pressure)) press[i] = p;
    if (filled && cv && (output & tq::specific_heat)) cv[i] = c;
    if (filled && bmod && (output & tq::bulk_modulus)) bmod[i] = b;
    if (lp) {
      lp[0] = lam.lRho;
      lp[1] = lam.lT;
    }
    if (errs) errs[i] = err;
  }
}

template <int KIND>
__global__ void __launch_bounds__(256)
    rhoe_PT_kernel(const EosDev *__restrict__ reg, int mat, const double *P, const double *T,
                   double *rho, double *sie, long n, int lstride, double *lambda, int *errs) {
  __shared__ EosDev es;
  stage_eos(es, reg + mat);
  const EosDev &ek = es;
  __builtin_assume(ek.kind == KIND);
  const long nthr = (long)gridDim.x * blockDim.x;
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += nthr) {
    Lam lam;
    double *lp = nullptr;
    if (lambda) {
      lp = lambda + (long)lstride * i;
      lam.present = true;
      lam.lRho = lp[0];
      lam.lT = lp[1];
    }
    int err = 0;
    Real r = rho[i], s = 0;
    eos_rhoe_PT(ek, P[i], T[i], r, s, lam, err);
    rho[i] = r;
    sie[i] = s;
    if (lp) {
      lp[0] = lam.lRho;
      lp[1] = lam.lT;
    }
    if (errs) errs[i] = err;
  }
}

int grid_for(long n, int per_thread) {
  static int ncu = -1;
  if (ncu < 0) ncu = device_cu_count();
  const long want = (n + 256L * per_thread - 1) / (256L * per_thread);
  const long cap = (long)ncu * 16;
  long g = want < cap ? want : cap;
  if (g < 1) g = 1;
  return (int)g;
}

template <int FUNC>
hipError_t launch_vec_f(int kind, const EosDev *d, int mat, const double *a, const double *b,
                        double *o, long n, int ls, double *lam, hipStream_t s) {
  if (lam) {
    const int g = grid_for(n, 1);
    switch (kind) {
    case KIND_IDEAL: vec_kernel_lambda<FUNC, KIND_IDEAL><<<g, 256, 0, s>>>(d, mat, a, b, o, n, ls, lam); break;
    case KIND_STIFF: vec_kernel_lambda<FUNC, KIND_STIFF><<<g, 256, 0, s>>>(d, mat, a, b, o, n, ls, lam); break;
    case KIND_GRUN: vec_kernel_lambda<FUNC, KIND_GRUN><<<g, 256, 0, s>>>(d, mat, a, b, o, n, ls, lam); break;
    case KIND_VINET: vec_kernel_lambda<FUNC, KIND_VINET><<<g, 256, 0, s>>>(d, mat, a, b, o, n, ls, lam); break;
    default: vec_kernel_lambda<FUNC, KIND_TABLE><<<g, 256, 0, s>>>(d, mat, a, b, o, n, ls, lam); break;
    }
  } else {
    const int g = grid_for(n, 2);
    switch (kind) {
    case KIND_IDEAL: vec_kernel<FUNC, KIND_IDEAL><<<g, 256, 0, s>>>(d, mat, a, b, o, n); break;
    case KIND_STIFF: vec_kernel<FUNC, KIND_STIFF><<<g, 256, 0, s>>>(d, mat, a, b, o, n); break;
    case KIND_GRUN: vec_kernel<FUNC, KIND_GRUN><<<g, 256, 0, s>>>(d, mat, a, b, o, n); break;
    case KIND_VINET: vec_kernel<FUNC, KIND_VINET><<<g, 256, 0, s>>>(d, mat, a, b, o, n); break;
    default: vec_kernel<FUNC, KIND_TABLE><<<g, 256, 0, s>>>(d, mat, a, b, o, n); break;
    }
  }
  return hipGetLastError();
}

} // namespace

int device_cu_count() {
  int dev = 0, ncu = 0;
  if (hipGetDevice(&dev) != hipSuccess) return 256;
  if (hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
      ncu <= 0)
    return 256;
  return ncu;
}

hipError_t launch_vec(const EosDev &eh, const EosDev *d, int mat, int func, const double *a,
                      const double *b, double *o, long n, int ls, double *lam, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  const int k = eh.kind;
  switch (func) {
  case SGM_F_P_RHO_SIE: return launch_vec_f<SGM_F_P_RHO_SIE>(k, d, mat, a, b, o, n, ls, lam, s);
  case SGM_F_T_RHO_SIE: return launch_vec_f<SGM_F_T_RHO_SIE>(k, d, mat, a, b, o, n, ls, lam, s);
  case SGM_F_SIE_RHO_T: return launch_vec_f<SGM_F_SIE_RHO_T>(k, d, mat, a, b, o, n, ls, lam, s);
  case SGM_F_P_RHO_T: return launch_vec_f<SGM_F_P_RHO_T>(k, d, mat, a, b, o, n, ls, lam, s);
  case SGM_F_CV_RHO_SIE: return launch_vec_f<SGM_F_CV_RHO_SIE>(k, d, mat, a, b, o, n, ls, lam, s);
  case SGM_F_CV_RHO_T: return launch_vec_f<SGM_F_CV_RHO_T>(k, d, mat, a, b, o, n, ls, lam, s);
  case SGM_F_BMOD_RHO_SIE:
    return launch_vec_f<SGM_F_BMOD_RHO_SIE>(k, d, mat, a, b, o, n, ls, lam, s);
  case SGM_F_BMOD_RHO_T: return launch_vec_f<SGM_F_BMOD_RHO_T>(k, d, mat, a, b, o, n, ls, lam, s);
  default: return launch_vec_f<SGM_F_MIN_SIE_RHO>(k, d, mat, a, b, o, n, ls, lam, s);
  }
}

hipError_t launch_fill(const EosDev &eh, const EosDev *d, int mat, double *rho, double *temp,
                       double *sie, double *press, double *cv, double *bmod, long n,
                       unsigned long output, int ls, double *lam, int *errs, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  const int g = grid_for(n, 1);
  switch (eh.kind) {
  case KIND_IDEAL: fill_kernel<KIND_IDEAL><<<g, 256, 0, s>>>(d, mat, rho, temp, sie, press, cv, bmod, n, output, ls, lam, errs); break;
  case KIND_STIFF: fill_kernel<KIND_STIFF><<<g, 256, 0, s>>>(d, mat, rho, temp, sie, press, cv, bmod, n, output, ls, lam, errs); break;
  case KIND_GRUN: fill_kernel<KIND_GRUN><<<g, 256, 0, s>>>(d, mat, rho, temp, sie, press, cv, bmod, n, output, ls, lam, errs); break;
  case KIND_VINET: fill_kernel<KIND_VINET><<<g, 256, 0, s>>>(d, mat, rho, temp, sie, press, cv, bmod, n, output, ls, lam, errs); break;
  default: fill_kernel<KIND_TABLE><<<g, 256, 0, s>>>(d, mat, rho, temp, sie, press, cv, bmod, n, output, ls, lam, errs); break;
  }
  return hipGetLastError();
}

hipError_t launch_rhoe_PT(const EosDev &eh, const EosDev *d, int mat, const double *P,
                          const double *T, double *rho, double *sie, long n, int ls,
                          double *lam, int *errs, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  const int g = grid_for(n, 1);
  switch (eh.kind) {
  case KIND_IDEAL: rhoe_PT_kernel<KIND_IDEAL><<<g, 256, 0, s>>>(d, mat, P, T, rho, sie, n, ls, lam, errs); break;
  case KIND_STIFF: rhoe_PT_kernel<KIND_STIFF><<<g, 256, 0, s>>>(d, mat, P, T, rho, sie, n, ls, lam, errs); break;
  case KIND_GRUN: rhoe_PT_kernel<KIND_GRUN><<<g, 256, 0, s>>>(d, mat, P, T, rho, sie, n, ls, lam, errs); break;
  case KIND_VINET: rhoe_PT_kernel<KIND_VINET><<<g, 256, 0, s>>>(d, mat, P, T, rho, sie, n, ls, lam, errs); break;
  default: rhoe_PT_kernel<KIND_TABLE><<<g, 256, 0, s>>>(d, mat, P, T, rho, sie, n, ls, lam, errs); break;
  }
  return hipGetLastError();
}

} // namespace sgm
