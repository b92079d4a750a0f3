// Plain per-call PTE kernels, method PTESolverFixedT.
#include "sgm_pte_kernels.inl"
namespace sgm {
hipError_t launch_pte_plain_fixedt(const PteArgs &a, const MixParamsDev &P, hipStream_t s) {
  return kern::plain_launch<2>(a, P, s);
}
hipError_t launch_pte_list_fixedt(const PteArgs &a, const MixParamsDev &P, const int *cells, long n,
                              const long *range, hipStream_t s) {
  return kern::list_launch<2>(a, P, cells, n, s, range);
}
hipError_t launch_pte_wave_fixedt(const PteArgs &a, const MixParamsDev &P, const int *cells, long n,
                              const long *range, hipStream_t s) {
  return kern::wave_launch<2>(a, P, cells, n, s, range);
}
} // namespace sgm
