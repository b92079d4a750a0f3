// Batched persistent PTE bin kernels (2..4 materials), method PTESolverRhoU.
#include "sgm_pte_kernels.inl"
namespace sgm {
hipError_t launch_pte_bin_rhou(int nm, const PteArgs &a, const MixParamsDev &P, const int *cells,
                               long n, const long *range, int *counter, int block, int cpw,
                               int grid, const ParkArgs *park, hipStream_t s) {
  return kern::bin_launch<PTE_RHOU>(nm, a, P, cells, n, counter, block, cpw, grid, s, range,
                                    park);
}
hipError_t launch_pte_tail_rhou(int nm, const PteArgs &a, const MixParamsDev &P,
                                const ParkArgs &park, hipStream_t s) {
  return kern::tail_launch<PTE_RHOU>(nm, a, P, park, s);
}
} // namespace sgm
