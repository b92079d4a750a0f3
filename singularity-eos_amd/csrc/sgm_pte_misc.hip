// Small helper kernels for the batched PTE plan: counter resets (fill
// kernels, so the resets can sit inside a captured hipGraph), per-cell sort
// keys (material count, then the material signature in cell order).
#include "sgm_internal.hpp"

namespace sgm {
namespace {

__global__ void zero_ints_kernel(int *p, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0;
}

// key = nmat << 56 | material ids in cell order (8 bits each; ids < 64, so bit
// 55 is free). With a cost array, bit 55 is set on cells below the
// threshold: within a material count, the costly cells sort first.
// Cells of two or more materials that hold a material with modifiers get
// count field nm | 0x40: past the register-resident bins; cells of 9..64
// materials (the one-wave-per-cell kernel's) count field 0x80.
__global__ void plan_keys_kernel(const int *offsets, const uint8_t *mat, long ncell, long nnz,
                                 const int *cost, int threshold, unsigned long long mod_mask,
                                 unsigned long long *keys, int *ids) {
  const long c = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= ncell) return;
  const long o0 = offsets[c], o1 = offsets[c + 1];
  unsigned long long k;
  if (o0 >= 0 && o1 >= o0 && o1 <= nnz && o1 - o0 <= MAX_MATS_PER_CELL) {
    const int nm = (int)(o1 - o0);
    bool modified = false;
    for (int m = 0; m < nm && mod_mask; ++m)
      modified = modified || (mat[o0 + m] < 64 && ((mod_mask >> mat[o0 + m]) & 1ull));
    k = (unsigned long long)((nm >= 2 && modified) ? (nm | 0x40) : nm) << 56;
    for (int m = 0; m < nm && m < 7; ++m) k |= (unsigned long long)mat[o0 + m] << (48 - 8 * m);
    if (cost && cost[c] < threshold) k |= 1ull << 55;
  } else if (o0 >= 0 && o1 >= o0 && o1 <= nnz && o1 - o0 <= MAX_MATS_WAVE) {
    k = 0x80ull << 56;
    for (int m = 0; m < 7; ++m) k |= (unsigned long long)mat[o0 + m] << (48 - 8 * m);
  } else {
    k = 0xFFull << 56;
  }
  keys[c] = k;
  ids[c] = (int)c;
}

} // namespace

hipError_t launch_zero_ints(int *p, int n, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  zero_ints_kernel<<<(n + 255) / 256, 256, 0, s>>>(p, n);
  return hipGetLastError();
}

namespace {
__global__ void compare_csr_kernel(const int *a, const int *b, long n, const uint8_t *c,
                                   const uint8_t *d, long m, int *flag) {
  const long stride = (long)gridDim.x * blockDim.x;
  bool diff = false;
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n || i < m; i += stride)
    diff = diff || (i < n && a[i] != b[i]) || (i < m && c[i] != d[i]);
  if (diff) flag[0] = 1;  // every writer stores the same value
}
} // namespace

hipError_t launch_compare_csr(const int *a, const int *b, long n, const uint8_t *c,
                              const uint8_t *d, long m, int *flag, hipStream_t s) {
  hipError_t e = launch_zero_ints(flag, 1, s);
  if (e != hipSuccess) return e;
  const long work = n > m ? n : m;
  const long blocks = (work + 255) / 256;
  const int g = (int)(blocks < 4096 ? (blocks > 0 ? blocks : 1) : 4096);
  compare_csr_kernel<<<g, 256, 0, s>>>(a, b, n, c, d, m, flag);
  return hipGetLastError();
}

hipError_t launch_plan_keys(const int *offsets, const uint8_t *mat, long ncell, long nnz,
                            unsigned long long *keys, int *ids, hipStream_t s, const int *cost,
                            int threshold, unsigned long long mod_mask) {
  if (ncell <= 0) return hipSuccess;
  plan_keys_kernel<<<(int)((ncell + 255) / 256), 256, 0, s>>>(offsets, mat, ncell, nnz, cost,
                                                                threshold, mod_mask, keys, ids);
  return hipGetLastError();
}

} // namespace sgm
