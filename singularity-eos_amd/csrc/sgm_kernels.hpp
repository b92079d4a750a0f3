// Declarations shared by the closure kernel translation units and the rest
// of the library: kernel argument records, method ids and the per-method
// launchers. Independent of the C ABI header, so that changes there do not
// rebuild the closure kernels.
#pragma once

#include "sgm_pte.hpp"
#include "sgm_pte_wave.hpp"

#include <hip/hip_runtime.h>

namespace sgm {

// materials per cell of the register-resident solvers (one lane per cell);
// cells of up to MAX_MATS_WAVE materials take one wave per cell
constexpr int MAX_MATS_PER_CELL = 8;
constexpr int MAX_MATS_WAVE = wave::MAX_MATS;
constexpr int MAX_REGISTRY = 64;
// status bit of a malformed cell (SGM_ERRB_INPUT of the C ABI)
constexpr int ERRB_INPUT = 32;

struct PteArgs {
  const EosDev *d_eos = nullptr;
  int nreg = 0;
  long ncell = 0, nnz = 0;
  const int *offsets = nullptr;
  const uint8_t *mat_ids = nullptr;
  double *rho = nullptr, *vfrac = nullptr, *sie = nullptr, *temp = nullptr, *press = nullptr;
  // per cell: vfrac_tot, and sie_tot = the specific internal energy (RhoT,
  // PT), the temperature (FixedT) or the pressure (FixedP); Tnorm: RhoT, PT
  const double *vfrac_tot = nullptr, *sie_tot = nullptr, *Tnorm = nullptr;
  double *lambda = nullptr;
  int *conv = nullptr, *niter = nullptr, *err = nullptr;
  double *resid = nullptr;
  // learned initial guess (batched kernels only): per entry vfrac, per cell T
  const double *guess_vfrac = nullptr, *guess_T = nullptr;
  int *used_guess = nullptr;
};
// Straggler parking for the batched bin kernels: a cell still iterating after
// `after` Newton iterations is saved (ParkSave, ParkWords(nm, nm + 1) FP64
// words per record) and its lane takes new work; after the bin, a tail kernel
// resumes the parked cells. Records are grouped by signature run (the
// stretch of sorted bin positions whose cells have the same materials in the
// same order), and each tail wave takes records of one run only, so a tail
// wave's EOS branches agree across its lanes.
struct ParkArgs {
  int after = 0;                    // 0: never park
  int stride = 0;                   // FP64 words per record
  double *rec = nullptr;            // records
  int *cell_of = nullptr;           // cell id per record
  const int *run_of = nullptr;      // per bin position: signature run
  const int *run_base = nullptr;    // per run: first record
  const int *run_cap = nullptr;     // per run: records available
  int *run_count = nullptr;         // per run: records claimed (claims past cap are not parked)
  const int *wave_run = nullptr;    // per tail wave: its run
  const int *wave_first = nullptr;  // per tail wave: its first record within the run
  int nwaves = 0;
  // the tail's second level: a cell still iterating `after2` iterations into
  // the first (one lane per cell) is saved again in its record for the
  // second, where TailGroup lanes share each cell
  int after2 = 0;                   // 0: one level
  int *l2_count = nullptr;          // cells saved for the second level
  int *l2_next = nullptr;           // second level (RhoT): next record item to take
};

// Closure formulations, by the method ids of the C ABI.
enum PteMethod {
  PTE_RHOT = 0,
  PTE_PT = 1,
  PTE_FIXED_T = 2,
  PTE_FIXED_P = 3,
  PTE_RHOU = 4,
  PTE_NMETHODS = 5
};

// Lanes per parked cell in the tail kernels: a group of 8 lanes runs each
// cell's line-search trials side by side (pte_iterate_group); the trials of
// PT (each P-T inversion starts from the previous trial's densities) and of
// RhoU (each T(rho, e) inversion starts from the lambda cache the previous
// trial left) are not independent, so those keep one lane per cell.
#ifndef SGM_TAIL_GROUP_MIN_NM
#define SGM_TAIL_GROUP_MIN_NM 2
#endif
constexpr int TailGroup(int method, int nm) {
  return (method != PTE_PT && method != PTE_RHOU && nm >= SGM_TAIL_GROUP_MIN_NM) ? 8 : 1;
}
// The largest material count with a register-resident bin kernel per method:
// RhoU's (2 nm)^2 Jacobian and per-material temperatures leave 5 and 6
// materials to the list kernel.
constexpr int BinMaxNM(int method) { return method == PTE_RHOU ? 4 : 6; }
// FP64 words of one straggler record (a bound on ParkWordCount of the
// method's solver: ParkWords for its equation count, plus RhoU's per-material
// temperatures)
constexpr int ParkStride(int method, int nm) {
  return method == PTE_RHOU ? ParkWords(nm, 2 * nm) + nm : ParkWords(nm, nm + 1);
}
// Whether a bin's tail may run its second level on whole waves (RhoT from 4
// materials: one material per lane pays for the wave's gathers only where
// there are enough materials to spread)
constexpr bool TailWide(int method, int nm) { return method == PTE_RHOT && nm >= 4; }

// Per-method launchers, one translation unit pair per method
// (sgm_pte_plain_<name>.hip: plain, list and wave kernels;
// sgm_pte_bins_<name>.hip: the batched bin kernels). range: NULL, or device
// {start, count} bounds into `cells` read by the kernel (n is then the host's
// upper bound on count, used to size the grid). The plain launcher runs the
// plain kernel and then the wave kernel over every cell; the list launcher
// the list kernel only; the wave launcher the wave kernel, which solves the
// listed cells of MAX_MATS_PER_CELL + 1 .. MAX_MATS_WAVE materials and
// leaves the others alone.
#define SGM_PTE_LAUNCHERS(name)                                                                \
  hipError_t launch_pte_plain_##name(const PteArgs &a, const MixParamsDev &P, hipStream_t s);  \
  hipError_t launch_pte_list_##name(const PteArgs &a, const MixParamsDev &P, const int *cells, \
                                    long n, const long *range, hipStream_t s);                 \
  hipError_t launch_pte_wave_##name(const PteArgs &a, const MixParamsDev &P, const int *cells, \
                                    long n, const long *range, hipStream_t s);                 \
  hipError_t launch_pte_bin_##name(int nm, const PteArgs &a, const MixParamsDev &P,            \
                                   const int *cells, long n, const long *range, int *counter,  \
                                   int block, int cpw, int grid, const ParkArgs *park,         \
                                   hipStream_t s);                                             \
  hipError_t launch_pte_tail_##name(int nm, const PteArgs &a, const MixParamsDev &P,           \
                                    const ParkArgs &park, hipStream_t s);
SGM_PTE_LAUNCHERS(rhot)
SGM_PTE_LAUNCHERS(pt)
SGM_PTE_LAUNCHERS(fixedt)
SGM_PTE_LAUNCHERS(fixedp)
SGM_PTE_LAUNCHERS(rhou)
#undef SGM_PTE_LAUNCHERS

// the plain per-call kernels: one lane per cell of 1..8 materials, one wave
// per cell of 9..64
inline hipError_t launch_pte_plain(int method, const PteArgs &a, const MixParamsDev &P,
                                   hipStream_t s) {
  switch (method) {
  case PTE_RHOT: return launch_pte_plain_rhot(a, P, s);
  case PTE_PT: return launch_pte_plain_pt(a, P, s);
  case PTE_FIXED_T: return launch_pte_plain_fixedt(a, P, s);
  case PTE_FIXED_P: return launch_pte_plain_fixedp(a, P, s);
  case PTE_RHOU: return launch_pte_plain_rhou(a, P, s);
  default: return hipErrorInvalidValue;
  }
}
// the one-wave-per-cell solver over an explicit cell list (cells of
// 9..64 materials; the others are left alone)
inline hipError_t launch_pte_wave(int method, const PteArgs &a, const MixParamsDev &P,
                                  const int *cells, long n, hipStream_t s,
                                  const long *range = nullptr) {
  switch (method) {
  case PTE_RHOT: return launch_pte_wave_rhot(a, P, cells, n, range, s);
  case PTE_PT: return launch_pte_wave_pt(a, P, cells, n, range, s);
  case PTE_FIXED_T: return launch_pte_wave_fixedt(a, P, cells, n, range, s);
  case PTE_FIXED_P: return launch_pte_wave_fixedp(a, P, cells, n, range, s);
  case PTE_RHOU: return launch_pte_wave_rhou(a, P, cells, n, range, s);
  default: return hipErrorInvalidValue;
  }
}
// the plain solvers over an explicit cell list: one lane per cell of 1..8
// materials, then (wide) one wave per cell of 9..64
inline hipError_t launch_pte_list(int method, const PteArgs &a, const MixParamsDev &P,
                                  const int *cells, long n, hipStream_t s,
                                  const long *range = nullptr, bool wide = true) {
  hipError_t e;
  switch (method) {
  case PTE_RHOT: e = launch_pte_list_rhot(a, P, cells, n, range, s); break;
  case PTE_PT: e = launch_pte_list_pt(a, P, cells, n, range, s); break;
  case PTE_FIXED_T: e = launch_pte_list_fixedt(a, P, cells, n, range, s); break;
  case PTE_FIXED_P: e = launch_pte_list_fixedp(a, P, cells, n, range, s); break;
  case PTE_RHOU: e = launch_pte_list_rhou(a, P, cells, n, range, s); break;
  default: return hipErrorInvalidValue;
  }
  if (e != hipSuccess || !wide) return e;
  return launch_pte_wave(method, a, P, cells, n, s, range);
}
// batched persistent bin (2..6 materials): cells all have nm materials
inline hipError_t launch_pte_bin(int method, int nm, const PteArgs &a, const MixParamsDev &P,
                                 const int *cells, long n, int *counter, int block, int cpw,
                                 int grid, hipStream_t s, const long *range = nullptr,
                                 const ParkArgs *park = nullptr) {
  switch (method) {
  case PTE_RHOT:
    return launch_pte_bin_rhot(nm, a, P, cells, n, range, counter, block, cpw, grid, park, s);
  case PTE_PT:
    return launch_pte_bin_pt(nm, a, P, cells, n, range, counter, block, cpw, grid, park, s);
  case PTE_FIXED_T:
    return launch_pte_bin_fixedt(nm, a, P, cells, n, range, counter, block, cpw, grid, park, s);
  case PTE_FIXED_P:
    return launch_pte_bin_fixedp(nm, a, P, cells, n, range, counter, block, cpw, grid, park, s);
  case PTE_RHOU:
    return launch_pte_bin_rhou(nm, a, P, cells, n, range, counter, block, cpw, grid, park, s);
  default: return hipErrorInvalidValue;
  }
}
// resumes a bin's parked cells (one 64-lane block per tail wave)
inline hipError_t launch_pte_tail(int method, int nm, const PteArgs &a, const MixParamsDev &P,
                                  const ParkArgs &park, hipStream_t s) {
  switch (method) {
  case PTE_RHOT: return launch_pte_tail_rhot(nm, a, P, park, s);
  case PTE_PT: return launch_pte_tail_pt(nm, a, P, park, s);
  case PTE_FIXED_T: return launch_pte_tail_fixedt(nm, a, P, park, s);
  case PTE_FIXED_P: return launch_pte_tail_fixedp(nm, a, P, park, s);
  case PTE_RHOU: return launch_pte_tail_rhou(nm, a, P, park, s);
  default: return hipErrorInvalidValue;
  }
}
hipError_t launch_zero_ints(int *p, int n, hipStream_t s);
// flag = 0 if a[0..n) == b[0..n) and c[0..m) == d[0..m), else 1
hipError_t launch_compare_csr(const int *a, const int *b, long n, const uint8_t *c,
                              const uint8_t *d, long m, int *flag, hipStream_t s);
// sort keys of CSR cells: material count, then the registry ids in cell
// order; mod_mask (bit i: registry entry i has modifiers) sends the cells of
// two or more materials that hold a modified material past the
// register-resident bins (count field | 0x40), to the list kernels, which
// apply modifiers
hipError_t launch_plan_keys(const int *offsets, const uint8_t *mat, long ncell, long nnz,
                            unsigned long long *keys, int *ids, hipStream_t s,
                            const int *cost = nullptr, int threshold = 0,
                            unsigned long long mod_mask = 0);
int device_cu_count();
} // namespace sgm
