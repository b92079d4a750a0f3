// PTE closure kernels (templates; instantiated per method -- RhoT, PT,
// FixedT, FixedP -- in the sgm_pte_*.hip translation units so the build
// compiles them in parallel).
//
//  * pte_plain_kernel: the per-call path. One lane per cell over the caller's
//    CSR order; the material count is read per cell and dispatched to the
//    register-resident solver instantiated for 1..8 materials (each a separate
//    function, called once per cell).
//  * pte_bin_kernel<NM>: the batched path. A persistent grid walks a bin of
//    cells that all have NM materials (sorted by material signature so the
//    EOS dispatch is wave-uniform); lanes retire finished cells and take new
//    ones between Newton iterations, from a per-wave reserve refilled
//    `chunk` (>= 64) cells at a time from a device counter.
//  * pte_wave_kernel: cells of 9..64 materials, one wave per cell
//    (sgm_pte_wave.hpp), after the plain kernel and beside the list kernel.
//
// All stage the material registry once per block into LDS (the per-material
// parameters every EOS evaluation touches) and keep all per-cell state in
// VGPRs (the wave kernel: the Newton system in LDS). Per cell: reads rho, vfrac, temp, press per material + sie_tot,
// vfrac_tot (+Tnorm); writes rho, vfrac, sie, temp, press + status.
#pragma once
#include "sgm_kernels.hpp"

#include <type_traits>

namespace sgm {
namespace kern {

__device__ __forceinline__ void stage_registry(EosDev *lds, const EosDev *reg, int nreg) {
  const double *s = reinterpret_cast<const double *>(reg);
  double *d = reinterpret_cast<double *>(lds);
  const int nd = nreg * (int)(sizeof(EosDev) / 8);
  for (int i = threadIdx.x; i < nd; i += blockDim.x) d[i] = s[i];
  __syncthreads();
}

// closure formulation by method id: 0 RhoT, 1 PT, 2 FixedT, 3 FixedP, 4 RhoU.
// MODS: whether the EOS calls apply the materials' modifiers. The batched
// bin and tail kernels are built without (the plans and the server send
// cells with a modified material to the list kernel), so their hot loops
// carry no modifier code; the plain and list kernels are built with.
template <int NM, int METHOD, bool MODS = true>
using SysOf = typename std::conditional<
    METHOD == 0, PteRhoT<NM, MODS>,
    typename std::conditional<
        METHOD == 1, PtePT<NM, MODS>,
        typename std::conditional<
            METHOD == 2, PteFixedT<NM, MODS>,
            typename std::conditional<METHOD == 3, PteFixedP<NM, MODS>,
                                      PteRhoU<NM, MODS>>::type>::type>::type>::type;

// formulations whose materials keep temperatures of their own (RhoU)
template <class S, class = void>
struct PerMaterialT : std::false_type {};
template <class S>
struct PerMaterialT<S, std::void_t<decltype(S::kPerMaterialT)>>
    : std::integral_constant<bool, S::kPerMaterialT> {};

template <int NM, class Sys>
__device__ __forceinline__ void load_cell(Sys &s, Real (&Tin)[NM], const EosDev *lds,
                                          const PteArgs &a, const MixParamsDev &P, long c,
                                          int off) {
#pragma unroll
  for (int m = 0; m < NM; ++m) {
    const int k = off + m;
    s.eidx[m] = a.mat_ids[k];
    s.rho[m] = a.rho[k];
    s.vfrac[m] = a.vfrac[k];
    s.press[m] = a.press[k];
    Tin[m] = a.temp[k];
    if (a.lambda) {
      s.lam[m].present = true;
      s.lam[m].lRho = a.lambda[3L * k + 0];
      s.lam[m].lT = a.lambda[3L * k + 1];
    } else {
      s.lam[m].present = false;
    }
  }
  s.SetCellInputs(a.vfrac_tot[c], a.sie_tot[c], a.Tnorm ? a.Tnorm[c] : 0.0);
  s.err = 0;
  s.P = P;
}

template <int NM, class Sys>
__device__ __forceinline__ void store_cell(const Sys &s, const CellStatus &st, const PteArgs &a,
                                           long c, int off) {
  const Real Tout = s.Tout();
#pragma unroll
  for (int m = 0; m < NM; ++m) {
    const int k = off + m;
    a.rho[k] = s.rho[m];
    a.vfrac[k] = s.vfrac[m];
    a.sie[k] = s.sie[m];
    if constexpr (PerMaterialT<Sys>::value)
      a.temp[k] = s.TempOut(m);
    else if (Sys::kStoresTemp)
      a.temp[k] = Tout;
    a.press[k] = s.Pout(m);
    if (a.lambda) {
      a.lambda[3L * k + 0] = s.lam[m].lRho;
      a.lambda[3L * k + 1] = s.lam[m].lT;
    }
  }
  if (a.conv) a.conv[c] = st.converged;
  if (a.niter) a.niter[c] = st.niter;
  if (a.resid) a.resid[c] = st.residual;
  if (a.err) a.err[c] = st.err;
}

template <int NM, int METHOD>
__device__ __forceinline__ void solve_cell(const EosDev *lds, const PteArgs &a,
                                           const MixParamsDev &P, long c, int off) {
  SysOf<NM, METHOD> s;
  Real Tin[NM];
  load_cell<NM>(s, Tin, lds, a, P, c, off);
  const CellStatus st = pte_solve<NM>(s, Tin);
  store_cell<NM>(s, st, a, c, off);
}

template <int NM, int METHOD>
__device__ __attribute__((noinline)) void solve_cell_call(const EosDev *lds, const PteArgs &a,
                                                          const MixParamsDev &P, long c,
                                                          int off) {
  solve_cell<NM, METHOD>(lds, a, P, c, off);
}

// validates a cell's CSR slice; returns nmat (1..MAX_MATS_WAVE) or -1
__device__ __forceinline__ int cell_slice(const PteArgs &a, long c, int &off) {
  const long o0 = a.offsets[c], o1 = a.offsets[c + 1];
  if (!(o0 >= 0 && o1 >= o0 && o1 <= a.nnz)) return -1;
  const int nm = (int)(o1 - o0);
  if (nm < 1 || nm > MAX_MATS_WAVE) return -1;
  for (int m = 0; m < nm; ++m)
    if (a.mat_ids[o0 + m] >= a.nreg) return -1;
  off = (int)o0;
  return nm;
}

__device__ __forceinline__ void mark_bad(const PteArgs &a, long c) {
  if (a.conv) a.conv[c] = 0;
  if (a.niter) a.niter[c] = 0;
  if (a.resid) a.resid[c] = 0;
  if (a.err) a.err[c] = ERRB_INPUT;
}

template <int METHOD>
__global__ void __launch_bounds__(256) pte_plain_kernel(PteArgs a, MixParamsDev P) {
  extern __shared__ EosDev lds[];
  stage_registry(lds, a.d_eos, a.nreg);
  const long nthr = (long)gridDim.x * blockDim.x;
  for (long c = (long)blockIdx.x * blockDim.x + threadIdx.x; c < a.ncell; c += nthr) {
    int off = 0;
    const int nm = cell_slice(a, c, off);
    switch (nm) {
    case 1: solve_cell_call<1, METHOD>(lds, a, P, c, off); break;
    case 2: solve_cell_call<2, METHOD>(lds, a, P, c, off); break;
    case 3: solve_cell_call<3, METHOD>(lds, a, P, c, off); break;
    case 4: solve_cell_call<4, METHOD>(lds, a, P, c, off); break;
    case 5: solve_cell_call<5, METHOD>(lds, a, P, c, off); break;
    case 6: solve_cell_call<6, METHOD>(lds, a, P, c, off); break;
    case 7: solve_cell_call<7, METHOD>(lds, a, P, c, off); break;
    case 8: solve_cell_call<8, METHOD>(lds, a, P, c, off); break;
    default:
      // cells of 9..64 materials are the wave kernel's
      if (!(nm > MAX_MATS_PER_CELL && nm <= wave::WaveMaxNM(METHOD))) mark_bad(a, c);
      break;
    }
  }
}

// Learned initial guess: restart the cell from the predicted volume
// fractions (densities follow from the fixed partial densities) and
// temperature (as Tnorm), and keep that start only if its initial residual is
// below the solver's own start and no condensed material (any model but the
// ideal gas) sits more than a factor 2 from its density in the solver's own
// start: condensed phases barely compress at these pressures, and a start far
// from their reference density can lead Newton to a second, unphysical
// equilibrium (an expanded Vinet branch with the gas collapsed).
template <int NM, class Sys>
__device__ __forceinline__ void start_from_guess(Sys &s, Real (&Tin)[NM], DriverState &d,
                                                 CellStatus &st, const EosDev *lds,
                                                 const PteArgs &a, const MixParamsDev &P, long c,
                                                 int off) {
  const Real gT = a.guess_T[c];
  int used = 0;
  if (gT > 0) {  // NaN: no guess for this cell
    const Real err_plain = d.err;
    Real rho_plain[NM];
#pragma unroll
    for (int m = 0; m < NM; ++m) rho_plain[m] = s.rho[m];
    load_cell<NM>(s, Tin, lds, a, P, c, off);
    bool ok = true;
#pragma unroll
    for (int m = 0; m < NM; ++m) {
      const Real g = a.guess_vfrac[off + m];
      ok = ok && (g > 0);
      const Real rb = s.rho[m] * s.vfrac[m];
      s.vfrac[m] = g;
      s.rho[m] = rb / g;
      if (s.E(m).kind != KIND_IDEAL)
        ok = ok && (s.rho[m] <= 2 * rho_plain[m]) && (2 * s.rho[m] >= rho_plain[m]);
      // the caller's pressure belongs to the caller's state: let the solver
      // evaluate it at the guessed state (a NaN input pressure is recomputed)
      s.press[m] = __builtin_nan("");
    }
    s.Tnorm = gT;
    if (ok) {
      pte_begin<NM>(s, Tin, d, st);
      used = d.err < err_plain ? 1 : 0;
    }
    if (!used) {
      load_cell<NM>(s, Tin, lds, a, P, c, off);
      pte_begin<NM>(s, Tin, d, st);
    }
  }
  if (a.used_guess) a.used_guess[c] = used;
}

// idle lanes that trigger a refill: refilling as soon as one lane is idle
// measured faster on the uneven 2^20-cell workload than waiting for 16
constexpr int REFILL_BATCH = 1;

// Batched bin kernel with lane refill: every lane owns one cell and runs one
// Newton iteration of it per trip of the loop; a lane whose cell has finished
// stores it and, once REFILL_BATCH lanes of the wave are idle (or none is
// busy), takes the next cell of the bin, so a wave does not idle behind its
// slowest cell. Cells come from a per-wave reserve that the wave
// refills `chunk` cells at a time with one atomic on the bin counter; the
// sorted bin order keeps a wave's cells on the same material signature.
// Two waves per SIMD (256 registers per lane): a wave's dependent FP64
// chains then hide behind the other wave's. For 3..6 materials a few of the
// cell's values spill to scratch (the 2-material state fits); measured on the
// uneven 2^20-cell workload this is faster and steadier than one wave per
// SIMD with the whole state in 512 registers.
#ifndef SGM_BIN_WAVES_PER_EU
#define SGM_BIN_WAVES_PER_EU 2
#endif
template <int NM, int METHOD>
__global__ void __launch_bounds__(256)
__attribute__((amdgpu_waves_per_eu(SGM_BIN_WAVES_PER_EU,
                                   SGM_BIN_WAVES_PER_EU))) pte_bin_kernel(PteArgs a, MixParamsDev P,
                                                      const int *__restrict__ cells,
                                                      long ncells, int *counter, int chunk,
                                                      const long *range, ParkArgs pk) {
  extern __shared__ EosDev lds[];
  stage_registry(lds, a.d_eos, a.nreg);
  if (range) {  // bin bounds computed on the device: {start, count} into cells
    cells += range[0];
    ncells = range[1];
  }
  const int lane = threadIdx.x & 63;
  const unsigned long long below = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
  if (chunk < 64) chunk = 64;  // one claim must cover a whole wave
  SysOf<NM, METHOD, false> s;
  DriverState d;
  CellStatus st;
  long c = -1, jpos = -1;
  int off = 0;
  bool park_ok = false;  // this lane may still park its cell (one attempt per cell)
  bool active = false;  // this lane holds a cell that is still iterating
  bool done = false;    // the bin is exhausted for this wave
  long rbase = 0, rend = 0;  // wave-uniform reserve [rbase, rend) of bin positions
  for (;;) {
    // hand out cells to idle lanes (REFILL_BATCH at a time)
    const unsigned long long need = __ballot(!active && !done);
    const int nidle = __popcll(need);
    const int nactive = __popcll(__ballot(active));
    if (need && (nidle >= REFILL_BATCH || nactive == 0)) {
      const long cnt = __popcll(need);
      const long rank = __popcll(need & below);
      long j = -1;
      if (rank < rend - rbase) {
        j = rbase + rank;
      }
      if (rend - rbase < cnt) {
        int b = 0;
        if (lane == 0) b = atomicAdd(counter, chunk);
        b = __shfl(b, 0, 64);
        const long avail = rend - rbase;
        if (rank >= avail) j = (long)b + (rank - avail);
        rbase = (long)b + (cnt - avail);
        rend = (long)b + chunk;
      } else {
        rbase += cnt;
      }
      if (rend > ncells) rend = ncells;
      if (rbase > rend) rbase = rend;
      if (!active && !done) {
        if (j >= 0 && j < ncells) {
          c = cells[j];
          jpos = j;
          const int nm = (c >= 0 && c < a.ncell) ? cell_slice(a, c, off) : -1;
          if (nm == NM) {
            Real Tin[NM];
            load_cell<NM>(s, Tin, lds, a, P, c, off);
            pte_begin<NM>(s, Tin, d, st);
            if (a.guess_T) start_from_guess<NM>(s, Tin, d, st, lds, a, P, c, off);
            active = true;
            park_ok = pk.after > 0;
          } else if (c >= 0 && c < a.ncell) {
            mark_bad(a, c);
          }
        } else {
          done = true;
        }
      }
    }
    if (!__any(active)) {
      if (__all(done)) break;
      continue;
    }
    if (active && pte_iterate<NM>(s, d, st)) {
      pte_end<NM>(s, d, st);
      store_cell<NM>(s, st, a, c, off);
      active = false;
    } else if (active && park_ok && d.niter >= pk.after) {
      // a straggler: save it in its signature run's records for the tail
      // kernel and free the lane
      const int r = pk.run_of[jpos];
      const int i = atomicAdd(pk.run_count + r, 1);
      if (i < pk.run_cap[r]) {
        const long slot = (long)pk.run_base[r] + i;
        ParkSave(s, d, st, pk.rec + slot * pk.stride);
        pk.cell_of[slot] = (int)c;
        active = false;
      }
      park_ok = false;
    }
  }
}

// Resumes the parked cells of one bin in two levels. Level 1 (LEVEL 1): a
// persistent grid of one-wave blocks walks the bin's tail wave slots (each
// on records of a single signature run), one lane per cell; a cell that
// reaches a fixed point of the Newton loop skips to its last pass, and a
// cell still iterating pk.after2 iterations later is saved back into its
// record for level 2 (finished cells clear their record's cell id). Level 2
// finishes the cells that keep wandering. When few remain (RhoT with 4+
// materials, TailWide, at most wide_max cells: a latency problem) each runs
// on a whole wave (LEVEL 3, tail_wide), cells taken in record order from a
// queue. Otherwise (LEVEL 2: FixedT, FixedP, RhoT with fewer materials, and
// RhoT with more cells, as after cold starts: a throughput problem)
// TailGroup lanes (G) share each cell and run its line-search trials side by
// side (pte_iterate_group) while those cells fit the grid at once, else one
// lane each. TailWide bins launch both kernels, each of which reads the
// level-2 count and returns at once when the other applies. Every cell runs
// to the end of its Newton loop.
#ifndef SGM_TAIL_WAVES_PER_EU_LARGE
#define SGM_TAIL_WAVES_PER_EU_LARGE 1
#endif
template <int NM>
constexpr int TailWavesL1() {
  return NM >= 5 ? SGM_TAIL_WAVES_PER_EU_LARGE : 2;
}

// Level 2 of a RhoT tail: the block's one wave takes record items (wave
// slot, position) from the bin's queue pk.l2_next until none is left, and
// runs each remaining cell with all 64 lanes (pte_iterate_wide). The
// fixed-point probe's copy sits in LDS.
// The loop keeps the wave converged: every lane executes every statement
// (the queue's atomic adds 1 from lane 0 and 0 from the others; all lanes
// write the same probe words and the same results), and every value that
// steers the loop (the item, whether the cell is done, whether to probe or
// skip) goes through readfirstlane, a scalar the compiler knows to be
// wave-uniform. A lane-0-only branch, or a value broadcast with a lane
// shuffle, would count as divergent, and the compiler may then split the
// loop so that some lanes run an iteration without the others, which the
// cross-lane gathers of each pass cannot allow.
template <int NM>
__device__ __forceinline__ void tail_wide(const PteArgs &a, const MixParamsDev &P,
                                          const ParkArgs &pk, int lane) {
  using Sys = SysOf<NM, PTE_RHOT, false>;
  constexpr int per_slot = 64 / TailGroup(PTE_RHOT, NM);
  constexpr int PW = ParkWordCount<Sys>();
  __shared__ double probe_lds[PW];
  const long items = (long)pk.nwaves * per_slot;
  // these waves carry the bin's longest dependency chains: let them issue
  // ahead of other kernels' waves on their CU
  __builtin_amdgcn_s_setprio(3);
  for (;;) {
    const int item = __builtin_amdgcn_readfirstlane(atomicAdd(pk.l2_next, lane == 0 ? 1 : 0));
    if (item >= items) break;
    const int w = item / per_slot;
    const int r = pk.wave_run[w];
    const int nr = min(pk.run_count[r], pk.run_cap[r]);
    const int i = pk.wave_first[w] + item % per_slot;
    if (i >= nr) continue;
    const long slot = (long)pk.run_base[r] + i;
    const long c = pk.cell_of[slot];
    if (c < 0) continue;  // finished in level 1
    Sys s;
    DriverState d;
    CellStatus st;
    ParkLoad(s, d, st, pk.rec + slot * pk.stride);
    s.P = P;
    fixed_point_probe(s, d, st, probe_lds);  // the copy of the loaded state
    ProbeGate gate;
    gate.start(d.err);
    for (;;) {
      const bool done = pte_iterate_wide<NM>(s, d, st, lane);
      if (__builtin_amdgcn_readfirstlane(done ? 1 : 0)) break;
      const bool due = gate.due(d.err);
      if (__builtin_amdgcn_readfirstlane(due ? 1 : 0)) {
        const bool same = fixed_point_probe(s, d, st, probe_lds);
        if (gate.skip(__builtin_amdgcn_readfirstlane(same ? 1 : 0) != 0)) skip_to_last_pass(d);
      }
    }
    pte_end<NM>(s, d, st);
    store_cell<NM>(s, st, a, c, a.offsets[c]);
  }
}

template <int NM, int METHOD, int WPE, int LEVEL>
__global__ void __launch_bounds__(64)
__attribute__((amdgpu_waves_per_eu(WPE, WPE))) pte_tail_kernel(PteArgs a, MixParamsDev P,
                                                                ParkArgs pk, int resident,
                                                                int wide_max) {
  constexpr int G = TailGroup(METHOD, NM);
  constexpr int per_slot = 64 / G;  // records per wave slot
  extern __shared__ EosDev lds[];
  stage_registry(lds, a.d_eos, a.nreg);
  const int lane = (int)threadIdx.x;
  const bool two_levels = G > 1 && pk.after2 > 0 && pk.l2_count && P.line_search_max_iter <= G;
  if constexpr (LEVEL == 3) {
    static_assert(TailWide(METHOD, NM), "whole-wave cells: RhoT, 4+ materials");
    if (two_levels && *pk.l2_count <= wide_max) tail_wide<NM>(a, P, pk, lane);
  } else {
    // level 2: groups while its cells fit the GPU's resident waves at once;
    // past that the level is throughput-bound and keeps one lane per cell
    bool grp = false;
    if constexpr (LEVEL == 2) {
      if (!two_levels) return;
      const int l2 = *pk.l2_count;
      if (TailWide(METHOD, NM) && l2 <= wide_max) return;  // the whole-wave kernel's
      grp = (long)l2 * G <= 64L * (long)resident;
    }
    const int lpc = grp ? G : 1;  // lanes per cell
    const int spb = grp ? 1 : G;  // wave slots per block step
    const int g = lane % lpc, base = lane - g;
    for (long W = blockIdx.x; W * spb < pk.nwaves; W += gridDim.x) {
      const long w = grp ? W : W * G + lane / per_slot;
      if (w >= pk.nwaves) continue;
      const int r = pk.wave_run[w];
      const int nr = min(pk.run_count[r], pk.run_cap[r]);
      const int i = pk.wave_first[w] + (grp ? lane / G : lane % per_slot);
      if (i >= nr) continue;  // in a group, the whole group skips
      const long slot = (long)pk.run_base[r] + i;
      const long c = pk.cell_of[slot];
      if (c < 0) continue;  // finished in level 1
      const int off = a.offsets[c];
      SysOf<NM, METHOD, false> s;
      DriverState d;
      CellStatus st;
      // the record, free once loaded, keeps the probe's copy of the state
      // when one lane runs the cell (a group keeps its copy in LDS)
      double *rec = pk.rec + slot * pk.stride;
      ParkLoad(s, d, st, rec);
      s.P = P;
      ProbeGate gate;
      gate.start(d.err);
      bool moved = false;
      if constexpr (G > 1 && LEVEL == 2) {
        if (grp) {
          // the group's probe copy in LDS: a probe's comparison and store
          // then cost LDS latency instead of a round trip to memory. Every
          // lane of the group probes (the same words, the same result), so
          // no lane branches on its position in the group
          constexpr int PW = ParkWordCount<SysOf<NM, METHOD, false>>();
          __shared__ double probe_lds[64 / G][PW];
          double *pw = probe_lds[lane / G];
          fixed_point_probe(s, d, st, pw);  // the copy of the loaded state
          while (!pte_iterate_group<NM, G>(s, d, st, g, base))
            if (gate.due(d.err) && gate.skip(fixed_point_probe(s, d, st, pw)))
              skip_to_last_pass(d);
        }
      }
      if (!grp) {
        const bool can_move = LEVEL == 1 && two_levels;
        const int stop = d.niter + pk.after2;
        while (!pte_iterate<NM>(s, d, st)) {
          if (gate.due(d.err) && gate.skip(fixed_point_probe(s, d, st, rec))) {
            skip_to_last_pass(d);
            continue;  // its last pass runs here, not in level 2
          }
          if (can_move && d.niter >= stop) {
            ParkSave(s, d, st, rec);
            atomicAdd(pk.l2_count, 1);
            moved = true;
            break;
          }
        }
      }
      if (moved) continue;
      pte_end<NM>(s, d, st);
      if (g == 0) {
        store_cell<NM>(s, st, a, c, off);
        if (LEVEL == 1 && two_levels) pk.cell_of[slot] = -1;
      }
    }
  }
}

template <int METHOD>
hipError_t wave_launch(const PteArgs &a, const MixParamsDev &P, const int *cells, long n,
                       hipStream_t s, const long *range = nullptr);

template <int METHOD>
hipError_t plain_launch(const PteArgs &a, const MixParamsDev &P, hipStream_t s) {
  if (a.ncell <= 0) return hipSuccess;
  const size_t lds = (size_t)a.nreg * sizeof(EosDev);
  const long blocks_needed = (a.ncell + 255) / 256;
  const long cap = (long)device_cu_count() * 32;
  const int g = (int)(blocks_needed < cap ? blocks_needed : cap);
  pte_plain_kernel<METHOD><<<g, 256, lds, s>>>(a, P);
  const hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  return wave_launch<METHOD>(a, P, nullptr, a.ncell, s);
}

template <int METHOD>
hipError_t bin_launch(int nm, const PteArgs &a, const MixParamsDev &P, const int *cells, long n,
                      int *counter, int block, int cpw, int grid, hipStream_t s,
                      const long *range = nullptr, const ParkArgs *park = nullptr) {
  if (n <= 0 && !range) return hipSuccess;
  const size_t lds = (size_t)a.nreg * sizeof(EosDev);
  const ParkArgs pk = park ? *park : ParkArgs{};
#define SGM_BIN(K)                                                                             \
  case K:                                                                                      \
    if constexpr (K <= BinMaxNM(METHOD)) {                                                     \
      static_assert(ParkWordCount<SysOf<K, METHOD, false>>() <= ParkStride(METHOD, K),         \
                    "park record");                                                            \
      if (pk.after > 0 && pk.stride < ParkStride(METHOD, K)) return hipErrorInvalidValue;      \
      pte_bin_kernel<K, METHOD><<<grid, block, lds, s>>>(a, P, cells, n, counter, cpw, range,  \
                                                         pk);                                  \
      break;                                                                                   \
    } else {                                                                                   \
      return hipErrorInvalidValue;                                                             \
    }
  switch (nm) {
    SGM_BIN(2)
    SGM_BIN(3)
    SGM_BIN(4)
    SGM_BIN(5)
    SGM_BIN(6)
  default: return hipErrorInvalidValue;
  }
#undef SGM_BIN
  return hipGetLastError();
}

template <int METHOD>
hipError_t tail_launch(int nm, const PteArgs &a, const MixParamsDev &P, const ParkArgs &pk,
                       hipStream_t s) {
  if (pk.nwaves <= 0) return hipSuccess;
  if (nm < 2 || nm > BinMaxNM(METHOD) || pk.stride < ParkStride(METHOD, nm))
    return hipErrorInvalidValue;
  const size_t lds = (size_t)a.nreg * sizeof(EosDev);
  // persistent: at most as many waves as the GPU holds at the kernel's
  // occupancy (one wave per SIMD from 5 materials, else two)
  const int ncu = device_cu_count();
  const int cap = (nm >= 5 ? SGM_TAIL_WAVES_PER_EU_LARGE : 2) * 4 * ncu;
  const int G = TailGroup(METHOD, nm);
  const int g1 = (pk.nwaves + G - 1) / G;
  const int g = g1 < cap ? g1 : cap;
  const bool l2 = G > 1 && pk.after2 > 0 && pk.l2_count &&
                  P.line_search_max_iter <= G && (!TailWide(METHOD, nm) || pk.l2_next);
  // level 2 on whole waves (RhoT): one wave per CU, so no two of the
  // kernel's long chains share a CU, each taking cells from the queue; in
  // groups: one block per wave slot, so that no block walks two slots that
  // both hold long-running cells (its few cells would then run one after
  // the other); empty slots exit at once
  const long items = (long)pk.nwaves * (64 / G);
  const int g3 = (int)(items < ncu ? items : ncu);
  // whole-wave level 2 while at most four cells per CU remain (the CU's
  // wave takes them one after another from the queue)
  const int wide_max = 4 * ncu;
  for (int level = 1; level <= (l2 ? 2 : 1); ++level) {
#define SGM_TAIL(K)                                                                            \
  case K:                                                                                      \
    if constexpr (K > BinMaxNM(METHOD)) {                                                      \
      return hipErrorInvalidValue;                                                             \
    } else if (level == 1) {                                                                   \
      pte_tail_kernel<K, METHOD, TailWavesL1<K>(), 1><<<g, 64, lds, s>>>(a, P, pk, cap,        \
                                                                         wide_max);           \
    } else {                                                                                   \
      if constexpr (TailWide(METHOD, K))                                                       \
        pte_tail_kernel<K, METHOD, 1, 3><<<g3, 64, lds, s>>>(a, P, pk, 4 * ncu, wide_max);     \
      pte_tail_kernel<K, METHOD, 1, 2><<<pk.nwaves, 64, lds, s>>>(a, P, pk, 4 * ncu,           \
                                                                  wide_max);                   \
    }                                                                                          \
    break;
    switch (nm) {
      SGM_TAIL(2)
      SGM_TAIL(3)
      SGM_TAIL(4)
      SGM_TAIL(5)
      SGM_TAIL(6)
    default: return hipErrorInvalidValue;
    }
#undef SGM_TAIL
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
  }
  return hipSuccess;
}

} // namespace kern
} // namespace sgm

namespace sgm {
namespace kern {
// the plain solver over an explicit cell list (cells whose material count has
// no register-resident bin kernel in the batched path: 1, 7, 8)
template <int METHOD>
__global__ void __launch_bounds__(256) pte_list_kernel(PteArgs a, MixParamsDev P,
                                                       const int *__restrict__ cells, long n,
                                                       const long *range) {
  extern __shared__ EosDev lds[];
  stage_registry(lds, a.d_eos, a.nreg);
  if (range) {
    cells += range[0];
    n = range[1];
  }
  const long nthr = (long)gridDim.x * blockDim.x;
  for (long j = (long)blockIdx.x * blockDim.x + threadIdx.x; j < n; j += nthr) {
    const long c = cells[j];
    if (c < 0 || c >= a.ncell) continue;
    int off = 0;
    const int nm = cell_slice(a, c, off);
    switch (nm) {
    case 1: solve_cell_call<1, METHOD>(lds, a, P, c, off); break;
    case 2: solve_cell_call<2, METHOD>(lds, a, P, c, off); break;
    case 3: solve_cell_call<3, METHOD>(lds, a, P, c, off); break;
    case 4: solve_cell_call<4, METHOD>(lds, a, P, c, off); break;
    case 5: solve_cell_call<5, METHOD>(lds, a, P, c, off); break;
    case 6: solve_cell_call<6, METHOD>(lds, a, P, c, off); break;
    case 7: solve_cell_call<7, METHOD>(lds, a, P, c, off); break;
    case 8: solve_cell_call<8, METHOD>(lds, a, P, c, off); break;
    default:
      // cells of 9..64 materials are the wave kernel's
      if (!(nm > MAX_MATS_PER_CELL && nm <= wave::WaveMaxNM(METHOD))) mark_bad(a, c);
      break;
    }
  }
}
// range (device {start, count}): n is then the host's upper bound on count
template <int METHOD>
hipError_t list_launch(const PteArgs &a, const MixParamsDev &P, const int *cells, long n,
                       hipStream_t s, const long *range = nullptr) {
  if (n <= 0) return hipSuccess;
  const size_t lds = (size_t)a.nreg * sizeof(EosDev);
  const long blocks_needed = (n + 255) / 256;
  const long cap = (long)device_cu_count() * 32;
  const int g = (int)(blocks_needed < cap ? blocks_needed : cap);
  pte_list_kernel<METHOD><<<g, 256, lds, s>>>(a, P, cells, n, range);
  return hipGetLastError();
}

// One cell of nm (9..64) materials on the whole wave: lane m loads material
// m, the wave runs the driver loop (pte_step) on the one-wave formulation,
// lane m stores material m and lane 0 the cell's status.
template <int METHOD>
__device__ __forceinline__ void wave_solve_cell(const PteArgs &a, const MixParamsDev &P, long c,
                                                int off, int nm, Real *work, int lane) {
  using Sys = wave::SysOf<METHOD>;
  Sys s;
  s.Setup(nm, Sys::Equations(nm), lane, work);
  const long k = off + (lane < nm ? lane : 0);
  s.eidx = a.mat_ids[k];
  s.rho = a.rho[k];
  s.vfrac = a.vfrac[k];
  s.press = a.press[k];
  const Real Tin = a.temp[k];
  if (a.lambda) {
    s.lam.present = true;
    s.lam.lRho = a.lambda[3L * k + 0];
    s.lam.lT = a.lambda[3L * k + 1];
  } else {
    s.lam.present = false;
  }
  s.SetCellInputs(a.vfrac_tot[c], a.sie_tot[c], a.Tnorm ? a.Tnorm[c] : 0.0);
  s.err = 0;
  s.P = P;
  CellStatus st;
  DriverState d;
  status_start(st);
  const Real err0 = s.Init(Tin);
  driver_start(d, err0, nm * P.pte_max_iter_per_mat);
  while (!wave::uni(pte_step(s, d, st))) {
  }
  s.Fixup();
  st.converged = d.converged ? 1 : 0;
  st.residual = d.err;
  st.err = s.ErrAll();
  if (lane < nm) {
    a.rho[k] = s.rho;
    a.vfrac[k] = s.vfrac;
    a.sie[k] = s.sie;
    if (Sys::kStoresTemp) a.temp[k] = s.Tout();
    a.press[k] = s.Pout();
    if (a.lambda) {
      a.lambda[3L * k + 0] = s.lam.lRho;
      a.lambda[3L * k + 1] = s.lam.lT;
    }
  }
  if (lane == 0) {
    if (a.conv) a.conv[c] = st.converged;
    if (a.niter) a.niter[c] = st.niter;
    if (a.resid) a.resid[c] = st.residual;
    if (a.err) a.err[c] = st.err;
  }
}

// Cells of 9..64 materials, one wave (the whole 64-lane workgroup) per cell:
// each wave reads 64 list entries at a time and solves the wide cells among
// them one after another (RhoU's up to 32 materials; it flags the wider
// ones as malformed); the others it leaves to the plain and list kernels. A wave that finds none never stages the registry. LDS: the
// registry, then the wave's Newton system. cells == nullptr: cells 0..n-1.
template <int METHOD>
__global__ void __launch_bounds__(64) pte_wave_kernel(PteArgs a, MixParamsDev P,
                                                      const int *__restrict__ cells, long n,
                                                      const long *range) {
  extern __shared__ EosDev lds[];
  if (range) {
    cells += range[0];
    n = range[1];
  }
  const int lane = threadIdx.x;
  Real *work = reinterpret_cast<Real *>(lds + a.nreg);
  bool staged = false;
  for (long j0 = (long)blockIdx.x * 64; j0 < n; j0 += (long)gridDim.x * 64) {
    const long j = j0 + lane;
    int c = -1, nm = 0, off = 0;
    if (j < n) {
      c = cells ? cells[j] : (int)j;
      if (c >= 0 && c < a.ncell) nm = cell_slice(a, c, off);
    }
    unsigned long long wide = __ballot(nm > MAX_MATS_PER_CELL);
    if (!wide) continue;
    if (!staged) {
      stage_registry(lds, a.d_eos, a.nreg);
      staged = true;
    }
    while (wide) {
      const int src = __builtin_ctzll(wide);
      wide &= wide - 1;
      const int k = wave::at(nm, src);
      if (k <= wave::WaveMaxNM(METHOD))
        wave_solve_cell<METHOD>(a, P, wave::at(c, src), wave::at(off, src), k, work, lane);
      else if (lane == 0)
        mark_bad(a, wave::at(c, src));  // past the formulation's one-wave size
    }
  }
}

template <int METHOD>
hipError_t wave_launch(const PteArgs &a, const MixParamsDev &P, const int *cells, long n,
                       hipStream_t s, const long *range) {
  if (n <= 0) return hipSuccess;
  const size_t lds = (size_t)a.nreg * sizeof(EosDev) + sizeof(Real) * wave::WORK_WORDS;
  const long blocks_needed = (n + 63) / 64;
  const long cap = (long)device_cu_count() * 8;
  const int g = (int)(blocks_needed < cap ? blocks_needed : cap);
  pte_wave_kernel<METHOD><<<g, 64, lds, s>>>(a, P, cells, n, range);
  return hipGetLastError();
}
} // namespace kern
} // namespace sgm
