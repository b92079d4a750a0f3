// Batched PTE plan: bins the cells of one CSR topology by material count and
// material signature (radix sort of a 64-bit key), picks a launch shape per
// bin (optionally by timing candidates on the device: the autotuner), and
// captures one hipGraph holding a counter reset plus one persistent kernel per
// bin. Replays of the graph solve every cell with a single graph launch.
#include "sgm_internal.hpp"

#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <cstring>
#include <map>
#include <vector>

using namespace sgm;

namespace {
struct Bin {
  int nm;          // 2..6 register-resident kernels, 0 = generic list kernel
  long start, n;   // range in the sorted cell list
  int block = 256;
  int cpw = 64;    // cells a wave claims per atomic on the bin counter (>= 64)
  int grid = 0;
  ParkArgs park;   // straggler records (register-resident bins only)
  int nruns = 0;   // signature runs in the bin
  bool wide = false;  // list bin of the cells of 9..64 materials (one wave per cell)
};

struct GraphKey {
  const void *ptrs[17];
  MixParamsDev P;
  bool operator==(const GraphKey &o) const { return std::memcmp(this, &o, sizeof(*this)) == 0; }
};
} // namespace

struct sgm_plan {
  sgm_ctx *ctx = nullptr;
  int method = 0;
  // the context's modified materials when the plan was made: their cells
  // sit in the list bins (executing after a change of modifiers is refused)
  unsigned long long mod_mask = 0;
  long ncell = 0, nnz = 0;
  int *d_offsets = nullptr;
  uint8_t *d_mat = nullptr;
  int *d_sorted = nullptr;  // cell ids grouped by bin
  int *d_counters = nullptr;
  // straggler parking: per sorted position its signature run; per run its
  // first record, capacity and claim counter; per tail wave its run and first
  // record (all bins in one array each, bins at their own offsets)
  int *d_run_of = nullptr, *d_run_base = nullptr, *d_run_cap = nullptr;
  int *d_run_count = nullptr, *d_wave_run = nullptr, *d_wave_first = nullptr;
  int nruns = 0;
  int *d_l2_counts = nullptr;    // per bin: cells saved for the tail's second level, its queue
  int *d_flag = nullptr;         // plan_match's result
  int park_after = 20;           // Newton iterations before a cell is parked; 0 = never
  std::vector<Bin> bins;
  hipStream_t cap_stream = nullptr;
  // fork streams, one per bin: each bin's tail kernel (its slowest cells)
  // runs beside the later bins' kernels instead of after them
  std::vector<hipStream_t> aux;
  std::vector<hipEvent_t> ev;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  bool have_key = false;
  GraphKey key{};
};

namespace {

// Default persistent grids: 8 one-wave blocks per CU, two waves per SIMD,
// the occupancy the bin kernels are built for. The bins' persistent kernels
// run one after another, each on the whole GPU, while the tail kernels of
// the bins before them finish their parked cells beside it (launch_bins).
// One-wave blocks: a SIMD whose wave has finished takes the next wave at
// once, instead of waiting for the other waves of a multi-wave block.
void default_grids(std::vector<Bin> &bins) {
  const int ncu = device_cu_count();
  for (Bin &b : bins) {
    if (b.nm < 2 || b.nm > 6) continue;
    b.block = 64;
    b.cpw = 64;
    const long need = (b.n + b.block - 1) / b.block;  // no more waves than cells fill
    b.grid = (int)std::max<long>(1, std::min<long>(need, 8L * ncu));
  }
}

// Straggler records for every register-resident bin, from the sort keys of
// the bin's positions: the cells of one signature (a "run"; contiguous in
// signature order, scattered after a cost reorder) get one record of
// ParkWords FP64 words each, and one tail wave slot per 64 / TailGroup
// records. Every cell can park: a straggler that found no record would run
// to the end of its Newton loop (up to 128 x nm iterations) in the bin
// kernel and hold its wave there. The records cost (11 nm + 25) x 8 bytes
// per cell, 0.5 GB for the uneven 2^20-cell workload.
hipError_t alloc_park(sgm_plan *p, const std::vector<unsigned long long> &keys) {
  std::vector<int> run_of(p->ncell, 0), run_base, run_cap, wave_run, wave_first;
  std::vector<size_t> run0(p->bins.size()), wave0(p->bins.size());
  std::vector<long> nrec(p->bins.size(), 0);
  for (size_t bi = 0; bi < p->bins.size(); ++bi) {
    Bin &b = p->bins[bi];
    run0[bi] = run_base.size();
    wave0[bi] = wave_run.size();
    if (b.nm < 2 || b.nm > 6) continue;
    std::map<unsigned long long, int> run_id;
    std::vector<long> len;
    unsigned long long last_k = ~0ull;
    int last_r = -1;
    for (long j = 0; j < b.n; ++j) {
      const unsigned long long k = keys[b.start + j] & ~(1ull << 55);  // signature only
      if (k == last_k) {  // runs are contiguous in signature order
        run_of[b.start + j] = last_r;
        ++len[last_r];
        continue;
      }
      auto it = run_id.find(k);
      int r;
      if (it == run_id.end()) {
        r = (int)len.size();
        run_id.emplace(k, r);
        len.push_back(0);
      } else {
        r = it->second;
      }
      run_of[b.start + j] = r;
      ++len[r];
      last_k = k;
      last_r = r;
    }
    for (int r = 0; r < (int)len.size(); ++r) {
      const long cap = len[r];
      run_base.push_back((int)nrec[bi]);
      run_cap.push_back((int)cap);
      const int per_wave = 64 / TailGroup(p->method, b.nm);
      for (long k = 0; k < cap; k += per_wave) {
        wave_run.push_back(r);
        wave_first.push_back((int)k);
      }
      nrec[bi] += cap;
    }
    b.nruns = (int)len.size();
  }
  p->nruns = (int)run_base.size();
  auto up = [](int **d, const std::vector<int> &h) {
    if (h.empty()) return hipSuccess;
    hipError_t e = hipMalloc(d, sizeof(int) * h.size());
    if (e == hipSuccess) e = hipMemcpy(*d, h.data(), sizeof(int) * h.size(), hipMemcpyHostToDevice);
    return e;
  };
  hipError_t e = up(&p->d_run_of, run_of);
  if (e == hipSuccess) e = up(&p->d_run_base, run_base);
  if (e == hipSuccess) e = up(&p->d_run_cap, run_cap);
  if (e == hipSuccess) e = up(&p->d_run_count, run_cap);  // contents reset per execution
  if (e == hipSuccess) e = up(&p->d_wave_run, wave_run);
  if (e == hipSuccess) e = up(&p->d_wave_first, wave_first);
  if (e == hipSuccess)
    e = hipMalloc(&p->d_l2_counts, sizeof(int) * 2 * std::max<size_t>(1, p->bins.size()));
  for (size_t bi = 0; bi < p->bins.size() && e == hipSuccess; ++bi) {
    Bin &b = p->bins[bi];
    if (b.nm < 2 || b.nm > 6 || nrec[bi] == 0) continue;
    ParkArgs &k = b.park;
    k.stride = ParkStride(p->method, b.nm);
    e = hipMalloc(&k.rec, sizeof(double) * k.stride * (size_t)nrec[bi]);
    if (e == hipSuccess) e = hipMalloc(&k.cell_of, sizeof(int) * nrec[bi]);
    k.l2_count = p->d_l2_counts + 2 * bi;
    k.l2_next = k.l2_count + 1;
    k.after2 = 4;
    k.run_of = p->d_run_of + b.start;
    k.run_base = p->d_run_base + run0[bi];
    k.run_cap = p->d_run_cap + run0[bi];
    k.run_count = p->d_run_count + run0[bi];
    k.wave_run = p->d_wave_run + wave0[bi];
    k.wave_first = p->d_wave_first + wave0[bi];
    k.nwaves = (int)((bi + 1 < p->bins.size() ? wave0[bi + 1] : wave_run.size()) - wave0[bi]);
  }
  return e;
}

// frees the straggler records and run tables and clears every pointer, so a
// later failed alloc_park or sgm_plan_destroy never frees them twice
void free_park(sgm_plan *p) {
  for (Bin &b : p->bins) {
    for (void *q : {(void *)b.park.rec, (void *)b.park.cell_of})
      if (q) (void)hipFree(q);
    b.park.rec = nullptr;
    b.park.cell_of = nullptr;
  }
  for (int **q : {&p->d_run_of, &p->d_run_base, &p->d_run_cap, &p->d_run_count, &p->d_wave_run,
                  &p->d_wave_first, &p->d_l2_counts}) {
    if (*q) (void)hipFree(*q);
    *q = nullptr;
  }
}

// Resets the bin counters and the straggler claim counters (fill kernels, so
// the resets can be captured).
hipError_t reset_counters(sgm_plan *p, hipStream_t s) {
  hipError_t e = launch_zero_ints(p->d_counters, (int)p->bins.size(), s);
  if (e == hipSuccess && p->park_after > 0 && p->nruns > 0)
    e = launch_zero_ints(p->d_run_count, p->nruns, s);
  // cleared with parking off too: sgm_plan_tail_counts then reports zeros
  if (e == hipSuccess && p->d_l2_counts)
    e = launch_zero_ints(p->d_l2_counts, 2 * (int)p->bins.size(), s);
  return e;
}

// A list bin: the plain solver over its cells, one lane per cell, or one
// wave per cell for the bin of 9..64 materials.
hipError_t launch_list_bin(const sgm_plan *p, const Bin &b, const PteArgs &a,
                           const MixParamsDev &P, const int *cells, hipStream_t s) {
  if (b.wide) return launch_pte_wave(p->method, a, P, cells, b.n, s);
  return launch_pte_list(p->method, a, P, cells, b.n, s, nullptr, false);
}

// One bin: the persistent kernel, then (with parking) the in-order listing of
// its parked cells and the tail kernel that finishes them.
hipError_t launch_one_bin(sgm_plan *p, size_t i, const PteArgs &a, const MixParamsDev &P,
                          hipStream_t s) {
  const Bin &b = p->bins[i];
  const int *cells = p->d_sorted + b.start;
  if (b.nm < 2 || b.nm > 6) return launch_list_bin(p, b, a, P, cells, s);
  ParkArgs pk = b.park;
  pk.after = pk.rec ? p->park_after : 0;
  hipError_t e = launch_pte_bin(p->method, b.nm, a, P, cells, b.n, p->d_counters + i, b.block,
                                b.cpw, b.grid, s, nullptr, &pk);
  if (e != hipSuccess || pk.after <= 0) return e;
  return launch_pte_tail(p->method, b.nm, a, P, pk, s);
}

// Order of the bins on the launch stream: the bins whose tail kernels run
// longest (5, 4 and 6 materials on the uneven workload: 14, 13 and 12 ms
// beside the other kernels) early, so their tails overlap the later bins'
// persistent kernels; then the list bins. The runtime spreads a captured
// graph's branches over the process's hardware queues (four by default),
// each bin's persistent kernel on the next queue with its tail behind it,
// so the fifth bin's persistent kernel waits behind the first bin's tail:
// the first bin is 3 materials, whose tail (10 ms) ends by the time the
// fifth (2 materials) is due.
std::vector<size_t> bin_order(const sgm_plan *p) {
  std::vector<size_t> order;
  for (int nm : {3, 5, 4, 6, 2})
    for (size_t i = 0; i < p->bins.size(); ++i)
      if (p->bins[i].nm == nm) order.push_back(i);
  for (size_t i = 0; i < p->bins.size(); ++i)
    if (p->bins[i].nm < 2 || p->bins[i].nm > 6) order.push_back(i);
  return order;
}

// Records the counter reset and every bin on stream s: the persistent bin
// kernels one after another on s; with fork streams (graph capture, or a
// direct launch that may overlap) each bin's tail kernel forks off to a
// stream of its own after its bin kernel and joins s at the end, so the
// tails run beside the later bins. Without forks everything stays on s.
hipError_t launch_bins(sgm_plan *p, const PteArgs &a, const MixParamsDev &P, hipStream_t s,
                       bool fork) {
  hipError_t e = reset_counters(p, s);
  if (e != hipSuccess) return e;
  const std::vector<size_t> order = bin_order(p);
  const size_t nb = p->bins.size();
  fork = fork && p->aux.size() >= nb && p->ev.size() >= 2 * nb;
  std::vector<size_t> joins;
  for (size_t i : order) {
    const Bin &b = p->bins[i];
    const int *cells = p->d_sorted + b.start;
    if (b.nm < 2 || b.nm > 6) {
      if ((e = launch_list_bin(p, b, a, P, cells, s)) != hipSuccess) return e;
      continue;
    }
    ParkArgs pk = b.park;
    pk.after = pk.rec ? p->park_after : 0;
    if ((e = launch_pte_bin(p->method, b.nm, a, P, cells, b.n, p->d_counters + i, b.block,
                            b.cpw, b.grid, s, nullptr, &pk)) != hipSuccess)
      return e;
    if (pk.after <= 0) continue;
    if (!fork) {
      if ((e = launch_pte_tail(p->method, b.nm, a, P, pk, s)) != hipSuccess) return e;
      continue;
    }
    if ((e = hipEventRecord(p->ev[2 * i], s)) != hipSuccess ||
        (e = hipStreamWaitEvent(p->aux[i], p->ev[2 * i], 0)) != hipSuccess ||
        (e = launch_pte_tail(p->method, b.nm, a, P, pk, p->aux[i])) != hipSuccess ||
        (e = hipEventRecord(p->ev[2 * i + 1], p->aux[i])) != hipSuccess)
      return e;
    joins.push_back(i);
  }
  for (size_t i : joins)
    if ((e = hipStreamWaitEvent(s, p->ev[2 * i + 1], 0)) != hipSuccess) return e;
  return hipSuccess;
}

PteArgs make_args(sgm_plan *p, double *rho, double *vfrac, double *sie, double *temp,
                  double *press, const double *vt, const double *st, const double *Tn,
                  double *lambda, sgm_pte_status status) {
  PteArgs a;
  a.d_eos = p->ctx->d_eos;
  a.nreg = p->ctx->nmat;
  a.ncell = p->ncell;
  a.nnz = p->nnz;
  a.offsets = p->d_offsets;
  a.mat_ids = p->d_mat;
  a.rho = rho;
  a.vfrac = vfrac;
  a.sie = sie;
  a.temp = temp;
  a.press = press;
  a.vfrac_tot = vt;
  a.sie_tot = st;
  a.Tnorm = Tn;
  a.lambda = lambda;
  a.conv = status.converged;
  a.niter = status.niter;
  a.err = status.err;
  a.resid = status.residual;
  return a;
}

MixParamsDev params_dev(const sgm_mix_params *params) {
  sgm_mix_params q;
  if (params)
    q = *params;
  else
    sgm_mix_params_default(&q);
  MixParamsDev d;
  std::memcpy(&d, &q, sizeof(d));
  return d;
}

// Sorts the plan's cells by key into d_sorted (radix sort of 64-bit keys);
// keys_host receives the sorted keys when given. Synchronizes s.
hipError_t sort_cells(sgm_plan *p, const int *cost, int threshold,
                      std::vector<unsigned long long> *keys_host, hipStream_t s) {
  const long ncell = p->ncell;
  unsigned long long *d_keys = nullptr, *d_keys_sorted = nullptr;
  int *d_ids = nullptr;
  void *d_tmp = nullptr;
  size_t tmp_bytes = 0;
  hipError_t e = hipMalloc(&d_keys, sizeof(unsigned long long) * ncell);
  if (e == hipSuccess) e = hipMalloc(&d_keys_sorted, sizeof(unsigned long long) * ncell);
  if (e == hipSuccess) e = hipMalloc(&d_ids, sizeof(int) * ncell);
  if (e == hipSuccess)
    e = launch_plan_keys(p->d_offsets, p->d_mat, ncell, p->nnz, d_keys, d_ids, s, cost,
                         threshold, p->mod_mask);
  if (e == hipSuccess)
    e = hipcub::DeviceRadixSort::SortPairs(nullptr, tmp_bytes, d_keys, d_keys_sorted, d_ids,
                                           p->d_sorted, (int)ncell, 0, 64, s);
  if (e == hipSuccess) e = hipMalloc(&d_tmp, tmp_bytes + 16);
  if (e == hipSuccess)
    e = hipcub::DeviceRadixSort::SortPairs(d_tmp, tmp_bytes, d_keys, d_keys_sorted, d_ids,
                                           p->d_sorted, (int)ncell, 0, 64, s);
  if (e == hipSuccess && keys_host) {
    keys_host->resize(ncell);
    e = hipMemcpyAsync(keys_host->data(), d_keys_sorted, sizeof(unsigned long long) * ncell,
                       hipMemcpyDeviceToHost, s);
  }
  const hipError_t es = hipStreamSynchronize(s);
  if (e == hipSuccess) e = es;
  for (void *q : {(void *)d_keys, (void *)d_keys_sorted, (void *)d_ids, d_tmp})
    if (q) (void)hipFree(q);
  return e;
}

} // namespace

namespace sgm {

int plan_create(sgm_ctx *ctx, long ncell, long nnz, const int *offsets, const uint8_t *mat_ids,
                int method, int min_nm, sgm_plan **out, hipStream_t s) {
  if (!ctx || !out || ncell < 1 || nnz < 1 || nnz > 2147483647L || ncell > 2147483647L ||
      !offsets || !mat_ids || method < 0 || method >= PTE_NMETHODS)
    return SGM_E_ARG;
  *out = nullptr;
  if (ctx->host_only) return SGM_E_NOGPU;
  auto *p = new sgm_plan();
  p->ctx = ctx;
  p->method = method;
  p->mod_mask = ctx_mod_mask(ctx);
  p->ncell = ncell;
  p->nnz = nnz;
  bool ok = hipMalloc(&p->d_offsets, sizeof(int) * (ncell + 1)) == hipSuccess &&
            hipMalloc(&p->d_mat, nnz) == hipSuccess &&
            hipMalloc(&p->d_sorted, sizeof(int) * ncell) == hipSuccess;
  std::vector<unsigned long long> keys;
  if (ok)
    ok = hipMemcpyAsync(p->d_offsets, offsets, sizeof(int) * (ncell + 1),
                        hipMemcpyDeviceToDevice, s) == hipSuccess &&
         hipMemcpyAsync(p->d_mat, mat_ids, nnz, hipMemcpyDeviceToDevice, s) == hipSuccess &&
         sort_cells(p, nullptr, 0, &keys, s) == hipSuccess;
  if (!ok) {
    sgm_plan_destroy(p);
    return SGM_E_HIP;
  }
  // bins: one per material count 2..BinMaxNM(method); everything else in one
  // generic list bin per count (cells with fewer than min_nm materials in none)
  long i = 0;
  while (i < ncell) {
    const int nm = (int)(keys[i] >> 56);
    long j = i;
    while (j < ncell && (int)(keys[j] >> 56) == nm) ++j;
    if (nm < min_nm) {
      // not solved
    } else if (nm >= 2 && nm <= BinMaxNM(method)) {
      Bin b{nm, i, j - i};
      p->bins.push_back(b);
    } else {
      // sorted keys put counts 1 before 2..6 and 7, 8, modified, 9..64
      // (count field 0x80), invalid after; keep each such run as its own
      // list bin
      Bin b{0, i, j - i};
      b.wide = nm == 0x80;
      p->bins.push_back(b);
    }
    i = j;
  }
  default_grids(p->bins);
  if (alloc_park(p, keys) != hipSuccess ||
      hipMalloc(&p->d_counters, sizeof(int) * std::max<size_t>(1, p->bins.size())) !=
          hipSuccess ||
      hipStreamCreateWithFlags(&p->cap_stream, hipStreamNonBlocking) != hipSuccess) {
    sgm_plan_destroy(p);
    return SGM_E_HIP;
  }
  for (size_t i = 0; i < p->bins.size(); ++i) {
    hipStream_t q = nullptr;
    if (hipStreamCreateWithFlags(&q, hipStreamNonBlocking) != hipSuccess) {
      sgm_plan_destroy(p);
      return SGM_E_HIP;
    }
    p->aux.push_back(q);
  }
  for (size_t i = 0; i < 2 * p->bins.size(); ++i) {  // per bin: tail fork and join
    hipEvent_t ev = nullptr;
    if (hipEventCreateWithFlags(&ev, hipEventDisableTiming) != hipSuccess) {
      sgm_plan_destroy(p);
      return SGM_E_HIP;
    }
    p->ev.push_back(ev);
  }
  *out = p;
  return SGM_OK;
}

int plan_match(sgm_plan *p, long ncell, long nnz, const int *offsets, const uint8_t *mat_ids,
               int method, hipStream_t s, bool &match) {
  match = false;
  if (!p || p->method != method || p->ncell != ncell || p->nnz != nnz ||
      p->mod_mask != ctx_mod_mask(p->ctx))
    return SGM_OK;
  if (!p->d_flag && hipMalloc(&p->d_flag, sizeof(int)) != hipSuccess) return SGM_E_HIP;
  int differ = 1;
  if (launch_compare_csr(p->d_offsets, offsets, ncell + 1, p->d_mat, mat_ids, nnz, p->d_flag,
                         s) != hipSuccess ||
      hipMemcpyAsync(&differ, p->d_flag, sizeof(int), hipMemcpyDeviceToHost, s) != hipSuccess ||
      hipStreamSynchronize(s) != hipSuccess)
    return SGM_E_HIP;
  match = differ == 0;
  return SGM_OK;
}

int plan_launch(sgm_plan *p, double *rho, double *vfrac, double *sie, double *temp,
                double *press, const double *vfrac_tot, const double *sie_tot, double *lambda,
                const sgm_mix_params *params, sgm_pte_status status, hipStream_t s) {
  const MixParamsDev P = params_dev(params);
  const PteArgs a = make_args(p, rho, vfrac, sie, temp, press, vfrac_tot, sie_tot, nullptr,
                              lambda, status);
  return launch_bins(p, a, P, s, true) == hipSuccess ? SGM_OK : SGM_E_HIP;
}

} // namespace sgm

extern "C" {

int sgm_plan_create(sgm_ctx *ctx, long ncell, long nnz, const int *offsets,
                    const uint8_t *mat_ids, int method, sgm_plan **out, sgm_stream_t stream) {
  return plan_create(ctx, ncell, nnz, offsets, mat_ids, method, 0, out, (hipStream_t)stream);
}

int sgm_plan_describe(const sgm_plan *p, long *out, int max_bins) {
  if (!p || !out) return SGM_E_ARG;
  int k = 0;
  for (const Bin &b : p->bins) {
    if (k >= max_bins) break;
    out[6 * k + 0] = b.nm;
    out[6 * k + 1] = b.n;
    out[6 * k + 2] = b.block;
    out[6 * k + 3] = b.cpw;
    out[6 * k + 4] = b.grid;
    out[6 * k + 5] = b.start;
    ++k;
  }
  return (int)p->bins.size();
}

int sgm_plan_tail_counts(const sgm_plan *p, int *out, int max_bins, sgm_stream_t stream) {
  if (!p || !out || max_bins < 0) return SGM_E_ARG;
  const int nb = (int)p->bins.size();
  std::vector<int> c(2 * (size_t)std::max(1, nb), 0);
  if (p->d_l2_counts) {
    hipStream_t s = (hipStream_t)stream;
    if (hipMemcpyAsync(c.data(), p->d_l2_counts, sizeof(int) * c.size(), hipMemcpyDeviceToHost,
                       s) != hipSuccess ||
        hipStreamSynchronize(s) != hipSuccess)
      return SGM_E_HIP;
  }
  for (int i = 0; i < nb && i < max_bins; ++i) out[i] = p->bins[i].park.rec ? c[2 * i] : 0;
  return nb;
}

// Puts, within every material count, the cells whose cost (typically the
// previous execution's Newton iteration counts) reaches 4x the mean cost
// first: the slowest cells then start with the bulk instead of trailing it.
// Results do not depend on the order.
int sgm_plan_reorder(sgm_plan *p, const int *cost, sgm_stream_t stream) {
  if (!p || !cost) return SGM_E_ARG;
  hipStream_t s = (hipStream_t)stream;
  std::vector<int> h(p->ncell);
  if (hipMemcpyAsync(h.data(), cost, sizeof(int) * p->ncell, hipMemcpyDeviceToHost, s) !=
          hipSuccess ||
      hipStreamSynchronize(s) != hipSuccess)
    return SGM_E_HIP;
  double mean = 0;
  for (int c : h) mean += c;
  mean /= (double)p->ncell;
  const int threshold = (int)std::max(1.0, 4.0 * mean);
  std::vector<unsigned long long> keys;
  if (sort_cells(p, cost, threshold, &keys, s) != hipSuccess) return SGM_E_HIP;
  // Within each bin the costly cells now lead; spread them one per refill
  // chunk (64 cells) so that each wave starts on one of them instead of a
  // few waves holding them all. Bin bounds are unchanged: the material count
  // still leads the key.
  std::vector<int> ids(p->ncell), out(p->ncell);
  if (hipMemcpy(ids.data(), p->d_sorted, sizeof(int) * p->ncell, hipMemcpyDeviceToHost) !=
      hipSuccess)
    return SGM_E_HIP;
  out = ids;
  for (const Bin &b : p->bins) {
    if (b.nm < 2 || b.n < 2) continue;
    long nh = 0;
    while (nh < b.n && !(keys[b.start + nh] & (1ull << 55))) ++nh;
    if (nh == 0 || nh == b.n) continue;
    const long stride = std::max<long>(1, std::min<long>(64, b.n / nh));
    long h = 0, l = nh, o = b.start;
    while (h < nh || l < b.n) {
      if (h < nh) out[o++] = ids[b.start + h++];
      for (long k = 1; k < stride && l < b.n; ++k) out[o++] = ids[b.start + l++];
      if (h >= nh)
        while (l < b.n) out[o++] = ids[b.start + l++];
    }
  }
  if (hipMemcpy(p->d_sorted, out.data(), sizeof(int) * p->ncell, hipMemcpyHostToDevice) !=
      hipSuccess)
    return SGM_E_HIP;
  // the straggler tables follow the positions: rebuild them for the new order
  std::vector<unsigned long long> key_of(p->ncell), fk(p->ncell);
  for (long j = 0; j < p->ncell; ++j) key_of[ids[j]] = keys[j];
  for (long j = 0; j < p->ncell; ++j) fk[j] = key_of[out[j]];
  free_park(p);
  for (Bin &b : p->bins) b.park = ParkArgs{};
  p->d_run_of = p->d_run_base = p->d_run_cap = p->d_run_count = nullptr;
  p->d_wave_run = p->d_wave_first = nullptr;
  if (alloc_park(p, fk) != hipSuccess) return SGM_E_HIP;
  p->have_key = false;
  return SGM_OK;
}

// Times each candidate launch shape for every register-resident bin on
// scratch copies of the inputs and keeps the fastest. Results of the solve do
// not depend on the launch shape (cells are independent), so any choice is
// exact; only throughput changes.
int sgm_plan_autotune(sgm_plan *p, const double *rho, const double *vfrac, const double *temp,
                      const double *press, const double *vfrac_tot, const double *sie_tot,
                      const sgm_mix_params *params, int reps, sgm_stream_t stream) {
  if (!p || !rho || !vfrac || !temp || !press || !vfrac_tot || !sie_tot) return SGM_E_ARG;
  if (reps < 1) reps = 1;
  hipStream_t s = (hipStream_t)stream;
  const size_t nb = sizeof(double) * p->nnz;
  double *w[5] = {nullptr, nullptr, nullptr, nullptr, nullptr};
  for (auto &q : w) {
    if (hipMalloc(&q, nb) != hipSuccess) {
      for (double *x : w)
        if (x) (void)hipFree(x);
      return SGM_E_HIP;
    }
  }
  const MixParamsDev P = params_dev(params);
  hipEvent_t e0, e1;
  hipEventCreate(&e0);
  hipEventCreate(&e1);
  const int ncu = device_cu_count();
  int rc = SGM_OK;
  for (size_t bi = 0; bi < p->bins.size() && rc == SGM_OK; ++bi) {
    Bin &b = p->bins[bi];
    if (b.nm < 2 || b.nm > 6 || b.n < 64) continue;
    // the bin's share of the GPU in waves (its default grid), scaled by
    // 1/2, 1, 2; block shapes and refill chunks around the default
    const int blocks[3] = {64, 128, 256};
    const int cpws[3] = {64, 128, 256};
    const int share_waves = b.grid * (b.block / 64);
    const int scales[3] = {1, 2, 4};  // quarter-steps: x0.5, x1, x2
    float best = 1e30f;
    Bin bestb = b;
    for (int bl : blocks)
      for (int cpw : cpws)
        for (int sc : scales) {
          const int wpb = bl / 64;
          const int grid = std::max(1, std::min(8 * ncu, share_waves * sc / 2 / wpb));
          float tot = 0;
          for (int r = 0; r < reps && rc == SGM_OK; ++r) {
            if (hipMemcpyAsync(w[0], rho, nb, hipMemcpyDeviceToDevice, s) != hipSuccess ||
                hipMemcpyAsync(w[1], vfrac, nb, hipMemcpyDeviceToDevice, s) != hipSuccess ||
                hipMemcpyAsync(w[2], temp, nb, hipMemcpyDeviceToDevice, s) != hipSuccess ||
                hipMemcpyAsync(w[3], press, nb, hipMemcpyDeviceToDevice, s) != hipSuccess ||
                launch_zero_ints(p->d_counters + bi, 1, s) != hipSuccess) {
              rc = SGM_E_HIP;
              break;
            }
            PteArgs a = make_args(p, w[0], w[1], w[4], w[2], w[3], vfrac_tot, sie_tot, nullptr,
                                  nullptr, sgm_pte_status{nullptr, nullptr, nullptr, nullptr});
            hipEventRecord(e0, s);
            const hipError_t e = launch_pte_bin(p->method, b.nm, a, P, p->d_sorted + b.start,
                                                b.n, p->d_counters + bi, bl, cpw, grid, s);
            hipEventRecord(e1, s);
            if (e != hipSuccess || hipEventSynchronize(e1) != hipSuccess) {
              rc = SGM_E_HIP;
              break;
            }
            float ms = 0;
            hipEventElapsedTime(&ms, e0, e1);
            tot += ms;
          }
          if (rc == SGM_OK && tot < best) {
            best = tot;
            bestb.block = bl;
            bestb.cpw = cpw;
            bestb.grid = grid;
          }
        }
    b = bestb;
  }
  hipEventDestroy(e0);
  hipEventDestroy(e1);
  for (auto &q : w) hipFree(q);
  // launch shapes changed: the captured graph is stale
  if (p->exec) {
    hipGraphExecDestroy(p->exec);
    p->exec = nullptr;
  }
  if (p->graph) {
    hipGraphDestroy(p->graph);
    p->graph = nullptr;
  }
  p->have_key = false;
  return rc;
}

} // extern "C"

namespace {
// The plan's graph for these arguments (captured when they change) launched
// on s. With an MLP, the graph first runs the network's guess kernel into
// guess_vfrac / guess_T, so inference and the guided solve are one launch.
int execute_common(sgm_plan *p, double *rho, double *vfrac, double *sie, double *temp,
                   double *press, const double *vfrac_tot, const double *sie_tot,
                   const double *Tnorm, double *lambda, const sgm_mix_params *params,
                   sgm_pte_status status, const double *guess_vfrac, const double *guess_T,
                   int *used_guess, hipStream_t s, sgm_mlp *mlp = nullptr) {
  const MixParamsDev P = params_dev(params);
  GraphKey k;
  std::memset(&k, 0, sizeof(k));
  const void *ptrs[17] = {rho,         vfrac,        sie,          temp,
                          status.residual, press,    vfrac_tot,    sie_tot,
                          Tnorm,       lambda,       status.converged, status.niter,
                          status.err,  guess_vfrac,  guess_T,      used_guess,
                          mlp};
  for (int i = 0; i < 17; ++i) k.ptrs[i] = ptrs[i];
  k.P = P;
  if (!p->have_key || !(p->key == k) || !p->exec) {
    if (p->exec) hipGraphExecDestroy(p->exec);
    if (p->graph) hipGraphDestroy(p->graph);
    p->exec = nullptr;
    p->graph = nullptr;
    PteArgs a = make_args(p, rho, vfrac, sie, temp, press, vfrac_tot, sie_tot, Tnorm, lambda,
                          status);
    a.guess_vfrac = guess_vfrac;
    a.guess_T = guess_T;
    a.used_guess = used_guess;
    if (hipStreamBeginCapture(p->cap_stream, hipStreamCaptureModeThreadLocal) != hipSuccess)
      return SGM_E_HIP;
    hipError_t le = hipSuccess;
    if (mlp &&
        sgm_mlp_guess(mlp, p->ctx, p->ncell, p->nnz, p->d_offsets, p->d_mat, rho, vfrac, temp,
                      vfrac_tot, sie_tot, const_cast<double *>(guess_vfrac),
                      const_cast<double *>(guess_T), (sgm_stream_t)p->cap_stream) != SGM_OK)
      le = hipErrorLaunchFailure;
    if (le == hipSuccess) le = launch_bins(p, a, P, p->cap_stream, true);
    hipGraph_t g = nullptr;
    const hipError_t ce = hipStreamEndCapture(p->cap_stream, &g);
    if (le != hipSuccess || ce != hipSuccess || !g) {
      if (g) hipGraphDestroy(g);
      return SGM_E_HIP;
    }
    if (hipGraphInstantiate(&p->exec, g, nullptr, nullptr, 0) != hipSuccess) {
      hipGraphDestroy(g);
      return SGM_E_HIP;
    }
    p->graph = g;
    p->key = k;
    p->have_key = true;
  }
  return hipGraphLaunch(p->exec, s) == hipSuccess ? SGM_OK : SGM_E_HIP;
}
} // namespace

extern "C" {
int sgm_plan_execute(sgm_plan *p, double *rho, double *vfrac, double *sie, double *temp,
                     double *press, const double *vfrac_tot, const double *sie_tot,
                     const double *Tnorm, double *lambda, const sgm_mix_params *params,
                     sgm_pte_status status, sgm_stream_t stream) {
  if (!p || !rho || !vfrac || !sie || !temp || !press || !vfrac_tot || !sie_tot) return SGM_E_ARG;
  if (p->mod_mask != ctx_mod_mask(p->ctx)) return SGM_E_ARG;  // modifiers changed
  return execute_common(p, rho, vfrac, sie, temp, press, vfrac_tot, sie_tot, Tnorm, lambda,
                        params, status, nullptr, nullptr, nullptr, (hipStream_t)stream);
}

int sgm_plan_execute_guided(sgm_plan *p, double *rho, double *vfrac, double *sie, double *temp,
                            double *press, const double *vfrac_tot, const double *sie_tot,
                            const double *guess_vfrac, const double *guess_T,
                            const sgm_mix_params *params, sgm_pte_status status,
                            int *used_guess, sgm_stream_t stream) {
  if (!p || !rho || !vfrac || !sie || !temp || !press || !vfrac_tot || !sie_tot ||
      !guess_vfrac || !guess_T)
    return SGM_E_ARG;
  if (p->mod_mask != ctx_mod_mask(p->ctx)) return SGM_E_ARG;  // modifiers changed
  // the guess replaces the volume fractions and Tnorm of the energy-closed
  // formulations; the fixed-T / fixed-P ones have no temperature guess
  if (p->method != PTE_RHOT && p->method != PTE_PT) return SGM_E_ARG;
  return execute_common(p, rho, vfrac, sie, temp, press, vfrac_tot, sie_tot, nullptr, nullptr,
                        params, status, guess_vfrac, guess_T, used_guess, (hipStream_t)stream);
}

int sgm_plan_execute_learned(sgm_plan *p, sgm_mlp *mlp, double *rho, double *vfrac, double *sie,
                             double *temp, double *press, const double *vfrac_tot,
                             const double *sie_tot, double *guess_vfrac, double *guess_T,
                             const sgm_mix_params *params, sgm_pte_status status,
                             int *used_guess, sgm_stream_t stream) {
  if (!p || !mlp || !rho || !vfrac || !sie || !temp || !press || !vfrac_tot || !sie_tot ||
      !guess_vfrac || !guess_T)
    return SGM_E_ARG;
  if (p->mod_mask != ctx_mod_mask(p->ctx)) return SGM_E_ARG;  // modifiers changed
  if (p->method != PTE_RHOT && p->method != PTE_PT) return SGM_E_ARG;
  return execute_common(p, rho, vfrac, sie, temp, press, vfrac_tot, sie_tot, nullptr, nullptr,
                        params, status, guess_vfrac, guess_T, used_guess, (hipStream_t)stream,
                        mlp);
}

int sgm_plan_set_shape(sgm_plan *p, int bin, int block, int cells_per_wave, int grid) {
  if (!p || bin < 0 || bin >= (int)p->bins.size() || grid < 1 || cells_per_wave < 64 ||
      (block != 64 && block != 128 && block != 256))
    return SGM_E_ARG;
  Bin &b = p->bins[bin];
  if (b.nm < 2 || b.nm > 6) return SGM_E_ARG;
  b.block = block;
  b.cpw = cells_per_wave;
  b.grid = grid;
  p->have_key = false;
  return SGM_OK;
}

int sgm_plan_set_straggler_threshold(sgm_plan *p, int iterations) {
  if (!p || iterations < 0) return SGM_E_ARG;
  if (iterations != p->park_after) {
    p->park_after = iterations;
    p->have_key = false;  // the captured graph holds the old setting
  }
  return SGM_OK;
}

// Launches one bin's kernels directly (no graph) — for per-kernel timing.
int sgm_plan_execute_bin(sgm_plan *p, int bin, double *rho, double *vfrac, double *sie,
                         double *temp, double *press, const double *vfrac_tot,
                         const double *sie_tot, const double *Tnorm, double *lambda,
                         const sgm_mix_params *params, sgm_pte_status status, float *kernel_ms,
                         sgm_stream_t stream) {
  if (!p || bin < 0 || bin >= (int)p->bins.size() || !rho || !vfrac || !sie || !temp ||
      !press || !vfrac_tot || !sie_tot)
    return SGM_E_ARG;
  if (p->mod_mask != ctx_mod_mask(p->ctx)) return SGM_E_ARG;  // modifiers changed
  hipStream_t s = (hipStream_t)stream;
  const MixParamsDev P = params_dev(params);
  PteArgs a = make_args(p, rho, vfrac, sie, temp, press, vfrac_tot, sie_tot, Tnorm, lambda,
                        status);
  hipError_t e = reset_counters(p, s);
  if (e != hipSuccess) return SGM_E_HIP;
  if (!kernel_ms) {
    e = launch_one_bin(p, (size_t)bin, a, P, s);
    return e == hipSuccess ? SGM_OK : SGM_E_HIP;
  }
  // the same launches as launch_one_bin, with an event between them
  const Bin &b = p->bins[(size_t)bin];
  const int *cells = p->d_sorted + b.start;
  hipEvent_t ev[3] = {nullptr, nullptr, nullptr};
  for (hipEvent_t &x : ev)
    if (e == hipSuccess) e = hipEventCreate(&x);
  if (e == hipSuccess) e = hipEventRecord(ev[0], s);
  ParkArgs pk = b.park;
  pk.after = (b.nm >= 2 && b.nm <= 6 && pk.rec) ? p->park_after : 0;
  if (e == hipSuccess)
    e = (b.nm < 2 || b.nm > 6)
            ? launch_list_bin(p, b, a, P, cells, s)
            : launch_pte_bin(p->method, b.nm, a, P, cells, b.n, p->d_counters + bin, b.block,
                             b.cpw, b.grid, s, nullptr, &pk);
  if (e == hipSuccess) e = hipEventRecord(ev[1], s);
  if (e == hipSuccess && pk.after > 0) e = launch_pte_tail(p->method, b.nm, a, P, pk, s);
  if (e == hipSuccess) e = hipEventRecord(ev[2], s);
  if (e == hipSuccess) e = hipEventSynchronize(ev[2]);
  if (e == hipSuccess) e = hipEventElapsedTime(&kernel_ms[0], ev[0], ev[1]);
  if (e == hipSuccess) e = hipEventElapsedTime(&kernel_ms[1], ev[1], ev[2]);
  if (e == hipSuccess && pk.after <= 0) kernel_ms[1] = 0.0f;
  for (hipEvent_t x : ev)
    if (x) (void)hipEventDestroy(x);
  return e == hipSuccess ? SGM_OK : SGM_E_HIP;
}

int sgm_plan_destroy(sgm_plan *p) {
  if (!p) return SGM_E_ARG;
  if (p->cap_stream) hipStreamSynchronize(p->cap_stream);
  if (p->exec) hipGraphExecDestroy(p->exec);
  if (p->graph) hipGraphDestroy(p->graph);
  if (p->d_offsets) hipFree(p->d_offsets);
  if (p->d_flag) hipFree(p->d_flag);
  if (p->d_mat) hipFree(p->d_mat);
  if (p->d_sorted) hipFree(p->d_sorted);
  if (p->d_counters) hipFree(p->d_counters);
  free_park(p);
  if (p->cap_stream) hipStreamDestroy(p->cap_stream);
  for (hipStream_t q : p->aux) hipStreamDestroy(q);
  for (hipEvent_t ev : p->ev) hipEventDestroy(ev);
  delete p;
  return SGM_OK;
}

} // extern "C"
