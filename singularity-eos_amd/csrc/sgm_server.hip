// Serving loop: device-resident EOS objects and tables, HBM-budget admission,
// continuous batching of closure and vector requests over several streams.
//
// Per launch group (one lane = one stream), everything below is asynchronous
// on the lane's stream and the scheduler thread only polls the lane's event:
//   H2D   request table, each closure request's raw offsets and per-entry /
//         per-cell inputs into one merged CSR, each vector request's inputs
//   ker   offsets rebase (merged CSR), sort keys, radix sort, bin bounds,
//         counter reset, one persistent bin kernel per material count 2..6
//         with straggler parking and the bin's tail kernels behind it, and
//         the list kernel for the rest (bin bounds read on the device);
//         one vector kernel per vector request
//   D2H   results straight into the clients' arrays
// Admission is first-come first-served: the queue head waits until its
// working set fits in a free stretch of the arena; a request larger than the
// whole arena is refused at submit.
#include "sgm_internal.hpp"

#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <thread>
#include <vector>

using namespace sgm;

namespace {

constexpr size_t ALIGN = 256;
// Newton iterations before a bin kernel parks a cell (as plans do)
constexpr int SERVER_PARK_AFTER = 20;
inline size_t align_up(size_t x) { return (x + ALIGN - 1) & ~(ALIGN - 1); }

// device-side description of one closure request inside a group
struct ReqDev {
  long cell_base;  // first merged cell
  long n;          // cells
  long nnz_base;   // first merged entry
  long raw_pos;    // position of its n+1 raw offsets in the raw staging array
};

__global__ void rebase_kernel(const ReqDev *__restrict__ reqs, int nreq,
                              const int *__restrict__ raw, int *__restrict__ off, long ncell,
                              long nnz) {
  const long g = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (g == 0) off[ncell] = (int)nnz;
  if (g >= ncell) return;
  int lo = 0, hi = nreq - 1;  // last request with cell_base <= g
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (reqs[mid].cell_base <= g)
      lo = mid;
    else
      hi = mid - 1;
  }
  const ReqDev r = reqs[lo];
  const long i = g - r.cell_base;
  const long r0 = raw[r.raw_pos], rn = raw[r.raw_pos + r.n], ri = raw[r.raw_pos + i];
  // offsets outside the request's own entry range mark the cell invalid
  // (the solver then flags it and touches no other request's entries)
  off[g] = (ri >= r0 && ri <= rn) ? (int)(r.nnz_base + (ri - r0)) : -1;
}

// bin bounds of sorted keys (nmat << 56 | signature): range[2k..] = {start,
// count} for k = 0: nmat < 2, 1..5: nmat = k+1, 6: nmat > 6 (incl. invalid)
__global__ void ranges_kernel(const unsigned long long *__restrict__ keys, long n,
                              long *__restrict__ range) {
  if (blockIdx.x != 0 || threadIdx.x != 0) return;
  long pos[7];
  for (int nm = 2; nm <= 8; ++nm) {
    const unsigned long long k = (unsigned long long)nm << 56;
    long lo = 0, hi = n;  // first index with keys[i] >= k
    while (lo < hi) {
      const long mid = (lo + hi) >> 1;
      if (keys[mid] < k)
        lo = mid + 1;
      else
        hi = mid;
    }
    pos[nm - 2] = lo;
  }
  range[0] = 0;
  range[1] = pos[0];
  for (int b = 1; b <= 5; ++b) {
    range[2 * b] = pos[b - 1];
    range[2 * b + 1] = pos[b] - pos[b - 1];
  }
  range[12] = pos[5];
  range[13] = n - pos[5];
}

// The straggler tables of a launch group. The group's bins are not split by
// material signature, so each bin is one run: every position maps to run 0
// (run_of), the tail's wave slots walk the bin's records in park order
// (wave_run 0, wave_first = slot x per_wave), and per bin `small` holds
// {run_base 0, run_cap R, run_count 0, level-2 count 0, level-2 queue 0}.
__global__ void park_tables_kernel(int *__restrict__ run_of, long ncell,
                                   int *__restrict__ wave_run, int *__restrict__ wave_first,
                                   int nw, int per_wave, int *__restrict__ small, int nsmall,
                                   int R) {
  const long g = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (g < ncell) run_of[g] = 0;
  if (g < nw) {
    wave_run[g] = 0;
    wave_first[g] = (int)g * per_wave;
  }
  if (g < nsmall) small[g] = (g % 5 == 1) ? R : 0;
}

struct Arena {
  size_t size = 0, used = 0, peak = 0;
  std::map<size_t, size_t> free_;  // offset -> length
  void init(size_t n) {
    size = n;
    free_.clear();
    free_[0] = n;
  }
  size_t largest() const {
    size_t m = 0;
    for (auto &kv : free_) m = std::max(m, kv.second);
    return m;
  }
  bool alloc(size_t n, size_t &off) {
    for (auto it = free_.begin(); it != free_.end(); ++it) {
      if (it->second >= n) {
        off = it->first;
        const size_t rest = it->second - n;
        free_.erase(it);
        if (rest) free_[off + n] = rest;
        used += n;
        peak = std::max(peak, used);
        return true;
      }
    }
    return false;
  }
  void release(size_t off, size_t n) {
    used -= n;
    auto it = free_.emplace(off, n).first;
    auto nx = std::next(it);
    if (nx != free_.end() && it->first + it->second == nx->first) {
      it->second += nx->second;
      free_.erase(nx);
    }
    if (it != free_.begin()) {
      auto pv = std::prev(it);
      if (pv->first + pv->second == it->first) {
        pv->second += it->second;
        free_.erase(it);
      }
    }
  }
};

struct Rec {
  sgm_request req{};
  long nnz = 0;
  double t_submit = 0, t_start = 0, t_done = 0;
  int status = 0;
  bool done = false;
};

// byte layout of a launch group inside its arena block
struct Layout {
  long ncell = 0, nnz = 0, nraw = 0;
  int npte = 0;
  size_t reqs, raw, off, mat, rho, vfrac, sie, temp, press, vt, st, conv, keys, keys2, ids,
      sorted, range, counters, sort_tmp, sort_tmp_bytes;
  // straggler parking of the batched bins (2..6 materials): R records per bin
  long park_R = 0;
  int park_nw = 0, park_per_wave = 0;
  size_t run_of = 0, wave_run = 0, wave_first = 0, park_small = 0;
  size_t rec[7] = {}, cell_of[7] = {};
  std::vector<size_t> vec_in, vec_out;  // per vector request: rho, sie | P, T
  size_t total = 0;
};

struct Lane {
  hipStream_t s = nullptr;
  hipEvent_t ev = nullptr;
  bool busy = false;
  std::vector<uint64_t> reqs;
  std::vector<Rec> snap;  // the group's requests, copied under the lock
  size_t arena_off = 0, arena_len = 0;
  ReqDev *h_reqs = nullptr;  // pinned staging for the request table
  int h_cap = 0;
};

} // namespace

struct sgm_server {
  sgm_ctx *ctx = nullptr;
  sgm_server_config cfg{};
  MixParamsDev P{};
  char *d_arena = nullptr;
  Arena arena;
  std::vector<Lane> lanes;
  std::vector<Rec> recs;  // indexed by ticket
  std::deque<uint64_t> queue;
  std::mutex mu;
  std::condition_variable cv_work, cv_done;
  std::thread worker;
  bool stop = false;
  long groups = 0, completed = 0, failed = 0;
  double cells_done = 0;
  std::chrono::steady_clock::time_point t0;
  int ncu = 0;

  double now() const {
    return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  }
};

namespace {

size_t sort_tmp_bytes(long n) {
  size_t b = 0;
  hipcub::DeviceRadixSort::SortPairs(nullptr, b, (unsigned long long *)nullptr,
                                     (unsigned long long *)nullptr, (int *)nullptr,
                                     (int *)nullptr, (int)std::max<long>(n, 1), 0, 64);
  return b + 16;
}

Layout make_layout(const sgm_server *sv, const std::vector<uint64_t> &ids) {
  Layout L;
  size_t o = 0;
  auto take = [&](size_t bytes) {
    const size_t at = o;
    o += align_up(std::max<size_t>(bytes, 1));
    return at;
  };
  long nvec = 0;
  for (uint64_t t : ids) {
    const Rec &r = sv->recs[t];
    if (r.req.kind & SGM_REQ_PTE) {
      L.ncell += r.req.n;
      L.nnz += r.nnz;
      L.nraw += r.req.n + 1;
      ++L.npte;
    } else {
      ++nvec;
    }
  }
  L.reqs = take(sizeof(ReqDev) * std::max(1, L.npte));
  if (L.npte) {
    L.raw = take(sizeof(int) * L.nraw);
    L.off = take(sizeof(int) * (L.ncell + 1));
    L.mat = take(L.nnz);
    L.rho = take(8 * L.nnz);
    L.vfrac = take(8 * L.nnz);
    L.sie = take(8 * L.nnz);
    L.temp = take(8 * L.nnz);
    L.press = take(8 * L.nnz);
    L.vt = take(8 * L.ncell);
    L.st = take(8 * L.ncell);
    L.conv = take(4 * L.ncell);
    L.keys = take(8 * L.ncell);
    L.keys2 = take(8 * L.ncell);
    L.ids = take(4 * L.ncell);
    L.sorted = take(4 * L.ncell);
    L.range = take(sizeof(long) * 14);
    L.counters = take(sizeof(int) * 8);
    L.sort_tmp_bytes = sort_tmp_bytes(L.ncell);
    L.sort_tmp = take(L.sort_tmp_bytes);
    if (!sv->cfg.plain) {
      // records for a sixteenth of the group's cells per bin (at least 1024):
      // a cell still iterating after SERVER_PARK_AFTER Newton iterations moves
      // to the bin's tail kernels while records last, and keeps its lane past
      // that
      L.park_R = std::min<long>(L.ncell, std::max<long>(1024, L.ncell / 16));
      L.park_per_wave = 64 / TailGroup(sv->cfg.method, 2);
      L.park_nw = (int)((L.park_R + L.park_per_wave - 1) / L.park_per_wave);
      L.run_of = take(4 * L.ncell);
      L.wave_run = take(4 * (size_t)L.park_nw);
      L.wave_first = take(4 * (size_t)L.park_nw);
      L.park_small = take(4 * 5 * 5);
      for (int nm = 2; nm <= 6; ++nm) {
        L.rec[nm] = take(8 * (size_t)ParkStride(sv->cfg.method, nm) * L.park_R);
        L.cell_of[nm] = take(4 * (size_t)L.park_R);
      }
    }
  }
  for (uint64_t t : ids) {
    const Rec &r = sv->recs[t];
    if (r.req.kind & SGM_REQ_PTE) continue;
    const size_t n8 = 8 * (size_t)r.req.n;
    L.vec_in.push_back(take(2 * n8));
    L.vec_out.push_back(take(2 * n8));
  }
  L.total = align_up(o);
  return L;
}

// a request's own working set (admission bound when alone)
size_t request_bytes(const sgm_server *sv, uint64_t t) { return make_layout(sv, {t}).total; }

int launch_group(sgm_server *sv, Lane &ln, const Layout &L) {
  hipStream_t s = ln.s;
  char *base = sv->d_arena + ln.arena_off;
  auto D = [&](size_t off) { return base + off; };
  hipError_t e = hipSuccess;
  auto ok = [&](hipError_t x) {
    if (x != hipSuccess && e == hipSuccess) e = x;
    return e == hipSuccess;
  };
  const sgm_ctx *ctx = sv->ctx;
  // closures: merged CSR
  if (L.npte) {
    if (ln.h_cap < L.npte) {
      if (ln.h_reqs) hipHostFree(ln.h_reqs);
      ln.h_reqs = nullptr;
      ln.h_cap = 0;
      if (!ok(hipHostMalloc((void **)&ln.h_reqs, sizeof(ReqDev) * L.npte, 0))) return SGM_E_HIP;
      ln.h_cap = L.npte;
    }
    long cb = 0, nb = 0, rp = 0;
    int k = 0;
    for (const Rec &r : ln.snap) {
      if (!(r.req.kind & SGM_REQ_PTE)) continue;
      const sgm_request &q = r.req;
      ln.h_reqs[k++] = ReqDev{cb, q.n, nb, rp};
      const size_t n = (size_t)q.n, z = (size_t)r.nnz;
      ok(hipMemcpyAsync(D(L.raw) + 4 * rp, q.offsets, 4 * (n + 1), hipMemcpyHostToDevice, s));
      ok(hipMemcpyAsync(D(L.mat) + nb, q.mat_ids, z, hipMemcpyHostToDevice, s));
      ok(hipMemcpyAsync(D(L.rho) + 8 * nb, q.rho_in, 8 * z, hipMemcpyHostToDevice, s));
      ok(hipMemcpyAsync(D(L.vfrac) + 8 * nb, q.vfrac_in, 8 * z, hipMemcpyHostToDevice, s));
      ok(hipMemcpyAsync(D(L.temp) + 8 * nb, q.temp_in, 8 * z, hipMemcpyHostToDevice, s));
      ok(hipMemcpyAsync(D(L.press) + 8 * nb, q.press_in, 8 * z, hipMemcpyHostToDevice, s));
      ok(hipMemcpyAsync(D(L.vt) + 8 * cb, q.vfrac_tot, 8 * n, hipMemcpyHostToDevice, s));
      ok(hipMemcpyAsync(D(L.st) + 8 * cb, q.sie_tot, 8 * n, hipMemcpyHostToDevice, s));
      cb += q.n;
      nb += r.nnz;
      rp += q.n + 1;
    }
    ok(hipMemcpyAsync(D(L.reqs), ln.h_reqs, sizeof(ReqDev) * L.npte, hipMemcpyHostToDevice, s));
    if (e != hipSuccess) return SGM_E_HIP;
    int *off = (int *)D(L.off);
    rebase_kernel<<<(int)((L.ncell + 255) / 256), 256, 0, s>>>(
        (const ReqDev *)D(L.reqs), L.npte, (const int *)D(L.raw), off, L.ncell, L.nnz);
    ok(hipGetLastError());
    PteArgs a{};
    a.d_eos = ctx->d_eos;
    a.nreg = ctx->nmat;
    a.ncell = L.ncell;
    a.nnz = L.nnz;
    a.offsets = off;
    a.mat_ids = (const uint8_t *)D(L.mat);
    a.rho = (double *)D(L.rho);
    a.vfrac = (double *)D(L.vfrac);
    a.sie = (double *)D(L.sie);
    a.temp = (double *)D(L.temp);
    a.press = (double *)D(L.press);
    a.vfrac_tot = (const double *)D(L.vt);
    a.sie_tot = (const double *)D(L.st);
    a.conv = (int *)D(L.conv);
    if (sv->cfg.plain) {
      ok(launch_pte_plain(sv->cfg.method, a, sv->P, s));
    } else {
      auto *keys = (unsigned long long *)D(L.keys), *keys2 = (unsigned long long *)D(L.keys2);
      int *ids = (int *)D(L.ids), *sorted = (int *)D(L.sorted);
      long *range = (long *)D(L.range);
      int *counters = (int *)D(L.counters);
      ok(launch_plan_keys(off, a.mat_ids, L.ncell, L.nnz, keys, ids, s, nullptr, 0,
                          ctx_mod_mask(ctx)));
      size_t tb = L.sort_tmp_bytes;
      ok(hipcub::DeviceRadixSort::SortPairs(D(L.sort_tmp), tb, keys, keys2, ids, sorted,
                                            (int)L.ncell, 0, 64, s));
      ranges_kernel<<<1, 64, 0, s>>>(keys2, L.ncell, range);
      ok(hipGetLastError());
      ok(launch_zero_ints(counters, 8, s));
      const int grid = std::max(1, sv->ncu * 2);  // 8 waves per CU of 256-thread blocks
      int *small = (int *)D(L.park_small);
      const long nt = std::max<long>(std::max<long>(L.ncell, L.park_nw), 25);
      park_tables_kernel<<<(int)((nt + 255) / 256), 256, 0, s>>>(
          (int *)D(L.run_of), L.ncell, (int *)D(L.wave_run), (int *)D(L.wave_first), L.park_nw,
          L.park_per_wave, small, 25, (int)L.park_R);
      ok(hipGetLastError());
      // each bin's persistent kernel parks its stragglers, then the bin's
      // tail kernels finish them (as a plan does, on the lane's stream), so
      // a few wandering cells no longer hold the bin kernel's whole grid
      for (int nm = 2; nm <= 6; ++nm) {
        const long *r = range + 2 * (nm - 1);
        ParkArgs pk;
        int *sm = small + 5 * (nm - 2);
        pk.after = SERVER_PARK_AFTER;
        pk.stride = ParkStride(sv->cfg.method, nm);
        pk.rec = (double *)D(L.rec[nm]);
        pk.cell_of = (int *)D(L.cell_of[nm]);
        pk.run_of = (const int *)D(L.run_of);
        pk.run_base = sm;
        pk.run_cap = sm + 1;
        pk.run_count = sm + 2;
        pk.wave_run = (const int *)D(L.wave_run);
        pk.wave_first = (const int *)D(L.wave_first);
        pk.nwaves = L.park_nw;
        pk.after2 = 4;
        pk.l2_count = sm + 3;
        pk.l2_next = sm + 4;
        ok(launch_pte_bin(sv->cfg.method, nm, a, sv->P, sorted, 0, counters + nm, 256, 64, grid,
                          s, r, &pk));
        ok(launch_pte_tail(sv->cfg.method, nm, a, sv->P, pk, s));
      }
      for (int side = 0; side < 2; ++side) {
        const long *r = range + (side ? 12 : 0);
        ok(launch_pte_list(sv->cfg.method, a, sv->P, sorted, L.ncell, s, r));
      }
    }
    if (e != hipSuccess) return SGM_E_HIP;
    cb = 0;
    nb = 0;
    for (const Rec &r : ln.snap) {
      if (!(r.req.kind & SGM_REQ_PTE)) continue;
      const sgm_request &q = r.req;
      const size_t n = (size_t)q.n, z = (size_t)r.nnz;
      ok(hipMemcpyAsync(q.rho_out, D(L.rho) + 8 * nb, 8 * z, hipMemcpyDeviceToHost, s));
      ok(hipMemcpyAsync(q.vfrac_out, D(L.vfrac) + 8 * nb, 8 * z, hipMemcpyDeviceToHost, s));
      ok(hipMemcpyAsync(q.sie_out, D(L.sie) + 8 * nb, 8 * z, hipMemcpyDeviceToHost, s));
      ok(hipMemcpyAsync(q.temp_out, D(L.temp) + 8 * nb, 8 * z, hipMemcpyDeviceToHost, s));
      ok(hipMemcpyAsync(q.press_out, D(L.press) + 8 * nb, 8 * z, hipMemcpyDeviceToHost, s));
      if (q.converged)
        ok(hipMemcpyAsync(q.converged, D(L.conv) + 4 * cb, 4 * n, hipMemcpyDeviceToHost, s));
      cb += q.n;
      nb += r.nnz;
    }
  }
  // vector requests
  size_t vi = 0;
  for (const Rec &r : ln.snap) {
    if (r.req.kind & SGM_REQ_PTE) continue;
    const sgm_request &q = r.req;
    const size_t n8 = 8 * (size_t)q.n;
    double *in = (double *)D(L.vec_in[vi]);
    double *out = (double *)D(L.vec_out[vi]);
    ++vi;
    ok(hipMemcpyAsync(in, q.rho, n8, hipMemcpyHostToDevice, s));
    ok(hipMemcpyAsync(in + q.n, q.sie, n8, hipMemcpyHostToDevice, s));
    const EosDev &eh = ctx->host_eos[q.matindex];
    if (q.kind & SGM_REQ_VEC_P) {
      ok(launch_vec(eh, ctx->d_eos, q.matindex, SGM_F_P_RHO_SIE, in, in + q.n, out, q.n, 0,
                    nullptr, s));
      ok(hipMemcpyAsync(q.P_out, out, n8, hipMemcpyDeviceToHost, s));
    }
    if (q.kind & SGM_REQ_VEC_T) {
      ok(launch_vec(eh, ctx->d_eos, q.matindex, SGM_F_T_RHO_SIE, in, in + q.n, out + q.n, q.n, 0,
                    nullptr, s));
      ok(hipMemcpyAsync(q.T_out, out + q.n, n8, hipMemcpyDeviceToHost, s));
    }
  }
  ok(hipEventRecord(ln.ev, s));
  return e == hipSuccess ? SGM_OK : SGM_E_HIP;
}

void finish_lane(sgm_server *sv, Lane &ln, int status) {
  const double t = sv->now();
  for (uint64_t id : ln.reqs) {
    Rec &r = sv->recs[id];
    r.t_done = t;
    r.status = status;
    r.done = true;
    if (status == SGM_OK) {
      ++sv->completed;
      sv->cells_done += (double)r.req.n;
    } else {
      ++sv->failed;
    }
  }
  ln.reqs.clear();
  ln.snap.clear();
  sv->arena.release(ln.arena_off, ln.arena_len);
  ln.busy = false;
}

void scheduler(sgm_server *sv) {
  hipSetDevice(sv->ctx->device);
  std::unique_lock<std::mutex> lk(sv->mu);
  for (;;) {
    bool progressed = false;
    // completions
    for (Lane &ln : sv->lanes) {
      if (!ln.busy) continue;
      const hipError_t q = hipEventQuery(ln.ev);
      if (q == hipErrorNotReady) continue;
      finish_lane(sv, ln, q == hipSuccess ? SGM_OK : SGM_E_HIP);
      progressed = true;
    }
    if (progressed) sv->cv_done.notify_all();
    // admission: first come, first served
    for (Lane &ln : sv->lanes) {
      if (ln.busy || sv->queue.empty()) continue;
      std::vector<uint64_t> grp;
      long cells = 0;
      size_t need = 0;
      const size_t room = sv->arena.largest();
      while (!sv->queue.empty() && (int)grp.size() < sv->cfg.max_group_requests) {
        const uint64_t t = sv->queue.front();
        const long n = sv->recs[t].req.n;
        if (!grp.empty() && cells + n > sv->cfg.max_group_cells) break;
        grp.push_back(t);
        const Layout Lc = make_layout(sv, grp);
        // merged offsets are int32: the group's entry count must fit
        if (Lc.total > room || Lc.nnz > 2147483647L || Lc.ncell > 2147483647L) {
          grp.pop_back();
          break;
        }
        need = Lc.total;
        cells += n;
        sv->queue.pop_front();
        if (sv->cfg.plain) break;
      }
      if (grp.empty()) break;  // the head waits for memory
      size_t off = 0;
      sv->arena.alloc(need, off);  // fits: need <= largest free stretch
      ln.busy = true;
      ln.reqs = grp;
      ln.snap.clear();
      for (uint64_t id : grp) ln.snap.push_back(sv->recs[id]);
      ln.arena_off = off;
      ln.arena_len = need;
      const double t = sv->now();
      for (uint64_t id : grp) sv->recs[id].t_start = t;
      ++sv->groups;
      const Layout L = make_layout(sv, grp);
      lk.unlock();
      const int rc = launch_group(sv, ln, L);
      lk.lock();
      if (rc != SGM_OK) {
        hipStreamSynchronize(ln.s);
        finish_lane(sv, ln, rc);
        sv->cv_done.notify_all();
      }
      progressed = true;
      if (sv->cfg.plain) break;  // one request in flight at a time
    }
    bool any_busy = false;
    for (Lane &ln : sv->lanes) any_busy |= ln.busy;
    if (sv->stop && sv->queue.empty() && !any_busy) break;
    if (!progressed) {
      if (any_busy)
        sv->cv_work.wait_for(lk, std::chrono::microseconds(20));
      else
        sv->cv_work.wait(lk, [&] { return sv->stop || !sv->queue.empty(); });
    }
  }
}

} // namespace

extern "C" {

int sgm_server_create(sgm_ctx *ctx, const sgm_server_config *cfg, const sgm_mix_params *params,
                      sgm_server **out) {
  if (!ctx || !cfg || !out) return SGM_E_ARG;
  *out = nullptr;
  if (ctx->host_only) return SGM_E_NOGPU;
  if (cfg->hbm_budget_bytes < (64ull << 20) || cfg->nlanes < 1 || cfg->nlanes > 16 ||
      cfg->max_group_requests < 1 || cfg->max_group_cells < 1 ||
      (cfg->method != 0 && cfg->method != 1))
    return SGM_E_ARG;
  auto *sv = new sgm_server();
  sv->ctx = ctx;
  sv->cfg = *cfg;
  if (sv->cfg.plain) sv->cfg.nlanes = 1;
  sgm_mix_params q;
  if (params)
    q = *params;
  else
    sgm_mix_params_default(&q);
  std::memcpy(&sv->P, &q, sizeof(sv->P));
  sv->t0 = std::chrono::steady_clock::now();
  sv->ncu = device_cu_count();
  hipSetDevice(ctx->device);
  if (hipMalloc(&sv->d_arena, cfg->hbm_budget_bytes) != hipSuccess) {
    delete sv;
    return SGM_E_HIP;
  }
  sv->arena.init(cfg->hbm_budget_bytes);
  sv->lanes.resize(sv->cfg.nlanes);
  for (Lane &ln : sv->lanes) {
    if (hipStreamCreateWithFlags(&ln.s, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreateWithFlags(&ln.ev, hipEventDisableTiming) != hipSuccess) {
      sv->stop = true;
      sgm_server_destroy(sv);
      return SGM_E_HIP;
    }
  }
  // the registry must be on the device before the first group
  {
    std::lock_guard<std::mutex> g(ctx->mu);
    if (ctx->dirty) {
      if (hipMemcpy(ctx->d_eos, ctx->host_eos.data(), sizeof(EosDev) * ctx->nmat,
                    hipMemcpyHostToDevice) != hipSuccess) {
        sgm_server_destroy(sv);
        return SGM_E_HIP;
      }
      ctx->dirty = false;
    }
  }
  sv->worker = std::thread(scheduler, sv);
  *out = sv;
  return SGM_OK;
}

int sgm_server_submit(sgm_server *sv, const sgm_request *q, uint64_t *ticket) {
  if (!sv || !q || !ticket || q->n < 0) return SGM_E_ARG;
  Rec r;
  r.req = *q;
  if (q->kind & SGM_REQ_PTE) {
    if (q->kind != SGM_REQ_PTE) return SGM_E_ARG;
    if (!q->offsets || !q->mat_ids || !q->rho_in || !q->vfrac_in || !q->temp_in ||
        !q->press_in || !q->vfrac_tot || !q->sie_tot || !q->rho_out || !q->vfrac_out ||
        !q->sie_out || !q->temp_out || !q->press_out)
      return SGM_E_ARG;
    const long nnz = (long)q->offsets[q->n] - (long)q->offsets[0];
    if (nnz < 0 || nnz > 2147483647L) return SGM_E_ARG;
    r.nnz = nnz;
  } else {
    if (!(q->kind & (SGM_REQ_VEC_P | SGM_REQ_VEC_T)) || (q->kind & ~3)) return SGM_E_ARG;
    if (q->matindex < 0 || q->matindex >= sv->ctx->nmat || !sv->ctx->is_set[q->matindex])
      return SGM_E_MAT;
    if (!q->rho || !q->sie || ((q->kind & SGM_REQ_VEC_P) && !q->P_out) ||
        ((q->kind & SGM_REQ_VEC_T) && !q->T_out))
      return SGM_E_ARG;
  }
  std::lock_guard<std::mutex> g(sv->mu);
  if (sv->stop) return SGM_E_ARG;
  const uint64_t t = sv->recs.size();
  r.t_submit = sv->now();
  sv->recs.push_back(r);
  if (request_bytes(sv, t) > sv->arena.size) {
    sv->recs[t].done = true;
    sv->recs[t].status = SGM_E_BUDGET;
    ++sv->failed;
    *ticket = t;
    return SGM_E_BUDGET;
  }
  sv->queue.push_back(t);
  *ticket = t;
  sv->cv_work.notify_one();
  return SGM_OK;
}

int sgm_server_wait(sgm_server *sv, uint64_t t, double *times) {
  if (!sv) return SGM_E_ARG;
  std::unique_lock<std::mutex> lk(sv->mu);
  if (t >= sv->recs.size()) return SGM_E_ARG;
  sv->cv_done.wait(lk, [&] { return sv->recs[t].done; });
  const Rec &r = sv->recs[t];
  if (times) {
    times[0] = r.t_submit;
    times[1] = r.t_start;
    times[2] = r.t_done;
  }
  return r.status;
}

int sgm_server_test(sgm_server *sv, uint64_t t) {
  if (!sv) return SGM_E_ARG;
  std::lock_guard<std::mutex> g(sv->mu);
  if (t >= sv->recs.size()) return SGM_E_ARG;
  return sv->recs[t].done ? 1 : 0;
}

double sgm_server_now(const sgm_server *sv) { return sv ? sv->now() : 0.0; }

int sgm_server_stats_get(sgm_server *sv, sgm_server_stats *o) {
  if (!sv || !o) return SGM_E_ARG;
  std::lock_guard<std::mutex> g(sv->mu);
  o->submitted = (long)sv->recs.size();
  o->completed = sv->completed;
  o->failed = sv->failed;
  o->groups = sv->groups;
  o->cells = sv->cells_done;
  o->arena_peak_bytes = sv->arena.peak;
  o->arena_bytes = sv->arena.size;
  return SGM_OK;
}

int sgm_server_destroy(sgm_server *sv) {
  if (!sv) return SGM_E_ARG;
  {
    std::lock_guard<std::mutex> g(sv->mu);
    sv->stop = true;
  }
  sv->cv_work.notify_all();
  if (sv->worker.joinable()) sv->worker.join();
  for (Lane &ln : sv->lanes) {
    if (ln.s) hipStreamSynchronize(ln.s);
    if (ln.ev) hipEventDestroy(ln.ev);
    if (ln.s) hipStreamDestroy(ln.s);
    if (ln.h_reqs) hipHostFree(ln.h_reqs);
  }
  if (sv->d_arena) hipFree(sv->d_arena);
  delete sv;
  return SGM_OK;
}

} // extern "C"
