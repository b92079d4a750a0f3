"""Benchmark: PTE closure cells/s on one MI355X.

Headline (`value`): batched uneven closure — 2^20 mixed cells with 2..6
materials from a pool of six (IdealGas x2, StiffGas, Gruneisen, Vinet, a
Gruneisen-sampled rho-T table), PTESolverRhoT at the 1e-10 tolerances of
singularity-eos/example/pte_2mat.cpp:169-180, FP64, one hipGraph launch per
step (persistent kernel per material-count bin). A step restores the solver
inputs (device-to-device copies; the closure works in place) and solves every
cell; inputs are resident in HBM before timing starts.

Reported beside it: the plain per-call path on the same cells and on the
65,536-cell two-material closure, the 2^24-query table lookups, the
roofline of the dominant bin kernel (HIP events on the launch stream) and the
oracle CPU baseline on a bounded sample of the same workload.
"""
import argparse
import importlib
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
for p in (ROOT, os.path.join(ROOT, "oracle")):
    if p not in sys.path:
        sys.path.insert(0, p)

METRIC = "PTE closure cells/s at 1e-10 tol + p99 request latency, served vs plain"
HBM_PEAK_GBS = 8000.0
FP64_PEAK_TFLOPS = 78.6

# Per material-count bin of the headline workload, one execution of the bin's
# kernels (pte_bin_kernel + both pte_tail_kernel levels): HBM bytes moved
# (rocprofv3 FETCH_SIZE + WRITE_SIZE, scaled by the counters' reading of a
# copy kernel of known size) and FP64 operations issued (SQ_INSTS_VALU_
# {ADD,MUL,TRANS}_F64 x 64 lanes + FMA_F64 x 128). Measured with the
# commands in profiles/pte_bins_uneven_2p20.md on the current kernels; None
# where not measured.
UNIT_HBM_BYTES = {2: 5.736e8, 3: 7.072e8, 4: 1.3336e9, 5: 2.3145e9, 6: 2.486e9}
UNIT_FP64_OPS = {2: 1.137e10, 3: 2.156e10, 4: 1.969e10, 5: 1.984e10, 6: 1.655e10}


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def bytes_per_cell(nm):
    """Algorithmic HBM bytes of one closure: reads rho, vfrac, temp, press per
    material, sie_tot, vfrac_tot, the CSR offset and material ids; writes rho,
    vfrac, sie, temp, press per material and the converged flag."""
    return 8 * 4 * nm + 16 + 4 + nm + 8 * 5 * nm + 4


def serve_bench(torch, ctx, a, params, nreq):
    """Serving loop on the same pool and closure workload: 64 clients, request
    sizes log-uniform in [1e3, 1e6], 70% vector {P, T}(rho, e) / 30% closure.
    Capacity = all requests at t = 0 (after an untimed warm-up trace of 16
    requests per mode); then Poisson arrivals at 50/80/95% of the served
    capacity. Served = continuous batching over 4 lanes; plain = one
    request at a time through the plain per-call kernels. Host arrays in, host
    arrays out: cells/s here include the PCIe copies."""
    serving = importlib.import_module("singularity-eos_amd.serving")
    pools = serving.Pools(torch, a)
    n = a["ncell"]
    out = {}
    sat = serving.make_trace(nreq, np.inf, 0x5EED0003, pool_cells=n)
    # a warm server: one untimed run of a short trace per mode first (the
    # first launch of each kernel in the process loads its code)
    warm = serving.make_trace(16, np.inf, 0x5EED0005, pool_cells=n)
    for plain in (False, True):
        serving.run(torch, ctx, pools, warm, plain=plain, params=params, ring_bytes=1 << 30)
    cap = {}
    for mode, plain in (("served", False), ("plain", True)):
        r = serving.run(torch, ctx, pools, sat, plain=plain, params=params, ring_bytes=1 << 30)
        s = serving.summarize(r)
        cap[mode] = nreq / r["makespan"]
        out[f"{mode}_capacity"] = dict(s, requests_per_s=cap[mode])
        log(f"[bench] serving {mode} capacity: {s['cells_per_s']:.4g} cells/s, "
            f"{cap[mode]:.3g} req/s, p99 {s['p99_ms']:.1f} ms")
    for load in (0.5, 0.8, 0.95):
        tr = serving.make_trace(nreq, load * cap["served"], 0x5EED0004, pool_cells=n)
        r = serving.run(torch, ctx, pools, tr, plain=False, params=params, ring_bytes=1 << 30)
        s = serving.summarize(r)
        out[f"served_at_{int(load * 100)}pct"] = dict(s, offered_req_per_s=load * cap["served"])
        log(f"[bench] served at {load:.0%}: p50 {s['p50_ms']:.1f} ms p99 {s['p99_ms']:.1f} ms")
    # plain at its own 50% load, and at the served loop's 50% load
    for tag, rate in (("plain_at_50pct", 0.5 * cap["plain"]),
                      ("plain_at_served_50pct", 0.5 * cap["served"])):
        tr = serving.make_trace(nreq, rate, 0x5EED0004, pool_cells=n)
        r = serving.run(torch, ctx, pools, tr, plain=True, params=params, ring_bytes=1 << 30)
        s = serving.summarize(r)
        out[tag] = dict(s, offered_req_per_s=rate)
        log(f"[bench] {tag}: p50 {s['p50_ms']:.1f} ms p99 {s['p99_ms']:.1f} ms")
    out["served_vs_plain_capacity"] = cap["served"] / cap["plain"]
    return out


def learned_bench(torch, ctx, params, ntrain, steps):
    """Learned initial guess: train on the device on ntrain solved cells of
    the uneven workload (another seed), evaluate on 2^16 held-out cells:
    Newton iterations with and without the guess at the same 1e-10
    tolerances, how often the guess was taken, and how many cells reached a
    different equilibrium than from the solver's own start."""
    wl = importlib.import_module("singularity-eos_amd.workloads")
    L = importlib.import_module("singularity-eos_amd.learned")
    t0 = time.perf_counter()
    tr = wl.uneven(ctx, ntrain, seed=wl.SEED_BASE + 4)
    held = wl.uneven(ctx, 1 << 16, seed=wl.SEED_BASE + 5)
    mlp = L.MLP(seed=7)
    X, tgt, rows = L.build_training_set(torch, mlp, ctx, tr, params)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    hist = L.train(torch, mlp, X, tgt, rows, steps=steps, batch=8192)
    torch.cuda.synchronize()
    t2 = time.perf_counter()
    d = L.to_device(torch, held)
    n, nnz = held["ncell"], held["nnz"]
    base, c0, it0, _ = L.solve(torch, ctx, d, n, nnz, params)
    gv, gT = mlp.guess(torch, ctx, d, n, nnz)
    out, c1, it1, used = L.solve(torch, ctx, d, n, nnz, params, guess=(gv, gT))
    # inference throughput (the guess kernel alone)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(5):
        mlp.guess(torch, ctx, d, n, nnz)
    e1.record()
    torch.cuda.synchronize()
    inf_s = e0.elapsed_time(e1) * 1e-3 / 5
    # one plan on the held-out cells: the plain execute against the guess and
    # the guided solve as one graph launch (execute_learned), each timed
    # after a warm-up that captures its graph
    plan = importlib.import_module("singularity-eos_amd").Plan(
        ctx, n, nnz, d["offsets"], d["mat_ids"], 0)
    o = {k: d[k].clone() for k in L.FIELDS}
    gv2, gT2 = torch.zeros_like(gv), torch.zeros_like(gT)

    def timed(run):
        best = 1e30
        for r in range(4):
            for k in L.FIELDS:
                o[k].copy_(d[k])
            a0, a1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            a0.record()
            assert run() == 0
            a1.record()
            torch.cuda.synchronize()
            if r:
                best = min(best, a0.elapsed_time(a1))
        return best

    plain_ms = timed(lambda: plan.execute(o["rho"], o["vfrac"], o["sie"], o["temp"], o["press"],
                                          d["vfrac_tot"], d["sie_tot"], params=params))
    fused_ms = timed(lambda: plan.execute_learned(mlp.h, o["rho"], o["vfrac"], o["sie"],
                                                  o["temp"], o["press"], d["vfrac_tot"],
                                                  d["sie_tot"], gv2, gT2, params=params))
    plan.close()
    it0, it1 = it0.cpu().numpy(), it1.cpu().numpy()
    c0, c1 = c0.cpu().numpy().astype(bool), c1.cpu().numpy().astype(bool)
    used = used.cpu().numpy().astype(bool)
    both = c0 & c1
    off = held["offsets"][:-1]
    T0 = base["temp"].cpu().numpy()[off]
    T1 = out["temp"].cpu().numpy()[off]
    diff = both & (np.abs(T1 - T0) > 1e-6 * np.abs(T0))
    flops = 2 * (64 * 128 + 2 * 128 * 128 + 128 * 16)
    res = {"train_cells": int(rows.numel()), "steps": steps, "batch": 8192,
           "loss_first": hist[0][1], "loss_last": hist[-1][1],
           "train_s": t2 - t1, "data_s": t1 - t0,
           "held_out_cells": n, "guess_used_frac": float(used.mean()),
           "plan_execute_ms": plain_ms, "learned_fused_execute_ms": fused_ms,
           "converged_plain": float(c0.mean()), "converged_guided": float(c1.mean()),
           "mean_iters_plain": float(it0[both].mean()), "mean_iters_guided": float(it1[both].mean()),
           "p99_iters_plain": float(np.percentile(it0[both], 99)),
           "p99_iters_guided": float(np.percentile(it1[both], 99)),
           "different_equilibrium_frac": float(diff.sum() / max(1, both.sum())),
           "inference_cells_per_s": n / inf_s,
           "inference_mfma_TFLOPs": n * flops / inf_s / 1e12}
    mlp.close()
    log(f"[bench] learned guess: iterations {res['mean_iters_plain']:.2f} -> "
        f"{res['mean_iters_guided']:.2f} (p99 {res['p99_iters_plain']:.0f} -> "
        f"{res['p99_iters_guided']:.0f}), used {res['guess_used_frac']:.3f}, "
        f"different equilibrium {res['different_equilibrium_frac']:.2e}")
    return res


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--ncell", type=int, default=1 << 20)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-extras", action="store_true")
    ap.add_argument("--autotune", action="store_true")
    ap.add_argument("--no-serve", action="store_true")
    ap.add_argument("--no-reorder", action="store_true")
    ap.add_argument("--no-learned", action="store_true")
    ap.add_argument("--learned-train-cells", type=int, default=1 << 22)
    ap.add_argument("--learned-steps", type=int, default=2000)
    ap.add_argument("--serve-requests", type=int, default=96)
    ap.add_argument("--serve-only", action="store_true",
                    help="run the serving loop's traces alone and print their summary "
                         "(for profiling the server); not the bench line")
    ap.add_argument("--copy-cal", action="store_true",
                    help="end with one elementwise kernel that reads 1 GiB and writes 1 GiB "
                         "(torch.mul; the scale for FETCH_SIZE / WRITE_SIZE counters)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build measures one GPU; the closure shards by cell "
                                   "ranges with no exchange (see DESIGN.md), --gpus must be 1"}))
        sys.exit(2)

    import torch
    sgm = importlib.import_module("singularity-eos_amd")
    wl = importlib.import_module("singularity-eos_amd.workloads")
    assert torch.cuda.is_available(), "bench needs a GPU"
    dev = torch.device("cuda:0")
    stream = torch.cuda.current_stream().cuda_stream
    params = sgm.mix_params(tight=True)

    ctx, table = wl.pool_context()
    log(f"[bench] building uneven workload: {args.ncell} cells")
    a = wl.uneven(ctx, args.ncell)
    n, nnz = a["ncell"], a["nnz"]
    nm_cell = np.diff(a["offsets"])
    fields = ("rho", "vfrac", "sie", "temp", "press")
    init = {k: torch.tensor(a[k], device=dev) for k in fields}
    work = {k: v.clone() for k, v in init.items()}
    vt = torch.tensor(a["vfrac_tot"], device=dev)
    st = torch.tensor(a["sie_tot"], device=dev)
    off = torch.tensor(a["offsets"], device=dev)
    mat = torch.tensor(a["mat_ids"], device=dev)
    conv = torch.zeros(n, dtype=torch.int32, device=dev)
    niter = torch.zeros(n, dtype=torch.int32, device=dev)

    if args.serve_only:
        print(json.dumps({"serving": serve_bench(torch, ctx, a, params, args.serve_requests)}))
        return

    # the plan (sorting the cells into material-count bins, straggler
    # records, graph capture on first execute) is built once per cell
    # topology, outside the timed steps; its cost is reported beside them
    torch.cuda.synchronize()
    tb = time.perf_counter()
    plan = sgm.Plan(ctx, n, nnz, off, mat, 0, stream=stream)
    torch.cuda.synchronize()
    plan_build_ms = (time.perf_counter() - tb) * 1e3
    if args.autotune:
        log("[bench] autotuning launch shapes")
        plan.autotune(init["rho"], init["vfrac"], init["temp"], init["press"], vt, st,
                      params=params, reps=1, stream=stream)
    bins = plan.describe()

    def restore():
        for k in fields:
            work[k].copy_(init[k], non_blocking=True)

    def step():
        restore()
        rc = plan.execute(work["rho"], work["vfrac"], work["sie"], work["temp"], work["press"], vt,
                          st, params=params, conv=conv, niter=niter, stream=stream)
        assert rc == 0, rc

    step_ms = []

    def timed(k):
        ev = [torch.cuda.Event(enable_timing=True) for _ in range(k + 1)]
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        ev[0].record()
        for i in range(k):
            step()
            ev[i + 1].record()
        torch.cuda.synchronize()
        dt = (time.perf_counter() - t0) / k
        step_ms[:] = [ev[i].elapsed_time(ev[i + 1]) for i in range(k)]
        return dt

    for _ in range(args.warmup):
        step()
    # the headline: cells in material-signature order
    dt = timed(args.steps)
    cells_per_s = n / dt
    conv_frac = float(conv.float().mean().item())
    mean_iter = float(niter.float().mean().item())
    log(f"[bench] batched: {dt * 1e3:.3f} ms/step, {cells_per_s:.4g} cells/s, "
        f"converged {conv_frac:.5f}, mean iters {mean_iter:.2f}")
    log(f"[bench] batched step times (ms): " + " ".join(f"{x:.1f}" for x in step_ms))
    step_stats = {"min": min(step_ms), "median": float(np.median(step_ms)), "max": max(step_ms)}

    # per-bin kernel timing (HIP events on the launch stream)
    per_bin = []
    for bi, b in enumerate(bins):
        restore()
        torch.cuda.synchronize()
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        reps = 3
        ms = 0.0
        kms = np.zeros(2)
        for _ in range(reps):
            restore()
            e0.record()
            km = []
            rc = plan.execute_bin(bi, work["rho"], work["vfrac"], work["sie"], work["temp"],
                                  work["press"], vt, st, params=params, conv=conv, niter=niter,
                                  stream=stream, kernel_ms=km)
            e1.record()
            assert rc == 0
            torch.cuda.synchronize()
            ms += e0.elapsed_time(e1)
            kms += km
        ms /= reps
        kms /= reps
        cells = b["ncells"]
        nm = b["nmat"]
        # ms: the bin from the caller's side (counter reset + kernels);
        # kernel_ms: the persistent kernel and its straggler tail kernel alone
        per_bin.append(dict(nmat=nm, cells=cells, ms=ms, bulk_kernel_ms=float(kms[0]),
                            tail_kernel_ms=float(kms[1]), block=b["block"],
                            cells_per_wave=b["cells_per_wave"], grid=b["grid"],
                            tail_level2_cells=plan.tail_counts(stream=stream)[bi]))
    extras = {"bins": per_bin, "step_ms": step_stats, "converged_frac": conv_frac,
              "plan_build_ms": plan_build_ms,
              "plan_build_note": "sgm_plan_create for the 2^20-cell topology (bins, straggler "
                                 "records), measured once before the timed steps and not part "
                                 "of them; a step with a new topology pays it once",
              "mean_newton_iters": mean_iter,
              "max_newton_iters": int(niter.max().item())}
    if not args.no_reorder:
        # the plan reused step after step: each bin ordered by the previous
        # step's iteration counts (the slowest cells spread one per wave)
        plan.reorder(niter, stream=stream)
        for _ in range(args.warmup):
            step()
        dt_r = timed(args.steps)
        extras["batched_cost_ordered_cells_per_s"] = n / dt_r
        log(f"[bench] batched, cost-ordered bins: {dt_r * 1e3:.3f} ms/step")
    dc = None  # get_sg_eos driver inputs (timed with the extras)
    if not args.no_extras:
        # plain per-call path, same cells
        restore()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        reps = 3
        for _ in range(reps):
            restore()
            rc = ctx.pte(0, n, nnz, off, mat, work["rho"], work["vfrac"], work["sie"],
                         work["temp"], work["press"], vt, st, params=params, conv=conv,
                         stream=stream)
            assert rc == 0
        torch.cuda.synchronize()
        extras["plain_uneven_cells_per_s"] = n * reps / (time.perf_counter() - t0)
        # two-material closure, plain per-call path (65,536 cells)
        c0 = wl.pte2(65536)
        c0["mat_ids"] = np.tile(np.array([0, 2], dtype=np.uint8), 65536)
        d0 = {k: torch.tensor(c0[k], device=dev) for k in fields + ("vfrac_tot", "sie_tot")}
        i0 = {k: d0[k].clone() for k in fields}
        o0 = torch.tensor(c0["offsets"], device=dev)
        m0 = torch.tensor(c0["mat_ids"], device=dev)
        c0_rates = {}
        for tol, prm in (("1e-10", params), ("default", sgm.mix_params())):
            for it in range(2):
                torch.cuda.synchronize()
                t0 = time.perf_counter()
                for _ in range(5):
                    for k in fields:
                        d0[k].copy_(i0[k])
                    ctx.pte(0, 65536, 131072, o0, m0, d0["rho"], d0["vfrac"], d0["sie"],
                            d0["temp"], d0["press"], d0["vfrac_tot"], d0["sie_tot"], params=prm,
                            stream=stream)
                torch.cuda.synchronize()
                c0_rates[tol] = 65536 * 5 / (time.perf_counter() - t0)
        extras["plain_pte2_65536_cells_per_s"] = c0_rates["1e-10"]
        extras["plain_pte2_65536_cells_per_s_by_tol"] = c0_rates
        # table lookups: 2^24 forward (P, e) and inverse T(rho, e)
        nq = 1 << 24
        rq, Tq = wl.table_queries(nq, table)
        rq = torch.tensor(rq, device=dev)
        Tq = torch.tensor(Tq, device=dev)
        Pq, eq, T2 = torch.empty_like(rq), torch.empty_like(rq), torch.empty_like(rq)
        for _ in range(2):
            ctx.fill_eos(5, rq, Tq, eq, Pq, None, None, nq, 2 | 4, stream=stream)
            ctx.vec_eval(5, 1, rq, eq, T2, stream=stream)
        torch.cuda.synchronize()
        e0, e1, e2 = (torch.cuda.Event(enable_timing=True) for _ in range(3))
        e0.record()
        ctx.fill_eos(5, rq, Tq, eq, Pq, None, None, nq, 2 | 4, stream=stream)
        e1.record()
        ctx.vec_eval(5, 1, rq, eq, T2, stream=stream)
        e2.record()
        torch.cuda.synchronize()
        fwd_ms, inv_ms = e0.elapsed_time(e1), e1.elapsed_time(e2)
        extras["table_forward_queries_per_s"] = nq / (fwd_ms * 1e-3)
        extras["table_forward_GBs"] = nq * 32 / (fwd_ms * 1e-3) / 1e9
        extras["table_inverse_queries_per_s"] = nq / (inv_ms * 1e-3)
        # analytic vector P(rho, e) on 2^24 points
        Pa = torch.empty_like(rq)
        ctx.vec_eval(2, 0, rq, eq, Pa, stream=stream)
        torch.cuda.synchronize()
        e0.record()
        ctx.vec_eval(2, 0, rq, eq, Pa, stream=stream)
        e1.record()
        torch.cuda.synchronize()
        vms = e0.elapsed_time(e1)
        extras["vector_P_rho_sie_GBs"] = nq * 24 / (vms * 1e-3) / 1e9
        # the get_sg_eos closure driver (rho-e input) over 2^20 cells built from
        # the same workload: every call includes its prologue/scan/fill/epilogue
        # passes, its temporary allocations and the PTE of the mixed cells
        dc = wl.driver_cells(ctx, n)
        nm_d = dc["nmat"]
        drv_in = {k: torch.tensor(dc[k], device=dev) for k in ("spvol", "sie", "temp", "press",
                                                                "frac_mass")}
        drv = {k: torch.zeros(n, dtype=torch.float64, device=dev)
               for k in ("pmax", "vol", "bmod", "dpde", "cv")}
        drv["vol"].fill_(1.0)
        for k in ("frac_vol", "frac_ie", "frac_bmod", "frac_dpde", "frac_cv"):
            drv[k] = torch.zeros(n * nm_d, dtype=torch.float64, device=dev)
        offs = torch.arange(1, n + 1, dtype=torch.int32, device=dev)

        def driver_call(mode, prm):
            for k, v in drv_in.items():
                drv[k] = v.clone()
            rc = sgm.get_sg_eos(ctx, mode, dc["eos_offsets"].tolist(), offs, drv, 1e-6,
                                params=prm, stream=stream)
            assert rc == 0, rc

        # rho-e (PTESolverRhoT), rho-T (FixedT), rho-P (FixedP: its absolute
        # pressure criterion cannot meet 1e-10, so the default tolerances),
        # P-T (per-material inversions, no PTE)
        drv_rates = {}
        for mode, name, prm in ((1, "rho_e", params), (-3, "rho_T", params),
                                (-2, "rho_P", sgm.mix_params()), (-1, "P_T", params)):
            driver_call(mode, prm)
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            reps = 3
            for _ in range(reps):
                driver_call(mode, prm)
            torch.cuda.synchronize()
            ddt = (time.perf_counter() - t0) / reps
            drv_rates[name] = n / ddt
            log(f"[bench] get_sg_eos ({name}) driver: {ddt * 1e3:.1f} ms per call of {n} cells")
        extras["get_sg_eos_rho_e_cells_per_s"] = drv_rates["rho_e"]
        extras["get_sg_eos_cells_per_s"] = drv_rates
        # the other closure formulations, plain per-call path, 65,536 uneven
        # cells: RhoU (cell energy), FixedT (the cell's first input
        # temperature), FixedP (its volume-weighted input pressure, default
        # tolerances)
        u = wl.uneven(ctx, 65536, seed=0x5EED0000 + 7)
        ucell = np.repeat(np.arange(u["ncell"]), np.diff(u["offsets"]))
        cin = {4: u["sie_tot"], 2: u["temp"][u["offsets"][:-1]].copy(),
               3: np.abs(np.bincount(ucell, weights=u["vfrac"] * u["press"])) + 1e5}
        du = {k: torch.tensor(u[k], device=dev) for k in fields + ("vfrac_tot",)}
        ou = torch.tensor(u["offsets"], device=dev)
        mu = torch.tensor(u["mat_ids"], device=dev)
        cu = torch.zeros(u["ncell"], dtype=torch.int32, device=dev)
        form_rates = {}
        for method, name in ((4, "RhoU"), (2, "FixedT"), (3, "FixedP")):
            ci = torch.tensor(cin[method], device=dev)
            prm = params if method != 3 else sgm.mix_params()
            ts = []
            for _ in range(3):
                w = {k: du[k].clone() for k in fields}
                torch.cuda.synchronize()
                t0 = time.perf_counter()
                rc = ctx.pte(method, u["ncell"], u["nnz"], ou, mu, w["rho"], w["vfrac"], w["sie"],
                             w["temp"], w["press"], du["vfrac_tot"], ci, params=prm, conv=cu,
                             stream=stream)
                assert rc == 0, rc
                torch.cuda.synchronize()
                ts.append(time.perf_counter() - t0)
            form_rates[name] = u["ncell"] / min(ts[1:])
            log(f"[bench] plain {name}, 65536 uneven cells: {form_rates[name]:.4g} cells/s, "
                f"converged {cu.float().mean().item():.4f}")
        extras["plain_formulations_65536_cells_per_s"] = form_rates

    if not args.no_serve:
        extras["serving"] = serve_bench(torch, ctx, a, params, args.serve_requests)
    if not args.no_learned:
        extras["learned_guess"] = learned_bench(torch, ctx, params, args.learned_train_cells,
                                                args.learned_steps)

    cpu = None
    flops_per_cell = None
    if not args.no_cpu:
        # oracle leg: CPU baseline timing, plus the algorithmic FP64 operation
        # count per cell (the restatement built with an op-counting Real)
        import oracle as O
        O.build()
        t = table
        otab, _, _ = O.gruneisen_table(wl.GRUN_C1, t.lRhoMin, t.lRhoMax, t.nRho, t.lTMin, t.lTMax,
                                       t.nT)
        olist = [O.ideal(*wl.IDEAL_A), O.ideal(*wl.IDEAL_M1), O.stiff(*wl.STIFF_B),
                 O.gruneisen(*wl.GRUN_C1), O.vinet(*wl.VINET_CU), otab.eos()]
        specs = [(0,) + wl.IDEAL_A, (0,) + wl.IDEAL_M1, (1,) + wl.STIFF_B, (2,) + wl.GRUN_C1,
                 (3,) + wl.VINET_CU, (4,)]
        ns = min(1 << 16, n)
        oo = a["offsets"][: ns + 1]
        k = oo[-1]
        me = a["mat_ids"][:k].astype(np.int32)
        threads = min(os.cpu_count() or 1, 16)

        def cpu_run(ncells, nthreads):
            kk = oo[ncells]
            r = {f: a[f][:kk].copy() for f in fields}
            t0 = time.perf_counter()
            O.pte_solve(0, olist, oo[: ncells + 1], me[:kk], r["rho"], r["vfrac"], r["sie"],
                        r["temp"], r["press"], a["vfrac_tot"][:ncells], a["sie_tot"][:ncells],
                        params.as_oracle_array(), nthreads=nthreads)
            return time.perf_counter() - t0

        cpu_run(ns, threads)  # warm-up
        cdt = float(np.median([cpu_run(ns, threads) for _ in range(10)]))
        n1 = min(8192, ns)
        c1dt = float(np.median([cpu_run(n1, 1) for _ in range(3)]))
        model = ""
        try:
            with open("/proc/cpuinfo") as f:
                model = next(l.split(":", 1)[1].strip() for l in f if l.startswith("model name"))
        except (OSError, StopIteration):
            pass
        cpu = {"value": ns / cdt, "unit": "cells/s", "cores": threads, "kind": "port",
               "sample": f"first {ns} cells of the same 2^20-cell uneven workload, "
                         f"PTESolverRhoT at 1e-10 tol, OpenMP over cells on {threads} of "
                         f"{os.cpu_count()} host threads ({model}), median of 10 runs "
                         f"({cdt:.3f} s each) after a warm-up; 1 thread on the first {n1} "
                         f"cells: {n1 / c1dt:.4g} cells/s",
               "single_core_value": n1 / c1dt}
        log(f"[bench] cpu baseline: {cpu['value']:.4g} cells/s on {threads} threads")
        # the two-material closure (65,536 cells) and the table lookups on a
        # 2^20-query sample, on the same host threads
        c0 = wl.pte2(65536)
        m0 = np.tile(np.array([0, 2], dtype=np.int32), 65536)
        c0_cpu = {}
        for tol, prm in (("1e-10", params), ("default", sgm.mix_params())):
            def c0_run():
                r = {f: c0[f].copy() for f in fields}
                t0 = time.perf_counter()
                O.pte_solve(0, olist, c0["offsets"], m0, r["rho"], r["vfrac"], r["sie"],
                            r["temp"], r["press"], c0["vfrac_tot"], c0["sie_tot"],
                            prm.as_oracle_array(), nthreads=threads)
                return time.perf_counter() - t0
            c0_run()
            c0_cpu[tol] = 65536 / float(np.median([c0_run() for _ in range(3)]))
        nq1 = 1 << 20
        rq1, Tq1 = wl.table_queries(nq1, table)
        te = otab.eos()
        F_P_RT, F_E_RT, F_T_RE = 3, 2, 1

        def c1_run():
            t0 = time.perf_counter()
            te.vec(F_P_RT, rq1, Tq1, nthreads=threads)
            eq1 = te.vec(F_E_RT, rq1, Tq1, nthreads=threads)
            t1 = time.perf_counter()
            te.vec(F_T_RE, rq1, eq1, nthreads=threads)
            return t1 - t0, time.perf_counter() - t1
        c1_run()
        c1 = [c1_run() for _ in range(3)]
        c1_fwd = nq1 / float(np.median([x[0] for x in c1]))
        c1_inv = nq1 / float(np.median([x[1] for x in c1]))
        extras["cpu_baselines_other_configs"] = {
            "pte2_65536_cells_per_s": c0_cpu,
            "table_forward_queries_per_s": c1_fwd,
            "table_inverse_queries_per_s": c1_inv,
            "cores": threads, "kind": "port",
            "sample": "two-material closure: all 65,536 cells (IdealGas + StiffGas), median of "
                      "3; table: 2^20 of the 2^24 log-uniform queries on the 256x512 "
                      "Gruneisen-sampled table, P(rho,T) and e(rho,T) forward, T(rho,e) "
                      "inverse, median of 3"}
        log(f"[bench] cpu two-material closure: {c0_cpu}; table forward {c1_fwd:.4g}, "
            f"inverse {c1_inv:.4g} queries/s")
        fl = O.count_pte_flops(0, specs, oo, me, a["rho"][:k], a["vfrac"][:k], a["temp"][:k],
                               a["press"][:k], a["vfrac_tot"][:ns], a["sie_tot"][:ns],
                               params.as_oracle_array(), table=otab)
        nms = nm_cell[:ns]
        flops_per_cell = {int(m): float(fl[nms == m].mean()) for m in np.unique(nms)}
        if dc is not None:
            # the oracle's get_sg_eos (rho-e) on the first 2^15 cells of the
            # driver workload, same threads
            nd = min(1 << 15, n)
            nmd = dc["nmat"]
            fm = dc["frac_mass"].reshape(nmd, n)[:, :nd]
            olist_d = [olist[i - 1] for i in dc["eos_offsets"]]

            def cpu_driver():
                arrs = {k: dc[k][:nd].copy() for k in ("spvol", "sie", "temp", "press")}
                for k in ("pmax", "bmod", "dpde", "cv"):
                    arrs[k] = np.zeros(nd)
                arrs["vol"] = np.ones(nd)
                arrs["frac_mass"] = np.ascontiguousarray(fm).ravel()
                for k in ("frac_vol", "frac_ie", "frac_bmod", "frac_dpde", "frac_cv"):
                    arrs[k] = np.zeros(nd * nmd)
                t0 = time.perf_counter()
                O.get_sg_eos(olist_d, 1, dc["eos_offsets"], np.arange(1, nd + 1), arrs, 1e-6,
                             params.as_oracle_array(), nthreads=threads)
                return time.perf_counter() - t0

            cpu_driver()
            ddt = float(np.median([cpu_driver() for _ in range(3)]))
            extras["get_sg_eos_rho_e_cpu_cells_per_s"] = nd / ddt
            extras["get_sg_eos_rho_e_cpu_sample"] = (
                f"oracle get_sg_eos (rho-e) on the first {nd} cells of the driver workload, "
                f"{threads} threads, median of 3")
            log(f"[bench] get_sg_eos cpu: {nd / ddt:.4g} cells/s on {threads} threads")

    # the dominant bin: its persistent kernel and its tail kernel together
    # solve the bin's cells (a cell parked by the first finishes in the
    # second), so the roofline is priced over the sum of their durations
    dom = max(per_bin, key=lambda r: r["bulk_kernel_ms"] + r["tail_kernel_ms"])
    nm = dom["nmat"] if dom["nmat"] else 4
    kern_ms = dom["bulk_kernel_ms"] + dom["tail_kernel_ms"]
    alg_bytes = dom["cells"] * bytes_per_cell(nm)
    gbs = alg_bytes / (kern_ms * 1e-3) / 1e9
    kname = f"pte_bin_kernel<{dom['nmat']}, 0> + pte_tail_kernel<{dom['nmat']}, 0> (RhoT)"
    split = {"bulk_kernel_ms": dom["bulk_kernel_ms"], "tail_kernel_ms": dom["tail_kernel_ms"]}
    traffic = UNIT_HBM_BYTES.get(nm)
    hbm = {"bound": "hbm", "achieved": gbs, "peak": HBM_PEAK_GBS, "unit": "GB/s",
           "frac": gbs / HBM_PEAK_GBS, "traffic": traffic, "kernel": kname,
           "bytes_per_cell": bytes_per_cell(nm)}
    if flops_per_cell and nm in flops_per_cell:
        tf = dom["cells"] * flops_per_cell[nm] / (kern_ms * 1e-3) / 1e12
        roofline = {"bound": "fp64_valu", "achieved": tf, "peak": FP64_PEAK_TFLOPS,
                    "unit": "TFLOP/s", "frac": tf / FP64_PEAK_TFLOPS, "traffic": traffic,
                    "kernel": kname, "cells": dom["cells"], "ms": kern_ms, **split,
                    "flops_per_cell": flops_per_cell[nm],
                    "algorithmic_bytes": alg_bytes,
                    "pipe": "FP64 on the vector ALUs (no MFMA in this kernel); 78.6 TFLOP/s "
                            "is MI355X's FP64 peak; achieved = the oracle's counted operations "
                            "per cell (every +,-,*,/ and math call once) x cells / the bin's "
                            "bulk + tail kernel time"}
        ops = UNIT_FP64_OPS.get(nm)
        if ops:
            roofline["rocprof_fp64_TFLOPs"] = ops / (kern_ms * 1e-3) / 1e12
            roofline["rocprof_fp64_frac"] = roofline["rocprof_fp64_TFLOPs"] / FP64_PEAK_TFLOPS
        extras["roofline_hbm"] = hbm
        extras["counted_flops_per_cell_by_nmat"] = flops_per_cell
    else:
        roofline = dict(hbm, cells=dom["cells"], ms=kern_ms, **split)

    out = {
        "metric": METRIC, "value": cells_per_s, "unit": "cells/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": dt * 1e3,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": "batched uneven PTE closure (PTESolverRhoT), 2^20 cells, "
                               "2-6 materials from a pool of 6, 1e-10 tolerance, hipGraph",
                   "cells": n, "nnz": nnz, "mean_nmat": float(nm_cell.mean()),
                   "parallelism": "single GPU"},
        "roofline": roofline, "cpu_baseline": cpu, "extras": extras,
    }
    print(json.dumps(out))
    if args.copy_cal:
        src = torch.ones(1 << 27, dtype=torch.float64, device=dev)  # 1 GiB
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        torch.mul(src, 2.0, out=dst)  # reads 1 GiB, writes 1 GiB
        torch.cuda.synchronize()
    plan.close()
    ctx.close()


if __name__ == "__main__":
    main()
