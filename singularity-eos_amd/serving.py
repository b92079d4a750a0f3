"""Serving-loop load generator: many mesh-block clients, Poisson arrivals,
request sizes log-uniform in [1e3, 1e6] cells, 70% vector {P, T}(rho, e)
requests on one pool material and 30% closure requests (slices of the
uneven 2..6-material workload). Open loop: each request is submitted at its
arrival time whatever the server's state; latency = completion - arrival.

Inputs come from pinned host pools built once (the client data); outputs go
to a pinned ring that is reused once a request has completed.
"""
import time

import numpy as np

from . import Server

FIELDS = ("rho", "vfrac", "sie", "temp", "press")


class Pools:
    """Pinned host pools: the uneven closure workload (per-entry and per-cell
    arrays, CSR offsets) and per-material (rho, sie) point sets for vector
    requests, each at least `npoint` long."""

    def __init__(self, torch, w, nmat=6, npoint=1 << 20):
        pin = dict(pin_memory=True)
        self.ncell = w["ncell"]
        self.offsets = torch.tensor(w["offsets"], dtype=torch.int32).pin_memory()
        self.mat_ids = torch.tensor(w["mat_ids"], dtype=torch.uint8).pin_memory()
        self.pte = {k: torch.tensor(w[k], dtype=torch.float64).pin_memory() for k in FIELDS}
        self.vfrac_tot = torch.tensor(w["vfrac_tot"], dtype=torch.float64).pin_memory()
        self.sie_tot = torch.tensor(w["sie_tot"], dtype=torch.float64).pin_memory()
        self.vec = []
        for m in range(nmat):
            sel = w["mat_ids"] == m
            r, e = w["rho"][sel], w["sie"][sel]
            reps = -(-npoint // max(1, len(r)))
            rr = torch.empty(len(r) * reps, dtype=torch.float64, **pin)
            ee = torch.empty(len(r) * reps, dtype=torch.float64, **pin)
            rr.copy_(torch.from_numpy(np.tile(r, reps)))
            ee.copy_(torch.from_numpy(np.tile(e, reps)))
            self.vec.append((rr, ee))


class Ring:
    """First-fit allocator over one pinned float64 buffer (outputs)."""

    def __init__(self, torch, nbytes):
        self.buf = torch.empty(nbytes // 8, dtype=torch.float64, pin_memory=True)
        self.free = [(0, self.buf.numel())]

    def alloc(self, n):
        n = max(1, int(n))
        for i, (o, l) in enumerate(self.free):
            if l >= n:
                if l == n:
                    self.free.pop(i)
                else:
                    self.free[i] = (o + n, l - n)
                return o
        return None

    def release(self, o, n):
        n = max(1, int(n))
        self.free.append((o, n))
        self.free.sort()
        merged = []
        for a, l in self.free:
            if merged and merged[-1][0] + merged[-1][1] == a:
                merged[-1] = (merged[-1][0], merged[-1][1] + l)
            else:
                merged.append((a, l))
        self.free = merged

    def view(self, o, n):
        return self.buf[o:o + n]


def make_trace(nreq, rate, seed, pool_cells, nmat=6, pte_frac=0.3, nmin=1e3, nmax=1e6,
               nclients=64):
    """Arrival times (s), kind (0 vector, 1 closure), size, material or cell
    start, client id. rate = requests/s (np.inf: all at t = 0)."""
    rng = np.random.default_rng(seed)
    gaps = np.zeros(nreq) if not np.isfinite(rate) else rng.exponential(1.0 / rate, nreq)
    t = np.cumsum(gaps)
    kind = (rng.random(nreq) < pte_frac).astype(int)
    size = np.exp(rng.uniform(np.log(nmin), np.log(nmax), nreq)).astype(np.int64)
    size = np.minimum(size, pool_cells)
    mat = rng.integers(0, nmat, nreq)
    start = (rng.random(nreq) * (pool_cells - size + 1)).astype(np.int64)
    client = rng.integers(0, nclients, nreq)
    return dict(t=t, kind=kind, size=size, mat=mat, start=start, client=client)


def run(torch, ctx, pools, trace, plain=False, budget=8 << 30, nlanes=4, params=None,
        ring_bytes=3 << 30, max_group_cells=1 << 22):
    """Replays a trace against a fresh server. Returns per-request latency
    (s), total cells, makespan and server stats."""
    srv = Server(ctx, budget_bytes=budget, nlanes=nlanes, plain=plain, params=params,
                 max_group_cells=max_group_cells)
    ring = Ring(torch, ring_bytes)
    off_np = pools.offsets.numpy()
    n = len(trace["t"])
    tickets, allocs, arrival = [], [], np.zeros(n)
    pending = []  # (ticket, ring offset, size) not yet known done
    t_base = srv.now()
    wall0 = time.perf_counter()

    def reap(block):
        keep = []
        for tk, o, l in pending:
            if srv.done(tk):
                ring.release(o, l)
            else:
                keep.append((tk, o, l))
        pending[:] = keep
        if block and pending:
            srv.wait(pending[0][0])

    for i in range(n):
        # wait for the arrival time (open loop)
        while True:
            dt = trace["t"][i] - (time.perf_counter() - wall0)
            if dt <= 0:
                break
            time.sleep(min(dt, 2e-4))
        arrival[i] = t_base + trace["t"][i]
        sz = int(trace["size"][i])
        if trace["kind"][i] == 0:
            m = int(trace["mat"][i])
            rr, ee = pools.vec[m]
            s0 = int(trace["start"][i]) % max(1, rr.numel() - sz)
            need = 2 * sz
            o = ring.alloc(need)
            while o is None:
                reap(True)
                o = ring.alloc(need)
            out = ring.view(o, need)
            tk = srv.submit_vec(m, rr[s0:s0 + sz], ee[s0:s0 + sz], P_out=out[:sz],
                                T_out=out[sz:])
        else:
            c0 = int(trace["start"][i])
            c1 = c0 + sz
            e0, e1 = int(off_np[c0]), int(off_np[c1])
            nnz = e1 - e0
            need = 5 * nnz + (sz + 1) // 2 + 1
            o = ring.alloc(need)
            while o is None:
                reap(True)
                o = ring.alloc(need)
            buf = ring.view(o, need)
            out = {k: buf[j * nnz:(j + 1) * nnz] for j, k in enumerate(FIELDS)}
            conv = buf[5 * nnz:].view(torch.int32)[:sz]
            tk = srv.submit_pte(pools.offsets[c0:c1 + 1], pools.mat_ids[e0:e1],
                                pools.pte["rho"][e0:e1], pools.pte["vfrac"][e0:e1],
                                pools.pte["temp"][e0:e1], pools.pte["press"][e0:e1],
                                pools.vfrac_tot[c0:c1], pools.sie_tot[c0:c1], out, conv)
        tickets.append(tk)
        pending.append((tk, o, need))
        reap(False)
    lat = np.zeros(n)
    status = np.zeros(n, dtype=int)
    t_done_max = 0.0
    for i, tk in enumerate(tickets):
        rc, (ts, tst, td) = srv.wait(tk)
        status[i] = rc
        lat[i] = td - arrival[i]
        t_done_max = max(t_done_max, td)
    stats = srv.stats()
    srv.close()
    cells = float(np.sum(trace["size"]))
    makespan = t_done_max - t_base
    return dict(latency=lat, status=status, cells=cells, makespan=makespan, stats=stats,
                lateness=None)


def summarize(res):
    lat = res["latency"]
    return {"cells_per_s": res["cells"] / res["makespan"], "p50_ms": float(np.percentile(lat, 50) * 1e3),
            "p99_ms": float(np.percentile(lat, 99) * 1e3), "requests": int(len(lat)),
            "failed": int(np.sum(res["status"] != 0)), "groups": int(res["stats"]["groups"]),
            "arena_peak_MiB": res["stats"]["arena_peak_bytes"] / 2**20,
            "makespan_s": res["makespan"]}
