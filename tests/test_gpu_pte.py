"""GPU parity of the PTE closure (plain per-call path and batched hipGraph
path) against the oracle, through the C ABI.

Bar: bit for bit — states, convergence flags and iteration counts — at both
the default MixParams and the 1e-10 tolerances of
singularity-eos/example/pte_2mat.cpp:169-180. Every operation on the path is
IEEE +,-,*,/,sqrt,floor or integer, compiled without FMA contraction on both
sides, with exp/cbrt (Vinet) evaluated by the same routine on host and device.

Convergence-rate checks count every cell; the few that stop at the iteration
cap or the small-step bailout do so identically in the oracle.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu
FIELDS = ("rho", "vfrac", "sie", "temp", "press")


def dev(torch, a, keys=FIELDS + ("vfrac_tot", "sie_tot")):
    d = {k: torch.tensor(a[k], device="cuda", dtype=torch.float64) for k in keys}
    d["offsets"] = torch.tensor(a["offsets"], device="cuda", dtype=torch.int32)
    d["mat_ids"] = torch.tensor(a["mat_ids"], device="cuda", dtype=torch.uint8)
    n = a["ncell"]
    d["conv"] = torch.zeros(n, dtype=torch.int32, device="cuda")
    d["niter"] = torch.zeros(n, dtype=torch.int32, device="cuda")
    d["resid"] = torch.zeros(n, dtype=torch.float64, device="cuda")
    d["err"] = torch.zeros(n, dtype=torch.int32, device="cuda")
    return d


def run_plain(torch, ctx, a, method, params):
    d = dev(torch, a)
    rc = ctx.pte(method, a["ncell"], a["nnz"], d["offsets"], d["mat_ids"], d["rho"], d["vfrac"],
                 d["sie"], d["temp"], d["press"], d["vfrac_tot"], d["sie_tot"], params=params,
                 conv=d["conv"], niter=d["niter"], resid=d["resid"], err=d["err"])
    assert rc == 0
    torch.cuda.synchronize()
    return {k: v.cpu().numpy() for k, v in d.items()}


def run_batched(torch, sgm, ctx, a, method, params, autotune=False, replays=1):
    d = dev(torch, a)
    plan = sgm.Plan(ctx, a["ncell"], a["nnz"], d["offsets"], d["mat_ids"], method)
    if autotune:
        plan.autotune(d["rho"], d["vfrac"], d["temp"], d["press"], d["vfrac_tot"], d["sie_tot"],
                      params=params, reps=1)
    init = {k: d[k].clone() for k in FIELDS}
    outs = []
    for r in range(replays):
        for k in FIELDS:
            d[k].copy_(init[k])
        if r:
            torch.randn(1 << 16, device="cuda").sum().item()  # unrelated eager work between replays
        rc = plan.execute(d["rho"], d["vfrac"], d["sie"], d["temp"], d["press"], d["vfrac_tot"],
                          d["sie_tot"], params=params, conv=d["conv"], niter=d["niter"],
                          resid=d["resid"], err=d["err"])
        assert rc == 0
        torch.cuda.synchronize()
        outs.append({k: v.cpu().numpy().copy() for k, v in d.items()})
    info = plan.describe()
    plan.close()
    return (outs if replays > 1 else outs[0]), info


def run_oracle(oracle, olist, a, method, params_arr):
    r = {k: a[k].copy() for k in FIELDS}
    nf, conv, nit, res, err = oracle.pte_solve(
        method, olist, a["offsets"], a["mat_ids"].astype(np.int32), r["rho"], r["vfrac"],
        r["sie"], r["temp"], r["press"], a["vfrac_tot"], a["sie_tot"], params_arr)
    r.update(conv=conv, niter=nit, resid=res, err=err)
    return r


def assert_identical(g, o, what=""):
    for k in FIELDS:
        a, b = g[k], o[k]
        ok = (a == b) | (np.isnan(a) & np.isnan(b))
        assert ok.all(), (what, k, int((~ok).sum()))
    assert np.array_equal(g["conv"], o["conv"]), what
    assert np.array_equal(g["niter"], o["niter"]), what


@pytest.fixture(scope="module")
def pool(sgm, workloads, oracle, torch_gpu):
    wl = workloads
    ctx, table = wl.pool_context()
    t = table
    otab, _, _ = oracle.gruneisen_table(wl.GRUN_C1, t.lRhoMin, t.lRhoMax, t.nRho, t.lTMin,
                                        t.lTMax, t.nT)
    o = [oracle.ideal(*wl.IDEAL_A), oracle.ideal(*wl.IDEAL_M1), oracle.stiff(*wl.STIFF_B),
         oracle.gruneisen(*wl.GRUN_C1), oracle.vinet(*wl.VINET_CU), otab.eos()]
    return ctx, o, otab


def c0_inputs(workloads, n):
    a = workloads.pte2(n)
    a["mat_ids"] = np.tile(np.array([0, 2], dtype=np.uint8), n)  # pool slots: IdealGas A, StiffGas B
    return a


@pytest.mark.parametrize("tight", [False, True])
def test_c0_plain_bit_exact_vs_oracle(sgm, workloads, oracle, pool, torch_gpu, tight):
    ctx, o, _ = pool
    a = c0_inputs(workloads, 65536)
    params = sgm.mix_params(tight=tight)
    g = run_plain(torch_gpu, ctx, a, 0, params)
    r = run_oracle(oracle, o, a, 0, params.as_oracle_array())
    assert_identical(g, r, f"c0 tight={tight}")
    assert g["conv"].all()


def test_c0_matches_closed_form(sgm, workloads, pool, torch_gpu):
    """IdealGas + StiffGas(q=0) equilibrium is a quadratic in alpha_B:
    P_inf (a+b+K) x^2 - ((a+b) u + (b+K) P_inf) x + b u = 0 with
    a = (g_A) rhobar_A Cv_A, b = (g_B) rhobar_B Cv_B, K = rhobar_A Cv_A + rhobar_B Cv_B,
    T = (u - P_inf x)/K. At the 1e-10 tolerance T and alpha agree to 1e-11 and
    P to 1e-9."""
    ctx, _, _ = pool
    n = 65536
    a = c0_inputs(workloads, n)
    g = run_plain(torch_gpu, ctx, a, 0, sgm.mix_params(tight=True))
    L = np.longdouble
    rb = (a["rho"] * a["vfrac"]).reshape(n, 2).astype(L)
    u = a["sie_tot"].astype(L) * rb.sum(1)
    gA, CA, gB, CB, Pi = L(0.4), L(7.176e6), L(3.4), L(3.6e7), L(6e9)
    aa, bb, K = gA * rb[:, 0] * CA, gB * rb[:, 1] * CB, rb[:, 0] * CA + rb[:, 1] * CB
    qa, qb, qc = Pi * (aa + bb + K), -((aa + bb) * u + (bb + K) * Pi), bb * u
    disc = np.sqrt(qb * qb - 4 * qa * qc)
    x1, x2 = (-qb - disc) / (2 * qa), (-qb + disc) / (2 * qa)
    x = np.where((x1 > 0) & (x1 < 1) & ((u - Pi * x1) > 0), x1, x2)
    T = (u - Pi * x) / K
    P = aa * T / (1 - x)
    Tg = g["temp"].reshape(n, 2)
    assert float(np.max(np.abs(Tg[:, 0] - T) / T)) < 1e-11
    assert float(np.max(np.abs(g["vfrac"].reshape(n, 2)[:, 1] - x) / x)) < 1e-11
    assert float(np.max(np.abs(g["press"].reshape(n, 2) - P[:, None]) / np.abs(P[:, None]))) < 1e-9


def compare_uneven(g, r, a):
    assert_identical(g, r, "uneven")
    assert np.array_equal(g["err"], r["err"])


@pytest.mark.parametrize("method", [0, 1])
def test_c2_plain_and_batched_vs_oracle(sgm, workloads, oracle, pool, torch_gpu, method):
    ctx, o, _ = pool
    n = 16384 if method == 0 else 4096
    a = workloads.uneven(ctx, n)
    params = sgm.mix_params(tight=True)
    g = run_plain(torch_gpu, ctx, a, method, params)
    r = run_oracle(oracle, o, a, method, params.as_oracle_array())
    compare_uneven(g, r, a)
    b, info = run_batched(torch_gpu, sgm, ctx, a, method, params)
    assert {i["nmat"] for i in info} >= {2, 3, 4, 5, 6}
    assert_identical(b, g, "batched vs plain")


def test_c2_default_tolerance_vs_oracle(sgm, workloads, oracle, pool, torch_gpu):
    ctx, o, _ = pool
    a = workloads.uneven(ctx, 8192, seed=0x5EED0000 + 22)
    params = sgm.mix_params()
    g = run_plain(torch_gpu, ctx, a, 0, params)
    r = run_oracle(oracle, o, a, 0, params.as_oracle_array())
    compare_uneven(g, r, a)
    assert g["conv"].mean() > 0.95


def test_batched_graph_replay_and_autotune_invariance(sgm, workloads, pool, torch_gpu):
    ctx, _, _ = pool
    a = workloads.uneven(ctx, 1 << 15, seed=0x5EED0000 + 23)
    params = sgm.mix_params(tight=True)
    outs, _ = run_batched(torch_gpu, sgm, ctx, a, 0, params, replays=3)
    for r in outs[1:]:
        assert_identical(r, outs[0], "replay")
    tuned, info = run_batched(torch_gpu, sgm, ctx, a, 0, params, autotune=True)
    assert_identical(tuned, outs[0], "autotuned")


def test_odd_material_counts_and_bad_cells(sgm, oracle, pool, torch_gpu):
    """1, 7 and 8 materials take the generic path; malformed cells (offsets
    out of order, more than 64 materials, a material the context lacks) are
    flagged SGM_ERRB_INPUT without touching memory outside their slice."""
    ctx, o, _ = pool
    rng = np.random.default_rng(5)
    counts = [1, 7, 8, 2, 3, 1, 8]
    mats, rho, vf, T, P, sie_tot, off = [], [], [], [], [], [], [0]
    for nm in counts:
        ids = rng.choice([0, 1, 2, 3, 5], size=nm, replace=True)
        a = rng.exponential(size=nm) + 1e-15
        a /= a.sum()
        rr = np.where(ids <= 1, 1e-3, np.where(ids == 2, 1.0, 2.7)) * rng.uniform(0.9, 1.1, nm)
        tt = rng.uniform(300, 3000, nm)
        ee = np.array([o[i].vec(2, [r_], [t_])[0] for i, r_, t_ in zip(ids, rr, tt)])
        pp = np.array([o[i].vec(3, [r_], [t_])[0] for i, r_, t_ in zip(ids, rr, tt)])
        mats += list(ids)
        rho += list(rr)
        vf += list(a)
        T += list(tt)
        P += list(pp)
        sie_tot.append(float((rr * a * ee).sum() / (rr * a).sum()))
        off.append(off[-1] + nm)
    A = dict(rho=np.array(rho), vfrac=np.array(vf), sie=np.zeros(len(rho)), temp=np.array(T),
             press=np.array(P), vfrac_tot=np.ones(len(counts)), sie_tot=np.array(sie_tot),
             offsets=np.array(off, dtype=np.int32), mat_ids=np.array(mats, dtype=np.uint8),
             ncell=len(counts), nnz=off[-1])
    params = sgm.mix_params(tight=True)
    g = run_plain(torch_gpu, ctx, A, 0, params)
    r = run_oracle(oracle, o, A, 0, params.as_oracle_array())
    assert_identical(g, r, "odd counts")
    b, _ = run_batched(torch_gpu, sgm, ctx, A, 0, params)
    assert_identical(b, g, "odd counts batched")
    # malformed: decreasing offsets, 68 materials (past the 64 of one wave),
    # unknown material index; the 9-material cell beside them is solved
    B = {k: (v.copy() if hasattr(v, "copy") else v) for k, v in A.items()}
    B["offsets"] = np.array([0, 3, 2, 70, 79, 81, 82, 90], dtype=np.int32)
    B["nnz"] = 90
    for k in ("rho", "vfrac", "sie", "temp", "press"):
        B[k] = np.resize(A[k], 90)
    B["mat_ids"] = np.resize(A["mat_ids"], 90)
    B["mat_ids"][79] = 60  # cell 4 names a material the context does not hold
    g = run_plain(torch_gpu, ctx, B, 0, params)
    assert g["err"][1] & 32 and g["err"][2] & 32 and g["err"][4] & 32
    assert g["err"][0] & 32 == 0 and g["err"][3] & 32 == 0


def test_empty_and_rejected_calls(sgm, pool, torch_gpu):
    ctx, _, _ = pool
    z = torch_gpu.zeros(1, dtype=torch_gpu.float64, device="cuda")
    zi = torch_gpu.zeros(2, dtype=torch_gpu.int32, device="cuda")
    zu = torch_gpu.zeros(1, dtype=torch_gpu.uint8, device="cuda")
    assert ctx.pte(0, 0, 0, zi, zu, z, z, z, z, z, z, z) == 0
    assert ctx.pte(0, 1, 1, None, zu, z, z, z, z, z, z, z) == -1
    assert ctx.vec_eval(99, 0, z, z, z, n=1) == -2
    assert ctx.vec_eval(0, 42, z, z, z, n=1) == -1


def test_c2_full_size_invariants(sgm, workloads, pool, torch_gpu):
    """2^20 uneven cells at 1e-10 tolerance, batched path: conservation and
    equilibrium properties that hold at any size, and batched == plain."""
    torch = torch_gpu
    ctx, _, _ = pool
    a = workloads.uneven(ctx, 1 << 20)
    params = sgm.mix_params(tight=True)
    b, _ = run_batched(torch, sgm, ctx, a, 0, params)
    g = run_plain(torch, ctx, a, 0, params)
    assert_identical(b, g, "full size batched vs plain")
    off = a["offsets"]
    conv = b["conv"].astype(bool)
    assert conv.mean() > 0.95
    cell = np.repeat(np.arange(a["ncell"]), np.diff(off))
    vsum = np.bincount(cell, weights=b["vfrac"])
    assert np.max(np.abs(vsum[conv] - 1.0)) < 1e-12
    rhobar_in = a["rho"] * a["vfrac"]
    assert np.max(np.abs(b["rho"] * b["vfrac"] - rhobar_in) / rhobar_in) < 1e-12
    u_in = a["sie_tot"] * np.bincount(cell, weights=rhobar_in)
    u_out = np.bincount(cell, weights=rhobar_in * b["sie"])
    assert np.max(np.abs(u_out[conv] - u_in[conv]) / np.abs(u_in[conv])) < 1e-9
    T = b["temp"]
    Tc = T[off[:-1]]
    assert np.max(np.abs(T - Tc[cell])) == 0.0


def test_reorder_by_cost_keeps_results(sgm, workloads, pool, torch_gpu):
    """Ordering each bin by a previous run's iteration counts changes only the
    schedule: every output is bit-identical before and after."""
    torch = torch_gpu
    ctx, _, _ = pool
    a = workloads.uneven(ctx, 20000, seed=0x5EED0000 + 23)
    params = sgm.mix_params(tight=True)
    d = dev(torch, a)
    plan = sgm.Plan(ctx, a["ncell"], a["nnz"], d["offsets"], d["mat_ids"], 0)
    outs = []
    for step in range(2):
        w = {k: d[k].clone() for k in FIELDS}
        assert plan.execute(w["rho"], w["vfrac"], w["sie"], w["temp"], w["press"],
                            d["vfrac_tot"], d["sie_tot"], params=params, conv=d["conv"],
                            niter=d["niter"]) == 0
        torch.cuda.synchronize()
        outs.append({k: w[k].cpu().numpy() for k in FIELDS})
        outs[-1]["niter"] = d["niter"].cpu().numpy().copy()
        if step == 0:
            plan.reorder(d["niter"])
    for k in FIELDS + ("niter",):
        assert np.array_equal(outs[0][k], outs[1][k], equal_nan=True), k
    plan.close()


@pytest.mark.parametrize("method", [0, 1, 2, 3])
def test_straggler_parking_keeps_results(sgm, workloads, pool, torch_gpu, method):
    """Parking cells between Newton iterations and resuming them in the tail
    kernels changes only the schedule: with every cell parked after 1 or 2
    iterations (so many that the tails' second level runs one lane per cell),
    after 20 (the default: the second level runs line-search groups) and 40,
    and with parking off, each output, flag and iteration count is
    bit-identical to the plain per-call path, whose loop runs every pass (the
    tails skip the passes of cells at a fixed point of the loop)."""
    torch = torch_gpu
    ctx, _, _ = pool
    a = workloads.uneven(ctx, 12000, seed=0x5EED0000 + 24)
    if method == 3:  # fixed pressure: the volume-weighted input pressure per cell
        cell = np.repeat(np.arange(a["ncell"]), np.diff(a["offsets"]))
        a["sie_tot"] = np.abs(np.bincount(cell, weights=a["vfrac"] * a["press"])) + 1e5
    if method == 2:  # fixed temperature: each cell's first input temperature
        a["sie_tot"] = a["temp"][a["offsets"][:-1]].copy()
    params = sgm.mix_params(tight=method != 3)
    g = run_plain(torch, ctx, a, method, params)
    d = dev(torch, a)
    plan = sgm.Plan(ctx, a["ncell"], a["nnz"], d["offsets"], d["mat_ids"], method)
    for after in (2, 20, 40, 0, 1):
        plan.set_straggler_threshold(after)
        w = {k: d[k].clone() for k in FIELDS}
        for k in ("conv", "niter", "resid", "err"):
            d[k].fill_(-7)
        assert plan.execute(w["rho"], w["vfrac"], w["sie"], w["temp"], w["press"],
                            d["vfrac_tot"], d["sie_tot"], params=params, conv=d["conv"],
                            niter=d["niter"], resid=d["resid"], err=d["err"]) == 0
        torch.cuda.synchronize()
        out = {k: w[k].cpu().numpy() for k in FIELDS}
        out.update({k: d[k].cpu().numpy() for k in ("conv", "niter", "resid", "err")})
        assert_identical(out, g, f"parking after {after}")
        assert np.array_equal(out["err"], g["err"])
        assert np.array_equal(out["resid"], g["resid"])
    plan.close()


def test_plan_tail_counts(sgm, workloads, pool, torch_gpu):
    """sgm_plan_tail_counts: per bin, the cells the last execution's tail
    resumed in its second level -- the same on a repeated execution, zero
    with parking off and for list bins, and at most the bin's cells whose
    Newton loop ran 4 iterations past the straggler threshold (each moved
    unless a fixed point let it finish in the first level)."""
    torch = torch_gpu
    ctx, _, _ = pool
    a = workloads.uneven(ctx, 1 << 18)
    d = dev(torch, a)
    nm_cell = np.diff(a["offsets"])
    plan = sgm.Plan(ctx, a["ncell"], a["nnz"], d["offsets"], d["mat_ids"], 0)
    params = sgm.mix_params(tight=True)
    init = {k: d[k].clone() for k in FIELDS}
    counts, niters = [], []
    for thr in (20, 20, 0):
        plan.set_straggler_threshold(thr)
        for k in FIELDS:
            d[k].copy_(init[k])
        rc = plan.execute(d["rho"], d["vfrac"], d["sie"], d["temp"], d["press"], d["vfrac_tot"],
                          d["sie_tot"], params=params, conv=d["conv"], niter=d["niter"])
        assert rc == 0
        torch.cuda.synchronize()
        counts.append(plan.tail_counts())
        niters.append(d["niter"].cpu().numpy())
    info = plan.describe()
    plan.close()
    assert counts[0] == counts[1]
    assert not any(counts[2])
    assert np.array_equal(niters[0], niters[2])  # the schedule changes nothing
    moved = 0
    for b, c in zip(info, counts[0]):
        if not 2 <= b["nmat"] <= 6:
            assert c == 0
            continue
        late = int((niters[0][nm_cell == b["nmat"]] >= 24).sum())
        assert 0 <= c <= late, (b["nmat"], c, late)
        moved += c
    assert moved > 0
