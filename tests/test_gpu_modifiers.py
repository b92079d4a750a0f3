"""ShiftedEOS / ScaledEOS in the device material switch (sgm_set_modifiers,
the init_sg_* enabled/vals layout), bit for bit against the oracle's
modifiers (pinned in test_oracle_modifiers.py by the reference's goldens):
every vector call and FillEos of every model under Scaled, Shifted and
Scaled<Shifted>, and the closures (RhoT and PT, 1e-10 tolerance, plain and
batched) on the uneven workload with modified pool materials."""
import numpy as np
import pytest

from test_gpu_pte import assert_identical, run_batched, run_oracle, run_plain

pytestmark = pytest.mark.gpu

MODS = [dict(scaled=True, scale=1.7), dict(shifted=True, shift=2.5e8),
        dict(scaled=True, shifted=True, scale=0.8, shift=-1.0e8)]


def oracle_pool(oracle, wl, t):
    otab, _, _ = oracle.gruneisen_table(wl.GRUN_C1, t.lRhoMin, t.lRhoMax, t.nRho, t.lTMin,
                                        t.lTMax, t.nT)
    return [oracle.ideal(*wl.IDEAL_A), oracle.ideal(*wl.IDEAL_M1), oracle.stiff(*wl.STIFF_B),
            oracle.gruneisen(*wl.GRUN_C1), oracle.vinet(*wl.VINET_CU), otab.eos()], otab


def test_golden_on_gpu(sgm, torch_gpu):
    """test_eos_modifiers.cpp:125-140 through the C ABI."""
    torch = torch_gpu
    ctx = sgm.Context(1)
    ctx.set_ideal(0, 0.5, 2.0)
    ctx.set_modifiers(0, scaled=True, shifted=True, scale=2.0, shift=0.1)
    r = torch.tensor([2.0], device="cuda", dtype=torch.float64)
    e = torch.tensor([0.5], device="cuda", dtype=torch.float64)
    P = torch.zeros(1, device="cuda", dtype=torch.float64)
    assert ctx.vec_eval(0, sgm.F_P_RHO_SIE, r, e, P) == 0
    torch.cuda.synchronize()
    assert P.item() == 0.3
    ctx.close()


@pytest.mark.parametrize("mi", range(len(MODS)))
def test_modified_vector_calls_and_fill(sgm, workloads, oracle, torch_gpu, mi):
    torch = torch_gpu
    wl = workloads
    ctx, t = wl.pool_context()
    o, otab = oracle_pool(oracle, wl, t)
    mod = MODS[mi]
    for m in range(6):
        ctx.set_modifiers(m, **mod)
        o[m].modify(**mod)
    rng = np.random.default_rng(30 + mi)
    n = 8192
    for m in range(6):
        if m in (0, 1):
            rho = np.exp(rng.normal(np.log(1e-3), 0.6, n))
        elif m == 2:
            rho = rng.uniform(0.5, 1.5, n)
        elif m in (3, 5):
            rho = rng.uniform(0.5, 12.0, n)
        else:
            rho = rng.uniform(4.0, 14.0, n)
        T = rng.uniform(50.0, 2e4, n)
        sie = o[m].vec(2, rho, T) * rng.uniform(0.9, 1.1, n)
        for f in range(9):
            in2 = T if f in (2, 3, 5, 7) else sie
            ref = o[m].vec(f, rho, in2)
            d1, d2 = torch.tensor(rho, device="cuda"), torch.tensor(in2, device="cuda")
            out = torch.empty_like(d1)
            assert ctx.vec_eval(m, f, d1, d2, out) == 0
            torch.cuda.synchronize()
            assert np.array_equal(out.cpu().numpy(), ref, equal_nan=True), (m, f)
        # FillEos from (rho, sie) and from (rho, T)
        for output in (8 | 4 | 16 | 32, 2 | 4 | 16 | 32):
            arrs = dict(rho=rho.copy(), temp=T.copy(), sie=sie.copy(), press=np.zeros(n),
                        cv=np.zeros(n), bmod=np.zeros(n))
            ref = {k: v.copy() for k, v in arrs.items()}
            errs = o[m].fill(ref["rho"], ref["temp"], ref["sie"], ref["press"], ref["cv"],
                             ref["bmod"], output)
            d = {k: torch.tensor(v, device="cuda") for k, v in arrs.items()}
            de = torch.zeros(n, dtype=torch.int32, device="cuda")
            assert ctx.fill_eos(m, d["rho"], d["temp"], d["sie"], d["press"], d["cv"], d["bmod"],
                                n, output, errs=de) == 0
            torch.cuda.synchronize()
            for k in arrs:
                assert np.array_equal(d[k].cpu().numpy(), ref[k], equal_nan=True), (m, output, k)
            assert np.array_equal(de.cpu().numpy() != 0, errs != 0), (m, output)
    ctx.close()


@pytest.mark.parametrize("method", [0, 1])
def test_modified_materials_in_the_closure(sgm, workloads, oracle, torch_gpu, method):
    wl = workloads
    ctx, t = wl.pool_context()
    o, otab = oracle_pool(oracle, wl, t)
    # scaled ideal gas, shifted stiffened gas, scaled+shifted Gruneisen,
    # shifted Vinet; the second ideal gas and the table unmodified
    plan = {0: MODS[0], 2: MODS[1], 3: MODS[2], 4: dict(shifted=True, shift=-3.0e8)}
    a = wl.uneven(ctx, 8192 if method == 0 else 4096, seed=0x5EED0000 + 51)
    for m, mod in plan.items():
        ctx.set_modifiers(m, **mod)
        o[m].modify(**mod)
    params = sgm.mix_params(tight=True)
    g = run_plain(torch_gpu, ctx, a, method, params)
    r = run_oracle(oracle, o, a, method, params.as_oracle_array())
    assert_identical(g, r, f"modified closure m{method}")
    assert np.array_equal(g["err"], r["err"])
    assert g["conv"].mean() > 0.8
    b, info = run_batched(torch_gpu, sgm, ctx, a, method, params)
    assert_identical(b, g, "modified batched vs plain")
    # the batched kernels are built without modifiers: cells with a modified
    # material run on the list kernel, the others in the bins
    listed = sum(i["ncells"] for i in info if i["nmat"] == 0)
    binned = sum(i["ncells"] for i in info if 2 <= i["nmat"] <= 6)
    assert listed > 0 and binned > 0 and listed + binned == a["ncell"]
    ctx.close()


def test_plan_refuses_execution_after_modifiers_change(sgm, workloads, torch_gpu):
    torch = torch_gpu
    wl = workloads
    ctx, _ = wl.pool_context()
    a = wl.uneven(ctx, 2048, seed=0x5EED0000 + 52)
    from test_gpu_pte import dev
    d = dev(torch, a)
    plan = sgm.Plan(ctx, a["ncell"], a["nnz"], d["offsets"], d["mat_ids"], 0)
    args = (d["rho"], d["vfrac"], d["sie"], d["temp"], d["press"], d["vfrac_tot"], d["sie_tot"])
    assert plan.execute(*args) == 0
    ctx.set_modifiers(3, shifted=True, shift=1e8)
    assert plan.execute(*args) == -1
    plan.close()
    plan = sgm.Plan(ctx, a["ncell"], a["nnz"], d["offsets"], d["mat_ids"], 0)
    assert plan.execute(*args) == 0
    torch.cuda.synchronize()
    plan.close()
    ctx.close()
