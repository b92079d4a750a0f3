"""GPU parity for the oracle pins that reach beyond the benchmark workloads,
through the C ABI, bit for bit against the oracle:

  * the reference's two-phase tin fixture (test/pte_test_2phaseVinetSn.hpp,
    both trial sets): ApproxTemperatureFromRhoMatU on the device, then
    PTESolverRhoT at the default MixParams with that temperature as Tnorm, on
    the plain and the batched path (the oracle's own equilibria are checked
    against the fixture's printed states in test_oracle_sn2phase.py);
  * Vinet with nonzero expansion coefficients d2..d40 (a few low orders,
    where the device skips the zero high orders of its Horner loops, and every
    order nonzero): every vector call, FillEos, and the closure (RhoT and PT,
    default and 1e-10 tolerances) on cells that hold that material;
  * PTESolverPT at the default MixParams on the uneven workload (plain and
    batched).
"""
import numpy as np
import pytest

import _sn2phase as SN
from test_gpu_pte import FIELDS, assert_identical, dev, run_batched, run_oracle, run_plain
from test_oracle_vinet_coeffs import D_ALL, D_LOW, VINET_CU

pytestmark = pytest.mark.gpu


def sn_context(sgm):
    ctx = sgm.Context(2)
    ctx.set_vinet(0, *SN.vinet_args("beta"))
    ctx.set_vinet(1, *SN.vinet_args("gamma"))
    return ctx


@pytest.mark.parametrize("which", ["long", "short"])
def test_sn_fixture_on_gpu(sgm, oracle, torch_gpu, which):
    import ctypes as C
    torch = torch_gpu
    ctx = sn_context(sgm)
    o = [oracle.vinet(*SN.vinet_args("beta")), oracle.vinet(*SN.vinet_args("gamma"))]
    a = SN.cells(which)
    n = a["ncell"]
    d = dev(torch, a)
    u = torch.tensor(a["u_tot"], device="cuda", dtype=torch.float64)
    Tg = torch.zeros(n, dtype=torch.float64, device="cuda")
    assert ctx.approx_temperature(n, a["nnz"], d["offsets"], d["mat_ids"], d["rho"],
                                  d["vfrac"], u, Tg) == 0
    torch.cuda.synchronize()
    Tg_h = Tg.cpu().numpy()
    handles = (C.c_void_p * 2)(o[0].h, o[1].h)
    ref = np.array([oracle.lib().sgo_approx_temperature(
        2, handles, float(a["u_tot"][t]), oracle.p(np.ascontiguousarray(a["rho"][2 * t:2 * t + 2])),
        oracle.p(np.ascontiguousarray(a["vfrac"][2 * t:2 * t + 2])), 0.0) for t in range(n)])
    assert np.array_equal(Tg_h, ref)
    # with a guess: the guess path of the bracket
    Tg2 = torch.zeros(n, dtype=torch.float64, device="cuda")
    guess = torch.tensor(ref * 1.3, device="cuda", dtype=torch.float64)
    assert ctx.approx_temperature(n, a["nnz"], d["offsets"], d["mat_ids"], d["rho"],
                                  d["vfrac"], u, Tg2, Tguess=guess) == 0
    torch.cuda.synchronize()
    ref2 = np.array([oracle.lib().sgo_approx_temperature(
        2, handles, float(a["u_tot"][t]), oracle.p(np.ascontiguousarray(a["rho"][2 * t:2 * t + 2])),
        oracle.p(np.ascontiguousarray(a["vfrac"][2 * t:2 * t + 2])), float(ref[t] * 1.3))
        for t in range(n)])
    assert np.array_equal(Tg2.cpu().numpy(), ref2)
    # the closure with Tnorm = the approximate temperature
    params = sgm.mix_params()
    r = {k: a[k].copy() for k in FIELDS}
    nf, conv, nit, res, err = oracle.pte_solve(0, o, a["offsets"], a["mat_ids"].astype(np.int32),
                                               r["rho"], r["vfrac"], r["sie"], r["temp"],
                                               r["press"], a["vfrac_tot"], a["sie_tot"],
                                               params.as_oracle_array(), Tnorm=ref.copy())
    r.update(conv=conv, niter=nit)
    d = dev(torch, a)
    rc = ctx.pte(0, n, a["nnz"], d["offsets"], d["mat_ids"], d["rho"], d["vfrac"], d["sie"],
                 d["temp"], d["press"], d["vfrac_tot"], d["sie_tot"], Tnorm=Tg, params=params,
                 conv=d["conv"], niter=d["niter"], resid=d["resid"], err=d["err"])
    assert rc == 0
    torch.cuda.synchronize()
    g = {k: v.cpu().numpy() for k, v in d.items()}
    assert_identical(g, r, "sn plain")
    assert g["conv"].all()
    # batched plan (the 2-material bin) with the same Tnorm
    d = dev(torch, a)
    plan = sgm.Plan(ctx, n, a["nnz"], d["offsets"], d["mat_ids"], 0)
    rc = plan.execute(d["rho"], d["vfrac"], d["sie"], d["temp"], d["press"], d["vfrac_tot"],
                      d["sie_tot"], Tnorm=Tg, params=params, conv=d["conv"], niter=d["niter"],
                      resid=d["resid"], err=d["err"])
    assert rc == 0
    torch.cuda.synchronize()
    plan.close()
    b = {k: v.cpu().numpy() for k, v in d.items()}
    assert_identical(b, r, "sn batched")
    ctx.close()


@pytest.fixture(scope="module", params=["low", "all"])
def vinet_coeff_pool(request, sgm, workloads, oracle, torch_gpu):
    """The uneven pool with the Vinet slot given expansion coefficients."""
    wl = workloads
    d = D_LOW if request.param == "low" else D_ALL
    ctx, t = wl.pool_context()
    ctx.set_vinet(4, *VINET_CU, d2to40=d)
    otab, _, _ = oracle.gruneisen_table(wl.GRUN_C1, t.lRhoMin, t.lRhoMax, t.nRho, t.lTMin,
                                        t.lTMax, t.nT)
    o = [oracle.ideal(*wl.IDEAL_A), oracle.ideal(*wl.IDEAL_M1), oracle.stiff(*wl.STIFF_B),
         oracle.gruneisen(*wl.GRUN_C1), oracle.vinet(*VINET_CU, d2to40=d), otab.eos()]
    return ctx, o, otab


def test_vinet_coeffs_vector_calls(vinet_coeff_pool, torch_gpu):
    torch = torch_gpu
    ctx, o, _ = vinet_coeff_pool
    rng = np.random.default_rng(17)
    n = 20000
    rho = rng.uniform(4.0, 14.0, n)
    T = rng.uniform(50.0, 2e4, n)
    sie = o[4].vec(2, rho, T) * rng.uniform(0.9, 1.1, n)
    for f in range(9):
        in2 = T if f in (2, 3, 5, 7) else sie
        ref = o[4].vec(f, rho, in2)
        d1 = torch.tensor(rho, device="cuda")
        d2 = torch.tensor(in2, device="cuda")
        out = torch.empty_like(d1)
        assert ctx.vec_eval(4, f, d1, d2, out) == 0
        torch.cuda.synchronize()
        got = out.cpu().numpy()
        assert np.array_equal(got, ref) or np.all((got == ref) | (np.isnan(got) & np.isnan(ref))), f
    # FillEos from (rho, sie): T, P, cv, bmod
    out = {k: np.zeros(n) for k in ("temp", "press", "cv", "bmod")}
    r_o, e_o = rho.copy(), sie.copy()
    errs = o[4].fill(r_o, out["temp"], e_o, out["press"], out["cv"], out["bmod"], 8 | 4 | 16 | 32)
    dd = {k: torch.zeros(n, dtype=torch.float64, device="cuda") for k in out}
    dr, de = torch.tensor(rho, device="cuda"), torch.tensor(sie, device="cuda")
    assert ctx.fill_eos(4, dr, dd["temp"], de, dd["press"], dd["cv"], dd["bmod"], n,
                        8 | 4 | 16 | 32) == 0
    torch.cuda.synchronize()
    for k in out:
        assert np.array_equal(dd[k].cpu().numpy(), out[k]), k


@pytest.mark.parametrize("method,tight", [(0, False), (0, True), (1, False), (1, True)])
def test_vinet_coeffs_closure(sgm, workloads, oracle, vinet_coeff_pool, torch_gpu, method, tight):
    ctx, o, _ = vinet_coeff_pool
    a = workloads.uneven(ctx, 8192 if method == 0 else 4096, seed=0x5EED0000 + 31)
    assert (a["mat_ids"] == 4).sum() > 1000
    params = sgm.mix_params(tight=tight)
    g = run_plain(torch_gpu, ctx, a, method, params)
    r = run_oracle(oracle, o, a, method, params.as_oracle_array())
    assert_identical(g, r, f"vinet coeffs m{method} tight={tight}")
    assert np.array_equal(g["err"], r["err"])
    b, _ = run_batched(torch_gpu, sgm, ctx, a, method, params)
    assert_identical(b, g, "batched vs plain")


def test_pt_default_tolerance_vs_oracle(sgm, workloads, oracle, torch_gpu):
    wl = workloads
    ctx, t = wl.pool_context()
    otab, _, _ = oracle.gruneisen_table(wl.GRUN_C1, t.lRhoMin, t.lRhoMax, t.nRho, t.lTMin,
                                        t.lTMax, t.nT)
    o = [oracle.ideal(*wl.IDEAL_A), oracle.ideal(*wl.IDEAL_M1), oracle.stiff(*wl.STIFF_B),
         oracle.gruneisen(*wl.GRUN_C1), oracle.vinet(*wl.VINET_CU), otab.eos()]
    a = wl.uneven(ctx, 8192, seed=0x5EED0000 + 32)
    params = sgm.mix_params()
    g = run_plain(torch_gpu, ctx, a, 1, params)
    r = run_oracle(oracle, o, a, 1, params.as_oracle_array())
    assert_identical(g, r, "PT default")
    assert np.array_equal(g["err"], r["err"])
    assert g["conv"].mean() > 0.9
    b, _ = run_batched(torch_gpu, sgm, ctx, a, 1, params)
    assert_identical(b, g, "PT default batched vs plain")
