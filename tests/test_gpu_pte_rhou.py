"""GPU parity of PTESolverRhoU (sgm_pte_rhoU: plain per-call kernels) against
the oracle's restatement, through the C ABI.

Bar: bit for bit -- per-material states (RhoU writes one temperature per
material), convergence flags, iteration counts and error bits -- at the
default and at the 1e-10 tolerances, on cells of the uneven 2..6-material
workload and on 1, 7 and 8-material cells; batched plans (register-resident
bins for 2..4 materials, the list kernel above) equal the plain path.
"""
import numpy as np
import pytest

from test_gpu_pte import assert_identical, pool, run_oracle, run_plain  # noqa: F401

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("tight", [False, True])
def test_rhou_plain_bit_exact(sgm, workloads, oracle, pool, torch_gpu, tight):
    ctx, o, _ = pool
    a = workloads.uneven(ctx, 3000, seed=0x5EED0000 + 60 + int(tight))
    params = sgm.mix_params(tight=tight)
    g = run_plain(torch_gpu, ctx, a, 4, params)
    r = run_oracle(oracle, o, a, 4, params.as_oracle_array())
    assert_identical(g, r, f"rhou tight={tight}")
    assert np.array_equal(g["err"], r["err"])
    assert g["conv"].mean() > 0.5


def test_rhou_odd_material_counts(sgm, oracle, pool, torch_gpu):
    ctx, o, _ = pool
    rng = np.random.default_rng(11)
    counts = [1, 7, 8, 2, 3, 1, 8]
    mats, rho, vf, T, sie, P, off = [], [], [], [], [], [], [0]
    for nm in counts:
        ids = rng.choice([0, 1, 2, 3, 4, 5], size=nm, replace=True)
        al = rng.exponential(size=nm) + 1e-3
        al /= al.sum()
        rr = np.where(ids <= 1, 1e-3, np.where(ids == 2, 1.0, np.where(ids == 4, 8.9, 2.7)))
        rr = rr * rng.uniform(0.95, 1.05, nm)
        tt = rng.uniform(300, 3000, nm)
        ee = np.array([o[i].vec(2, [r_], [t_])[0] for i, r_, t_ in zip(ids, rr, tt)])
        pp = np.array([o[i].vec(3, [r_], [t_])[0] for i, r_, t_ in zip(ids, rr, tt)])
        mats += list(ids)
        rho += list(rr)
        vf += list(al)
        T += list(tt)
        sie += list(ee)
        P += list(pp)
        off.append(off[-1] + nm)
    rho, vf, sie = np.array(rho), np.array(vf), np.array(sie)
    cell = np.repeat(np.arange(len(counts)), counts)
    rb = rho * vf
    sie_tot = np.bincount(cell, weights=rb * sie) / np.bincount(cell, weights=rb)
    A = dict(rho=rho, vfrac=vf, sie=sie, temp=np.array(T), press=np.array(P),
             vfrac_tot=np.ones(len(counts)), sie_tot=sie_tot,
             offsets=np.array(off, dtype=np.int32), mat_ids=np.array(mats, dtype=np.uint8),
             ncell=len(counts), nnz=off[-1])
    params = sgm.mix_params()
    g = run_plain(torch_gpu, ctx, A, 4, params)
    r = run_oracle(oracle, o, A, 4, params.as_oracle_array())
    assert_identical(g, r, "rhou odd counts")


@pytest.mark.parametrize("tight", [False, True])
def test_rhou_batched_equals_plain_and_oracle(sgm, workloads, oracle, pool, torch_gpu, tight):
    """Batched plans run RhoU in register-resident bins for 2..4 materials
    and on the list kernel for 5 and 6; with straggler parking at several
    thresholds (every cell parked after 1 iteration, the default 20, off)
    every output, flag and iteration count equals the plain per-call path
    and the oracle bit for bit."""
    torch = torch_gpu
    ctx, o, _ = pool
    a = workloads.uneven(ctx, 4000, seed=0x5EED0000 + 63 + int(tight))
    params = sgm.mix_params(tight=tight)
    g = run_plain(torch, ctx, a, 4, params)
    r = run_oracle(oracle, o, a, 4, params.as_oracle_array())
    assert_identical(g, r, "rhou plain")
    from test_gpu_pte import FIELDS, dev
    d = dev(torch, a)
    plan = sgm.Plan(ctx, a["ncell"], a["nnz"], d["offsets"], d["mat_ids"], 4)
    bins = plan.describe()
    assert sorted(b["nmat"] for b in bins if b["nmat"] >= 2) == [2, 3, 4]
    assert sum(b["ncells"] for b in bins) == a["ncell"]
    for after in (1, 20, 0):
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
        assert_identical(out, g, f"rhou batched after={after}")
        assert np.array_equal(out["err"], g["err"])
    plan.close()
