"""GPU parity of the fixed-temperature and fixed-pressure closures
(sgm_pte_fixedT / sgm_pte_fixedP, and plans of method 2 / 3) against the
oracle's PTESolverFixedT / PTESolverFixedP, through the C ABI.

Bar: bit for bit -- states, convergence flags, iteration counts and error
bits -- as for the energy-closed formulations. The uneven 2..6-material
workload supplies the cells; the given temperature is each cell's first input
temperature and the given pressure its volume-weighted input pressure, so
cells start away from equilibrium. The fixed-temperature closure writes no
temperatures: the test checks the array comes back untouched.
"""
import numpy as np
import pytest

from test_gpu_pte import FIELDS, assert_identical, pool, run_batched, run_oracle, run_plain  # noqa: F401

pytestmark = pytest.mark.gpu


def fixed_inputs(workloads, ctx, method, n, seed):
    a = workloads.uneven(ctx, n, seed=seed)
    off = a["offsets"]
    if method == 2:
        a["sie_tot"] = a["temp"][off[:-1]].copy()
    else:
        cell = np.repeat(np.arange(a["ncell"]), np.diff(off))
        P = np.bincount(cell, weights=a["vfrac"] * a["press"], minlength=a["ncell"])
        a["sie_tot"] = np.abs(P) + 1e5
    return a


@pytest.mark.parametrize("method,tight", [(2, False), (2, True), (3, False)])
def test_fixed_plain_and_batched_bit_exact(sgm, workloads, oracle, pool, torch_gpu, method,
                                           tight):
    ctx, o, _ = pool
    a = fixed_inputs(workloads, ctx, method, 8192, 0x5EED0000 + 50 + method)
    params = sgm.mix_params(tight=tight)
    g = run_plain(torch_gpu, ctx, a, method, params)
    r = run_oracle(oracle, o, a, method, params.as_oracle_array())
    assert_identical(g, r, f"fixed method={method} tight={tight}")
    assert np.array_equal(g["err"], r["err"])
    if method == 2:
        assert np.array_equal(g["temp"], a["temp"])
    b, info = run_batched(torch_gpu, sgm, ctx, a, method, params)
    assert {i["nmat"] for i in info} >= {2, 3, 4, 5, 6}
    assert_identical(b, g, f"fixed method={method} batched vs plain")
    assert g["conv"].mean() > (0.9 if method == 2 else 0.3)


@pytest.mark.parametrize("method", [2, 3])
def test_fixed_odd_material_counts(sgm, oracle, pool, torch_gpu, method):
    """1, 7 and 8 materials (the generic kernel) next to 2 and 3."""
    ctx, o, _ = pool
    rng = np.random.default_rng(7 + method)
    counts = [1, 7, 8, 2, 3, 1, 8, 7]
    mats, rho, vf, T, P, cin, off = [], [], [], [], [], [], [0]
    for nm in counts:
        ids = rng.choice([0, 1, 2, 3, 4, 5], size=nm, replace=True)
        al = rng.exponential(size=nm) + 1e-3
        al /= al.sum()
        rr = np.where(ids <= 1, 1e-3, np.where(ids == 2, 1.0, np.where(ids == 4, 8.9, 2.7)))
        rr = rr * rng.uniform(0.95, 1.05, nm)
        tt = rng.uniform(300, 3000, nm)
        pp = np.array([o[i].vec(3, [r_], [t_])[0] for i, r_, t_ in zip(ids, rr, tt)])
        mats += list(ids)
        rho += list(rr)
        vf += list(al)
        T += list(tt)
        P += list(pp)
        cin.append(float(tt[0]) if method == 2 else float(abs((al * pp).sum()) + 1e5))
        off.append(off[-1] + nm)
    A = dict(rho=np.array(rho), vfrac=np.array(vf), sie=np.zeros(len(rho)), temp=np.array(T),
             press=np.array(P), vfrac_tot=np.ones(len(counts)), sie_tot=np.array(cin),
             offsets=np.array(off, dtype=np.int32), mat_ids=np.array(mats, dtype=np.uint8),
             ncell=len(counts), nnz=off[-1])
    params = sgm.mix_params()
    g = run_plain(torch_gpu, ctx, A, method, params)
    r = run_oracle(oracle, o, A, method, params.as_oracle_array())
    assert_identical(g, r, f"odd counts method={method}")
    b, _ = run_batched(torch_gpu, sgm, ctx, A, method, params)
    assert_identical(b, g, f"odd counts batched method={method}")


def test_fixed_full_size_properties(sgm, workloads, pool, torch_gpu):
    """2^20 uneven cells, fixed temperature, batched path at 1e-10: volume
    fractions sum to one, partial densities are kept, the material pressures
    agree, and batched == plain."""
    torch = torch_gpu
    ctx, _, _ = pool
    a = fixed_inputs(workloads, ctx, 2, 1 << 20, 0x5EED0000 + 2)
    params = sgm.mix_params(tight=True)
    b, _ = run_batched(torch, sgm, ctx, a, 2, params)
    g = run_plain(torch, ctx, a, 2, params)
    assert_identical(b, g, "fixedT full size batched vs plain")
    off = a["offsets"]
    conv = b["conv"].astype(bool)
    assert conv.mean() > 0.9
    cell = np.repeat(np.arange(a["ncell"]), np.diff(off))
    vsum = np.bincount(cell, weights=b["vfrac"])
    assert np.max(np.abs(vsum[conv] - 1.0)) < 1e-12
    rb = a["rho"] * a["vfrac"]
    assert np.max(np.abs(b["rho"] * b["vfrac"] - rb) / rb) < 1e-12
    Pm = np.bincount(cell, weights=b["press"] * b["vfrac"])
    rel = np.abs(b["press"] - Pm[cell]) / np.abs(Pm[cell])
    assert np.max(rel[conv[cell]]) < 1e-9
    assert np.array_equal(b["temp"], a["temp"])


def test_guided_execute_rejects_fixed_methods(sgm, workloads, pool, torch_gpu):
    torch = torch_gpu
    ctx, _, _ = pool
    a = fixed_inputs(workloads, ctx, 2, 256, 0x5EED0000 + 60)
    d = {k: torch.tensor(a[k], device="cuda") for k in FIELDS + ("vfrac_tot", "sie_tot")}
    offs = torch.tensor(a["offsets"], device="cuda", dtype=torch.int32)
    mats = torch.tensor(a["mat_ids"], device="cuda", dtype=torch.uint8)
    plan = sgm.Plan(ctx, a["ncell"], a["nnz"], offs, mats, 2)
    gT = torch.full((a["ncell"],), 1000.0, device="cuda", dtype=torch.float64)
    rc = plan.execute_guided(d["rho"], d["vfrac"], d["sie"], d["temp"], d["press"],
                             d["vfrac_tot"], d["sie_tot"], d["vfrac"].clone(), gT)
    assert rc == -1
    plan.close()
