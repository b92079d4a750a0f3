"""Pins the oracle's fixed-temperature and fixed-pressure closures
(PTESolverFixedT / PTESolverFixedP, singularity-eos/closure/mixed_cell_models.hpp
:1451-1921, restated in oracle/sgo_pte.hpp) on the CPU against what the
mathematics fixes: closed-form equilibria of ideal and stiffened gases, and
agreement with the independently pinned PTESolverRhoT on the uneven
2..6-material workload (FixedT at the RhoT temperature, FixedP at the RhoT
pressure must land on the same equilibrium). A planted input error shows the
agreement check is sharp enough to notice a 1e-6 shift."""
import numpy as np
import pytest

FIELDS = ("rho", "vfrac", "sie", "temp", "press")
GM1 = [0.4, 2.0 / 3.0, 0.3, 1.1, 0.55]
CV = [7.176e6, 3.1e7, 1.2e7, 5.0e6, 2.2e7]


def solve(oracle, method, eos, offsets, mat, a, cell_in, params):
    r = {k: np.array(a[k], dtype=np.float64) for k in FIELDS}
    nf, conv, nit, res, err = oracle.pte_solve(method, eos, offsets, mat, r["rho"], r["vfrac"],
                                               r["sie"], r["temp"], r["press"], a["vfrac_tot"],
                                               np.ascontiguousarray(cell_in, dtype=np.float64),
                                               params)
    r.update(conv=conv, niter=nit, err=err)
    return r


def gas_cells(n, k, seed):
    rng = np.random.default_rng(seed)
    al = rng.exponential(size=(n, k)) + 0.05
    al /= al.sum(1, keepdims=True)
    rho = np.exp(rng.normal(np.log(1e-2), 1.0, (n, k)))
    T = rng.uniform(300, 3000, (n, k))
    a = dict(rho=rho.ravel(), vfrac=al.ravel(), sie=np.zeros(n * k), temp=T.ravel(),
             press=np.zeros(n * k), vfrac_tot=np.ones(n))
    return a, rho * al


def test_fixedT_ideal_gases_closed_form(oracle):
    """N ideal gases at a given T: P = T sum(rb (g-1) Cv) / V and
    alpha_m = rb_m (g_m-1) Cv_m T / P."""
    n, k = 400, len(GM1)
    a, rb = gas_cells(n, k, 11)
    T = np.random.default_rng(12).uniform(200, 5000, n)
    eos = [oracle.ideal(g, c) for g, c in zip(GM1, CV)]
    off = np.arange(0, k * n + 1, k, dtype=np.int32)
    mat = np.tile(np.arange(k, dtype=np.int32), n)
    r = solve(oracle, 2, eos, off, mat, a, T, oracle.tight_params())
    assert r["conv"].all()
    g = rb * np.array(GM1) * np.array(CV)
    P = T * g.sum(1)
    al = g * T[:, None] / P[:, None]
    assert np.max(np.abs(r["press"].reshape(n, k) - P[:, None]) / P[:, None]) < 1e-9
    assert np.max(np.abs(r["vfrac"].reshape(n, k) - al) / al) < 1e-9
    assert np.max(np.abs(r["rho"] * r["vfrac"] - rb.ravel()) / rb.ravel()) < 1e-14
    # e_m = Cv_m T, and the temperature array is left as it was
    assert np.max(np.abs(r["sie"].reshape(n, k) - np.array(CV) * T[:, None])
                  / (np.array(CV) * T[:, None])) < 1e-15
    assert np.array_equal(r["temp"], a["temp"])


@pytest.mark.parametrize("stiff,eight", [(False, False), (True, False), (True, True)])
def test_fixedP_closed_form(oracle, stiff, eight):
    """Ideal gases (+ a stiffened gas) at a given P: alpha_m =
    rb_m (g_m-1) Cv_m T / (P + Pinf_m), so T = V / sum(rb (g-1) Cv / (P + Pinf)).
    FixedP compares the absolute pressure error with the tolerances, as the
    reference does, so the default MixParams are used. `eight`: 8 materials,
    a 9 x 9 Newton system (the largest)."""
    n = 400
    gm1, cv, pinf = list(GM1), list(CV), [0.0] * len(GM1)
    if eight:
        gm1, cv, pinf = gm1 + [0.25, 0.8], cv + [9e6, 1.5e7], pinf + [0.0, 0.0]
    eos = [oracle.ideal(g, c) for g, c in zip(gm1, cv)]
    if stiff:
        gm1.append(3.4)
        cv.append(3.6e7)
        pinf.append(6e9)
        eos.append(oracle.stiff(3.4, 3.6e7, 6e9, 0.0))
    k = len(eos)
    a, rb = gas_cells(n, k, 13)
    P = np.exp(np.random.default_rng(14).uniform(np.log(1e8), np.log(1e11), n))
    a["press"] = np.repeat(P, k)
    off = np.arange(0, k * n + 1, k, dtype=np.int32)
    mat = np.tile(np.arange(k, dtype=np.int32), n)
    r = solve(oracle, 3, eos, off, mat, a, P, oracle.default_params())
    assert r["conv"].mean() > 0.99
    c = r["conv"].astype(bool)
    w = rb * np.array(gm1) * np.array(cv) / (P[:, None] + np.array(pinf))
    T = 1.0 / w.sum(1)
    al = w * T[:, None]
    rel = lambda x, y: np.abs(x - y) / np.abs(y)
    assert np.max(rel(r["temp"].reshape(n, k), T[:, None])[c]) < 1e-12
    assert np.max(rel(r["vfrac"].reshape(n, k), al)[c]) < 1e-12
    assert np.max(rel(r["press"].reshape(n, k), P[:, None])[c]) < 1e-12


@pytest.fixture(scope="module")
def c2eq(oracle, sgm, workloads):
    """The uneven workload and its PTESolverRhoT equilibrium (1e-10)."""
    wl = workloads
    ctx, t = wl.pool_context(host_only=True)
    a = wl.uneven(ctx, 2048, seed=0x5EED0000 + 31)
    otab, _, _ = oracle.gruneisen_table(wl.GRUN_C1, t.lRhoMin, t.lRhoMax, t.nRho, t.lTMin,
                                        t.lTMax, t.nT)
    eos = [oracle.ideal(*wl.IDEAL_A), oracle.ideal(*wl.IDEAL_M1), oracle.stiff(*wl.STIFF_B),
           oracle.gruneisen(*wl.GRUN_C1), oracle.vinet(*wl.VINET_CU), otab.eos()]
    mat = a["mat_ids"].astype(np.int32)
    r0 = solve(oracle, 0, eos, a["offsets"], mat, a, a["sie_tot"],
               sgm.mix_params(tight=True).as_oracle_array())
    yield a, eos, mat, r0
    ctx.close()


def agreement(a, r0, r, keys):
    cell = np.repeat(np.arange(a["ncell"]), np.diff(a["offsets"]))
    both = (r0["conv"] & r["conv"]).astype(bool)
    out = {}
    for k in keys:
        rel = np.abs(r[k] - r0[k]) / np.abs(r0[k])
        out[k] = float(np.percentile(rel[both[cell]], 99))
    return both.mean(), out


def test_fixedT_lands_on_the_rhoT_equilibrium(oracle, sgm, c2eq):
    a, eos, mat, r0 = c2eq
    T_star = r0["temp"][a["offsets"][:-1]]
    r = solve(oracle, 2, eos, a["offsets"], mat, a, T_star,
              sgm.mix_params(tight=True).as_oracle_array())
    frac, d = agreement(a, r0, r, ("press", "vfrac", "sie"))
    assert frac > 0.9
    assert max(d.values()) < 1e-7, d
    # planted: the same check at a temperature 1e-6 off sees the shift
    r_bad = solve(oracle, 2, eos, a["offsets"], mat, a, T_star * (1 + 1e-6),
                  sgm.mix_params(tight=True).as_oracle_array())
    _, d_bad = agreement(a, r0, r_bad, ("press",))
    assert d_bad["press"] > 1e-7


def test_fixedP_lands_on_the_rhoT_equilibrium(oracle, sgm, c2eq):
    """FixedP at the RhoT pressure returns the RhoT temperature and volume
    fractions where the temperature is well conditioned: with a gas holding
    vfrac > 0.3, a volume error dV moves T by about dV / vfrac_gas. (Cells of
    condensed materials only leave T nearly free within the 1e-6 volume
    tolerance, and the reference's FixedP convergence test compares the
    absolute pressure error with 1e-6, so many cells run to the iteration
    cap while already at the equilibrium.)"""
    a, eos, mat, r0 = c2eq
    off = a["offsets"]
    cell = np.repeat(np.arange(a["ncell"]), np.diff(off))
    P_star = r0["press"][off[:-1]]
    start = dict(a)
    start["press"] = np.repeat(P_star, np.diff(off))
    r = solve(oracle, 3, eos, off, mat, start, P_star, oracle.default_params())
    gas = np.bincount(cell, weights=np.where(mat <= 1, r0["vfrac"], 0.0), minlength=a["ncell"])
    well = (r0["conv"] & r["conv"]).astype(bool) & (gas > 0.3)
    assert well.sum() > 200
    T0, T1 = r0["temp"][off[:-1]], r["temp"][off[:-1]]
    assert np.max(np.abs(T1 - T0)[well] / T0[well]) < 1e-9
    rv = np.abs(r["vfrac"] - r0["vfrac"]) / r0["vfrac"]
    assert np.percentile(rv[well[cell]], 99) < 1e-9
    vsum = np.bincount(cell, weights=r["vfrac"], minlength=a["ncell"])
    assert np.max(np.abs(vsum - 1)[r["conv"].astype(bool)]) < 1e-12
    # planted: a pressure 1e-5 off moves those temperatures visibly
    r_bad = solve(oracle, 3, eos, off, mat, start, P_star * (1 + 1e-5), oracle.default_params())
    Tb = r_bad["temp"][off[:-1]]
    assert np.median(np.abs(Tb - T0)[well] / T0[well]) > 1e-7
