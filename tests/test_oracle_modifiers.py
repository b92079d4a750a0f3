"""Pins the oracle's ShiftedEOS / ScaledEOS (singularity-eos/eos/modifiers/
shifted_eos.hpp, scaled_eos.hpp, composed ScaledEOS<ShiftedEOS<model>> as
singularity_eos_init_utils.hpp:25-41 builds them):

  * the reference's golden: ScaledEOS(ShiftedEOS(IdealGas(gm1=0.5, Cv=2), 0.1),
    2) gives P(rho=2, sie=0.5) == 0.3 exactly (test_eos_modifiers.cpp:125-140;
    the other order of the two modifiers would give 0.4);
  * compare_two_eoss (test_eos_modifiers.cpp:218-227,
    eos_unit_test_helpers.hpp:107-160): a scale of 1 or a shift of 0 leaves
    every function of the model unchanged;
  * bounds (test_eos_modifiers.cpp:334-359): minimum and maximum density and
    RhoPmin divide by the scale, the other bounds are unchanged;
  * the closure: shifting each material's energy by s_m and the cell's
    specific energy by sum y_m s_m (y the mass fractions) leaves the
    equilibrium unchanged, and ScaledEOS(IdealGas(g, Cv), c) is IdealGas(g,
    c Cv), so closures over them agree with closures over the plain models
    (to 1e-9 at the 1e-10 tolerance, as rounding differs).
"""
import numpy as np
import pytest

F = dict(P_RE=0, T_RE=1, E_RT=2, P_RT=3, CV_RE=4, CV_RT=5, B_RE=6, B_RT=7, MIN_E=8)


def test_scaled_shifted_ideal_gas_golden(oracle):
    e = oracle.ideal(0.5, 2.0).modify(scaled=True, shifted=True, scale=2.0, shift=0.1)
    assert e.vec(F["P_RE"], np.array([2.0]), np.array([0.5]))[0] == 0.3


def test_identity_modifiers_change_nothing(oracle, workloads):
    wl = workloads
    makers = [lambda: oracle.ideal(*wl.IDEAL_A), lambda: oracle.stiff(*wl.STIFF_B),
              lambda: oracle.gruneisen(*wl.GRUN_C1), lambda: oracle.vinet(*wl.VINET_CU)]
    rng = np.random.default_rng(2)
    for mk in makers:
        ref = mk()
        for mod in (dict(scaled=True, scale=1.0), dict(shifted=True, shift=0.0),
                    dict(scaled=True, shifted=True, scale=1.0, shift=0.0)):
            e = mk().modify(**mod)
            for rho, x in ((np.ones(1), np.ones(1)),
                           (rng.uniform(0.5, 12, 500), rng.uniform(100, 5000, 500))):
                for f in F.values():
                    a, b = e.vec(f, rho, x), ref.vec(f, rho, x)
                    assert np.array_equal(a, b, equal_nan=True), (mod, f)


def test_shift_and_scale_relations(oracle, workloads):
    """Definitions the golden above fixes the order of, on every model:
    e = c (e_model(rho, T) + s), P(rho, e) = P_model(c rho, e / c - s),
    T(rho, e) = T_model(c rho, e / c - s), P(rho, T) = P_model(c rho, T)."""
    wl = workloads
    rng = np.random.default_rng(4)
    c, s = 1.7, 2.5e8
    for mk, rr in ((lambda: oracle.stiff(*wl.STIFF_B), (0.8, 1.2)),
                   (lambda: oracle.gruneisen(*wl.GRUN_C1), (1.5, 4.0)),
                   (lambda: oracle.vinet(*wl.VINET_CU), (5.0, 12.0))):
        base = mk()
        e = mk().modify(scaled=True, shifted=True, scale=c, shift=s)
        rho = rng.uniform(*rr, 300)
        T = rng.uniform(200, 4000, 300)
        assert np.array_equal(e.vec(F["E_RT"], rho, T), c * (base.vec(F["E_RT"], rho, T) + s))
        assert np.array_equal(e.vec(F["P_RT"], rho, T), base.vec(F["P_RT"], c * rho, T))
        sie = e.vec(F["E_RT"], rho, T)
        assert np.array_equal(e.vec(F["P_RE"], rho, sie),
                              base.vec(F["P_RE"], c * rho, (1.0 / c) * sie - s))
        assert np.array_equal(e.vec(F["T_RE"], rho, sie),
                              base.vec(F["T_RE"], c * rho, (1.0 / c) * sie - s))


def closure(oracle, eos, a, tight=True):
    r = {k: a[k].copy() for k in ("rho", "vfrac", "sie", "temp", "press")}
    p = oracle.tight_params() if tight else oracle.default_params()
    nf, conv, nit, res, err = oracle.pte_solve(0, eos, a["offsets"], a["mat_ids"].astype(np.int32),
                                               r["rho"], r["vfrac"], r["sie"], r["temp"],
                                               r["press"], a["vfrac_tot"], a["sie_tot"], p)
    r["conv"] = conv
    return r


def test_closure_shift_invariance(oracle, workloads):
    wl = workloads
    a = wl.pte2(2000)
    a["mat_ids"] = np.tile(np.array([0, 1], dtype=np.uint8), 2000)
    plain = [oracle.ideal(*wl.IDEAL_A), oracle.stiff(*wl.STIFF_B)]
    shifts = (3.0e9, -1.0e9)
    mod = [oracle.ideal(*wl.IDEAL_A).modify(shifted=True, shift=shifts[0]),
           oracle.stiff(*wl.STIFF_B).modify(shifted=True, shift=shifts[1])]
    rb = (a["rho"] * a["vfrac"]).reshape(-1, 2)
    y = rb / rb.sum(1, keepdims=True)
    b = dict(a)
    b["sie_tot"] = a["sie_tot"] + y[:, 0] * shifts[0] + y[:, 1] * shifts[1]
    # the pressure-preferring ideal gas starts from the input pressures, which
    # the shift does not change; the stiffened gas's energies carry the shift
    r0 = closure(oracle, plain, a)
    r1 = closure(oracle, mod, b)
    assert r0["conv"].all() and r1["conv"].all()
    for k in ("rho", "vfrac", "temp", "press"):
        rel = np.abs(r1[k] - r0[k]) / np.abs(r0[k])
        assert rel.max() < 1e-9, (k, rel.max())
    sh = np.tile(np.array(shifts), 2000)
    assert np.max(np.abs(r1["sie"] - sh - r0["sie"]) / np.abs(r0["sie"] + 1e8)) < 1e-9


def test_closure_scaled_ideal_gas_is_ideal_gas_with_scaled_cv(oracle):
    rng = np.random.default_rng(6)
    n = 2000
    g = [0.4, 2.0 / 3.0, 0.3]
    cv = [7.176e6, 3.1e7, 1.2e7]
    c = [1.5, 0.7, 2.0]
    al = rng.exponential(size=(n, 3)) + 0.05
    al /= al.sum(1, keepdims=True)
    rho = np.exp(rng.normal(np.log(1e-2), 1.0, (n, 3)))
    T = rng.uniform(300, 3000, (n, 3))
    e = np.array(cv)[None, :] * np.array(c)[None, :] * T
    P = np.array(g)[None, :] * rho * e
    rb = rho * al
    a = dict(rho=rho.ravel(), vfrac=al.ravel(), sie=e.ravel(), temp=T.ravel(), press=P.ravel(),
             vfrac_tot=np.ones(n), sie_tot=(rb * e).sum(1) / rb.sum(1),
             offsets=np.arange(0, 3 * n + 1, 3, dtype=np.int32),
             mat_ids=np.tile(np.arange(3, dtype=np.uint8), n))
    scaled = [oracle.ideal(gi, ci).modify(scaled=True, scale=si) for gi, ci, si in zip(g, cv, c)]
    equiv = [oracle.ideal(gi, ci * si) for gi, ci, si in zip(g, cv, c)]
    r0 = closure(oracle, equiv, a)
    r1 = closure(oracle, scaled, a)
    assert r0["conv"].all() and r1["conv"].all()
    for k in ("rho", "vfrac", "sie", "temp", "press"):
        rel = np.abs(r1[k] - r0[k]) / np.abs(r0[k])
        assert rel.max() < 1e-9, (k, rel.max())
