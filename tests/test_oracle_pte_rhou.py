"""Pins the oracle's PTESolverRhoU (oracle/sgo_pte.hpp, restating
singularity-eos/closure/mixed_cell_models.hpp:1924-2183 and the base's
GetIdealPTE / TryIdealPTE, :672-763) on the CPU.

RhoU solves for every material's volume fraction and energy density, with
temperatures from T(rho, e), so the equilibrium is the one RhoT finds:
* ideal-gas mixtures: one temperature T = e_tot / sum(x Cv), every material
  at the common pressure P = T sum(rhobar gm1 Cv) / vfrac_tot (closed form);
* an ideal gas with a stiffened gas: the same state as PTESolverRhoT on the
  same cells;
* every solve conserves volume and energy (sum vfrac = vfrac_tot,
  sum rhobar e = rho_tot e_tot).
A planted mistake (a perturbed temperature of one material) fails the
equal-temperature check, so the closed-form test can see a wrong answer."""
import numpy as np
import pytest

GM1 = np.array([0.4, 2.0 / 3.0, 0.3, 1.1])
CV = np.array([7.176e6, 3.1e7, 1.2e7, 5.0e6])


def ideal_cells(n, k, seed):
    rng = np.random.default_rng(seed)
    alpha = rng.uniform(0.1, 1.0, (n, k))
    alpha /= alpha.sum(1, keepdims=True)
    rho = 10 ** rng.uniform(-3, -1, (n, k))
    T = rng.uniform(300, 3000, (n, k))
    sie = CV[:k] * T
    press = GM1[:k] * rho * sie
    return alpha, rho, T, sie, press


def solve(oracle, method, eos, alpha, rho, T, sie, press, params, sie_tot=None):
    n, k = alpha.shape
    offs = np.arange(0, n * k + 1, k).astype(np.int32)
    me = np.tile(np.arange(k), n).astype(np.int32)
    rb = rho * alpha
    if sie_tot is None:
        sie_tot = (rb * sie).sum(1) / rb.sum(1)
    arr = dict(rho=rho.ravel().copy(), vfrac=alpha.ravel().copy(), sie=sie.ravel().copy(),
               temp=T.ravel().copy(), press=press.ravel().copy())
    nf, conv, nit, res, err = oracle.pte_solve(method, eos, offs, me, arr["rho"], arr["vfrac"],
                                               arr["sie"], arr["temp"], arr["press"],
                                               np.ones(n), sie_tot, params)
    return {kk: v.reshape(n, k) for kk, v in arr.items()}, conv, nit, err, sie_tot


@pytest.mark.parametrize("k", [1, 2, 3, 4])
def test_rhou_ideal_gas_closed_form(oracle, k):
    eos = [oracle.ideal(g, c) for g, c in zip(GM1[:k], CV[:k])]
    alpha, rho, T, sie, press = ideal_cells(300, k, 10 + k)
    rb = rho * alpha
    out, conv, nit, err, sie_tot = solve(oracle, 4, eos, alpha, rho, T, sie, press,
                                         oracle.tight_params())
    assert conv.all() and (err == 0).all()
    Teq = sie_tot * rb.sum(1) / (rb * CV[:k]).sum(1)
    Peq = (rb * GM1[:k] * CV[:k]).sum(1) * Teq
    assert np.max(np.abs(out["temp"] - Teq[:, None]) / Teq[:, None]) < 1e-9
    assert np.max(np.abs(out["press"] - Peq[:, None]) / Peq[:, None]) < 1e-9
    assert np.max(np.abs(out["vfrac"].sum(1) - 1.0)) < 1e-12
    e_out = (out["rho"] * out["vfrac"] * out["sie"]).sum(1)
    assert np.max(np.abs(e_out - rb.sum(1) * sie_tot) / (rb.sum(1) * sie_tot)) < 1e-12
    # a planted mistake is seen
    bad = out["temp"].copy()
    bad[0, 0] *= 1 + 1e-6
    assert np.max(np.abs(bad - Teq[:, None]) / Teq[:, None]) > 1e-9


def test_rhou_matches_rhot_with_stiffened_gas(oracle):
    """IdealGas + StiffGas: RhoU and RhoT reach the same equilibrium."""
    eos = [oracle.ideal(0.4, 7.176e6), oracle.stiff(3.4, 3.6e7, 6e9, 0.0)]
    rng = np.random.default_rng(3)
    n = 400
    a0 = rng.uniform(0.05, 0.95, n)
    alpha = np.stack([a0, 1 - a0], 1)
    rho = np.stack([np.exp(np.log(1e-3) + 0.3 * rng.standard_normal(n)),
                    rng.uniform(0.9, 1.1, n)], 1)
    T = rng.uniform(300, 3000, (n, 2))
    sie = np.stack([7.176e6 * T[:, 0], 3.6e7 * T[:, 1] + 6e9 / rho[:, 1]], 1)
    press = np.stack([0.4 * rho[:, 0] * sie[:, 0], 3.4 * rho[:, 1] * sie[:, 1] - 4.4 * 6e9], 1)
    p = oracle.tight_params()
    u, cu, _, _, st = solve(oracle, 4, eos, alpha, rho, T, sie, press, p)
    t, ct, _, _, _ = solve(oracle, 0, eos, alpha, rho, T, sie, press, p, sie_tot=st)
    ok = cu.astype(bool) & ct.astype(bool)
    assert ok.mean() > 0.95
    for f in ("temp", "press", "vfrac", "rho"):
        rel = np.abs(u[f][ok] - t[f][ok]) / np.abs(t[f][ok])
        assert np.max(rel) < 1e-7, f
