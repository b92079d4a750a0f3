"""Pins the oracle's Vinet with nonzero expansion coefficients d2..d40
(singularity-eos/eos/eos_vinet.hpp:199-262: the energy coefficients f_n
from the d_n by the recurrence of InitializeVinet, and the three Horner sums
of Vinet_F_DT_func), which the reference's tests only run with all d_n = 0.

What the mathematics fixes, for any coefficients:
  * (de/drho)_T = (P - T A0 B0) / rho^2 (the energy is the integral of the
    pressure along an isotherm; e and P use different Horner sums, f_n and
    d_n, so a wrong f_n recurrence breaks it);
  * rho (dP/drho)_T = B_S - T (A0 B0)^2 / (rho Cv0), with B_S the bulk
    modulus the model returns (its own sum over n d_n);
  * e(rho, T) - e(rho0, T) = integral_rho0^rho (P(r, T) - T A0 B0) / r^2 dr.
Central differences in rho (h = 1e-4 rho, truncation ~1e-9 relative) and an
adaptive quadrature check them to 1e-6 relative. A planted mistake (the
pressure of a Vinet whose d2 is 10 % off, on the CPU) fails the first.
"""
import numpy as np
import pytest

VINET_CU = (8.93, 298.0, 1.37e12, 5.0, 5.19245e-5, 3.83e6, 0.0, 5.05e8)
# two coefficient sets: a few low orders (the device skips the zero high
# orders of its Horner loops) and every order nonzero (no skip). The
# recurrence f_n = d_n - (n / eta) f_(n+1) grows like n! / eta^n, so with
# slowly decaying d_n the energy sum cancels catastrophically (in the
# reference's formula as here); the second set decays geometrically, which
# keeps every f_n below 2.
D_LOW = [0.5, -0.3, 0.2, 0.1, -0.05] + [0.0] * 34
D_ALL = [(-1.0) ** n * 0.3 / n * 0.25 ** (n - 2) for n in range(2, 41)]
F_P_RT, F_E_RT, F_B_RT = 3, 2, 7


@pytest.mark.parametrize("d", [D_LOW, D_ALL], ids=["low", "all"])
def test_vinet_coeffs_thermodynamic_identities(oracle, d):
    eos = oracle.vinet(*VINET_CU, d2to40=d)
    rho0, _, B0, _, A0, Cv0 = VINET_CU[:6]
    rng = np.random.default_rng(3)
    rho = rng.uniform(0.7 * rho0, 1.5 * rho0, 400)
    T = rng.uniform(100.0, 5000.0, 400)
    h = 1e-4 * rho
    e_p = eos.vec(F_E_RT, rho + h, T)
    e_m = eos.vec(F_E_RT, rho - h, T)
    P = eos.vec(F_P_RT, rho, T)
    dedr = (e_p - e_m) / (2 * h)
    ident = (P - T * A0 * B0) / rho ** 2
    scale = np.abs(ident) + 1e-3 * B0 / rho0 ** 2
    assert np.max(np.abs(dedr - ident) / scale) < 1e-6
    # planted mistake: the same identity against a Vinet with d2 10 % off
    d_bad = list(d)
    d_bad[0] *= 1.1
    P_bad = oracle.vinet(*VINET_CU, d2to40=d_bad).vec(F_P_RT, rho, T)
    bad = (P_bad - T * A0 * B0) / rho ** 2
    assert np.max(np.abs(dedr - bad) / scale) > 1e-5
    # bulk modulus against the pressure's own density derivative
    dPdr = (eos.vec(F_P_RT, rho + h, T) - eos.vec(F_P_RT, rho - h, T)) / (2 * h)
    BS = eos.vec(F_B_RT, rho, T)
    BT = BS - T * (A0 * B0) ** 2 / (rho * Cv0)
    assert np.max(np.abs(rho * dPdr - BT) / (np.abs(BT) + 1e-3 * B0)) < 1e-6


@pytest.mark.parametrize("d", [D_LOW, D_ALL], ids=["low", "all"])
def test_vinet_coeffs_energy_is_pressure_integral(oracle, d):
    from scipy.integrate import quad
    eos = oracle.vinet(*VINET_CU, d2to40=d)
    rho0, _, B0, _, A0 = VINET_CU[:5]
    for rho, T in ((6.5, 300.0), (8.0, 1200.0), (11.0, 298.0), (13.0, 4000.0)):
        f = lambda r: (eos.vec(F_P_RT, np.array([r]), np.array([T]))[0] - T * A0 * B0) / r ** 2
        integral, _ = quad(f, rho0, rho, epsabs=0.0, epsrel=1e-12, limit=200)
        de = (eos.vec(F_E_RT, np.array([rho]), np.array([T]))[0]
              - eos.vec(F_E_RT, np.array([rho0]), np.array([T]))[0])
        assert abs(de - integral) <= 1e-6 * (abs(integral) + 1e-3 * B0 / rho0), (rho, T)


def test_vinet_zero_coeffs_reduce_to_plain_vinet(oracle):
    """With every d_n = 0 the model is the plain Vinet curve, whose goldens
    (test_eos_vinet.cpp) pin the oracle; an explicit zero list and the
    default agree bit for bit."""
    a = oracle.vinet(*VINET_CU)
    b = oracle.vinet(*VINET_CU, d2to40=[0.0] * 39)
    rho = np.linspace(5.0, 14.0, 101)
    T = np.linspace(100.0, 6000.0, 101)
    for f in (F_P_RT, F_E_RT, F_B_RT):
        assert np.array_equal(a.vec(f, rho, T), b.vec(f, rho, T))
