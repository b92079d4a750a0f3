"""Pins the CPU oracle before it is trusted as the parity checker.

Every function of the oracle is checked against (a) the known-answer values of
singularity-eos's own tests (tests/golden/reference_goldens.json, sources
cited there), or (b) what the mathematics fixes: closed forms, invariants and
brute force on small inputs. These tests do not use the GPU.
"""
import json
import math
import os

import numpy as np
import pytest

GOLD = json.load(open(os.path.join(os.path.dirname(__file__), "golden",
                                   "reference_goldens.json")))

F_P_RE, F_T_RE, F_E_RT, F_P_RT, F_CV_RE, F_CV_RT, F_B_RE, F_B_RT, F_MINE = range(9)
Q_DENSITY, Q_SIE, Q_PRESSURE, Q_TEMPERATURE, Q_CV, Q_BMOD = 1, 2, 4, 8, 16, 32


def is_close(a, b, eps):
    """test/eos_unit_test_helpers.hpp:61-63"""
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    return np.abs(b - a) / (np.abs(a + b) + 1e-20) <= eps


# ---------------------------------------------------------------- fast math
def test_nqt_constants(oracle):
    import ctypes as C
    ints = (C.c_longlong * 3)()
    dbls = (C.c_double * 2)()
    oracle.lib().sgo_nqt_constants(ints, dbls)
    g = GOLD["nqt_constants"]
    assert ints[0] == g["one_as_int"]
    assert ints[1] == g["mantissa_mask"]
    assert ints[2] == g["low_mask"]
    assert dbls[0] == g["scale_down"]
    assert dbls[1] == g["scale_up"]


def test_nqt_roundtrip(oracle):
    g = GOLD["nqt_roundtrip"]
    lx = g["lxmin"] + np.arange(g["nx"]) * ((g["lxmax"] - g["lxmin"]) / (g["nx"] - 1))
    x = 10.0 ** lx
    y = oracle.nqt(3, oracle.nqt(2, x))
    eps = np.finfo(np.float64).eps
    rel = 2.0 * np.abs(x - y) / (np.abs(x) + np.abs(y) + eps)
    assert np.all(rel <= g["acceptable_err_eps"] * eps)


def test_nqt_log_is_close_to_log2_and_exact_at_powers_of_two(oracle):
    # order-2 NQT is the quadratic e + (4f - f^2)/3 for x = 2^e (1 + f)
    # (arXiv:2206.08957; the integer version drops the low*low product, an
    # error below 2^-52). Powers of two are exact up to the decimal rounding
    # of the 2^-52 scale constant (2.22044604925031e-16).
    k = np.arange(-40, 41)
    x = 2.0 ** k
    assert np.all(np.abs(oracle.nqt(0, x) - k) <= 1e-14 * np.maximum(1, np.abs(k)))
    r = np.random.default_rng(3).uniform(1e-6, 1e6, 20000)
    m, e = np.frexp(r)  # r = m 2^e, m in [0.5, 1)
    f = 2 * m - 1
    closed = (e - 1) + (4 * f - f * f) / 3
    assert np.max(np.abs(oracle.nqt(0, r) - closed)) < 1e-13
    # and the quadratic stays within 0.0099 of log2
    assert np.max(np.abs(oracle.nqt(0, r) - np.log2(r))) < 0.0099


def test_portable_exp_cbrt_within_one_ulp_of_libm(oracle):
    rng = np.random.default_rng(4)
    x = np.concatenate([rng.uniform(-745, 709.7, 100000), rng.uniform(-1, 1, 100000)])
    y, r = oracle.nqt(5, x), np.exp(x)
    ok = r > 0
    assert np.max(np.abs(y[ok] - r[ok]) / np.spacing(r[ok])) <= 1.0
    assert oracle.nqt(5, [710.0])[0] == np.inf and oracle.nqt(5, [-746.0])[0] == 0.0
    x = np.concatenate([10.0 ** rng.uniform(-310, 308, 100000), -(10.0 ** rng.uniform(-3, 3, 1000))])
    y, r = oracle.nqt(6, x), np.cbrt(x)
    assert np.max(np.abs(y - r) / np.spacing(np.abs(r))) <= 1.0
    assert np.array_equal(oracle.nqt(6, [0.0, 8.0, -27.0, 0.125]), [0.0, 2.0, -3.0, 0.5])


# ---------------------------------------------------------------- root finding
def test_regula_falsi_atan(oracle):
    import ctypes as C
    g = GOLD["regula_falsi_atan"]
    root = C.c_double()
    it = C.c_int()
    st = oracle.lib().sgo_regula_falsi_atan(g["shift"], g["scale"], g["offset"], g["guess"],
                                            g["a"], g["b"], g["xtol"], g["ytol"],
                                            C.byref(root), C.byref(it))
    assert st == 0
    assert is_close(root.value, g["root"], g["tol"])
    # the closed form: 2 atan(x-1) + 0.5 = 0 -> x = 1 - tan(0.25)
    assert abs(root.value - (1.0 - math.tan(0.25))) < 1e-9


def test_findroot_cubic_matches_cbrt(oracle):
    import ctypes as C
    root = C.c_double()
    for c in (2.0, 27.0, 1234.5):
        st = oracle.lib().sgo_findroot_cubic(c, 1.0, 1e-3, 50.0, 1e-14, 1e-14, C.byref(root))
        assert st == 0
        assert abs(root.value - c ** (1 / 3)) <= 1e-12 * c ** (1 / 3)


# ---------------------------------------------------------------- analytic EOS
def test_ideal_gas_vector_goldens(oracle):
    g = GOLD["ideal_vector"]
    e = oracle.ideal(g["gm1"], g["Cv"])
    rho, sie = np.array(g["rho"]), np.array(g["sie"])
    assert np.all(is_close(e.vec(F_T_RE, rho, sie), g["T"], g["tol"]))
    assert np.all(is_close(e.vec(F_P_RE, rho, sie), g["P"], g["tol"]))
    assert np.all(is_close(e.vec(F_B_RE, rho, sie), g["bmod"], g["tol"]))
    assert np.all(e.vec(F_MINE, rho) == 0.0)
    T = np.array(g["T_in"])
    assert np.all(is_close(e.vec(F_E_RT, rho, T), g["sie_rhoT"], g["tol"]))
    assert np.all(is_close(e.vec(F_P_RT, rho, T), g["P_rhoT"], g["tol"]))
    assert np.all(is_close(e.vec(F_B_RT, rho, T), g["bmod_rhoT"], g["tol"]))
    # vector FillEos: T, P, cv, bmod from (rho, sie)
    t, p, cv, b = (np.full(3, -1.0) for _ in range(4))
    errs = e.fill(rho.copy(), t, sie.copy(), p, cv, b, Q_TEMPERATURE | Q_PRESSURE | Q_CV | Q_BMOD)
    assert not errs.any()
    assert np.all(is_close(t, g["T"], g["tol"])) and np.all(is_close(p, g["P"], g["tol"]))
    assert np.all(cv == g["Cv"]) and np.all(is_close(b, g["bmod"], g["tol"]))


@pytest.mark.parametrize("name", ["stiff_1"])
def test_stiff_gas_goldens(oracle, name):
    g = GOLD[name]
    e = oracle.stiff(g["gm1"], g["Cv"], g["Pinf"], g["qq"])
    rho, sie = np.array(g["rho"]), np.array(g["sie"])
    assert np.all(is_close(e.vec(F_P_RE, rho, sie), g["P"], g["tol"]))
    assert np.all(is_close(e.vec(F_B_RE, rho, sie), g["bmod"], g["tol"]))
    assert np.all(is_close(e.vec(F_T_RE, rho, sie), g["T"], g["tol"]))


def test_stiff_gas_sound_speed(oracle):
    g = GOLD["stiff_sound_speed"]
    e = oracle.stiff(g["gm1"], g["Cv"], g["Pinf"], g["qq"])
    b = e.vec(F_B_RE, [g["rho"]], [g["sie"]])[0]
    assert is_close(math.sqrt(b / g["rho"]), g["cs"], g["tol"])


def _grun(g, rho_max=0.0):
    return [g[k] for k in ("C0", "s1", "s2", "s3", "G0", "b", "rho0", "T0", "P0", "Cv")]


def test_gruneisen_cu_goldens(oracle):
    g = GOLD["gruneisen_cu"]
    e = oracle.gruneisen(*_grun(g))
    rho, sie = np.array(g["rho"]), np.array(g["sie"])
    P0c = g["P0"] - g["C0"] * g["C0"] * g["rho0"]
    Pt = np.array(g["P"][:3] + [P0c])
    Bt = np.array(g["bmod"][:3] + [g["G0"] * P0c])
    assert np.all(is_close(e.vec(F_P_RE, rho, sie), Pt, g["tol"]))
    assert np.all(is_close(e.vec(F_B_RE, rho, sie), Bt, g["tol"]))
    assert np.all(is_close(e.vec(F_T_RE, rho, sie), g["T"], g["tol"]))


def test_gruneisen_al_goldens_and_pt_roundtrip(oracle):
    g = GOLD["gruneisen_al"]
    e = oracle.gruneisen(*_grun(g))
    rho, sie = g["rho"], g["sie"]
    assert rho < oracle.lib().sgo_gruneisen_compute_rho_max(g["s1"], g["s2"], g["s3"], g["rho0"])
    P = e.vec(F_P_RE, [rho], [sie])[0]
    assert is_close(P / 1e12, g["P_Mbar"], g["tol"])
    B = e.vec(F_B_RE, [rho], [sie])[0]
    assert is_close(math.sqrt(B / rho) / 1e6, g["cs_cm_per_us"], g["tol"])
    T = e.vec(F_T_RE, [rho], [sie])[0]
    r2, s2, errs = e.rhoe_PT([P], [T], [0.0])
    assert errs[0] == 0
    assert is_close(r2[0], rho, g["tol"]) and is_close(s2[0], sie, g["tol"])


def test_gruneisen_rho_max_closed_form(oracle):
    # linear Us-up: eta* = 1/s1 -> rho_max = rho0 / (1 - 1/s1 + EPS)
    eps10 = 10 * np.finfo(np.float64).eps
    for s1, rho0 in ((1.34, 2.7), (1.489, 8.93)):
        r = oracle.lib().sgo_gruneisen_compute_rho_max(s1, 0.0, 0.0, rho0)
        assert r == rho0 / ((1 - 1.0 / s1) + eps10)
    # quadratic: root of 1 - s1 eta - s2 eta^2 in (0,1)
    s1, s2 = 1.2, 0.3
    eta = (-s1 + math.sqrt(s1 * s1 + 4 * s2)) / (2 * s2)
    r = oracle.lib().sgo_gruneisen_compute_rho_max(s1, s2, 0.0, 2.0)
    assert abs(r - 2.0 / (1 - eta)) < 1e-12 * r
    # cubic (regula-falsi branch)
    s1, s2, s3 = 1.0, 0.5, 0.2
    roots = [x.real for x in np.roots([-s3, -s2, -s1, 1.0]) if abs(x.imag) < 1e-12 and 0 < x.real < 1]
    r = oracle.lib().sgo_gruneisen_compute_rho_max(s1, s2, s3, 2.0)
    assert abs(r - 2.0 / (1 - min(roots))) < 1e-9 * r


def _vinet(g):
    return [g[k] for k in ("rho0", "T0", "B0", "BP0", "A0", "Cv0", "E0", "S0")]


def test_vinet_goldens(oracle):
    g = GOLD["vinet_cu_rho_sie"]
    e = oracle.vinet(*_vinet(g))
    rho, sie = np.array(g["rho"]), np.array(g["sie"])
    T = e.vec(F_T_RE, rho, sie)
    assert np.all(is_close(T, g["T"], g["tol"]))
    P = e.vec(F_P_RE, rho, sie)
    assert np.all(is_close(P[:3], g["P"][:3], g["tol"])) and abs(P[3]) < 1e-3
    B = e.vec(F_B_RE, rho, sie)
    bt = g["bmod"][:3] + [g["B0"] + g["B0"] ** 2 * g["A0"] ** 2 / g["Cv0"] / g["rho0"] * g["T0"]]
    assert np.all(is_close(B, bt, g["tol"]))
    h = GOLD["vinet_cu_rho_T"]
    rho, T = np.array(h["rho"]), np.array(h["T"])
    s = e.vec(F_E_RT, rho, T)
    assert np.all(is_close(s[:3], h["sie"][:3], h["tol"])) and abs(s[3]) < 1e-3
    P = e.vec(F_P_RT, rho, T)
    assert np.all(is_close(P[:3], h["P"][:3], h["tol"])) and abs(P[3]) < 1e-3


# ---------------------------------------------------------------- sums and solves
def test_kahan(oracle):
    g = GOLD["kahan"]
    A = np.full(g["N"], 1 - g["delta"])
    s = oracle.lib().sgo_sum_neumaier(A.ctypes.data, g["N"] - 1, 1, -1)
    assert is_close(s, (g["N"] - 1) * (1 - g["delta"]), 10 * np.finfo(float).eps)
    # skip index and catastrophic cancellation
    d = np.array([1e16, 1.0, -1e16, 3.0])
    assert oracle.lib().sgo_sum_neumaier(d.ctypes.data, 4, 0, -1) == 4.0
    assert oracle.lib().sgo_sum_neumaier(d.ctypes.data, 4, 0, 1) == 3.0


def _solve(oracle, A, b):
    a = np.ascontiguousarray(A, dtype=np.float64).ravel().copy()
    x = np.ascontiguousarray(b, dtype=np.float64).copy()
    ok = oracle.lib().sgo_solve(len(x), a.ctypes.data, x.ctypes.data)
    return ok, x


def test_solve_2x2_and_3x3(oracle):
    d = GOLD["solve_2x2"]["delta"]
    ok, x = _solve(oracle, [[1, 1 + d], [1, 1]], [1, 2])
    assert ok
    assert is_close(x[0], 2 * (1 + d) / d - 1 / d, math.sqrt(d))
    assert is_close(x[1], -1 / d, math.sqrt(d))
    d = GOLD["solve_3x3"]["delta"]
    A = np.ones((3, 3)) + d * np.eye(3)
    ok, x = _solve(oracle, A, [1, 1, 2])
    den = d * (3 + d)
    assert ok
    assert np.all(is_close(x, [(d - 1) / den, (d - 1) / den, 2 * (1 + d) / den], math.sqrt(d)))


def test_solve_random_systems_against_lu(oracle):
    rng = np.random.default_rng(7)
    for n in (1, 2, 3, 4, 5, 6):
        for _ in range(50):
            A = rng.normal(size=(n, n)) + n * np.eye(n)
            b = rng.normal(size=n)
            ok, x = _solve(oracle, A, b)
            assert ok
            xr = np.linalg.solve(A, b)
            assert np.max(np.abs(x - xr)) <= 1e-12 * (1 + np.max(np.abs(xr)))


def test_qr_pivots_and_r_match_lapack_geqp3(oracle):
    """The reference solves n >= 3 systems with Eigen's colPivHouseholderQr
    (mixed_cell_models.hpp:265-273), which follows LAPACK's xGEQP3 (largest
    remaining column norm, first on ties, norm downdating, beta =
    -sign(alpha) ||x||). scipy's qr(pivoting=True) runs LAPACK dgeqp3: on
    random systems whose column scales spread over six decades, the oracle's
    pivot order must be LAPACK's and its R LAPACK's R, signs included. The
    pivot order after the first step depends on the downdated norms: ordering
    by the initial norms instead (a dropped downdate) would fail on about a
    third of these systems (107 of 320), as the count below checks."""
    import scipy.linalg as sl
    rng = np.random.default_rng(11)
    lib = oracle.lib()
    naive_differs = total = 0
    for n in (3, 4, 5, 6, 7, 8, 12, 16):
        for _ in range(40):
            A = rng.normal(size=(n, n)) * 10.0 ** rng.uniform(-3, 3, size=n)
            b = rng.normal(size=n)
            a = np.ascontiguousarray(A).ravel().copy()
            bb = b.copy()
            x = np.zeros(n)
            perm = np.zeros(n, dtype=np.int32)
            lib.sgo_qr_colpiv(n, a.ctypes.data, bb.ctypes.data, x.ctypes.data, perm.ctypes.data)
            _, R, P = sl.qr(A, pivoting=True)
            assert np.array_equal(perm, P), (n, perm, P)
            Ro = np.triu(a.reshape(n, n))
            assert np.allclose(Ro, R, rtol=1e-10, atol=1e-12 * np.abs(R).max()), n
            xr = np.linalg.solve(A, b)
            assert np.allclose(x, xr, rtol=1e-8, atol=1e-10 * np.abs(xr).max())
            naive = np.argsort(-np.linalg.norm(A, axis=0), kind="stable")
            naive_differs += int(not np.array_equal(naive, P))
            total += 1
    assert naive_differs > total // 4


def test_solve_singular_rows(oracle):
    ok, _ = _solve(oracle, [[0, 0, 0], [1, 2, 3], [4, 5, 6]], [1, 2, 3])
    assert not ok
    # rank-deficient but non-zero rows: pivoted QR returns a finite solution
    ok, x = _solve(oracle, [[1, 2, 3], [2, 4, 6], [1, 0, 1]], [1, 2, 1])
    assert ok and np.all(np.isfinite(x))
    assert np.allclose(np.array([[1, 2, 3], [2, 4, 6], [1, 0, 1]]) @ x, [1, 2, 1], atol=1e-12)


def test_enforce_sums(oracle):
    nmat, tot_rho, tot_u = 3, 2.5, 1.23e12
    tot_sie = tot_u / tot_rho
    mu = np.array([0.025, 0.9, 0.075])
    vf = np.array([0.55, 0.05, 0.4])
    rhobar = mu * tot_rho
    rho = rhobar / vf + np.array([-0.01, 0.004, -0.002])
    vf = rhobar / (rho + 10 * np.finfo(float).tiny)
    sie = np.array([0.5 * tot_u / rhobar[0] + 0.01 * tot_sie, -0.0025 * tot_sie,
                    0.5 * tot_u / rhobar[2] - 0.06 * tot_sie])
    assert oracle.lib().sgo_enforce_mass_volumes(nmat, 1.0, rho.ctypes.data, vf.ctypes.data)
    assert is_close(vf.sum(), 1.0, 1e-12) and is_close((vf * rho).sum(), tot_rho, 1e-12)
    assert oracle.lib().sgo_enforce_energies(nmat, tot_rho, tot_sie, rho.ctypes.data,
                                             vf.ctypes.data, sie.ctypes.data)
    assert is_close((vf * rho * sie).sum(), tot_u, 1e-12)


# ---------------------------------------------------------------- tables
def _linear_table(oracle, nR=17, nT=23):
    lrmin, lrmax, ltmin, ltmax = -3.0, 2.0, 2.0, 6.0
    x = lrmin + np.arange(nR) * ((lrmax - lrmin) / (nR - 1))
    y = ltmin + np.arange(nT) * ((ltmax - ltmin) / (nT - 1))
    X, Y = np.meshgrid(x, y, indexing="ij")
    c0, c1, c2, c3 = 1.5, -0.25, 0.75, 0.125
    F = c0 + c1 * X + c2 * Y + c3 * X * Y
    fields = {k: F.ravel().copy() for k in ("P", "sie", "bMod", "dPdRho", "dPdE", "dEdRho", "dEdT")}
    fields["dPdE"] = np.ones(nR * nT)
    fields["dEdRho"] = np.zeros(nR * nT)
    cold = {"PCold": F[:, 0].copy(), "sieCold": F[:, 0].copy(), "bModCold": F[:, 0].copy()}
    t = oracle.Table(lrmin, lrmax, nR, ltmin, ltmax, nT, fields, cold)
    f = lambda a, b: c0 + c1 * a + c2 * b + c3 * a * b
    return t, f, (lrmin, lrmax, ltmin, ltmax), (x, y)


def test_table_bilinear_reproduces_bilinear_functions(oracle):
    t, f, (a0, a1, b0, b1), (x, y) = _linear_table(oracle)
    rng = np.random.default_rng(11)
    lr = rng.uniform(a0, a1, 5000)
    lt = rng.uniform(b0, b1, 5000)
    v = t.interp(0, lr, lt)
    assert np.max(np.abs(v - f(lr, lt)) / np.abs(f(lr, lt))) < 1e-14
    # nodes exactly
    X, Y = np.meshgrid(x, y, indexing="ij")
    assert np.max(np.abs(t.interp(0, X.ravel(), Y.ravel()) - f(X.ravel(), Y.ravel()))) < 1e-14
    # outside the grid: linear extension of the edge cell
    lr = np.array([a0 - 0.3, a1 + 0.2])
    lt = np.array([b0 - 0.1, b1 + 0.4])
    assert np.allclose(t.interp(0, lr, lt), f(lr, lt), rtol=1e-13)


def test_table_inverse_roundtrip(oracle):
    t, f, (a0, a1, b0, b1), _ = _linear_table(oracle)
    rng = np.random.default_rng(12)
    lr = rng.uniform(a0, a1, 2000)
    lt = rng.uniform(b0 + 0.01, b1 - 0.01, 2000)
    sie = t.interp(1, lr, lt)
    lt2, it, wh = t.lT_from_sie(lr, sie)
    on = wh == 0
    assert on.mean() > 0.95
    assert np.max(np.abs(lt2[on] - lt[on])) < 1e-12


def test_table_derived_curves(oracle, workloads):
    wl = workloads
    nq = __import__("importlib").import_module("singularity-eos_amd.nqt")
    lr0, lr1, lt0, lt1 = nq.log10(1e-3), nq.log10(1e2), nq.log10(1e2), nq.log10(1e6)
    t, f, c = oracle.gruneisen_table(wl.GRUN_C1, lr0, lr1, 64, lt0, lt1, 96)
    P = f["P"].reshape(64, 96)
    assert np.array_equal(t.export(0), P[:, -1])                 # PMax = top isotherm
    assert np.array_equal(t.export(1), f["sie"].reshape(64, 96)[:, -1])
    assert t.scalar(0) == min(0.0, P.min())                      # PMin
    rap = t.export(4)                                            # rho at P min per T
    for i in range(96):
        j = int(np.argmin(P[:, i]))
        expect = 0.0 if P[:, i].min() > 0 else oracle.nqt(3, [lr0 + j * ((lr1 - lr0) / 63)])[0]
        assert rap[i] == expect


def test_gruneisen_table_b_choice(oracle, workloads):
    """The sampled table's density of minimum pressure, per isotherm: with the
    workload's b = 0 the minimum sits in tension at the lowest grid density;
    with the Al test's b = 0.48 the Hugoniot term turns negative above
    ~8.2 g/cc and the minimum moves to ~10.9 g/cc at every T (which is why the
    workload does not use it)."""
    wl = workloads
    nq = __import__("importlib").import_module("singularity-eos_amd.nqt")
    lr0, lr1, lt0, lt1 = nq.log10(1e-3), nq.log10(1e2), nq.log10(1e2), nq.log10(1e6)
    t, _, _ = oracle.gruneisen_table(wl.GRUN_C1, lr0, lr1, 256, lt0, lt1, 64)
    assert np.all(t.export(4) < 1.01e-3)
    g = list(wl.GRUN_C1)
    g[5] = wl.GRUN_AL_TEST_B
    t2, _, _ = oracle.gruneisen_table(tuple(g), lr0, lr1, 256, lt0, lt1, 64)
    assert np.all(np.abs(t2.export(4) - 10.9) < 0.3)
