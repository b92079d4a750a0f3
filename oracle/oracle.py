"""ORACLE — test infrastructure only. ctypes wrapper of liboracle.so, the CPU
restatement of singularity-eos used as the parity checker (tests/) and as the
reported host baseline (bench.py cpu_baseline). The shipped library never
imports this module.
"""
import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(BUILD, "liboracle.so")
SRCS = ["oracle_capi.cpp"]
HDRS = ["sgo_math.hpp", "sgo_eos.hpp", "sgo_pte.hpp", "sgo_driver.hpp"]
CXXFLAGS = ["-O3", "-std=c++17", "-ffp-contract=off", "-fno-fast-math", "-fopenmp", "-fPIC",
            "-shared"]


FLOPS_LIB = os.path.join(BUILD, "liboracle_flops.so")
# the same restatement with the host libm's exp / cbrt in Vinet in place of
# the portable routines the device shares (loaded when SGO_ORACLE_LIBM=1)
LIBM_LIB = os.path.join(BUILD, "liboracle_libm.so")


def build(force=False):
    """Compiles liboracle.so with g++ (no FMA contraction, no fast-math) and
    liboracle_flops.so, the same restatement with an operation-counting Real."""
    os.makedirs(BUILD, exist_ok=True)
    deps = [os.path.join(HERE, s) for s in SRCS + ["flops_count.cpp"]]
    deps += [os.path.join(HERE, h) for h in HDRS + ["sgo_real.hpp"]]
    newest = max(os.path.getmtime(d) for d in deps)
    for out, src, extra in ((LIB, SRCS, []), (FLOPS_LIB, ["flops_count.cpp"], ["-DSGO_COUNT_FLOPS"]),
                            (LIBM_LIB, SRCS, ["-DSGO_LIBM_EXP_CBRT"])):
        if not force and os.path.exists(out) and os.path.getmtime(out) >= newest:
            continue
        tmp = out + ".tmp"
        subprocess.check_call(["g++"] + CXXFLAGS + extra + [os.path.join(HERE, s) for s in src]
                              + ["-o", tmp])
        os.replace(tmp, out)
    return LIB


def count_pte_flops(method, eos_specs, offsets, mat_eos, rho, vfrac, temp, press, vfrac_tot,
                    sie_tot, params_arr, table=None):
    """Algorithmic FP64 operations per cell (every +,-,*,/ and math call once).
    eos_specs: list of (kind, params...) tuples; kind 4 = `table`
    (an oracle.Table-like object with grid and field arrays)."""
    L = C.CDLL(FLOPS_LIB)
    f = L.sgo_count_pte_flops
    f.restype = None
    f.argtypes = [ci, cl, vp, vp, vp, ci, vp, vp, vp, vp, vp, vp, vp, vp, vp, vp]
    ep = np.zeros(16 * len(eos_specs))
    for i, spec in enumerate(eos_specs):
        ep[16 * i:16 * i + len(spec)] = spec
    ncell = len(offsets) - 1
    out = np.zeros(ncell)
    offsets = np.ascontiguousarray(offsets, dtype=np.int32)
    mat_eos = np.ascontiguousarray(mat_eos, dtype=np.int32)
    arrs = [np.ascontiguousarray(x, dtype=np.float64) for x in
            (rho, vfrac, temp, press, vfrac_tot, sie_tot, params_arr)]
    tg = tf = None
    keep = []
    if table is not None:
        tg = np.array([table.lRhoMin, table.lRhoMax, table.nRho, table.lTMin, table.lTMax, table.nT],
                      dtype=np.float64)
        fl = [table.f[k] for k in Table.FIELDS] + [table.c[k] for k in Table.COLD]
        keep = [np.ascontiguousarray(x, dtype=np.float64) for x in fl]
        tf = (C.c_void_p * 10)(*[x.ctypes.data for x in keep])
    f(method, ncell, p(offsets), p(mat_eos), p(ep), len(eos_specs), *[p(a) for a in arrs],
      p(tg), tf, p(out))
    return out


vp, ci, cl, cd = C.c_void_p, C.c_int, C.c_long, C.c_double
_SIGS = {
    "sgo_eos_ideal": (vp, [cd, cd]),
    "sgo_eos_stiff": (vp, [cd, cd, cd, cd]),
    "sgo_eos_gruneisen": (vp, [cd] * 11),
    "sgo_gruneisen_compute_rho_max": (cd, [cd] * 4),
    "sgo_eos_vinet": (vp, [cd] * 8 + [vp]),
    "sgo_table_create": (vp, [cd, cd, ci, cd, cd, ci] + [vp] * 10 + [cd, cd]),
    "sgo_table_free": (None, [vp]),
    "sgo_eos_table": (vp, [vp]),
    "sgo_eos_free": (None, [vp]),
    "sgo_eos_set_modifiers": (None, [vp, ci, ci, cd, cd]),
    "sgo_table_export": (cl, [vp, ci, vp]),
    "sgo_table_scalar": (cd, [vp, ci]),
    "sgo_table_interp": (None, [vp, ci, vp, vp, cl, vp]),
    "sgo_table_lT_from_sie": (None, [vp, vp, vp, cl, vp, vp, vp]),
    "sgo_nqt": (None, [ci, vp, vp, cl]),
    "sgo_nqt_constants": (None, [vp, vp]),
    "sgo_regula_falsi_atan": (ci, [cd] * 8 + [vp, vp]),
    "sgo_findroot_cubic": (ci, [cd] * 6 + [vp]),
    "sgo_solve": (ci, [ci, vp, vp]),
    "sgo_qr_colpiv": (None, [ci, vp, vp, vp, vp]),
    "sgo_sum_neumaier": (cd, [vp, ci, ci, ci]),
    "sgo_enforce_mass_volumes": (ci, [ci, cd, vp, vp]),
    "sgo_enforce_energies": (ci, [ci, cd, cd, vp, vp, vp]),
    "sgo_approx_temperature": (cd, [ci, vp, cd, vp, vp, cd]),
    "sgo_vec_eval": (None, [vp, ci, vp, vp, vp, cl, vp, ci, ci]),
    "sgo_vec_fill": (None, [vp, vp, vp, vp, vp, vp, vp, cl, C.c_ulong, vp, ci, vp, ci]),
    "sgo_vec_rhoe_PT": (None, [vp, vp, vp, vp, vp, cl, vp, ci, vp, ci]),
    "sgo_build_gruneisen_table": (None, [vp, cd, cd, ci, cd, cd, ci] + [vp] * 10),
    "sgo_get_sg_eos": (cl, [ci, ci, ci, ci, vp, vp, ci, vp] + [vp] * 15 + [cd, vp, vp, vp, ci]),
    "sgo_pte_solve": (cl, [ci, cl, vp, vp, vp, ci, vp, vp, vp, vp, vp, vp, vp, vp, vp, vp, vp,
                           vp, vp, vp, ci]),
    "sgo_max_threads": (ci, []),
}

_lib = None


def lib():
    global _lib
    if _lib is None:
        path = LIBM_LIB if os.environ.get("SGO_ORACLE_LIBM") == "1" else LIB
        if not os.path.exists(path):
            build()
        L = C.CDLL(path)
        for n, (r, a) in _SIGS.items():
            f = getattr(L, n)
            f.restype = r
            f.argtypes = a
        _lib = L
    return _lib


def p(a):
    return None if a is None else a.ctypes.data


class Eos:
    """Owning handle to one oracle EOS object."""

    def __init__(self, h, keep=None):
        self.h = h
        self.keep = keep

    def __del__(self):
        try:
            lib().sgo_eos_free(self.h)
        except Exception:
            pass

    def modify(self, scaled=False, shifted=False, scale=1.0, shift=0.0):
        """ScaledEOS<ShiftedEOS<...>> (either may be off), in place; returns self."""
        lib().sgo_eos_set_modifiers(self.h, int(scaled), int(shifted), float(scale), float(shift))
        return self

    def vec(self, func, in1, in2=None, lam=None, nthreads=0):
        in1 = np.ascontiguousarray(in1, dtype=np.float64)
        in2 = in1 if in2 is None else np.ascontiguousarray(in2, dtype=np.float64)
        out = np.zeros_like(in1)
        lib().sgo_vec_eval(self.h, func, p(in1), p(in2), p(out), len(in1), p(lam), 3, nthreads)
        return out

    def fill(self, rho, temp, sie, press, cv, bmod, output, lam=None, nthreads=0):
        errs = np.zeros(len(rho), dtype=np.int32)
        lib().sgo_vec_fill(self.h, p(rho), p(temp), p(sie), p(press), p(cv), p(bmod), len(rho),
                           output, p(lam), 3, p(errs), nthreads)
        return errs

    def rhoe_PT(self, P, T, rho_guess, lam=None, nthreads=0):
        rho = np.array(rho_guess, dtype=np.float64)
        sie = np.zeros_like(rho)
        errs = np.zeros(len(rho), dtype=np.int32)
        P = np.ascontiguousarray(P, dtype=np.float64)  # keep alive across the call
        T = np.ascontiguousarray(T, dtype=np.float64)
        lib().sgo_vec_rhoe_PT(self.h, p(P), p(T), p(rho), p(sie), len(rho), p(lam), 3, p(errs),
                              nthreads)
        return rho, sie, errs


def ideal(gm1, Cv):
    return Eos(lib().sgo_eos_ideal(gm1, Cv))


def stiff(gm1, Cv, Pinf, qq):
    return Eos(lib().sgo_eos_stiff(gm1, Cv, Pinf, qq))


def gruneisen(C0, s1, s2, s3, G0, b, rho0, T0, P0, Cv, rho_max=0.0):
    return Eos(lib().sgo_eos_gruneisen(C0, s1, s2, s3, G0, b, rho0, T0, P0, Cv, rho_max))


def vinet(rho0, T0, B0, BP0, A0, Cv0, E0, S0, d2to40=None):
    arr = np.zeros(39) if d2to40 is None else np.asarray(d2to40, dtype=np.float64)
    return Eos(lib().sgo_eos_vinet(rho0, T0, B0, BP0, A0, Cv0, E0, S0, p(arr)))


class Table:
    FIELDS = ("P", "sie", "bMod", "dPdRho", "dPdE", "dEdRho", "dEdT")
    COLD = ("PCold", "sieCold", "bModCold")

    def __init__(self, lRhoMin, lRhoMax, nRho, lTMin, lTMax, nT, f, c, rho_off=0.0, T_off=0.0):
        self.lRhoMin, self.lRhoMax, self.nRho = lRhoMin, lRhoMax, nRho
        self.lTMin, self.lTMax, self.nT = lTMin, lTMax, nT
        self.f = {k: np.ascontiguousarray(f[k], dtype=np.float64) for k in self.FIELDS}
        self.c = {k: np.ascontiguousarray(c[k], dtype=np.float64) for k in self.COLD}
        self.h = lib().sgo_table_create(lRhoMin, lRhoMax, nRho, lTMin, lTMax, nT,
                                        *[p(self.f[k]) for k in self.FIELDS],
                                        *[p(self.c[k]) for k in self.COLD], rho_off, T_off)

    def __del__(self):
        try:
            lib().sgo_table_free(self.h)
        except Exception:
            pass

    def eos(self):
        return Eos(lib().sgo_eos_table(self.h), keep=self)

    def export(self, which):
        n = lib().sgo_table_export(self.h, which, None)
        out = np.zeros(n)
        lib().sgo_table_export(self.h, which, p(out))
        return out

    def scalar(self, which):
        return lib().sgo_table_scalar(self.h, which)

    def interp(self, field, lRho, lT):
        lRho = np.ascontiguousarray(lRho, dtype=np.float64)
        lT = np.ascontiguousarray(lT, dtype=np.float64)
        out = np.zeros_like(lRho)
        lib().sgo_table_interp(self.h, field, p(lRho), p(lT), len(lRho), p(out))
        return out

    def lT_from_sie(self, lRho, sie):
        lRho = np.ascontiguousarray(lRho, dtype=np.float64)
        sie = np.ascontiguousarray(sie, dtype=np.float64)
        lT = np.zeros_like(lRho)
        it = np.zeros(len(lRho), dtype=np.int32)
        wh = np.zeros(len(lRho), dtype=np.int32)
        lib().sgo_table_lT_from_sie(self.h, p(lRho), p(sie), len(lRho), p(lT), p(it), p(wh))
        return lT, it, wh


def gruneisen_table(grun_params, lRhoMin, lRhoMax, nRho, lTMin, lTMax, nT):
    """The oracle's own Gruneisen-sampled table (independent of the library's
    generator)."""
    g = gruneisen(*grun_params)
    f = {k: np.zeros(nRho * nT) for k in Table.FIELDS}
    c = {k: np.zeros(nRho) for k in Table.COLD}
    lib().sgo_build_gruneisen_table(g.h, lRhoMin, lRhoMax, nRho, lTMin, lTMax, nT,
                                    *[p(f[k]) for k in Table.FIELDS], *[p(c[k]) for k in Table.COLD])
    return Table(lRhoMin, lRhoMax, nRho, lTMin, lTMax, nT, f, c), f, c


def nqt(which, x):
    x = np.ascontiguousarray(x, dtype=np.float64)
    y = np.zeros_like(x)
    lib().sgo_nqt(which, p(x), p(y), len(x))
    return y


def pte_solve(method, eos_list, offsets, mat_eos, rho, vfrac, sie, temp, press, vfrac_tot,
              sie_tot, params_arr, Tnorm=None, lam=None, nthreads=0):
    """Runs the oracle closure over a CSR cell list; arrays are updated in
    place. Returns (nfail, converged, niter, residual, err)."""
    ncell = len(offsets) - 1
    handles = (C.c_void_p * len(eos_list))(*[e.h for e in eos_list])
    conv = np.zeros(ncell, dtype=np.int32)
    nit = np.zeros(ncell, dtype=np.int32)
    res = np.zeros(ncell)
    err = np.zeros(ncell, dtype=np.int32)
    pa = np.asarray(params_arr, dtype=np.float64)
    offsets = np.ascontiguousarray(offsets, dtype=np.int32)
    mat_eos = np.ascontiguousarray(mat_eos, dtype=np.int32)
    nfail = lib().sgo_pte_solve(method, ncell, p(offsets), p(mat_eos), handles, len(eos_list),
                                p(rho), p(vfrac), p(sie), p(temp), p(press), p(vfrac_tot),
                                p(sie_tot), p(Tnorm), p(pa), p(lam), p(conv), p(nit), p(res),
                                p(err), nthreads)
    return nfail, conv, nit, res, err


SG_FIELDS = ("press", "pmax", "vol", "spvol", "sie", "temp", "bmod", "dpde", "cv")
SG_FRAC = ("frac_mass", "frac_vol", "frac_ie", "frac_bmod", "frac_dpde", "frac_cv")


def get_sg_eos(eos_list, input_mode, eos_offsets, offsets, arrays, cutoff, params_arr,
               nthreads=0):
    """The get_sg_eos driver on the CPU. arrays: dict of the per-cell arrays
    (SG_FIELDS, length cell_dim) and column-major per-material arrays
    (SG_FRAC, cell_dim * nmat; the three optional ones may be absent), updated
    in place. Returns (nfail, err[ncell], converged[ncell])."""
    nmat = len(eos_offsets)
    ncell = len(offsets)
    cell_dim = len(arrays["press"])
    handles = (C.c_void_p * len(eos_list))(*[e.h for e in eos_list])
    eo = np.ascontiguousarray(eos_offsets, dtype=np.int32)
    off = np.ascontiguousarray(offsets, dtype=np.int32)
    err = np.zeros(ncell, dtype=np.int32)
    conv = np.zeros(ncell, dtype=np.int32)
    pa = np.asarray(params_arr, dtype=np.float64)
    ptrs = [p(arrays[k]) for k in SG_FIELDS] + [p(arrays.get(k)) for k in SG_FRAC]
    nfail = lib().sgo_get_sg_eos(nmat, ncell, cell_dim, input_mode, p(eo), handles,
                                 len(eos_list), p(off), *ptrs, cutoff, p(pa), p(err), p(conv),
                                 nthreads)
    return nfail, err, conv


def default_params():
    """MixParams() defaults packed in the oracle order."""
    # the _sufficient values are products (1.e2 * 1.0e-6 is not 1e-4 in binary)
    return [3e-6, 1e-6, 1e-4, 1e-6, 1e-6, 1e-6, 1e-4, 0.99, 1.e2 * 1.0e-6, 10 * 1.0e-4,
            1.e2 * 1.0e-6, 1.e2 * 1.0e-6, 1.e2 * 1.0e-6, 10 * 1.0e-4,
            128, 1e-2, 6, 0.5, 0.95, 1e15, 300.0, 1e-16, 2, 1e-16, -1e-8, -1, 1]


def tight_params():
    """MixParams of singularity-eos/example/pte_2mat.cpp:169-180."""
    p_ = default_params()
    p_[1] = p_[2] = p_[3] = 1e-10
    p_[8] = p_[9] = p_[10] = 1e-10
    p_[5] = p_[6] = p_[12] = p_[13] = 0.0
    return p_
