"""Python binding of libsgm.so — the C ABI declared in include/sgm.h.

Device memory is passed as raw pointers (torch tensors' ``data_ptr()`` or
ints); streams as ``torch.cuda.current_stream().cuda_stream`` or 0. This
module is plumbing for tests and benchmarks: all computation happens in the
HIP kernels of libsgm.so, and loading fails loudly when the library is absent.
"""
import ctypes as C
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libsgm.so")

# EOS kinds / function selectors / thermalqs bits (include/sgm.h)
IDEAL, STIFF, GRUNEISEN, VINET, TABLE_RHOT = 0, 1, 2, 3, 4
F_P_RHO_SIE, F_T_RHO_SIE, F_SIE_RHO_T, F_P_RHO_T = 0, 1, 2, 3
F_CV_RHO_SIE, F_CV_RHO_T, F_BMOD_RHO_SIE, F_BMOD_RHO_T, F_MIN_SIE_RHO = 4, 5, 6, 7, 8
Q_DENSITY, Q_SIE, Q_PRESSURE, Q_TEMPERATURE, Q_CV, Q_BMOD = 1, 2, 4, 8, 16, 32
ERRB_GRUN_MAXRHO, ERRB_GRUN_PT, ERRB_FILLEOS, ERRB_LINSOLVE, ERRB_SMALLSTEP, ERRB_INPUT = (
    1, 2, 4, 8, 16, 32)


class MixParams(C.Structure):
    """sgm_mix_params (MixParams, mixed_cell_models.hpp:43-78)."""
    _fields_ = [(n, C.c_double) for n in (
        "derivative_eps", "pte_rel_tolerance_p", "pte_rel_tolerance_t", "pte_rel_tolerance_e",
        "pte_rel_tolerance_v", "pte_abs_tolerance_p", "pte_abs_tolerance_t",
        "pte_slow_convergence_thresh", "pte_rel_tolerance_p_sufficient",
        "pte_rel_tolerance_t_sufficient", "pte_rel_tolerance_e_sufficient",
        "pte_rel_tolerance_v_sufficient", "pte_abs_tolerance_p_sufficient",
        "pte_abs_tolerance_t_sufficient", "line_search_alpha", "line_search_fac",
        "vfrac_safety_fac", "temperature_limit", "default_tguess", "min_dtde",
        "pte_small_step_thresh", "pte_max_dpdv")] + [(n, C.c_int) for n in (
        "pte_max_iter_per_mat", "line_search_max_iter", "pte_small_step_tries",
        "pte_reduced_system_exclude_idx", "iterate_t_guess", "pad")]

    # packed-double order used by the oracle's C entry points
    ORACLE_ORDER = (
        "derivative_eps", "pte_rel_tolerance_p", "pte_rel_tolerance_t", "pte_rel_tolerance_e",
        "pte_rel_tolerance_v", "pte_abs_tolerance_p", "pte_abs_tolerance_t",
        "pte_slow_convergence_thresh", "pte_rel_tolerance_p_sufficient",
        "pte_rel_tolerance_t_sufficient", "pte_rel_tolerance_e_sufficient",
        "pte_rel_tolerance_v_sufficient", "pte_abs_tolerance_p_sufficient",
        "pte_abs_tolerance_t_sufficient", "pte_max_iter_per_mat", "line_search_alpha",
        "line_search_max_iter", "line_search_fac", "vfrac_safety_fac", "temperature_limit",
        "default_tguess", "min_dtde", "pte_small_step_tries", "pte_small_step_thresh",
        "pte_max_dpdv", "pte_reduced_system_exclude_idx", "iterate_t_guess")

    def as_oracle_array(self):
        return [float(getattr(self, k)) for k in self.ORACLE_ORDER]


class PteStatus(C.Structure):
    _fields_ = [("converged", C.c_void_p), ("niter", C.c_void_p), ("residual", C.c_void_p),
                ("err", C.c_void_p)]


class TableDesc(C.Structure):
    _fields_ = [("lRhoMin", C.c_double), ("lRhoMax", C.c_double), ("nRho", C.c_int),
                ("lTMin", C.c_double), ("lTMax", C.c_double), ("nT", C.c_int),
                ("rho_offset", C.c_double), ("T_offset", C.c_double)] + [
        (n, C.c_void_p) for n in ("P", "sie", "bMod", "dPdRho", "dPdE", "dEdRho", "dEdT",
                                  "PCold", "sieCold", "bModCold")] + [
        ("reproducible", C.c_int), ("pmin_vapor_dome", C.c_int)]


class ServerConfig(C.Structure):
    _fields_ = [("hbm_budget_bytes", C.c_size_t), ("nlanes", C.c_int),
                ("max_group_requests", C.c_int), ("max_group_cells", C.c_long),
                ("plain", C.c_int), ("method", C.c_int)]


REQ_VEC_P, REQ_VEC_T, REQ_PTE = 1, 2, 4


class Request(C.Structure):
    _fields_ = [("kind", C.c_int), ("matindex", C.c_int), ("n", C.c_long)] + [
        (k, C.c_void_p) for k in ("rho", "sie", "P_out", "T_out", "offsets", "mat_ids",
                                  "rho_in", "vfrac_in", "temp_in", "press_in", "vfrac_tot",
                                  "sie_tot", "rho_out", "vfrac_out", "sie_out", "temp_out",
                                  "press_out", "converged")]


class ServerStats(C.Structure):
    _fields_ = [("submitted", C.c_long), ("completed", C.c_long), ("failed", C.c_long),
                ("groups", C.c_long), ("cells", C.c_double), ("arena_peak_bytes", C.c_size_t),
                ("arena_bytes", C.c_size_t)]


_lib = None
vp, ci, cl, cd = C.c_void_p, C.c_int, C.c_long, C.c_double

_SIGS = {
    "sgm_init": (ci, [ci, C.POINTER(vp)]),
    "sgm_init_host": (ci, [ci, C.POINTER(vp)]),
    "sgm_finalize": (ci, [vp]),
    "sgm_set_IdealGas": (ci, [vp, ci, cd, cd]),
    "sgm_set_StiffGas": (ci, [vp, ci, cd, cd, cd, cd]),
    "sgm_set_Gruneisen": (ci, [vp, ci] + [cd] * 11),
    "sgm_set_Vinet": (ci, [vp, ci] + [cd] * 8 + [vp]),
    "sgm_set_TableRhoT": (ci, [vp, ci, C.POINTER(TableDesc)]),
    "sgm_get_kind": (ci, [vp, ci]),
    "sgm_table_query": (cl, [vp, ci, ci, vp]),
    "sgm_vec_P_from_rho_sie": (ci, [vp, ci, vp, vp, vp, cl, ci, vp, vp]),
    "sgm_vec_T_from_rho_sie": (ci, [vp, ci, vp, vp, vp, cl, ci, vp, vp]),
    "sgm_vec_eval": (ci, [vp, ci, ci, vp, vp, vp, cl, ci, vp, vp]),
    "sgm_vec_fill_eos": (ci, [vp, ci, vp, vp, vp, vp, vp, vp, cl, C.c_ulong, ci, vp, vp, vp]),
    "sgm_vec_rhoe_from_PT": (ci, [vp, ci, vp, vp, vp, vp, cl, ci, vp, vp, vp]),
    "sgm_pte_rhoT": (ci, [vp, cl, cl, vp, vp, vp, vp, vp, vp, vp, vp, vp, vp, vp,
                          C.POINTER(MixParams), PteStatus, vp]),
    "sgm_pte_PT": (ci, [vp, cl, cl, vp, vp, vp, vp, vp, vp, vp, vp, vp, vp, vp,
                        C.POINTER(MixParams), PteStatus, vp]),
    "sgm_pte_rhoU": (ci, [vp, cl, cl, vp, vp, vp, vp, vp, vp, vp, vp, vp, vp, vp,
                          C.POINTER(MixParams), PteStatus, vp]),
    "sgm_pte_fixedT": (ci, [vp, cl, cl, vp, vp, vp, vp, vp, vp, vp, vp, vp, vp,
                            C.POINTER(MixParams), PteStatus, vp]),
    "sgm_pte_fixedP": (ci, [vp, cl, cl, vp, vp, vp, vp, vp, vp, vp, vp, vp, vp,
                            C.POINTER(MixParams), PteStatus, vp]),
    "sgm_set_modifiers": (ci, [vp, ci, vp, vp]),
    "sgm_get_sg_eos_noalloc": (ci, [vp, ci, ci, ci, ci, vp, vp] + [vp] * 15 +
                               [cd, C.POINTER(MixParams), vp, vp]),
    "sgm_release_driver_cache": (ci, [vp]),
    "sgm_driver_cache_info": (ci, [vp, vp]),
    "sgm_device_registry": (ci, [vp, C.POINTER(vp), C.POINTER(ci)]),
    "sgm_approx_temperature": (ci, [vp, cl, cl, vp, vp, vp, vp, vp, vp, vp, vp, vp]),
    "sgm_get_sg_eos": (ci, [vp, ci, ci, ci, ci, vp, vp] + [vp] * 15 +
                       [cd, C.POINTER(MixParams), vp, vp]),
    "sgm_plan_create": (ci, [vp, cl, cl, vp, vp, ci, C.POINTER(vp), vp]),
    "sgm_plan_set_straggler_threshold": (ci, [vp, ci]),
    "sgm_plan_set_shape": (ci, [vp, ci, ci, ci, ci]),
    "sgm_plan_autotune": (ci, [vp, vp, vp, vp, vp, vp, vp, C.POINTER(MixParams), ci, vp]),
    "sgm_plan_execute": (ci, [vp, vp, vp, vp, vp, vp, vp, vp, vp, vp, C.POINTER(MixParams),
                              PteStatus, vp]),
    "sgm_plan_execute_bin": (ci, [vp, ci, vp, vp, vp, vp, vp, vp, vp, vp, vp,
                                  C.POINTER(MixParams), PteStatus, vp, vp]),
    "sgm_plan_describe": (ci, [vp, C.POINTER(C.c_long), ci]),
    "sgm_plan_tail_counts": (ci, [vp, C.POINTER(C.c_int), ci, vp]),
    "sgm_plan_reorder": (ci, [vp, vp, vp]),
    "sgm_plan_destroy": (ci, [vp]),
    "sgm_mix_params_default": (None, [C.POINTER(MixParams)]),
    "sgm_mix_params_tight": (None, [C.POINTER(MixParams)]),
    "sgm_device_info": (ci, [C.POINTER(ci), C.POINTER(ci), C.POINTER(cd)]),
    "sgm_version": (C.c_char_p, []),
    "sgm_build_gruneisen_table": (ci, [cd] * 10 + [cd, cd, ci, cd, cd, ci] + [vp] * 10),
    "sgm_workload_pte2": (ci, [C.c_uint64, cl, vp, vp, vp, vp, vp, vp]),
    "sgm_mlp_create": (ci, [C.c_uint64, C.POINTER(vp)]),
    "sgm_mlp_destroy": (ci, [vp]),
    "sgm_mlp_nparam": (ci, []),
    "sgm_mlp_params": (ci, [vp, vp, vp]),
    "sgm_mlp_set_norm": (ci, [vp, vp, vp, C.c_float, C.c_float]),
    "sgm_mlp_features": (ci, [vp, cl, cl, vp, vp, vp, vp, vp, vp, vp, vp, vp, vp]),
    "sgm_mlp_targets": (ci, [vp, cl, cl, vp, vp, vp, vp, vp, vp, vp, vp]),
    "sgm_mlp_gradients": (ci, [vp, vp, vp, vp, cl, C.c_float, vp, vp, vp]),
    "sgm_mlp_train_step": (ci, [vp, vp, vp, vp, cl, C.c_float, C.c_float, vp, vp]),
    "sgm_mlp_forward": (ci, [vp, vp, vp, cl, vp, vp]),
    "sgm_mlp_guess": (ci, [vp, vp, cl, cl, vp, vp, vp, vp, vp, vp, vp, vp, vp, vp]),
    "sgm_plan_execute_guided": (ci, [vp, vp, vp, vp, vp, vp, vp, vp, vp, vp,
                                     C.POINTER(MixParams), PteStatus, vp, vp]),
    "sgm_plan_execute_learned": (ci, [vp, vp, vp, vp, vp, vp, vp, vp, vp, vp, vp,
                                      C.POINTER(MixParams), PteStatus, vp, vp]),
    "sgm_server_create": (ci, [vp, C.POINTER(ServerConfig), C.POINTER(MixParams),
                               C.POINTER(vp)]),
    "sgm_server_submit": (ci, [vp, C.POINTER(Request), C.POINTER(C.c_uint64)]),
    "sgm_server_wait": (ci, [vp, C.c_uint64, C.POINTER(C.c_double)]),
    "sgm_server_test": (ci, [vp, C.c_uint64]),
    "sgm_server_now": (cd, [vp]),
    "sgm_server_stats_get": (ci, [vp, C.POINTER(ServerStats)]),
    "sgm_server_destroy": (ci, [vp]),
    "sgm_workload_uneven": (cl, [vp, C.c_uint64, cl, vp, vp, vp, vp, vp, vp, vp, vp, vp]),
}


def lib():
    """Loads libsgm.so once; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"libsgm.so not built at {LIB_PATH}; run __graft_entry__.build()")
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in _SIGS.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def exported_symbols():
    return list(_SIGS)


def ptr(x):
    """Device/host pointer of a torch tensor, numpy array, int or None."""
    if x is None:
        return None
    if isinstance(x, int):
        return x
    if hasattr(x, "data_ptr"):
        return x.data_ptr()
    if hasattr(x, "ctypes"):
        return x.ctypes.data
    raise TypeError(type(x))


def mix_params(tight=False, **over):
    p = MixParams()
    (lib().sgm_mix_params_tight if tight else lib().sgm_mix_params_default)(C.byref(p))
    for k, v in over.items():
        setattr(p, k, v)
    return p


def check(rc, what="sgm call"):
    if rc != 0:
        raise RuntimeError(f"{what} failed with code {rc}")
    return rc


class Context:
    """A material registry on the current GPU (sgm_ctx). host_only=True makes
    a registry with no device (setters, table queries and the workload
    builders only; launching calls fail with SGM_E_NOGPU)."""

    def __init__(self, nmat, host_only=False):
        self.h = vp()
        if host_only:
            check(lib().sgm_init_host(nmat, C.byref(self.h)), "sgm_init_host")
        else:
            check(lib().sgm_init(nmat, C.byref(self.h)), "sgm_init")
        self.nmat = nmat
        self.host_only = host_only
        self._keep = []

    def close(self):
        if self.h:
            lib().sgm_finalize(self.h)
            self.h = vp()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # materials
    def set_ideal(self, m, gm1, Cv):
        return check(lib().sgm_set_IdealGas(self.h, m, gm1, Cv), "set_IdealGas")

    def set_stiff(self, m, gm1, Cv, Pinf, qq):
        return check(lib().sgm_set_StiffGas(self.h, m, gm1, Cv, Pinf, qq), "set_StiffGas")

    def set_gruneisen(self, m, C0, s1, s2, s3, G0, b, rho0, T0, P0, Cv, rho_max=0.0):
        return check(lib().sgm_set_Gruneisen(self.h, m, C0, s1, s2, s3, G0, b, rho0, T0, P0, Cv,
                                             rho_max), "set_Gruneisen")

    def set_vinet(self, m, rho0, T0, B0, BP0, A0, Cv0, E0, S0, d2to40=None):
        arr = (C.c_double * 39)(*(d2to40 if d2to40 is not None else [0.0] * 39))
        return check(lib().sgm_set_Vinet(self.h, m, rho0, T0, B0, BP0, A0, Cv0, E0, S0,
                                         C.cast(arr, vp)), "set_Vinet")

    def set_modifiers(self, m, scaled=False, shifted=False, scale=1.0, shift=0.0):
        """ScaledEOS<ShiftedEOS<...>> around material m (init_sg_* enabled/vals)."""
        en = (C.c_int * 4)(1 if scaled else 0, 1 if shifted else 0, 0, 0)
        va = (C.c_double * 6)(scale, shift, 0, 0, 0, 0)
        return check(lib().sgm_set_modifiers(self.h, m, C.cast(en, vp), C.cast(va, vp)),
                     "set_modifiers")

    def set_table(self, m, desc):
        return check(lib().sgm_set_TableRhoT(self.h, m, C.byref(desc)), "set_TableRhoT")

    def kind(self, m):
        return lib().sgm_get_kind(self.h, m)

    def table_query(self, m, which):
        import numpy as np
        n = lib().sgm_table_query(self.h, m, which, None)
        if n < 0:
            raise RuntimeError(f"table_query {n}")
        out = np.zeros(n)
        lib().sgm_table_query(self.h, m, which, out.ctypes.data)
        return out

    # vector calls
    def vec_eval(self, m, func, in1, in2, out, n=None, lam=None, lam_stride=3, stream=0):
        n = out.numel() if n is None else n
        return lib().sgm_vec_eval(self.h, m, func, ptr(in1), ptr(in2), ptr(out), n,
                                  lam_stride if lam is not None else 0, ptr(lam), stream)

    def fill_eos(self, m, rho, temp, sie, press, cv, bmod, n, output, lam=None, lam_stride=3,
                 errs=None, stream=0):
        return lib().sgm_vec_fill_eos(self.h, m, ptr(rho), ptr(temp), ptr(sie), ptr(press),
                                      ptr(cv), ptr(bmod), n, output,
                                      lam_stride if lam is not None else 0, ptr(lam), ptr(errs),
                                      stream)

    def rhoe_from_PT(self, m, P, T, rho, sie, n, lam=None, lam_stride=3, errs=None, stream=0):
        return lib().sgm_vec_rhoe_from_PT(self.h, m, ptr(P), ptr(T), ptr(rho), ptr(sie), n,
                                          lam_stride if lam is not None else 0, ptr(lam),
                                          ptr(errs), stream)

    def pte(self, method, ncell, nnz, offsets, mat_ids, rho, vfrac, sie, temp, press, vfrac_tot,
            sie_tot, Tnorm=None, lam=None, params=None, conv=None, niter=None, resid=None,
            err=None, stream=0):
        """method 0 PTESolverRhoT, 1 PTESolverPT, 4 PTESolverRhoU (sie_tot:
        per-cell specific internal energy), 2 PTESolverFixedT (sie_tot: per-cell
        temperature), 3 PTESolverFixedP (sie_tot: per-cell pressure); Tnorm:
        methods 0, 1, 4."""
        st = PteStatus(ptr(conv), ptr(niter), ptr(resid), ptr(err))
        pp = C.byref(params) if params is not None else None
        if method in (2, 3):
            f = lib().sgm_pte_fixedT if method == 2 else lib().sgm_pte_fixedP
            return f(self.h, ncell, nnz, ptr(offsets), ptr(mat_ids), ptr(rho), ptr(vfrac),
                     ptr(sie), ptr(temp), ptr(press), ptr(vfrac_tot), ptr(sie_tot), ptr(lam), pp,
                     st, stream)
        f = {0: lib().sgm_pte_rhoT, 1: lib().sgm_pte_PT, 4: lib().sgm_pte_rhoU}[method]
        return f(self.h, ncell, nnz, ptr(offsets), ptr(mat_ids), ptr(rho), ptr(vfrac), ptr(sie),
                 ptr(temp), ptr(press), ptr(vfrac_tot), ptr(sie_tot), ptr(Tnorm), ptr(lam), pp,
                 st, stream)

    def approx_temperature(self, ncell, nnz, offsets, mat_ids, rho, vfrac, u_tot, T_out,
                           Tguess=None, lam=None, stream=0):
        """ApproxTemperatureFromRhoMatU per CSR cell (u_tot: rho*e per cell)."""
        return lib().sgm_approx_temperature(self.h, ncell, nnz, ptr(offsets), ptr(mat_ids),
                                            ptr(rho), ptr(vfrac), ptr(u_tot), ptr(Tguess),
                                            ptr(lam), ptr(T_out), stream)


SG_FIELDS = ("press", "pmax", "vol", "spvol", "sie", "temp", "bmod", "dpde", "cv")
SG_FRAC = ("frac_mass", "frac_vol", "frac_ie", "frac_bmod", "frac_dpde", "frac_cv")


def get_sg_eos(ctx, input_mode, eos_offsets, offsets, arrays, cutoff, params=None,
               converged=None, stream=0, noalloc=False):
    """sgm_get_sg_eos: eos_offsets is a host sequence of 1-based context
    material indices; offsets (1-based cell indices) and every array in
    `arrays` (SG_FIELDS per cell, SG_FRAC column-major per material, the last
    three optional) are device tensors, updated in place."""
    eo = (C.c_int * len(eos_offsets))(*[int(x) for x in eos_offsets])
    ptrs = [ptr(arrays[k]) for k in SG_FIELDS] + [ptr(arrays.get(k)) for k in SG_FRAC]
    f = lib().sgm_get_sg_eos_noalloc if noalloc else lib().sgm_get_sg_eos
    return f(ctx.h, len(eos_offsets), int(offsets.numel()),
                                int(arrays["press"].numel()), int(input_mode), eo, ptr(offsets),
                                *ptrs, float(cutoff),
                                C.byref(params) if params is not None else None,
                                ptr(converged), stream)


class Plan:
    """Batched persistent PTE plan over one CSR topology (sgm_plan)."""

    def __init__(self, ctx, ncell, nnz, offsets, mat_ids, method=0, stream=0):
        self.h = vp()
        self.ctx = ctx
        check(lib().sgm_plan_create(ctx.h, ncell, nnz, ptr(offsets), ptr(mat_ids), method,
                                    C.byref(self.h), stream), "plan_create")

    def autotune(self, rho, vfrac, temp, press, vfrac_tot, sie_tot, params=None, reps=1,
                 stream=0):
        return check(lib().sgm_plan_autotune(self.h, ptr(rho), ptr(vfrac), ptr(temp), ptr(press),
                                             ptr(vfrac_tot), ptr(sie_tot),
                                             C.byref(params) if params is not None else None,
                                             reps, stream), "plan_autotune")

    def execute(self, rho, vfrac, sie, temp, press, vfrac_tot, sie_tot, Tnorm=None, lam=None,
                params=None, conv=None, niter=None, resid=None, err=None, stream=0):
        st = PteStatus(ptr(conv), ptr(niter), ptr(resid), ptr(err))
        return lib().sgm_plan_execute(self.h, ptr(rho), ptr(vfrac), ptr(sie), ptr(temp),
                                      ptr(press), ptr(vfrac_tot), ptr(sie_tot), ptr(Tnorm),
                                      ptr(lam), C.byref(params) if params is not None else None,
                                      st, stream)

    def execute_guided(self, rho, vfrac, sie, temp, press, vfrac_tot, sie_tot, guess_vfrac,
                       guess_T, params=None, conv=None, niter=None, resid=None, err=None,
                       used=None, stream=0):
        st = PteStatus(ptr(conv), ptr(niter), ptr(resid), ptr(err))
        return lib().sgm_plan_execute_guided(self.h, ptr(rho), ptr(vfrac), ptr(sie), ptr(temp),
                                             ptr(press), ptr(vfrac_tot), ptr(sie_tot),
                                             ptr(guess_vfrac), ptr(guess_T),
                                             C.byref(params) if params is not None else None,
                                             st, ptr(used), stream)

    def execute_learned(self, mlp, rho, vfrac, sie, temp, press, vfrac_tot, sie_tot,
                        guess_vfrac, guess_T, params=None, conv=None, niter=None, resid=None,
                        err=None, used=None, stream=0):
        """The network's guess and the guided solve in one graph launch."""
        st = PteStatus(ptr(conv), ptr(niter), ptr(resid), ptr(err))
        return lib().sgm_plan_execute_learned(self.h, mlp, ptr(rho), ptr(vfrac), ptr(sie),
                                              ptr(temp), ptr(press), ptr(vfrac_tot),
                                              ptr(sie_tot), ptr(guess_vfrac), ptr(guess_T),
                                              C.byref(params) if params is not None else None,
                                              st, ptr(used), stream)

    def execute_bin(self, b, rho, vfrac, sie, temp, press, vfrac_tot, sie_tot, Tnorm=None,
                    lam=None, params=None, conv=None, niter=None, resid=None, err=None,
                    stream=0, kernel_ms=None):
        """One bin's kernels without the graph. kernel_ms: a list that receives
        [persistent kernel ms, tail kernel ms] (HIP events; the call syncs)."""
        st = PteStatus(ptr(conv), ptr(niter), ptr(resid), ptr(err))
        ms = (C.c_float * 2)() if kernel_ms is not None else None
        rc = lib().sgm_plan_execute_bin(self.h, b, ptr(rho), ptr(vfrac), ptr(sie), ptr(temp),
                                        ptr(press), ptr(vfrac_tot), ptr(sie_tot), ptr(Tnorm),
                                        ptr(lam),
                                        C.byref(params) if params is not None else None, st,
                                        ms, stream)
        if kernel_ms is not None:
            kernel_ms[:] = [ms[0], ms[1]]
        return rc

    def reorder(self, cost, stream=0):
        return check(lib().sgm_plan_reorder(self.h, ptr(cost), stream), "plan_reorder")

    def set_shape(self, bin, block, cells_per_wave, grid):
        """Launch shape of one register-resident bin (as the autotuner sets it)."""
        return check(lib().sgm_plan_set_shape(self.h, bin, block, cells_per_wave, grid),
                     "plan_set_shape")

    def set_straggler_threshold(self, iterations):
        """Newton iterations after which a cell is parked for the tail kernel
        (0 = never)."""
        return check(lib().sgm_plan_set_straggler_threshold(self.h, int(iterations)),
                     "plan_set_straggler_threshold")

    def describe(self):
        buf = (C.c_long * (6 * 16))()
        n = lib().sgm_plan_describe(self.h, buf, 16)
        return [dict(nmat=buf[6 * i], ncells=buf[6 * i + 1], block=buf[6 * i + 2],
                     cells_per_wave=buf[6 * i + 3], grid=buf[6 * i + 4], start=buf[6 * i + 5])
                for i in range(min(n, 16))]

    def tail_counts(self, stream=0):
        """Per bin: cells the last execution's tail resumed in its second level."""
        buf = (C.c_int * 16)()
        n = lib().sgm_plan_tail_counts(self.h, buf, 16, stream)
        if n < 0:
            check(n, "plan_tail_counts")
        return [buf[i] for i in range(min(n, 16))]

    def close(self):
        if self.h:
            lib().sgm_plan_destroy(self.h)
            self.h = vp()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Server:
    """Serving loop over a context (sgm_server). Request arrays are HOST
    numpy arrays or CPU tensors (pinned memory copies asynchronously); they
    must stay alive until the request is done (the Server keeps a reference)."""

    def __init__(self, ctx, budget_bytes=8 << 30, nlanes=4, max_group_requests=64,
                 max_group_cells=1 << 22, plain=False, method=0, params=None):
        self.h = vp()
        self.ctx = ctx
        cfg = ServerConfig(budget_bytes, nlanes, max_group_requests, max_group_cells,
                           1 if plain else 0, method)
        check(lib().sgm_server_create(ctx.h, C.byref(cfg),
                                      C.byref(params) if params is not None else None,
                                      C.byref(self.h)), "server_create")
        self._live = {}

    @staticmethod
    def _p(x):
        if x is None:
            return None
        if hasattr(x, "data_ptr"):
            return x.data_ptr()
        return x.ctypes.data

    def submit_vec(self, m, rho, sie, P_out=None, T_out=None):
        kind = (REQ_VEC_P if P_out is not None else 0) | (REQ_VEC_T if T_out is not None else 0)
        r = Request()
        r.kind, r.matindex, r.n = kind, m, len(rho)
        r.rho, r.sie = self._p(rho), self._p(sie)
        r.P_out, r.T_out = self._p(P_out), self._p(T_out)
        return self._submit(r, (rho, sie, P_out, T_out))

    def submit_pte(self, offsets, mat_ids, rho, vfrac, temp, press, vfrac_tot, sie_tot, out,
                   converged=None):
        """out: dict with rho, vfrac, sie, temp, press host arrays [nnz]."""
        r = Request()
        r.kind, r.n = REQ_PTE, len(offsets) - 1
        r.offsets, r.mat_ids = self._p(offsets), self._p(mat_ids)
        r.rho_in, r.vfrac_in, r.temp_in, r.press_in = (self._p(x) for x in (rho, vfrac, temp,
                                                                            press))
        r.vfrac_tot, r.sie_tot = self._p(vfrac_tot), self._p(sie_tot)
        r.rho_out, r.vfrac_out, r.sie_out, r.temp_out, r.press_out = (
            self._p(out[k]) for k in ("rho", "vfrac", "sie", "temp", "press"))
        r.converged = self._p(converged)
        return self._submit(r, (offsets, mat_ids, rho, vfrac, temp, press, vfrac_tot, sie_tot,
                                out, converged))

    def _submit(self, r, keep):
        t = C.c_uint64()
        rc = lib().sgm_server_submit(self.h, C.byref(r), C.byref(t))
        if rc != 0:
            raise RuntimeError(f"server_submit failed with code {rc}")
        self._live[t.value] = keep
        return t.value

    def wait(self, ticket):
        """-> (status, (t_submit, t_start, t_done))"""
        tm = (C.c_double * 3)()
        rc = lib().sgm_server_wait(self.h, ticket, tm)
        self._live.pop(ticket, None)
        return rc, (tm[0], tm[1], tm[2])

    def done(self, ticket):
        return lib().sgm_server_test(self.h, ticket) == 1

    def now(self):
        return lib().sgm_server_now(self.h)

    def stats(self):
        st = ServerStats()
        check(lib().sgm_server_stats_get(self.h, C.byref(st)), "server_stats")
        return {k: getattr(st, k) for k, _ in ServerStats._fields_}

    def close(self):
        if self.h:
            lib().sgm_server_destroy(self.h)
            self.h = vp()
            self._live.clear()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
