/*
 * sgm — MI355X-native EOS evaluation and mixed-cell PTE closure, C ABI.
 *
 * Every array argument is a DEVICE pointer owned by the caller (HIP memory on
 * the context's GPU) unless the comment says "host". Calls are asynchronous on
 * the given hipStream_t (0 = the default stream); results are ready after the
 * stream is synchronised. Functions return 0 on success and a negative
 * SGM_E_* code on a rejected argument, before anything is launched. Each
 * declaration names the singularity-eos interface it replaces.
 */
#ifndef SGM_H_
#define SGM_H_

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct ihipStream_t *sgm_stream_t; /* == hipStream_t */
typedef struct sgm_ctx sgm_ctx;
typedef struct sgm_plan sgm_plan;
typedef struct sgm_server sgm_server;
typedef struct sgm_mlp sgm_mlp;

enum {
  SGM_OK = 0,
  SGM_E_ARG = -1,      /* null/invalid argument or size */
  SGM_E_MAT = -2,      /* material index out of range or not set */
  SGM_E_HIP = -3,      /* HIP runtime error */
  SGM_E_KIND = -4,     /* operation not defined for this EOS kind */
  SGM_E_NOGPU = -5,    /* no usable GPU */
  SGM_E_BUDGET = -6    /* request larger than the server's memory budget */
};

/* EOS kinds (the material models of the singularity::Variant on this path) */
enum { SGM_IDEAL = 0, SGM_STIFF = 1, SGM_GRUNEISEN = 2, SGM_VINET = 3, SGM_TABLE_RHOT = 4 };

/* thermalqs bit flags (singularity-eos/base/constants.hpp:22-35) */
enum {
  SGM_Q_DENSITY = 1 << 0,
  SGM_Q_SIE = 1 << 1,
  SGM_Q_PRESSURE = 1 << 2,
  SGM_Q_TEMPERATURE = 1 << 3,
  SGM_Q_CV = 1 << 4,
  SGM_Q_BMOD = 1 << 5
};

/* vector function selectors for sgm_vec_eval */
enum {
  SGM_F_P_RHO_SIE = 0,   /* PressureFromDensityInternalEnergy */
  SGM_F_T_RHO_SIE = 1,   /* TemperatureFromDensityInternalEnergy */
  SGM_F_SIE_RHO_T = 2,   /* InternalEnergyFromDensityTemperature */
  SGM_F_P_RHO_T = 3,     /* PressureFromDensityTemperature */
  SGM_F_CV_RHO_SIE = 4,  /* SpecificHeatFromDensityInternalEnergy */
  SGM_F_CV_RHO_T = 5,    /* SpecificHeatFromDensityTemperature */
  SGM_F_BMOD_RHO_SIE = 6,/* BulkModulusFromDensityInternalEnergy */
  SGM_F_BMOD_RHO_T = 7,  /* BulkModulusFromDensityTemperature */
  SGM_F_MIN_SIE_RHO = 8  /* MinInternalEnergyFromDensity (in2 unused) */
};

/* MixParams (singularity-eos/closure/mixed_cell_models.hpp:43-78). Fill with
 * sgm_mix_params_default() and override fields. */
typedef struct sgm_mix_params {
  double derivative_eps;
  double pte_rel_tolerance_p, pte_rel_tolerance_t, pte_rel_tolerance_e, pte_rel_tolerance_v;
  double pte_abs_tolerance_p, pte_abs_tolerance_t;
  double pte_slow_convergence_thresh;
  double pte_rel_tolerance_p_sufficient, pte_rel_tolerance_t_sufficient;
  double pte_rel_tolerance_e_sufficient, pte_rel_tolerance_v_sufficient;
  double pte_abs_tolerance_p_sufficient, pte_abs_tolerance_t_sufficient;
  double line_search_alpha, line_search_fac, vfrac_safety_fac, temperature_limit;
  double default_tguess, min_dtde, pte_small_step_thresh, pte_max_dpdv;
  int pte_max_iter_per_mat, line_search_max_iter, pte_small_step_tries;
  int pte_reduced_system_exclude_idx, iterate_t_guess, pad;
} sgm_mix_params;

/* Per-cell solver status (SolverStatus, mixed_cell_models.hpp:80-87, as SoA).
 * Any pointer may be NULL. err holds SGM_ERRB_* bits. */
typedef struct sgm_pte_status {
  int *converged;
  int *niter;
  double *residual;
  int *err;
} sgm_pte_status;
enum {
  SGM_ERRB_GRUN_MAXRHO = 1, SGM_ERRB_GRUN_PT = 2, SGM_ERRB_FILLEOS = 4,
  SGM_ERRB_LINSOLVE = 8, SGM_ERRB_SMALLSTEP = 16, SGM_ERRB_INPUT = 32
};

/* In-memory rho-T table (replaces the SP5/HDF5 file of
 * SpinerEOSDependsRhoT(filename, matid), eos_spiner_rho_temp.hpp:373-420).
 * HOST pointers. Grids are uniform in NQT log10 space; 2-D fields are
 * [nRho][nT] with T fastest; cold-curve fields are [nRho]. */
typedef struct sgm_table_desc {
  double lRhoMin, lRhoMax;
  int nRho;
  double lTMin, lTMax;
  int nT;
  double rho_offset, T_offset;
  const double *P, *sie, *bMod, *dPdRho, *dPdE, *dEdRho, *dEdT;
  const double *PCold, *sieCold, *bModCold;
  int reproducible;
  int pmin_vapor_dome;
} sgm_table_desc;

/* ---- context and materials --------------------------------------------- */
/* replaces init_sg_eos (singularity-eos/eos/singularity_eos.hpp:28) */
int sgm_init(int nmat, sgm_ctx **ctx);
/* A context with no device (no counterpart in the reference): material setters,
 * table queries and the seeded workload builders work on the host; every call
 * that would launch returns SGM_E_NOGPU. For input generation and host tests. */
int sgm_init_host(int nmat, sgm_ctx **ctx);
/* replaces finalize_sg_eos (singularity_eos.hpp:134) */
int sgm_finalize(sgm_ctx *ctx);
/* replaces init_sg_IdealGas (singularity_eos.hpp:30) */
int sgm_set_IdealGas(sgm_ctx *ctx, int matindex, double gm1, double Cv);
/* replaces init_sg_StiffGas (singularity_eos.hpp:60) */
int sgm_set_StiffGas(sgm_ctx *ctx, int matindex, double gm1, double Cv, double Pinf,
                     double qq);
/* replaces init_sg_Gruneisen (singularity_eos.hpp:36-39); rho_max <= 0 selects
 * the automatic 0.99*ComputeRhoMax of the reference constructor */
int sgm_set_Gruneisen(sgm_ctx *ctx, int matindex, double C0, double s1, double s2, double s3,
                      double G0, double b, double rho0, double T0, double P0, double Cv,
                      double rho_max);
/* replaces Vinet(rho0,T0,B0,BP0,A0,Cv0,E0,S0,expconsts) (eos_vinet.hpp:44-51);
 * d2to40: HOST array of 39 coefficients */
int sgm_set_Vinet(sgm_ctx *ctx, int matindex, double rho0, double T0, double B0, double BP0,
                  double A0, double Cv0, double E0, double S0, const double *d2to40);
/* replaces init_sg_SpinerDependsRhoT (singularity_eos.hpp:80) */
int sgm_set_TableRhoT(sgm_ctx *ctx, int matindex, const sgm_table_desc *desc);
/* Wraps material matindex (set beforehand) in the reference's modifiers, as
 * the init_sg_* functions' enabled / vals arguments do
 * (singularity_eos.hpp:30-93, singularity_eos_init_utils.hpp:25-41, 75-81):
 * enabled[0] == 1 ScaledEOS by vals[0], enabled[1] == 1 ShiftedEOS by
 * vals[1], composed as ScaledEOS<ShiftedEOS<model>> (eos/modifiers/
 * scaled_eos.hpp, shifted_eos.hpp). HOST arrays enabled[4], vals[6]; the
 * bilinear ramp (enabled[2], enabled[3]) is refused with SGM_E_KIND. Every
 * vector call, FillEos, the closures and get_sg_eos then see the modified
 * material. A later setter call on matindex clears the modifiers. */
int sgm_set_modifiers(sgm_ctx *ctx, int matindex, const int *enabled, const double *vals);
/* The context's material registry as device code reads it (no counterpart in
 * the reference, where EOS objects are copied to the device by the caller):
 * *d_registry receives the device pointer of the context's material records
 * and *n their count, for kernels written against include/sgm_device.hpp
 * (their stage_registry copies the records into LDS). Valid until the next
 * material setter or sgm_finalize. */
int sgm_device_registry(sgm_ctx *ctx, const void **d_registry, int *n);
/* host-side query: kind of a material, or SGM_E_MAT */
int sgm_get_kind(const sgm_ctx *ctx, int matindex);
/* host-side query: derived table curve (0 PMax 1 sielTMax 2 dEdTMax 3 gm1Max
 * 4 rho_at_pmin 5 lTColdCrit 6 bModCold 7 dPdECold 8 dEdTCold 9 bMod[2-D]
 * 10 scalars {PMin, rhoMax, TMax}); out is HOST, returns element count */
long sgm_table_query(const sgm_ctx *ctx, int matindex, int which, double *out);

/* ---- vector EOS calls --------------------------------------------------- */
/* replaces get_sg_PressureFromDensityInternalEnergy (singularity_eos.hpp:101-104) and
 * Variant::PressureFromDensityInternalEnergy(rhos, sies, pressures, num, lambdas)
 * (eos_variant.hpp:543-545). lambda: NULL or [n*lambda_stride] with
 * lambda[i*stride+0] = log10 rho, [+1] = log10 T (table EOS cache). */
int sgm_vec_P_from_rho_sie(sgm_ctx *ctx, int matindex, const double *rho, const double *sie,
                           double *P, long n, int lambda_stride, double *lambda,
                           sgm_stream_t stream);
/* replaces Variant::TemperatureFromDensityInternalEnergy (eos_variant.hpp:543) */
int sgm_vec_T_from_rho_sie(sgm_ctx *ctx, int matindex, const double *rho, const double *sie,
                           double *T, long n, int lambda_stride, double *lambda,
                           sgm_stream_t stream);
/* the remaining SG_EOS_VEC_2IN_1OUT vector calls (eos_base.hpp:169-230), by selector */
int sgm_vec_eval(sgm_ctx *ctx, int matindex, int func, const double *in1, const double *in2,
                 double *out, long n, int lambda_stride, double *lambda, sgm_stream_t stream);
/* replaces Variant::FillEos(rhos, temps, energies, presses, cvs, bmods, num, output,
 * lambdas) (eos_variant.hpp:652-708). Arrays not read or written for `output`
 * may be NULL. errs: NULL or [n] SGM_ERRB_* bits. A point whose output set
 * the model cannot fill (the reference's EOS_ERROR) gets SGM_ERRB_FILLEOS and
 * keeps its array values. */
int sgm_vec_fill_eos(sgm_ctx *ctx, int matindex, double *rho, double *temp, double *sie,
                     double *press, double *cv, double *bmod, long n, unsigned long output,
                     int lambda_stride, double *lambda, int *errs, sgm_stream_t stream);
/* DensityEnergyFromPressureTemperature over arrays (eos_base.hpp:630-671 and
 * overrides); rho is the in/out initial guess */
int sgm_vec_rhoe_from_PT(sgm_ctx *ctx, int matindex, const double *P, const double *T,
                         double *rho, double *sie, long n, int lambda_stride, double *lambda,
                         int *errs, sgm_stream_t stream);

/* ---- PTE closure (plain per-call path) ----------------------------------- */
/* One launch over a CSR list of mixed cells: cell c owns entries
 * [offsets[c], offsets[c+1]) of the per-(cell,material) arrays; mat_ids holds
 * context material indices; 1..64 materials per cell (1..8 one lane per cell
 * with the state in registers, 9..64 one wave per cell with the Newton system
 * in LDS; sgm_pte_rhoU, whose 2 nm equations must fit that workspace, up to
 * 32 and flags wider cells SGM_ERRB_INPUT). Replaces constructing
 * PTESolverRhoT / PTESolverPT(nmat, eos, vfrac_tot, sie_tot, rho, vfrac, sie,
 * temp, press, lambda, scratch, Tnorm, params) and calling PTESolver
 * (mixed_cell_models.hpp:922-935, 1245-1253, 2188-2291) once per cell.
 * Inputs: rho, vfrac (guess), temp and press (guesses; press used by
 * pressure-preferring EOS), vfrac_tot[c], sie_tot[c], Tnorm[c] (NULL = 0).
 * Outputs overwrite rho, vfrac, sie, temp, press. lambda: NULL or [nnz*3].
 * params: HOST pointer (NULL = defaults). */
int sgm_pte_rhoT(sgm_ctx *ctx, long ncell, long nnz, const int *offsets, const uint8_t *mat_ids,
                 double *rho, double *vfrac, double *sie, double *temp, double *press,
                 const double *vfrac_tot, const double *sie_tot, const double *Tnorm,
                 double *lambda, const sgm_mix_params *params, sgm_pte_status status,
                 sgm_stream_t stream);
int sgm_pte_PT(sgm_ctx *ctx, long ncell, long nnz, const int *offsets, const uint8_t *mat_ids,
               double *rho, double *vfrac, double *sie, double *temp, double *press,
               const double *vfrac_tot, const double *sie_tot, const double *Tnorm,
               double *lambda, const sgm_mix_params *params, sgm_pte_status status,
               sgm_stream_t stream);

/* Replaces PTESolverFixedT(nmat, eos, vfrac_tot, T_true, rho, vfrac, sie, temp,
 * press, lambda, scratch, params) + PTESolver (mixed_cell_models.hpp:1451-1669,
 * 2188-2291), as get_sg_eos_rho_t drives it (eos/get_sg_eos_rho_t.cpp:69-72),
 * once per cell: the temperature T_true[c] is given; the volume fractions
 * that equalize the material pressures at that temperature are solved for.
 * Inputs rho, vfrac (guess), vfrac_tot[c], T_true[c]; outputs rho, vfrac, sie,
 * press. temp is neither read nor written (the reference scales it by 1). */
int sgm_pte_fixedT(sgm_ctx *ctx, long ncell, long nnz, const int *offsets,
                   const uint8_t *mat_ids, double *rho, double *vfrac, double *sie, double *temp,
                   double *press, const double *vfrac_tot, const double *T_true, double *lambda,
                   const sgm_mix_params *params, sgm_pte_status status, sgm_stream_t stream);
/* Replaces PTESolverFixedP(nmat, eos, vfrac_tot, P, rho, vfrac, sie, temp,
 * press, lambda, scratch, params) + PTESolver (mixed_cell_models.hpp:1674-1921,
 * 2188-2291), as get_sg_eos_rho_p drives it (eos/get_sg_eos_rho_p.cpp:67-70):
 * the pressure P[c] is given; volume fractions and the common temperature are
 * solved for. Inputs rho, vfrac (guess), temp (guess: max(default_tguess,
 * temp) starts the temperature search), press (used by pressure-preferring
 * EOS in that search), vfrac_tot[c], P[c]; outputs rho, vfrac, sie, temp,
 * press. The reference reads two scale factors before assigning them in its
 * temperature guess (mixed_cell_models.hpp:522-537, called from :1756); here
 * both are zero there. Convergence compares the absolute pressure error with
 * pte_rel_tolerance_p and pte_abs_tolerance_p, as the reference does. */
int sgm_pte_fixedP(sgm_ctx *ctx, long ncell, long nnz, const int *offsets,
                   const uint8_t *mat_ids, double *rho, double *vfrac, double *sie, double *temp,
                   double *press, const double *vfrac_tot, const double *P, double *lambda,
                   const sgm_mix_params *params, sgm_pte_status status, sgm_stream_t stream);

/* Replaces PTESolverRhoU(nmat, eos, vfrac_tot, sie_tot, rho, vfrac, sie, temp,
 * press, lambda, scratch, Tnorm, params) + PTESolver
 * (mixed_cell_models.hpp:1924-2183, 2188-2291): the volume fractions and the
 * energy densities of every material are solved for, so that the materials
 * share one pressure and one temperature (each from T(rho, e) of its EOS).
 * Arguments as sgm_pte_rhoT; outputs overwrite rho, vfrac, sie, temp (per
 * material), press. Batched plans take method 4 (bins for 2..4 materials). */
int sgm_pte_rhoU(sgm_ctx *ctx, long ncell, long nnz, const int *offsets, const uint8_t *mat_ids,
                 double *rho, double *vfrac, double *sie, double *temp, double *press,
                 const double *vfrac_tot, const double *sie_tot, const double *Tnorm,
                 double *lambda, const sgm_mix_params *params, sgm_pte_status status,
                 sgm_stream_t stream);

/* Replaces ApproxTemperatureFromRhoMatU(nmat, eos, u_tot, rho, vfrac, Tguess,
 * lambda) (mixed_cell_models.hpp:790-847), once per CSR cell: a rough
 * temperature whose energy sum sum_m rho_m vfrac_m e_m(rho_m, T) brackets
 * u_tot[c] (the cell's internal energy density rho*e, erg/cm^3), from
 * bisection in log T over [1e-9, 1e9] K. Tguess: NULL (= 0) or [ncell];
 * lambda: NULL or [nnz*3] (table caches, updated); T_out: [ncell] (NaN for
 * a malformed cell). The usual start of PTESolverRhoT's Tnorm argument
 * (test/test_pte_2phaseVinetSn.cpp:187-196). */
int sgm_approx_temperature(sgm_ctx *ctx, long ncell, long nnz, const int *offsets,
                           const uint8_t *mat_ids, const double *rho, const double *vfrac,
                           const double *u_tot, const double *Tguess, double *lambda,
                           double *T_out, sgm_stream_t stream);

/* ---- closure driver ----------------------------------------------------- */
/* Replaces get_sg_eos(nmat, ncell, cell_dim, input_int, eos_offsets, eos,
 * offsets, press, pmax, vol, spvol, sie, temp, bmod, dpde, cv, frac_mass,
 * frac_vol, frac_ie, frac_bmod, frac_dpde, frac_cv, mass_frac_cutoff)
 * (eos/singularity_eos.hpp:114-131, eos/get_sg_eos.cpp:35-280): the
 * host-code caller of the closure. The EOS array of the reference is the
 * context: eos_offsets (HOST, [nmat]) holds 1-based context material indices.
 * Every other array is DEVICE memory: offsets [ncell] 1-based cell indices
 * into the cell_dim-long per-cell arrays; frac_* are column-major
 * [cell_dim x nmat] (frac(i, m) = frac[i + m * cell_dim]). input_int: -3
 * rho-T (PTESolverFixedT), -2 rho-P (PTESolverFixedP), -1 P-T (per-material
 * P-T inversion), 0 / 1 rho-e (PTESolverRhoT). Temperatures are in eV (K =
 * eV x 1.160451930280894026e4, get_sg_eos.hpp:68). Mass fractions are
 * normalised per cell, materials at or below mass_frac_cutoff take no part,
 * and the outputs are the reference's cell averages and per-material
 * fractions; frac_bmod / frac_dpde may be NULL and, as the reference decides
 * (get_sg_eos.cpp:61-62), frac_cv is written exactly when frac_dpde is given
 * (frac_dpde without frac_cv is rejected). Extensions: params (NULL = the
 * reference's defaults) and pte_converged (NULL or [ncell] per loop index).
 * Differences: each cell's lambda cache has storage of its own (the
 * reference places it inside the PTE solver scratch, where with three or
 * more materials it overlaps the solver's last arrays; only tabulated EOS
 * write the cache, so results without tables are the same). Cells of 9..64
 * participating materials are solved one wave per cell. The
 * context keeps the driver's work buffers and the batched plan of the last
 * call's cell topology until sgm_finalize, and reuses the plan while calls
 * bring the same topology; calls on one context are serialized, and each
 * synchronizes `stream` before it returns. */
int sgm_get_sg_eos(sgm_ctx *ctx, int nmat, int ncell, int cell_dim, int input_int,
                   const int *eos_offsets, const int *offsets, double *press, double *pmax,
                   double *vol, double *spvol, double *sie, double *temp, double *bmod,
                   double *dpde, double *cv, double *frac_mass, double *frac_vol,
                   double *frac_ie, double *frac_bmod, double *frac_dpde, double *frac_cv,
                   double mass_frac_cutoff, const sgm_mix_params *params,
                   int *pte_converged, sgm_stream_t stream);

/* No counterpart in the reference: for callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc). sgm_get_sg_eos with the same arguments and the same
 * results, which calls no hipMalloc, hipFree or allocating library routine:
 * its work arrays live in one static device buffer of 64 MiB per GPU (held
 * until the process ends), laid out for ncell x nmat participating entries
 * (about 40 B per loop index + 69 B per entry, rounded to 256 B per array:
 * e.g. 65,536 cells of 6 materials fit). Larger calls are refused with
 * SGM_E_BUDGET before anything is launched. The closure runs on the plain
 * per-cell kernels (no batched plan). Asynchronous on `stream` (no
 * synchronization); every call on one GPU shares that GPU's buffer, so calls
 * that overlap on one GPU (other streams, other threads, other contexts)
 * race: order them on one stream. */
int sgm_get_sg_eos_noalloc(sgm_ctx *ctx, int nmat, int ncell, int cell_dim, int input_int,
                           const int *eos_offsets, const int *offsets, double *press,
                           double *pmax, double *vol, double *spvol, double *sie, double *temp,
                           double *bmod, double *dpde, double *cv, double *frac_mass,
                           double *frac_vol, double *frac_ie, double *frac_bmod,
                           double *frac_dpde, double *frac_cv, double mass_frac_cutoff,
                           const sgm_mix_params *params, int *pte_converged,
                           sgm_stream_t stream);
/* No reference counterpart: frees the work buffers and the batched plan that
 * sgm_get_sg_eos keeps in the context between calls (for the 2^20-cell
 * uneven workload about 0.5 GB of straggler records plus the arenas); the
 * next call rebuilds them. This memory is outside any server's
 * hbm_budget_bytes. sgm_finalize releases it too. */
int sgm_release_driver_cache(sgm_ctx *ctx);
/* No reference counterpart: out[4] (HOST) = {calls that built the driver's
 * plan, calls that reused it (same topology), bytes of the work buffers,
 * 1 if a plan is held}. */
int sgm_driver_cache_info(sgm_ctx *ctx, long *out);

/* ---- PTE closure (batched, persistent, hipGraph path) -------------------- */
/* No reference counterpart: a plan bins the cells of a fixed CSR topology by
 * material count (2..6 run register-resident templated kernels; 2..4 for
 * RhoU; 9..64 one wave per cell) and by material signature, autotunes block
 * shape and cells-per-wave
 * per bin, and captures one hipGraph whose persistent kernel nodes solve
 * every bin. method: 0 = RhoT, 1 = PT, 2 = FixedT, 3 = FixedP (the per-cell
 * sie_tot array of execute then holds the cell temperature / pressure, and
 * Tnorm is ignored), 4 = RhoU. Results are identical to sgm_pte_rhoT /
 * sgm_pte_PT / sgm_pte_fixedT / sgm_pte_fixedP / sgm_pte_rhoU on the same
 * inputs. The plan keeps device copies of offsets and mat_ids. */
int sgm_plan_create(sgm_ctx *ctx, long ncell, long nnz, const int *offsets,
                    const uint8_t *mat_ids, int method, sgm_plan **plan, sgm_stream_t stream);
/* autotune (optional; uses the given arrays as scratch-copied input) */
int sgm_plan_autotune(sgm_plan *plan, const double *rho, const double *vfrac,
                      const double *temp, const double *press, const double *vfrac_tot,
                      const double *sie_tot, const sgm_mix_params *params, int reps,
                      sgm_stream_t stream);
int sgm_plan_execute(sgm_plan *plan, double *rho, double *vfrac, double *sie, double *temp,
                     double *press, const double *vfrac_tot, const double *sie_tot,
                     const double *Tnorm, double *lambda, const sgm_mix_params *params,
                     sgm_pte_status status, sgm_stream_t stream);
/* No reference counterpart: sets one register-resident bin's launch shape
 * (block 64/128/256 threads, cells claimed per wave >= 64, persistent grid),
 * as sgm_plan_autotune would. Results do not change. */
int sgm_plan_set_shape(sgm_plan *plan, int bin, int block, int cells_per_wave, int grid);
/* No reference counterpart. Cells of the register-resident bins that are
 * still iterating after `iterations` Newton iterations are parked (their
 * solver state saved) and finished by a per-bin tail kernel whose waves hold
 * only such cells, so the slowest cells neither stall the bulk of the bin nor
 * hold its compute units. Results do not change. 0 disables; default 20. */
int sgm_plan_set_straggler_threshold(sgm_plan *plan, int iterations);
/* one bin's kernels without the graph (per-kernel timing and profiling).
 * kernel_ms (may be NULL): receives the durations of the bin's persistent
 * kernel and of its straggler tail kernel (0 without one), from HIP events
 * recorded on `stream` around each launch; the call then synchronizes. */
int sgm_plan_execute_bin(sgm_plan *plan, int bin, double *rho, double *vfrac, double *sie,
                         double *temp, double *press, const double *vfrac_tot,
                         const double *sie_tot, const double *Tnorm, double *lambda,
                         const sgm_mix_params *params, sgm_pte_status status,
                         float *kernel_ms, sgm_stream_t stream);
/* Reorders cells within each material count so that cells whose cost (e.g.
 * the niter status of a previous execute) reaches 4x the mean come first:
 * the slowest cells start with the bulk instead of trailing it. Results do
 * not depend on the order. Synchronizes the stream. */
int sgm_plan_reorder(sgm_plan *plan, const int *cost, sgm_stream_t stream);
/* host: per-bin description, out[6*i..] = {nmat, ncells, block, cells_per_wave, grid, start} */
int sgm_plan_describe(const sgm_plan *plan, long *out, int max_bins);
/* No reference counterpart (diagnostics of the straggler tails). After an
 * execution on `stream`: out[i] = the number of bin i's cells that its tail
 * resumed in the second level (still iterating after the first level), 0 for
 * bins without parking. Synchronizes the stream; returns the bin count. */
int sgm_plan_tail_counts(const sgm_plan *plan, int *out, int max_bins, sgm_stream_t stream);
int sgm_plan_destroy(sgm_plan *plan);


/* ---- serving loop ---------------------------------------------------------
 * No reference counterpart. A server keeps the context's EOS objects and
 * tables resident on the device and serves closure and vector requests from
 * many clients (mesh blocks). Requests name HOST arrays (pinned memory gives
 * asynchronous copies); the client keeps them alive until the request is
 * done. A scheduler thread admits queued requests in arrival order while
 * their device working set fits the HBM budget, merges the admitted ones into
 * a launch group (one merged CSR for the closures, binned by material count
 * and signature on the device), and runs up to `nlanes` groups at once on
 * separate streams, so requests that arrive while others run are started as
 * soon as a lane and memory free up. Results equal the plain calls'
 * bit for bit, alone or in any mix. */
typedef struct {
  size_t hbm_budget_bytes; /* device arena for request working sets (>= 64 MiB) */
  int nlanes;              /* launch groups in flight, 1..16 */
  int max_group_requests;  /* requests merged into one launch group, >= 1 */
  long max_group_cells;    /* points + cells per group (a larger request runs alone) */
  int plain;               /* 1: one request at a time through the plain per-call kernels */
  int method;              /* closure formulation: 0 PTESolverRhoT, 1 PTESolverPT */
} sgm_server_config;

enum { SGM_REQ_VEC_P = 1, SGM_REQ_VEC_T = 2, SGM_REQ_PTE = 4 };

typedef struct {
  int kind;      /* SGM_REQ_VEC_P and/or SGM_REQ_VEC_T, or SGM_REQ_PTE */
  int matindex;  /* vector requests: context material */
  long n;        /* points (vector) or cells (closure) */
  /* vector: P and/or T from (rho, sie): host rho[n], sie[n] -> P_out[n], T_out[n] */
  const double *rho, *sie;
  double *P_out, *T_out;
  /* closure, like sgm_pte_rhoT: host offsets[n+1] (any starting value; the
   * cell's entries are offsets[c]-offsets[0] .. offsets[c+1]-offsets[0] of the
   * per-entry arrays below), mat_ids/rho_in/vfrac_in/temp_in/press_in [nnz],
   * vfrac_tot/sie_tot [n]; results to rho_out/vfrac_out/sie_out/temp_out/
   * press_out [nnz] and converged [n] (may be NULL) */
  const int *offsets;
  const uint8_t *mat_ids;
  const double *rho_in, *vfrac_in, *temp_in, *press_in, *vfrac_tot, *sie_tot;
  double *rho_out, *vfrac_out, *sie_out, *temp_out, *press_out;
  int *converged;
} sgm_request;

typedef struct {
  long submitted, completed, failed;
  long groups;               /* launch groups run */
  double cells;              /* points + cells completed */
  size_t arena_peak_bytes;   /* largest arena use seen */
  size_t arena_bytes;
} sgm_server_stats;

int sgm_server_create(sgm_ctx *ctx, const sgm_server_config *cfg, const sgm_mix_params *params,
                      sgm_server **server);
/* Thread-safe, non-blocking. SGM_E_BUDGET if the request alone can never fit. */
int sgm_server_submit(sgm_server *server, const sgm_request *req, uint64_t *ticket);
/* Blocks until the request is done: its status (SGM_OK or an error code).
 * times (may be NULL): {submit, start, done} in seconds on the server's clock. */
int sgm_server_wait(sgm_server *server, uint64_t ticket, double *times);
/* Non-blocking: 1 done, 0 pending, < 0 unknown ticket. */
int sgm_server_test(sgm_server *server, uint64_t ticket);
/* seconds on the server's clock (steady) */
double sgm_server_now(const sgm_server *server);
int sgm_server_stats_get(sgm_server *server, sgm_server_stats *out);
/* drains the queue and stops the scheduler */
int sgm_server_destroy(sgm_server *server);



/* ---- learned initial guess ------------------------------------------------
 * No reference counterpart. A 64 -> 128 -> 128 -> 128 -> 16 MLP (GELU; BF16
 * weights, FP32 accumulation on the matrix cores) predicts a cell's
 * equilibrium volume fractions (masked softmax over its materials) and ln T
 * from its inputs. Features per material slot s < 6: one-hot(material id),
 * alpha, ln rho, ln T; per cell: asinh(sie_tot / 1e10), ln(rho_total). Cells
 * with other material counts (or ids >= 6) get no guess (guess_T = NaN).
 * Training runs on the device (forward, loss, backward, Adam kernels). */
int sgm_mlp_create(uint64_t seed, sgm_mlp **mlp);
int sgm_mlp_destroy(sgm_mlp *mlp);
int sgm_mlp_nparam(void);
/* host copies of the FP32 master parameters: out and/or in may be NULL */
int sgm_mlp_params(sgm_mlp *mlp, float *out, const float *in);
/* feature standardization (host arrays [64]) and the ln T standardization */
int sgm_mlp_set_norm(sgm_mlp *mlp, const float *mu, const float *sigma, float lnT_mu,
                     float lnT_sigma);
/* features of CSR cells: raw_out [ncell*64] FP32 (may be NULL), X_out
 * [ncell*64] bf16 standardized (may be NULL), nmat_out [ncell] (0 = no guess) */
int sgm_mlp_features(sgm_mlp *mlp, long ncell, long nnz, const int *offsets,
                     const uint8_t *mat_ids, const double *rho, const double *vfrac,
                     const double *temp, const double *sie_tot, float *raw_out, void *X_out,
                     int *nmat_out, sgm_stream_t stream);
/* training targets [ncell*8] from a solved closure (alphas, standardized ln T, nmat) */
int sgm_mlp_targets(sgm_mlp *mlp, long ncell, long nnz, const int *offsets,
                    const double *vfrac, const double *temp, const double *vfrac_tot,
                    const int *nmat, const int *converged, float *targets, sgm_stream_t stream);
/* loss gradient over rows idx[0..batch) of X/targets (no update); grad_out
 * [nparam], loss_out [batch] (either may be NULL) */
int sgm_mlp_gradients(sgm_mlp *mlp, const void *X, const float *targets, const int *idx,
                      long batch, float wT, float *grad_out, float *loss_out,
                      sgm_stream_t stream);
/* one Adam step (lr) on the same loss */
int sgm_mlp_train_step(sgm_mlp *mlp, const void *X, const float *targets, const int *idx,
                       long batch, float wT, float lr, float *loss_out, sgm_stream_t stream);
/* network outputs [batch*16] for rows idx of X */
int sgm_mlp_forward(sgm_mlp *mlp, const void *X, const int *idx, long batch, float *Y_out,
                    sgm_stream_t stream);
/* guesses for a closure call: guess_vfrac [nnz], guess_T [ncell] */
int sgm_mlp_guess(sgm_mlp *mlp, sgm_ctx *ctx, long ncell, long nnz, const int *offsets,
                  const uint8_t *mat_ids, const double *rho, const double *vfrac,
                  const double *temp, const double *vfrac_tot, const double *sie_tot,
                  double *guess_vfrac, double *guess_T, sgm_stream_t stream);
/* sgm_plan_execute starting each cell from the guess where it lowers the
 * initial residual of the closure, from the solver's own guess elsewhere
 * (cells with guess_T NaN or <= 0 always use their own) */
int sgm_plan_execute_guided(sgm_plan *plan, double *rho, double *vfrac, double *sie,
                            double *temp, double *press, const double *vfrac_tot,
                            const double *sie_tot, const double *guess_vfrac,
                            const double *guess_T, const sgm_mix_params *params,
                            sgm_pte_status status, int *used_guess, sgm_stream_t stream);
/* No reference counterpart: sgm_mlp_guess and sgm_plan_execute_guided as one
 * launch -- the plan's graph runs the network's guess kernel (into the
 * caller's guess_vfrac [nnz] / guess_T [ncell]) and then the guided bins.
 * Results equal sgm_mlp_guess followed by sgm_plan_execute_guided. RhoT and
 * PT plans. */
int sgm_plan_execute_learned(sgm_plan *plan, sgm_mlp *mlp, double *rho, double *vfrac,
                             double *sie, double *temp, double *press, const double *vfrac_tot,
                             const double *sie_tot, double *guess_vfrac, double *guess_T,
                             const sgm_mix_params *params, sgm_pte_status status,
                             int *used_guess, sgm_stream_t stream);

/* ---- defaults and misc -------------------------------------------------- */
void sgm_mix_params_default(sgm_mix_params *p);
/* MixParams of singularity-eos/example/pte_2mat.cpp:169-180 (1e-10 tolerances) */
void sgm_mix_params_tight(sgm_mix_params *p);
int sgm_device_info(int *ncu, int *arch_gfx, double *hbm_gbs);
const char *sgm_version(void);

/* ---- synthetic workloads (host memory; deterministic mt19937_64 streams) - */
/* Gruneisen-sampled rho-T table on NQT log grids; fields as sgm_table_desc,
 * caller-provided HOST buffers [nRho*nT] and [nRho]. */
int sgm_build_gruneisen_table(double C0, double s1, double s2, double s3, double G0, double b,
                              double rho0, double T0, double P0, double Cv, double lRhoMin,
                              double lRhoMax, int nRho, double lTMin, double lTMax, int nT,
                              double *P, double *sie, double *bMod, double *dPdRho,
                              double *dPdE, double *dEdRho, double *dEdT, double *PCold,
                              double *sieCold, double *bModCold);
/* Two-material closure inputs (IdealGas A, StiffGas B), HOST arrays:
 * rho/vfrac/sie/temp/press [2n], sie_tot [n]. */
int sgm_workload_pte2(uint64_t seed, long n, double *rho, double *vfrac, double *sie,
                      double *temp, double *press, double *sie_tot);
/* Uneven closure inputs (2..6 materials from a pool of 6), HOST arrays sized
 * for the worst case: offsets[n+1], mat_ids[6n], rho/vfrac/sie/temp/press[6n],
 * sie_tot[n]. pool_matindex[6] maps pool slot -> context material index.
 * Returns nnz. Requires the pool materials to be set on ctx (sie/press drawn
 * from the context's EOS on the host). */
long sgm_workload_uneven(sgm_ctx *ctx, uint64_t seed, long n, const int *pool_matindex,
                         int *offsets, uint8_t *mat_ids, double *rho, double *vfrac,
                         double *sie, double *temp, double *press, double *sie_tot);

#ifdef __cplusplus
}
#endif
#endif /* SGM_H_ */
