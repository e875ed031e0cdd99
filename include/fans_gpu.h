/*
 * fans_gpu.h -- C ABI of the MI355X (gfx950) implementation of the FANS
 * per-iteration solver path.  libfans_gpu.so exports exactly these symbols.
 *
 * One context = one load case of the reference: the pair (MaterialManager,
 * Solver<howmany,n_str>) that src/main.cpp:15-16 creates per load case.  The
 * context owns every device buffer; host arrays passed in or out are caller
 * owned and copied.  All functions return 0 on success and a negative
 * FANS_GPU_E* code otherwise; fans_gpu_last_error() gives the message.  The
 * reference reports configuration errors as C++ exceptions
 * (MaterialManager.h:69-108, mixedBCs.h:63-128); no exception crosses this ABI.
 * A context is not thread safe (the reference is single threaded per rank);
 * use one host thread per context.  There is no CPU fallback: without a HIP
 * device fans_gpu_create() fails with FANS_GPU_ENODEV.
 *
 * Field layouts are the reference's in-memory layouts without ghost planes or
 * FFT padding: nodal fields [x][y][z][howmany], microstructure [x][y][z]
 * (Reader::ms after ReadMS, reader.cpp:380-399), element fields [x][y][z][n_str],
 * Gauss point fields [x][y][z][n_gp][n_str].  n_str is 3 (thermal), 6 (small
 * strain, Mandel) or 9 (finite strain: F_11 F_12 ... F_33 and P likewise).
 */
#ifndef FANS_GPU_H
#define FANS_GPU_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define FANS_GPU_ABI_VERSION 1

enum {
    FANS_GPU_OK       = 0,
    FANS_GPU_EINVAL   = -1, /* bad argument / configuration (reference: exception)   */
    FANS_GPU_ENODEV   = -2, /* no HIP device                                         */
    FANS_GPU_EHIP     = -3, /* HIP runtime error                                     */
    FANS_GPU_ESTATE   = -4, /* call out of order (e.g. solve before set_materials)   */
    FANS_GPU_EUNSUP   = -5  /* valid for the reference but not implemented here      */
};

/* FE_type (reader.cpp:120-128) */
enum { FANS_FE_HEX8 = 0, FANS_FE_HEX8R = 1, FANS_FE_BBAR = 2 };
/* method (setup.h:75-85) */
enum { FANS_METHOD_CG = 0, FANS_METHOD_FP = 1 };
/* error_parameters.measure / .type (solver.h:414-452) */
enum { FANS_MEASURE_L1 = 0, FANS_MEASURE_L2 = 1, FANS_MEASURE_LINF = 2 };
enum { FANS_ERR_ABSOLUTE = 0, FANS_ERR_RELATIVE = 1 };
/* arithmetic of the operator apply */
enum {
    FANS_PREC_FP64  = 0, /* everything FP64: the drop-in form                         */
    FANS_PREC_MIXED = 1  /* FP32 operator inside an FP64 defect-correction loop       */
};
/* material models (setup.h:21-61); per-phase parameters in brackets, in the order
 * of the reference's material_properties keys */
enum {
    FANS_MODEL_THERMAL_ISO    = 1, /* LinearThermalIsotropic [conductivity]              */
    FANS_MODEL_THERMAL_TRI    = 2, /* LinearThermalTriclinic [K_11 K_12 K_13 K_22 K_23 K_33] */
    FANS_MODEL_ELASTIC_ISO    = 3, /* LinearElasticIsotropic [bulk_modulus shear_modulus] */
    FANS_MODEL_ELASTIC_TRI    = 4, /* LinearElasticTriclinic [C_11 .. C_66, 21 values]   */
    FANS_MODEL_J2_NEW         = 5, /* J2PlasticityNew_LinearIsotropicHardening
                                      [bulk_modulus shear_modulus yield_stress
                                       isotropic_hardening_parameter]                     */
    FANS_MODEL_PSEUDO_PLASTIC = 6, /* PseudoPlasticLinearHardening
                                      [bulk_modulus shear_modulus yield_stress hardening_parameter] */
    FANS_MODEL_SVK            = 7, /* SaintVenantKirchhoff [bulk_modulus shear_modulus]
                                      (finite strain, setup.h:62-73)                       */
    FANS_MODEL_NEO_HOOKE      = 8, /* CompressibleNeoHookean [bulk_modulus shear_modulus]
                                      (finite strain)                                      */
    FANS_MODEL_PSEUDO_PLASTIC_NL = 9, /* PseudoPlasticNonLinearHardening
                                      [bulk_modulus shear_modulus yield_stress
                                       hardening_exponent eps_0]                          */
    FANS_MODEL_J2_VISCO_LIN   = 10, /* J2ViscoPlastic_LinearIsotropicHardening
                                      [bulk_modulus shear_modulus yield_stress
                                       isotropic_hardening_parameter
                                       kinematic_hardening_parameter viscosity time_step]
                                      (J2Plasticity.h:162-178; time_step repeated per phase) */
    FANS_MODEL_J2_VISCO_NL    = 11  /* J2ViscoPlastic_NonLinearIsotropicHardening
                                      [the seven above, saturation_stress saturation_exponent]
                                      (J2Plasticity.h:182-243)                             */
};
/* fields for fans_gpu_get_field / fans_gpu_set_field.  The nodal fields
 * (U, R, U_PREV) may be read into and written from host or device memory; a
 * device pointer must come from the HIP runtime the library is loaded with
 * (with PyTorch in the process: import torch first, so that both share its
 * runtime).  The Gauss-point and element fields go to host memory. */
enum {
    FANS_FIELD_U              = 0, /* displacement fluctuation v_u, double [N][h]      */
    FANS_FIELD_R              = 1, /* residual v_r, double [N][h]                      */
    FANS_FIELD_U_PREV         = 2, /* v_u_prev, double [N][h]                          */
    FANS_FIELD_PLASTIC_STRAIN = 3, /* committed plastic strain, double [N][n_gp][6]    */
    FANS_FIELD_HARDENING      = 4, /* committed isotropic hardening q, double [N][n_gp] */
    FANS_FIELD_PLASTIC_FLAG   = 5, /* PseudoPlastic::postprocess plastic_flag at the current
                                      u (PseudoPlastic.h:56-64), double [N]; get only   */
    FANS_FIELD_KIN_HARDENING  = 6  /* committed kinematic variable psi_bar of the
                                      J2ViscoPlastic models (J2Plasticity.h:150),
                                      double [N][n_gp][6]; get only                     */
};

typedef struct {
    int32_t n_x, n_y, n_z;   /* grid (Reader::dims)                                  */
    int32_t howmany;         /* 1 thermal, 3 small-strain mechanics                  */
    int32_t fe_type;         /* FANS_FE_*                                            */
    int32_t method;          /* FANS_METHOD_*                                        */
    int32_t measure;         /* FANS_MEASURE_*                                       */
    int32_t err_type;        /* FANS_ERR_*                                           */
    int32_t n_it;            /* max iterations (Reader::n_it)                        */
    int32_t ls_max_iter;     /* linesearch_parameters.max_iter (default 5)           */
    double  L[3];            /* microstructure.L                                     */
    double  tol;             /* error_parameters.tolerance                           */
    double  ls_tol;          /* linesearch_parameters.tol (default 1e-2)             */
    int32_t precision;       /* FANS_PREC_*                                          */
    int32_t device;          /* HIP device ordinal                                   */
    int32_t large_strain;    /* 1: strain_type "large" (main.cpp:71-72): n_str = 9,
                                gradient and stress are F and P, row major           */
    int32_t reserved[5];     /* must be zero                                         */
} fans_gpu_desc;

typedef struct {
    int32_t iterations;       /* CG/FP iterations of the last solve (Solver::iter)   */
    int32_t inner_iterations; /* FP32 inner iterations (mixed form), else = iterations */
    int32_t outer_iterations; /* FP64 residual refreshes (mixed form), else 0        */
    int32_t converged;        /* final error <= tolerance                            */
    double  solve_ms;         /* device time of internalSolve (HIP events)           */
    double  final_error;      /* last entry of the error history                     */
} fans_gpu_stats;

typedef struct fans_gpu_ctx fans_gpu_ctx;

int         fans_gpu_abi_version(void);
const char *fans_gpu_last_error(const fans_gpu_ctx *ctx);

/* Solver + Reader setup (solver.h:105-142, reader.cpp:227-411).  Fails with
 * FANS_GPU_ENODEV when no HIP device is present. */
int fans_gpu_create(fans_gpu_ctx **out, const fans_gpu_desc *desc);
int fans_gpu_destroy(fans_gpu_ctx *ctx);

/* Reader::ms (reader.cpp:380-399): n_x*n_y*n_z phase ids in [x][y][z] order */
int fans_gpu_set_microstructure(fans_gpu_ctx *ctx, const uint16_t *ms_xyz, int64_t count);

/* MaterialManager ctor (MaterialManager.h:38-159): n_groups material groups, group
 * g of model kind[g] owning nphase[g] phases (listed consecutively in phase_ids)
 * with nphase[g] * nparams(kind[g]) parameters.  ref_material (n_str x n_str row
 * major, MaterialManager.h:179-197) may be NULL.  Builds the per-phase tables, the
 * reference stiffness and the Green operator (solver.h:144-204) on the device. */
int fans_gpu_set_materials(fans_gpu_ctx *ctx, int32_t n_groups, const int32_t *kind, const int32_t *nphase,
                           const int32_t *phase_ids, const double *params, int64_t n_params,
                           const double *ref_material);

/* MaterialManager::set_gradient (MaterialManager.h:216-221) */
int fans_gpu_set_gradient(fans_gpu_ctx *ctx, const double *g0, int32_t n_str);
int fans_gpu_get_gradient(fans_gpu_ctx *ctx, double *g0, int32_t n_str);

/* Per-voxel isotropic moduli replacing the phase table: mechanics k = bulk and
 * g = shear modulus per voxel, thermal k = conductivity (g unused, may be NULL);
 * [x][y][z] order, n_x*n_y*n_z values each.  No counterpart in the reference,
 * whose uint16 phase ids with per-phase parameters cannot express per-voxel
 * moduli (SIMP densities, pyfans differentiation); the reference stiffness
 * follows the rules of LinearElastic.h:55-69 / LinearThermal.h:19-45 as if
 * every voxel were a phase (unless set_materials got ref_material), and the
 * Green operator is rebuilt.  Requires isotropic linear phases.  k and g may be
 * host or device arrays (as for the nodal fields); the check that every value
 * is positive and the reference-stiffness statistics run on the device. */
int fans_gpu_set_voxel_moduli(fans_gpu_ctx *ctx, const double *k, const double *g, int64_t count);

/* Sensitivity of the homogenized response to per-voxel moduli.  For the
 * context's current solution (u, g0) and a second solution ub for the macro
 * gradient g0b ([N][h], host or device, and n_str, host), writes per element
 * (sens: host or device) the Gauss
 * integrals sens[e] = (int_e eps_a^T dC/dK eps_b, int_e eps_a^T dC/dG eps_b)
 * (thermal: int_e eps_a . eps_b, one value), eps_a = B u + g0, eps_b = B ub + g0b.
 * When both solve their load cases, d(g0b . Sigma_bar(g0)) / dK_e = sens[e][0] / V
 * (V = L_x L_y L_z): the adjoint of v . Sigma_bar is the solve for macro gradient
 * v, the operator being self-adjoint.  No counterpart in the reference (pyfans
 * differentiation). */
int fans_gpu_sensitivity(fans_gpu_ctx *ctx, const double *ub, const double *g0b, double *sens, int64_t count);

/* Solver::enableMixedBC / disableMixedBC (solver.h:71-78, mixedBCs.h:181-226):
 * strain path [n_steps][n_E], stress path [n_steps][n_F] */
int fans_gpu_enable_mixed_bc(fans_gpu_ctx *ctx, int32_t n_E, const int32_t *idx_E, int32_t n_F, const int32_t *idx_F,
                             int32_t n_steps, const double *strain_path, const double *stress_path, int32_t step);
int fans_gpu_disable_mixed_bc(fans_gpu_ctx *ctx);

/* Solver::solve (solver.h:282-300): internalSolve + update_internal_variables */
int fans_gpu_solve(fans_gpu_ctx *ctx, fans_gpu_stats *stats);
/* Solver::err_all[0..iter] */
int fans_gpu_error_history(fans_gpu_ctx *ctx, double *out, int32_t cap, int32_t *count);

/* Solver::get_homogenized_stress (solver.h:707-737) */
int fans_gpu_homogenized_stress(fans_gpu_ctx *ctx, double *out, int32_t n_str);
/* Solver::get_homogenized_tangent (solver.h:739-778): out n_str x n_str row major */
int fans_gpu_homogenized_tangent(fans_gpu_ctx *ctx, double pert, double *out, int32_t n_str);
/* Solver::extrapolateDisplacement (solver.h:302-311) */
int fans_gpu_extrapolate_displacement(fans_gpu_ctx *ctx);
/* State of a freshly constructed Solver (solver.h:133-139): u = u_prev = 0,
 * internal variables zero, mixed BC off.  Keeps the Green operator and tables
 * (the reference rebuilds them per load case, main.cpp:15-16, with equal result). */
int fans_gpu_reset_state(fans_gpu_ctx *ctx);

/* element pass of Solver::postprocess (solver.h:503-574); any output may be NULL.
 * phase_* arrays have n_slots rows (phase ids 0..n_slots-1). */
int fans_gpu_postprocess(fans_gpu_ctx *ctx, double *strain_elem, double *stress_elem, double *strain_gp,
                         double *stress_gp, double *stress_avg, double *strain_avg, double *phase_stress_avg,
                         double *phase_strain_avg, int32_t *phase_counts, int32_t n_slots);

int fans_gpu_get_field(fans_gpu_ctx *ctx, int32_t which, double *host, int64_t count);
int fans_gpu_set_field(fans_gpu_ctx *ctx, int32_t which, const double *host, int64_t count);

/* single operators (Solver::compute_residual, solver.h:228-280, and
 * Solver::convolution, solver.h:387-412) on host arrays [N][h]; linear_kd=1 is
 * the linear CG matvec of solverCG.h:97-103 (no macroscopic gradient) */
int fans_gpu_residual(fans_gpu_ctx *ctx, const double *u, double *r, int32_t linear_kd);
int fans_gpu_convolve(fans_gpu_ctx *ctx, const double *r, double *out);
/* Green operator, [n_x][n_y][n_z/2+1][h(h+1)/2] packed (00 01 02 11 12 22) */
int fans_gpu_get_green(fans_gpu_ctx *ctx, double *out, int64_t count);
/* reference stiffness kapparef_mat (MaterialManager.h:177-205), n_str x n_str */
int fans_gpu_get_reference_stiffness(fans_gpu_ctx *ctx, double *out, int32_t n_str);

/* standalone 3-D r2c / c2r of a [nx][ny][nz][h] real field (unnormalised, like
 * the FFTW plans of solver.h:221-222); spectrum [nx][ny][nz/2+1][h] interleaved */
int fans_gpu_fft_r2c(int32_t device, int32_t nx, int32_t ny, int32_t nz, int32_t h, const double *in, double *out_cplx);
int fans_gpu_fft_c2r(int32_t device, int32_t nx, int32_t ny, int32_t nz, int32_t h, const double *in_cplx, double *out);
/* fans_gpu_fft_r2c without allocation -- no counterpart in the reference; for
 * callers that cannot allocate device memory (code under hipGraph capture,
 * serving loops).  `in` and `out_cplx` are device buffers of the layouts above,
 * `stream` a hipStream_t (NULL: the null stream); the call only enqueues work
 * on it.  The twiddle tables live in a file-scope device array of 16384 complex
 * doubles per device, so n_x + n_y + n_z must stay below about 16300
 * (FANS_GPU_EINVAL otherwise, message via fans_gpu_last_error(NULL)).  Calls
 * that overlap on one device (other streams or threads) race on that array. */
int fans_gpu_fft_r2c_noalloc(int32_t device, int32_t nx, int32_t ny, int32_t nz, int32_t h, const double *in,
                             double *out_cplx, void *stream);

/* Batched contexts (no reference counterpart: the reference solves one RVE per
 * run, main.cpp:10-41, and a batch of them is that run repeated).  A batched
 * context holds `slots` RVEs (1..4096) of one creation setting (grid, physics,
 * element, CG method, tolerances), each with its own microstructure, phase
 * table (at most 256 phases, linear models), Green operator, gradient and
 * solution.  The single-RVE entry points above address the selected slot
 * (fans_gpu_select_slot; slot 0 after creation), except fans_gpu_solve, which
 * solves every slot at once: each kernel launch covers every slot still
 * running, and each slot takes the iterations and gets the results of a
 * solve of that RVE alone.  Mixed boundary conditions, the tangent, per-voxel
 * moduli, the sensitivity and per-kernel timing are refused on such a context.
 * fans_gpu_solve's statistics are then the largest iteration counts and final
 * error over the slots (converged: all slots). */
int fans_gpu_create_batch(fans_gpu_ctx **out, const fans_gpu_desc *desc, int32_t slots);
int fans_gpu_select_slot(fans_gpu_ctx *ctx, int32_t slot);
/* per slot of the last solve (count = slots; any output may be NULL): CG
 * iterations (mixed form: FP32 inner iterations), converged flag, final error */
int fans_gpu_batch_results(fans_gpu_ctx *ctx, int32_t *iterations, int32_t *converged, double *final_error,
                           int32_t count);
/* Solver::get_homogenized_stress (solver.h:707-737) of every slot in one
 * launch: out [count][n_str] */
int fans_gpu_batch_homogenized_stress(fans_gpu_ctx *ctx, double *out, int32_t n_str, int32_t count);

/* measurement helper (no reference counterpart): average milliseconds of one
 * device-to-device copy of `bytes` (>= 1 MiB) by the library's own streaming
 * kernel (16 bytes per lane, contiguous chunks per workgroup), over `reps`
 * copies after one warm-up -- the attainable-bandwidth reference for the
 * streaming passes (2 * bytes moved per copy) */
int fans_gpu_stream_bandwidth(int32_t device, int64_t bytes, int32_t reps, double *ms_per_copy);

/* options (no reference counterpart): FANS_OPT_TIMING=1 brackets every launch
 * with HIP events on the context's stream (replaces the clock() counters of
 * solver.h:286-297); FANS_OPT_FFT_CHUNKS=k runs the y / x+Green / y^-1 passes
 * of the convolution on k kz-chunks in turn (k >= 1); FANS_OPT_GRAPHS=0/1
 * launches the linear CG iterations eagerly / replays them from a captured
 * hipGraph (default 1); FANS_OPT_ASM_PLANES=p sets the x planes each assembly
 * workgroup marches through (default 16, or n_x/4 below 64);
 * FANS_OPT_OPERATOR_FP32=1 makes fans_gpu_convolve and the linear matvec of
 * fans_gpu_residual (linear_kd=1) run the FP32 kernels of the mixed form's
 * inner solve (input rounded to FP32, output widened; contexts created with
 * FANS_PREC_MIXED only); FANS_OPT_FFT_LDS_KIB=k sets the LDS budget (8..64
 * KiB, default 40) of one workgroup of the generic FFT passes, i.e. their
 * lines or columns per workgroup (lengths other than the powers of two
 * 16..256, whose kernels have fixed tiles); FANS_OPT_MIXED_RESTART=1 makes a
 * single-RVE context's mixed form restart its FP32 CG in every outer round
 * (default 0: one FP32 CG whose residual is replaced by the true FP64 one at
 * each outer check, as batched contexts also do). */
enum { FANS_OPT_TIMING = 1, FANS_OPT_FFT_CHUNKS = 2, FANS_OPT_GRAPHS = 3, FANS_OPT_ASM_PLANES = 4,
       FANS_OPT_OPERATOR_FP32 = 5, FANS_OPT_FFT_LDS_KIB = 6, FANS_OPT_MIXED_RESTART = 7 };
int fans_gpu_set_option(fans_gpu_ctx *ctx, int32_t opt, int64_t value);
/* per kernel family of the last solve: out[2k] = total ms, out[2k+1] = launches */
int         fans_gpu_kernel_times(fans_gpu_ctx *ctx, double *out, int32_t cap, int32_t *count);
const char *fans_gpu_kernel_family_name(int32_t family);

#ifdef __cplusplus
}
#endif
#endif
