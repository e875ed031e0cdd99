/*
 * fans_oracle.h -- CPU FP64 restatement of the FANS per-iteration solver path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the parity checker for the HIP
 * implementation in fans_amd/.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load it; the product path never does.
 *
 * It restates, for a single rank (no x-slab decomposition), the reference
 * algorithm of FANS 0.6.2:
 *   include/matmodel.h:104-253       B matrices, element residual, reference Ke
 *   include/material_models/ (*.h)  LinearThermal/LinearElastic (isotropic,
 *                                    triclinic), J2PlasticityNew, PseudoPlastic (linear and
 *                                    power-law hardening)
 *                                    linear hardening
 *   include/MaterialManager.h:177-205 reference stiffness
 *   include/solver.h:144-204         Green operator (fundamental solution)
 *   include/solver.h:228-270,313-385 residual assembly / element loop
 *   include/solver.h:387-452         convolution, error norms
 *   include/solver.h:707-778         homogenized stress / tangent
 *   include/solverCG.h, solverFP.h   CG (linear + secant line search), FP
 *   include/mixedBCs.h               mixed strain/stress control
 *   include/LargeStrainMechModel.h   finite strain kinematics (F, P = F S),
 *   material_models/SaintVenantKirchhoff.h, CompressibleNeoHookean.h
 *
 * Parity pins (see tests/test_oracle_*.py): numpy.fft for the FFT, brute-force
 * assembled sparse operators and scipy direct solves on tiny grids, closed
 * forms (partition of unity, homogeneous RVE, laminates, radial return), and
 * the golden macro strains embedded in the reference's test_MixedBCs.json.
 */
#ifndef FANS_ORACLE_H
#define FANS_ORACLE_H

#include <stdint.h>
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* finite element types (reader.cpp:120-128) */
enum { OR_FE_HEX8 = 0, OR_FE_HEX8R = 1, OR_FE_BBAR = 2 };
/* solver methods (setup.h:75-85) */
enum { OR_METHOD_CG = 0, OR_METHOD_FP = 1 };
/* error measures and types (solver.h:414-452) */
enum { OR_MEASURE_L1 = 0, OR_MEASURE_L2 = 1, OR_MEASURE_LINF = 2 };
enum { OR_ERR_ABSOLUTE = 0, OR_ERR_RELATIVE = 1 };

/* material model kinds; per-phase parameter lists in brackets */
enum {
    OR_MODEL_THERMAL_ISO     = 1, /* [conductivity]                                  */
    OR_MODEL_THERMAL_TRI     = 2, /* [K11 K12 K13 K22 K23 K33]                       */
    OR_MODEL_ELASTIC_ISO     = 3, /* [bulk shear]                                    */
    OR_MODEL_ELASTIC_TRI     = 4, /* [C11..C16 C22..C26 C33..C36 C44..C46 C55 C56 C66] */
    OR_MODEL_J2_NEW          = 5, /* [bulk shear yield_stress iso_hardening]         */
    OR_MODEL_PSEUDO_PLASTIC  = 6, /* [bulk shear yield_stress hardening]             */
    OR_MODEL_SVK             = 7, /* SaintVenantKirchhoff [bulk shear] (large strain) */
    OR_MODEL_NEO_HOOKE       = 8, /* CompressibleNeoHookean [bulk shear] (large strain) */
    OR_MODEL_PSEUDO_NL       = 9, /* PseudoPlasticNonLinearHardening
                                     [bulk shear yield_stress hardening_exponent eps_0] */
    OR_MODEL_J2_VISCO_LIN    = 10, /* J2ViscoPlastic_LinearIsotropicHardening [bulk shear
                                      yield_stress iso_hardening kin_hardening viscosity time_step] */
    OR_MODEL_J2_VISCO_NL     = 11  /* J2ViscoPlastic_NonLinearIsotropicHardening [the 7 above,
                                      saturation_stress saturation_exponent] */
};

typedef struct {
    int32_t n_x, n_y, n_z;
    int32_t howmany;     /* 1: thermal, 3: small-strain mechanics */
    int32_t fe_type;
    int32_t method;
    int32_t measure;
    int32_t err_type;
    int32_t n_it;
    int32_t ls_max_iter;
    double  L[3];
    double  tol;
    double  ls_tol;
    int32_t large_strain; /* mechanics with strain_type "large": n_str = 9 (F row major) */
} or_desc;

typedef struct or_solver or_solver;

int         or_model_nparams(int kind);
or_solver  *or_create(const or_desc *d);
void        or_destroy(or_solver *S);
/* ms in x-major (x,y,z) order, n_x*n_y*n_z entries */
int         or_set_microstructure(or_solver *S, const uint16_t *ms);
/* Material groups like the "materials" JSON array (MaterialManager.h:38-110):
 * group g has kind[g], nphase[g] phases listed consecutively in phase_ids and
 * nphase[g]*or_model_nparams(kind[g]) consecutive params.  ref_material
 * (n_str*n_str, row-major) may be NULL.  Builds B, Ke, C_ref and the Green
 * operator. */
int         or_set_materials(or_solver *S, int n_groups, const int32_t *kind, const int32_t *nphase,
                             const int32_t *phase_ids, const double *params, const double *ref_material);
void        or_set_threads(int n);

void        or_set_gradient(or_solver *S, const double *g0);
int         or_enable_mixed_bc(or_solver *S, int n_E, const int32_t *idx_E, int n_F, const int32_t *idx_F,
                               int n_steps, const double *F_E_path, const double *P_F_path, int step);
void        or_disable_mixed_bc(or_solver *S);
int         or_mixed_active(or_solver *S);
void        or_get_macro_gradient(or_solver *S, double *g0);

int         or_solve(or_solver *S);               /* internalSolve + update internal vars; returns iterations */
int         or_error_history(or_solver *S, double *out, int cap);
void        or_homogenized_stress(or_solver *S, double *out);
int         or_homogenized_tangent(or_solver *S, double pert, double *out);
void        or_extrapolate_displacement(or_solver *S);

void        or_get_u(or_solver *S, double *out);
void        or_set_u(or_solver *S, const double *in);
void        or_get_r(or_solver *S, double *out);
void        or_get_state(or_solver *S, double *plastic_strain, double *q, int committed);
/* J2ViscoPlastic kinematic variable psi_bar, [N][n_gp][6] (zeros for other models) */
void        or_get_kinematic_state(or_solver *S, double *psi_bar, int committed);

/* PseudoPlastic::postprocess (PseudoPlastic.h:56-64): per element the mean of
 * the Gauss-point plastic flags (0 for elements of other models) */
void        or_plastic_flag(or_solver *S, double *flag);
/* element pass of postprocess (solver.h:503-556): element averages and GP
 * values of strain/stress (any pointer may be NULL) + phase sums */
void        or_element_fields(or_solver *S, double *strain_elem, double *stress_elem,
                              double *strain_gp, double *stress_gp,
                              double *stress_avg, double *strain_avg,
                              double *phase_stress_avg, double *phase_strain_avg,
                              int32_t *phase_counts, int n_phase_slots);

/* single operators */
void        or_residual(or_solver *S, const double *u, double *r, int linear_kd);
void        or_convolve(or_solver *S, const double *r, double *out);
void        or_get_green(or_solver *S, double *packed);
void        or_get_reference_stiffness(or_solver *S, double *kref);
void        or_get_ref_element_stiffness(or_solver *S, double *ke);
void        or_get_B(or_solver *S, double *B);
int         or_get_phase_stiffness(or_solver *S, int phase, double *ke);
int         or_n_gp(or_solver *S);

/* standalone transforms, layout [x][y][z][h] <-> [x][y][kz][h] */
void        or_fft_r2c(int nx, int ny, int nz, int h, const double *in, double *out_cplx);
void        or_fft_c2r(int nx, int ny, int nz, int h, const double *in_cplx, double *out);
/* one J2Plasticity::get_sigma call (J2Plasticity.h:65-108): psi / psi_bar accumulate */
void        or_visco_point(int kind, const double *params, const double *eps, const double *ep_t, double psi_t,
                           const double *pb_t, double *sigma, double *ep, double *psi, double *pb);
/* single Gauss point stress update for closed-form checks */
void        or_material_point(int kind, const double *params, const double *eps, const double *ep_t, double q_t,
                              double *sigma, double *ep, double *q);

#ifdef __cplusplus
}
#endif
#endif
