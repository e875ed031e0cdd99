// C++ host classes of the drop-in CLI over the C ABI of include/fans_gpu.h.
//
// They keep the reference's names and call order so that runSolver
// (src/main.cpp:10-41) reads the same:
//   Matmodel<howmany, n_str>          include/matmodel.h:10-75 (n_mat, n_gp, FE_type,
//                                     setGradient, macroscale_loading; the
//                                     constitutive update itself runs on the GPU)
//   MaterialManager<howmany, n_str>   include/MaterialManager.h (models, all_linear,
//                                     set_gradient, phases)
//   Solver<howmany, n_str>            include/solver.h:11-103 (solve, internalSolve,
//                                     postprocess, enableMixedBC,
//                                     extrapolateDisplacement,
//                                     get_homogenized_stress / _tangent)
//   SolverCG / SolverFP, createSolver include/solverCG.h:6-34, solverFP.h:6-23,
//                                     setup.h:75-85
// One Solver owns one GPU context (one load case, like main.cpp:15-16).  Errors
// of the GPU library are raised as std::runtime_error with its message, like the
// reference's exceptions.
#pragma once

#include "fans_gpu.h"
#include "reader.h"

#include <cstdio>
#include <initializer_list>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

namespace fans_host {

inline void check(fans_gpu_ctx *ctx, int rc, const char *what)
{
    if (rc != FANS_GPU_OK)
        throw std::runtime_error(std::string(what) + ": " + fans_gpu_last_error(ctx) + " (code " +
                                 std::to_string(rc) + ")");
}

inline int model_kind(const std::string &m)
{
    if (m == "LinearThermalIsotropic")
        return FANS_MODEL_THERMAL_ISO;
    if (m == "LinearThermalTriclinic")
        return FANS_MODEL_THERMAL_TRI;
    if (m == "LinearElasticIsotropic")
        return FANS_MODEL_ELASTIC_ISO;
    if (m == "LinearElasticTriclinic")
        return FANS_MODEL_ELASTIC_TRI;
    if (m == "J2PlasticityNew_LinearIsotropicHardening")
        return FANS_MODEL_J2_NEW;
    if (m == "PseudoPlasticLinearHardening")
        return FANS_MODEL_PSEUDO_PLASTIC;
    if (m == "PseudoPlasticNonLinearHardening")
        return FANS_MODEL_PSEUDO_PLASTIC_NL;
    if (m == "SaintVenantKirchhoff")
        return FANS_MODEL_SVK;
    if (m == "CompressibleNeoHookean")
        return FANS_MODEL_NEO_HOOKE;
    if (m == "J2ViscoPlastic_LinearIsotropicHardening")
        return FANS_MODEL_J2_VISCO_LIN;
    if (m == "J2ViscoPlastic_NonLinearIsotropicHardening")
        return FANS_MODEL_J2_VISCO_NL;
    throw std::invalid_argument(m + " is not a supported matmodel");
}

inline bool model_is_linear(int kind)
{
    return kind == FANS_MODEL_THERMAL_ISO || kind == FANS_MODEL_THERMAL_TRI || kind == FANS_MODEL_ELASTIC_ISO ||
           kind == FANS_MODEL_ELASTIC_TRI;
}

template <int howmany, int n_str>
class Solver;

// Matmodel (include/matmodel.h:10-75): one entry of the input's "materials"
// array.  The host object keeps what the reference's model exposes at the
// solver boundary -- its phases (n_mat), Gauss points (n_gp from FE_type),
// macroscale_loading and linearity -- while get_sigma, the element residual
// and the internal variables live on the device (the per-phase tables built
// from `params`, in the order of the model's material_properties keys).
template <int howmany, int n_str>
class Matmodel {
  public:
    static constexpr int num_str = n_str;
    Matmodel(const Reader &reader, const MaterialGroup &g)
        : name(g.matmodel), kind(model_kind(g.matmodel)), n_mat((int) g.phases.size()),
          n_gp(reader.FE_type == "HEX8R" ? 1 : 8), FE_type(reader.FE_type), phases(g.phases), params(g.params)
    {
        // the per-(howmany, n_str) name checks of createMatmodel (setup.h:21-73)
        const bool thermal = kind == FANS_MODEL_THERMAL_ISO || kind == FANS_MODEL_THERMAL_TRI;
        const bool large   = kind == FANS_MODEL_SVK || kind == FANS_MODEL_NEO_HOOKE;
        if (howmany == 1 && !thermal)
            throw std::invalid_argument(name + " is not a valid matmodel for thermal problem");
        if (n_str == 6 && (thermal || large))
            throw std::invalid_argument(name + " is not a valid small strain material model");
        if (n_str == 9 && !large)
            throw std::invalid_argument(name + " is not a valid large strain material model");
    }
    virtual ~Matmodel() = default;
    void setGradient(std::vector<double> g0) { macroscale_loading = std::move(g0); }
    bool is_linear() const { return model_is_linear(kind); }

    std::string         name;
    int                 kind;
    int                 n_mat;  // number of phases of this model
    int                 n_gp;
    std::string         FE_type;
    std::vector<int>    phases;
    std::vector<double> params;  // [n_mat][keys]
    std::vector<double> macroscale_loading;
};

// createMatmodel / createMaterialManager (include/setup.h:21-73, 87-91): the
// model names of setup.h; an unknown name throws invalid_argument
template <int howmany, int n_str>
std::unique_ptr<Matmodel<howmany, n_str>> createMatmodel(const Reader &reader, const MaterialGroup &g)
{
    return std::make_unique<Matmodel<howmany, n_str>>(reader, g);
}

// MaterialManager (include/MaterialManager.h): material groups -> per-phase
// tables, reference stiffness and Green operator (built on the device)
template <int howmany, int n_str>
class MaterialManager {
  public:
    explicit MaterialManager(const Reader &reader) : reader_(reader)
    {
        for (const auto &g : reader.materials) {
            models.push_back(createMatmodel<howmany, n_str>(reader, g));
            all_linear = all_linear && models.back()->is_linear();
        }
    }
    // MaterialManager::set_gradient: every model gets the macroscale loading
    void set_gradient(const std::vector<double> &g0)
    {
        if ((int) g0.size() != n_str)
            throw std::invalid_argument("set_gradient: expected " + std::to_string(n_str) + " components");
        g0_ = g0;
        for (auto &m : models)
            m->setGradient(g0);
        if (solver_)
            check(solver_, fans_gpu_set_gradient(solver_, g0.data(), n_str), "set_gradient");
    }

    std::vector<std::unique_ptr<Matmodel<howmany, n_str>>> models;
    bool                                                   all_linear = true;

    bool has_kind(std::initializer_list<int> kinds) const
    {
        for (const auto &m : models)
            for (int k : kinds)
                if (m->kind == k)
                    return true;
        return false;
    }
    bool has_j2() const { return has_kind({FANS_MODEL_J2_NEW, FANS_MODEL_J2_VISCO_LIN, FANS_MODEL_J2_VISCO_NL}); }
    bool has_visco() const { return has_kind({FANS_MODEL_J2_VISCO_LIN, FANS_MODEL_J2_VISCO_NL}); }
    bool has_pseudo_plastic() const { return has_kind({FANS_MODEL_PSEUDO_PLASTIC, FANS_MODEL_PSEUDO_PLASTIC_NL}); }

  private:
    friend class Solver<howmany, n_str>;
    void attach(fans_gpu_ctx *ctx)
    {
        solver_ = ctx;
        std::vector<int32_t> kind, nph, pids;
        std::vector<double>  params;
        for (const auto &m : models) {
            kind.push_back(m->kind);
            nph.push_back((int32_t) m->n_mat);
            for (int p : m->phases)
                pids.push_back(p);
            params.insert(params.end(), m->params.begin(), m->params.end());
        }
        check(ctx,
              fans_gpu_set_materials(ctx, (int32_t) kind.size(), kind.data(), nph.data(), pids.data(), params.data(),
                                     (int64_t) params.size(),
                                     reader_.has_ref_material ? reader_.ref_material.data() : nullptr),
              "MaterialManager");
        if (!g0_.empty())
            check(ctx, fans_gpu_set_gradient(ctx, g0_.data(), n_str), "set_gradient");
    }
    const Reader       &reader_;
    fans_gpu_ctx       *solver_ = nullptr;
    std::vector<double> g0_;
};

template <int howmany, int n_str>
class Solver {
  public:
    // method: FANS_METHOD_CG or FANS_METHOD_FP (SolverCG / SolverFP below)
    Solver(Reader &reader, MaterialManager<howmany, n_str> &mat, int precision = FANS_PREC_FP64,
           int method = -1)
        : reader_(reader), mat_(mat)
    {
        fans_gpu_desc d{};
        d.n_x = reader.dims[0], d.n_y = reader.dims[1], d.n_z = reader.dims[2];
        d.howmany      = howmany;
        d.large_strain = n_str == 9;
        d.fe_type     = reader.FE_type == "HEX8R" ? FANS_FE_HEX8R : (reader.FE_type == "BBAR" ? FANS_FE_BBAR
                                                                                                : FANS_FE_HEX8);
        d.method      = method >= 0 ? method : (reader.method == "fp" ? FANS_METHOD_FP : FANS_METHOD_CG);
        d.measure     = reader.err_measure == "L1" ? FANS_MEASURE_L1
                        : reader.err_measure == "L2" ? FANS_MEASURE_L2
                                                     : FANS_MEASURE_LINF;
        d.err_type    = reader.err_type == "relative" ? FANS_ERR_RELATIVE : FANS_ERR_ABSOLUTE;
        d.n_it        = reader.n_it;
        d.ls_max_iter = reader.ls_max_iter;
        d.ls_tol      = reader.ls_tol;
        d.tol         = reader.TOL;
        for (int i = 0; i < 3; ++i)
            d.L[i] = reader.L[i];
        d.precision = precision;
        const int rc = fans_gpu_create(&ctx_, &d);
        if (rc != FANS_GPU_OK) {
            const std::string msg = rc == FANS_GPU_ENODEV ? "no HIP device"
                                    : ctx_                ? fans_gpu_last_error(ctx_)
                                                          : "context creation failed";
            fans_gpu_destroy(ctx_);
            ctx_ = nullptr;
            throw std::runtime_error("Solver: " + msg + " (code " + std::to_string(rc) + ")");
        }
        check(ctx_, fans_gpu_set_microstructure(ctx_, reader.ms.data(), (int64_t) reader.ms.size()),
              "set_microstructure");
        mat_.attach(ctx_);
    }
    virtual ~Solver() { fans_gpu_destroy(ctx_); }
    Solver(const Solver &)            = delete;
    Solver &operator=(const Solver &) = delete;

    // Solver::solve (solver.h:282-300): internalSolve, then the internal
    // variables are committed -- on the device, inside fans_gpu_solve
    void solve()
    {
        internalSolve();
        std::printf("# Complete FANS - Number of iterations: %d (%.3f ms on the GPU)\n", iter, stats_.solve_ms);
    }
    // the iteration of the context's method (SolverCG / SolverFP)
    virtual void internalSolve()
    {
        check(ctx_, fans_gpu_solve(ctx_, &stats_), "solve");
        iter = stats_.iterations;
    }
    bool isMixedBCActive() const { return mixed_; }
    void enableMixedBC(const MixedBC &m, size_t step)
    {
        check(ctx_,
              fans_gpu_enable_mixed_bc(ctx_, (int32_t) m.idx_E.size(), m.idx_E.data(), (int32_t) m.idx_F.size(),
                                       m.idx_F.data(), m.n_steps, m.strain_path.data(), m.stress_path.data(),
                                       (int32_t) step),
              "enableMixedBC");
        mixed_ = true;
    }
    void disableMixedBC()
    {
        check(ctx_, fans_gpu_disable_mixed_bc(ctx_), "disableMixedBC");
        mixed_ = false;
    }
    void extrapolateDisplacement() { check(ctx_, fans_gpu_extrapolate_displacement(ctx_), "extrapolateDisplacement"); }
    std::vector<double> get_homogenized_stress()
    {
        std::vector<double> s(n_str);
        check(ctx_, fans_gpu_homogenized_stress(ctx_, s.data(), n_str), "get_homogenized_stress");
        return s;
    }
    std::vector<double> get_homogenized_tangent(double pert)
    {
        std::vector<double> C(n_str * n_str);
        check(ctx_, fans_gpu_homogenized_tangent(ctx_, pert, C.data(), n_str), "get_homogenized_tangent");
        return C;
    }
    std::vector<double> error_history()
    {
        int32_t             n = 0;
        std::vector<double> h((size_t) reader_.n_it + 2);
        check(ctx_, fans_gpu_error_history(ctx_, h.data(), (int32_t) h.size(), &n), "error_history");
        h.resize(n);
        return h;
    }

    // Solver::postprocess (solver.h:454-705) + the model postprocess of
    // J2PlasticityNew (J2PlasticityNew.h:155-208)
    void postprocess(Reader &reader, int load_idx, int time_idx);

    int iter = 0;

  protected:
    Reader                          &reader_;
    MaterialManager<howmany, n_str> &mat_;
    fans_gpu_ctx                    *ctx_ = nullptr;
    fans_gpu_stats                   stats_{};
    bool                             mixed_ = false;
};

// SolverCG (solverCG.h:6-34): Green-preconditioned CG, linear fast path and
// secant line search for nonlinear models
template <int howmany, int n_str>
class SolverCG : public Solver<howmany, n_str> {
  public:
    SolverCG(Reader &reader, MaterialManager<howmany, n_str> &mat, int precision = FANS_PREC_FP64)
        : Solver<howmany, n_str>(reader, mat, precision, FANS_METHOD_CG)
    {
    }
};

// SolverFP (solverFP.h:6-23): the fixed-point (Moulinec-Suquet) iteration
template <int howmany, int n_str>
class SolverFP : public Solver<howmany, n_str> {
  public:
    SolverFP(Reader &reader, MaterialManager<howmany, n_str> &mat, int precision = FANS_PREC_FP64)
        : Solver<howmany, n_str>(reader, mat, precision, FANS_METHOD_FP)
    {
    }
};

// createSolver (setup.h:75-85)
template <int howmany, int n_str>
std::unique_ptr<Solver<howmany, n_str>> createSolver(Reader &reader, MaterialManager<howmany, n_str> &mat,
                                                     int precision = FANS_PREC_FP64)
{
    if (reader.method == "fp")
        return std::make_unique<SolverFP<howmany, n_str>>(reader, mat, precision);
    if (reader.method == "cg")
        return std::make_unique<SolverCG<howmany, n_str>>(reader, mat, precision);
    throw std::invalid_argument(reader.method + " is not a valid method");
}

template <int howmany, int n_str>
void Solver<howmany, n_str>::postprocess(Reader &reader, int load_idx, int time_idx)
{
    const int     n_gp = reader.FE_type == "HEX8R" ? 1 : 8;
    const size_t  N    = reader.ms.size();
    const int     nm   = reader.n_mat;
    const bool    w_s = reader.wants("stress"), w_e = reader.wants("strain");
    const bool    w_sg = reader.wants("stress_gp"), w_eg = reader.wants("strain_gp");
    std::vector<double>  strain_elem(w_e ? N * n_str : 0), stress_elem(w_s ? N * n_str : 0);
    std::vector<double>  strain_gp(w_eg ? N * n_gp * n_str : 0), stress_gp(w_sg ? N * n_gp * n_str : 0);
    std::vector<double>  sa(n_str), ea(n_str), psa((size_t) nm * n_str), pea((size_t) nm * n_str);
    std::vector<int32_t> pc(nm);
    // the element pass runs only when a result needs it (solver.h:460-479):
    // otherwise the averages stay zero, as in the reference, and no material
    // evaluation happens (which the J2ViscoPlastic state would count)
    bool need_compute = w_s || w_e || w_sg || w_eg || reader.wants("stress_average") || reader.wants("strain_average");
    for (int m = 0; m < nm && !need_compute; ++m) {
        char nmn[64];
        std::snprintf(nmn, sizeof(nmn), "phase_stress_average_phase%d", m);
        need_compute = reader.wants(nmn);
    }
    if (need_compute)
        check(ctx_,
              fans_gpu_postprocess(ctx_, w_e ? strain_elem.data() : nullptr, w_s ? stress_elem.data() : nullptr,
                                   w_eg ? strain_gp.data() : nullptr, w_sg ? stress_gp.data() : nullptr, sa.data(),
                                   ea.data(), psa.data(), pea.data(), pc.data(), nm),
              "postprocess");
    std::printf("# Effective Stress .. (");
    for (int i = 0; i < n_str; ++i)
        std::printf("%+.12f ", sa[i]);
    std::printf(") \n# Effective Strain .. (");
    for (int i = 0; i < n_str; ++i)
        std::printf("%+.12f ", ea[i]);
    std::printf(") \n\n");

    std::vector<double> u(N * howmany), r(N * howmany);
    check(ctx_, fans_gpu_get_field(ctx_, FANS_FIELD_U, u.data(), (int64_t) u.size()), "get_field u");
    check(ctx_, fans_gpu_get_field(ctx_, FANS_FIELD_R, r.data(), (int64_t) r.size()), "get_field r");
    // u_total = g0 . X + u (solver.h:586-648), X relative to the cell centre
    std::vector<double> ut(N * howmany);
    const double        dx = reader.l_e[0], dy = reader.l_e[1], dz = reader.l_e[2];
    const double        Lx2 = reader.L[0] / 2, Ly2 = reader.L[1] / 2, Lz2 = reader.L[2] / 2;
    const double        rs2 = 0.7071067811865475;
    size_t              n   = 0;
    for (int ix = 0; ix < reader.dims[0]; ++ix) {
        const double x = ix * dx - Lx2;
        for (int iy = 0; iy < reader.dims[1]; ++iy) {
            const double y = iy * dy - Ly2;
            for (int iz = 0; iz < reader.dims[2]; ++iz, ++n) {
                const double z = iz * dz - Lz2;
                if (howmany == 3 && n_str == 9) {
                    // u = (F - I) X with F = strain_average (solver.h:623-638)
                    for (int i = 0; i < 3; ++i)
                        ut[3 * n + i] = u[3 * n + i] + (ea[3 * i] - (i == 0)) * x + (ea[3 * i + 1] - (i == 1)) * y +
                                        (ea[3 * i + 2] - (i == 2)) * z;
                } else if (howmany == 3) {
                    const double g11 = ea[0], g22 = ea[1], g33 = ea[2];
                    const double g12 = ea[3] * rs2, g13 = ea[4] * rs2, g23 = ea[5] * rs2;
                    ut[3 * n]     = u[3 * n] + g11 * x + g12 * y + g13 * z;
                    ut[3 * n + 1] = u[3 * n + 1] + g12 * x + g22 * y + g23 * z;
                    ut[3 * n + 2] = u[3 * n + 2] + g13 * x + g23 * y + g33 * z;
                } else {
                    ut[n] = u[n] + ea[0] * x + ea[1] * y + ea[2] * z;
                }
            }
        }
    }
    hsize_t d1[1] = {(hsize_t) n_str};
    reader.writeData("stress_average", load_idx, time_idx, sa.data(), d1, 1);
    reader.writeData("strain_average", load_idx, time_idx, ea.data(), d1, 1);
    for (int m = 0; m < nm; ++m) {
        char sn[64], en[64];
        std::snprintf(sn, sizeof(sn), "phase_stress_average_phase%d", m);
        std::snprintf(en, sizeof(en), "phase_strain_average_phase%d", m);
        reader.writeData(sn, load_idx, time_idx, &psa[(size_t) m * n_str], d1, 1);
        reader.writeData(en, load_idx, time_idx, &pea[(size_t) m * n_str], d1, 1);
    }
    const std::vector<double> err = error_history();
    hsize_t                   de[1] = {(hsize_t) err.size()};
    reader.writeData("absolute_error", load_idx, time_idx, err.data(), de, 1);

    const std::vector<int> rank_field(N, 0);
    reader.writeSlab("mpi_rank", load_idx, time_idx, rank_field.data(), {1});
    reader.writeSlab("microstructure", load_idx, time_idx, reader.ms.data(), {1});
    reader.writeSlab("displacement_fluctuation", load_idx, time_idx, u.data(), {howmany});
    reader.writeSlab("displacement", load_idx, time_idx, ut.data(), {howmany});
    // compact [x][y][z][h] layout (the reference writes its padded buffer as if
    // compact, which scrambles this field; the compact layout is kept here)
    reader.writeSlab("residual", load_idx, time_idx, r.data(), {howmany});
    if (w_e)
        reader.writeSlab("strain", load_idx, time_idx, strain_elem.data(), {n_str});
    if (w_s)
        reader.writeSlab("stress", load_idx, time_idx, stress_elem.data(), {n_str});
    if (w_eg)
        reader.writeSlab("strain_gp", load_idx, time_idx, strain_gp.data(), {n_gp, n_str});
    if (w_sg)
        reader.writeSlab("stress_gp", load_idx, time_idx, stress_gp.data(), {n_gp, n_str});

    // J2PlasticityNew::postprocess: committed state, element means and per GP
    if (howmany == 3 && mat_.has_j2()) {
        const bool w_p = reader.wants("plastic_strain"), w_pg = reader.wants("plastic_strain_gp");
        const bool w_q = reader.wants("isotropic_hardening_variable"),
                   w_qg = reader.wants("isotropic_hardening_variable_gp");
        if (w_p || w_pg || w_q || w_qg) {
            std::vector<double> ep(N * n_gp * 6), q(N * n_gp);
            check(ctx_, fans_gpu_get_field(ctx_, FANS_FIELD_PLASTIC_STRAIN, ep.data(), (int64_t) ep.size()),
                  "get_field plastic_strain");
            check(ctx_, fans_gpu_get_field(ctx_, FANS_FIELD_HARDENING, q.data(), (int64_t) q.size()),
                  "get_field hardening");
            if (w_p || w_q) {
                std::vector<double> pm(N * 6, 0.0), qm(N, 0.0);
                for (size_t e = 0; e < N; ++e)
                    for (int g = 0; g < n_gp; ++g) {
                        for (int c = 0; c < 6; ++c)
                            pm[e * 6 + c] += ep[(e * n_gp + g) * 6 + c] / n_gp;
                        qm[e] += q[e * n_gp + g] / n_gp;
                    }
                if (w_p)
                    reader.writeSlab("plastic_strain", load_idx, time_idx, pm.data(), {6});
                if (w_q)
                    reader.writeSlab("isotropic_hardening_variable", load_idx, time_idx, qm.data(), {1});
            }
            if (w_pg)
                reader.writeSlab("plastic_strain_gp", load_idx, time_idx, ep.data(), {n_gp, 6});
            if (w_qg)
                reader.writeSlab("isotropic_hardening_variable_gp", load_idx, time_idx, q.data(), {n_gp});
        }
        // J2Plasticity::postprocess (J2Plasticity.h:246-318): the kinematic
        // variable psi_bar as well
        const bool w_k = reader.wants("kinematic_hardening_variable"),
                   w_kg = reader.wants("kinematic_hardening_variable_gp");
        if (mat_.has_visco() && (w_k || w_kg)) {
            std::vector<double> kb(N * n_gp * 6);
            check(ctx_, fans_gpu_get_field(ctx_, FANS_FIELD_KIN_HARDENING, kb.data(), (int64_t) kb.size()),
                  "get_field kinematic_hardening");
            if (w_k) {
                std::vector<double> km(N * 6, 0.0);
                for (size_t e = 0; e < N; ++e)
                    for (int g = 0; g < n_gp; ++g)
                        for (int c = 0; c < 6; ++c)
                            km[e * 6 + c] += kb[(e * n_gp + g) * 6 + c] / n_gp;
                reader.writeSlab("kinematic_hardening_variable", load_idx, time_idx, km.data(), {6});
            }
            if (w_kg)
                reader.writeSlab("kinematic_hardening_variable_gp", load_idx, time_idx, kb.data(), {n_gp, 6});
        }
    }

    // PseudoPlastic::postprocess (PseudoPlastic.h:56-64): element means of the
    // Gauss-point plastic flags at the converged u, as float
    if (howmany == 3 && n_str == 6 && mat_.has_pseudo_plastic() && reader.wants("plastic_flag")) {
        std::vector<double> fl(N);
        check(ctx_, fans_gpu_get_field(ctx_, FANS_FIELD_PLASTIC_FLAG, fl.data(), (int64_t) N), "get_field plastic_flag");
        std::vector<float> ff(fl.begin(), fl.end());
        reader.writeSlab("plastic_flag", load_idx, time_idx, ff.data(), {1});
    }

    if (reader.wants("homogenized_tangent")) {
        const std::vector<double> C = get_homogenized_tangent(1e-6);
        std::printf("# Homogenized tangent: \n");
        for (int i = 0; i < n_str; ++i) {
            for (int j = 0; j < n_str; ++j)
                std::printf("%.12g ", C[i * n_str + j]);
            std::printf("\n");
        }
        std::printf("\n");
        hsize_t d2[2] = {(hsize_t) n_str, (hsize_t) n_str};
        reader.writeData("homogenized_tangent", load_idx, time_idx, C.data(), d2, 2);
        // the tangent switches the shared error parameters to relative with
        // TOL >= 1e-6 for every later solve, later load cases included
        // (solver.h:749-750); the library did so for this context already
        reader.err_type = "relative";
        reader.TOL      = reader.TOL > 1e-6 ? reader.TOL : 1e-6;
    }
}

}  // namespace fans_host
