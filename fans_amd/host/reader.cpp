// Reader of the drop-in CLI (see reader.h for the reference lines followed).
#include "reader.h"

#include "h5dyn.h"

#include <json.hpp>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <map>
#include <stdexcept>

using json = nlohmann::json;

namespace fans_host {

namespace {

bool has(const json &j, const char *key) { return j.is_object() && j.count(key) > 0; }

// material_properties keys per model, in the parameter order of the C ABI
// (createMatmodel, include/setup.h:21-61; model constructors in
// include/material_models/*.h)
const std::map<std::string, std::vector<std::string>> &model_keys()
{
    static const std::map<std::string, std::vector<std::string>> m = [] {
        std::map<std::string, std::vector<std::string>> k;
        k["LinearThermalIsotropic"] = {"conductivity"};
        k["LinearThermalTriclinic"] = {"K_11", "K_12", "K_13", "K_22", "K_23", "K_33"};
        k["LinearElasticIsotropic"] = {"bulk_modulus", "shear_modulus"};
        std::vector<std::string> c;
        for (int i = 1; i <= 6; ++i)
            for (int j = i; j <= 6; ++j)
                c.push_back("C_" + std::to_string(i) + std::to_string(j));
        k["LinearElasticTriclinic"]                   = c;
        k["J2PlasticityNew_LinearIsotropicHardening"] = {"bulk_modulus", "shear_modulus", "yield_stress",
                                                         "isotropic_hardening_parameter"};
        k["PseudoPlasticLinearHardening"] = {"bulk_modulus", "shear_modulus", "yield_stress", "hardening_parameter"};
        k["PseudoPlasticNonLinearHardening"] = {"bulk_modulus", "shear_modulus", "yield_stress", "hardening_exponent",
                                                "eps_0"};
        k["SaintVenantKirchhoff"]         = {"bulk_modulus", "shear_modulus"};
        k["CompressibleNeoHookean"]       = {"bulk_modulus", "shear_modulus"};
        // J2Plasticity.h:12-18 (+ :187-188 for the nonlinear law)
        k["J2ViscoPlastic_LinearIsotropicHardening"] = {
            "bulk_modulus", "shear_modulus", "yield_stress", "isotropic_hardening_parameter",
            "kinematic_hardening_parameter", "viscosity", "time_step"};
        k["J2ViscoPlastic_NonLinearIsotropicHardening"] = {
            "bulk_modulus", "shear_modulus", "yield_stress", "isotropic_hardening_parameter",
            "kinematic_hardening_parameter", "viscosity", "time_step", "saturation_stress", "saturation_exponent"};
        return k;
    }();
    return m;
}

bool thermal_model(const std::string &m) { return m == "LinearThermalIsotropic" || m == "LinearThermalTriclinic"; }
// createMatmodel<3,9> (include/setup.h:62-73)
bool large_strain_model(const std::string &m) { return m == "SaintVenantKirchhoff" || m == "CompressibleNeoHookean"; }

// MixedBC::from_json (include/mixedBCs.h:51-133)
MixedBC mixed_from_json(const json &jc, int n_str)
{
    MixedBC b;
    if (!has(jc, "strain_indices") || !has(jc, "stress_indices"))
        throw std::invalid_argument("mixed BC: strain_indices or stress_indices missing");
    b.idx_E = jc["strain_indices"].get<std::vector<int>>();
    b.idx_F = jc["stress_indices"].get<std::vector<int>>();
    std::vector<int> present(n_str, 0);
    for (int k : b.idx_E) {
        if (k < 0 || k >= n_str)
            throw std::invalid_argument("strain index out of range");
        present[k] = 1;
    }
    for (int k : b.idx_F) {
        if (k < 0 || k >= n_str)
            throw std::invalid_argument("stress index out of range");
        if (present[k])
            throw std::invalid_argument("index appears in both strain_indices and stress_indices");
        present[k] = 1;
    }
    for (int k = 0; k < n_str; ++k)
        if (!present[k])
            throw std::invalid_argument("each component must be either strain- or stress-controlled");
    std::vector<std::vector<double>> E, F;
    if (!b.idx_E.empty()) {
        if (!has(jc, "strain"))
            throw std::invalid_argument("strain array missing");
        E = jc["strain"].get<std::vector<std::vector<double>>>();
    }
    if (!b.idx_F.empty()) {
        if (!has(jc, "stress"))
            throw std::invalid_argument("stress array missing");
        F = jc["stress"].get<std::vector<std::vector<double>>>();
    }
    b.n_steps = (int) std::max(E.size(), F.size());
    if (b.n_steps == 0)
        throw std::invalid_argument("mixed BC: at least one of strain/stress must have timesteps");
    if (E.empty())
        E.assign(b.n_steps, {});
    if (F.empty())
        F.assign(b.n_steps, {});
    if ((int) E.size() != b.n_steps || (int) F.size() != b.n_steps)
        throw std::invalid_argument("strain and stress paths differ in length");
    for (int t = 0; t < b.n_steps; ++t) {
        if (E[t].size() != b.idx_E.size())
            throw std::invalid_argument("strain row length mismatch");
        if (F[t].size() != b.idx_F.size())
            throw std::invalid_argument("stress row length mismatch");
        b.strain_path.insert(b.strain_path.end(), E[t].begin(), E[t].end());
        b.stress_path.insert(b.stress_path.end(), F[t].begin(), F[t].end());
    }
    return b;
}

}  // namespace

Reader::~Reader() { CloseResultsFile(); }

void Reader::ReadInputFile(const char *input_fn)
{
    std::ifstream in(input_fn);
    if (!in)
        throw std::runtime_error(std::string("cannot open input file '") + input_fn + "'");
    json j;
    in >> j;

    const json &mi = j.at("microstructure");
    ms_filename    = mi.at("filepath").get<std::string>();
    std::string ds = mi.at("datasetname").get<std::string>();
    if (ds.empty())
        throw std::invalid_argument("datasetname must not be empty and must refer to a valid HDF5 path");
    ms_datasetname = (ds.front() == '/' ? "" : "/") + ds;
    L              = mi.at("L").get<std::vector<double>>();
    if (L.size() != 3)
        throw std::invalid_argument("microstructure.L must have 3 entries");
    // a relative microstructure path is taken relative to the working directory,
    // like the reference (H5Fopen on the string as given)

    results_prefix = has(j, "results_prefix") ? j["results_prefix"].get<std::string>() : "";
    dataset_name   = ms_datasetname + "_results/" + results_prefix;

    const json &ep = j.at("error_parameters");
    TOL            = ep.at("tolerance").get<double>();
    err_type       = ep.at("type").get<std::string>();
    err_measure    = ep.at("measure").get<std::string>();
    n_it           = j.at("n_it").get<int>();
    extrapolate_displacement = j.value("extrapolate_displacement", extrapolate_displacement);
    if (has(j, "linesearch_parameters")) {
        ls_max_iter = j["linesearch_parameters"].value("max_iter", ls_max_iter);
        ls_tol      = j["linesearch_parameters"].value("tol", ls_tol);
        if (ls_max_iter < 1 || ls_tol <= 0.0)
            throw std::invalid_argument("linesearch_parameters: max_iter >= 1 and tol > 0 required");
    }
    problemType = j.at("problem_type").get<std::string>();
    method      = j.at("method").get<std::string>();
    if (has(j, "strain_type")) {
        strain_type = j["strain_type"].get<std::string>();
        if (strain_type != "small" && strain_type != "large")
            throw std::invalid_argument("strain_type must be either 'small' or 'large'");
    }
    if (has(j, "FE_type")) {
        FE_type = j["FE_type"].get<std::string>();
        if (FE_type != "HEX8" && FE_type != "HEX8R" && FE_type != "BBAR")
            throw std::invalid_argument("FE_type must be one of: 'HEX8', 'HEX8R', or 'BBAR'");
    }
    resultsToWrite = j.at("results").get<std::vector<std::string>>();

    if (problemType == "thermal") {
        howmany = 1, n_str = 3;
    } else if (problemType == "mechanical" && strain_type == "small") {
        howmany = 3, n_str = 6;
    } else if (problemType == "mechanical" && strain_type == "large") {
        howmany = 3, n_str = 9;  // F_11 F_12 ... F_33 (reader.cpp:141-142)
    } else {
        throw std::invalid_argument(problemType + " is not a valid problem type");
    }
    if (method != "cg" && method != "fp")
        throw std::invalid_argument(method + " is not a valid method");
    if (err_measure != "L1" && err_measure != "L2" && err_measure != "Linfinity")
        throw std::invalid_argument("Unknown measure type: " + err_measure);
    if (err_type != "absolute" && err_type != "relative")
        throw std::invalid_argument("Unknown error type: " + err_type);

    load_cases.clear();
    const json &ml = j.at("macroscale_loading");
    if (!ml.is_array())
        throw std::runtime_error("macroscale_loading must be an array");
    for (const auto &entry : ml) {
        LoadCase lc;
        if (entry.is_array()) {  // legacy pure strain
            lc.mixed   = false;
            lc.g0_path = entry.get<std::vector<std::vector<double>>>();
            lc.n_steps = (int) lc.g0_path.size();
            for (const auto &row : lc.g0_path)
                if ((int) row.size() != n_str)
                    throw std::invalid_argument("Invalid length of loading vector: expected " +
                                                std::to_string(n_str) + " components but got " +
                                                std::to_string(row.size()));
        } else {
            lc.mixed   = true;
            lc.mbc     = mixed_from_json(entry, n_str);
            lc.n_steps = lc.mbc.n_steps;
        }
        load_cases.push_back(std::move(lc));
    }

    // material groups (MaterialManager.h:38-110, legacy top-level format included)
    json mats;
    if (has(j, "matmodel") && has(j, "material_properties")) {
        json g;
        g["matmodel"]            = j["matmodel"];
        g["material_properties"] = j["material_properties"];
        int n_m                  = 0;
        for (auto it = j["material_properties"].begin(); it != j["material_properties"].end(); ++it)
            if (it.value().is_array()) {
                n_m = (int) it.value().size();
                break;
            }
        std::vector<int> ph(n_m);
        for (int i = 0; i < n_m; ++i)
            ph[i] = i;
        g["phases"] = ph;
        mats        = json::array({g});
    } else {
        mats = j.at("materials");
    }
    if (!mats.is_array() || mats.empty())
        throw std::runtime_error("MaterialManager: 'materials' must be non-empty array");
    materials.clear();
    int max_phase = -1;
    for (const auto &mg : mats) {
        if (!has(mg, "phases") || !has(mg, "matmodel") || !has(mg, "material_properties"))
            throw std::runtime_error("MaterialManager: material group missing required fields");
        MaterialGroup g;
        g.matmodel = mg["matmodel"].get<std::string>();
        g.phases   = mg["phases"].get<std::vector<int>>();
        auto kit   = model_keys().find(g.matmodel);
        if (kit == model_keys().end())
            throw std::invalid_argument(g.matmodel + " is not a supported matmodel");
        if (thermal_model(g.matmodel) != (howmany == 1))
            throw std::invalid_argument(g.matmodel + " is not a valid matmodel for this problem type");
        if (howmany == 3 && large_strain_model(g.matmodel) != (n_str == 9))
            throw std::invalid_argument(g.matmodel + (n_str == 9 ? " is not a valid large strain material model"
                                                                 : " is not a valid small strain material model"));
        g.keys                 = kit->second;
        const json &props      = mg["material_properties"];
        const size_t np        = g.phases.size();
        g.params.assign(np * g.keys.size(), 0.0);
        for (size_t k = 0; k < g.keys.size(); ++k) {
            if (!has(props, g.keys[k].c_str()))
                throw std::invalid_argument("Missing material property '" + g.keys[k] + "' for " + g.matmodel);
            const json         &v = props[g.keys[k]];
            std::vector<double> vals =
                v.is_array() ? v.get<std::vector<double>>() : std::vector<double>{v.get<double>()};
            if (g.keys[k] == "time_step") {
                // one number for the whole model (J2Plasticity.h:19), repeated per phase
                if (!v.is_number())
                    throw std::invalid_argument("Material property 'time_step' must be a number");
                vals.assign(np, v.get<double>());
            }
            if (vals.size() != np)
                throw std::invalid_argument("Inconsistent size for material property: " + g.keys[k]);
            for (size_t p = 0; p < np; ++p)
                g.params[p * g.keys.size() + k] = vals[p];
        }
        for (int p : g.phases)
            max_phase = std::max(max_phase, p);
        materials.push_back(std::move(g));
    }
    n_mat = max_phase + 1;
    if (n_mat == 0)
        throw std::runtime_error("MaterialManager: No phases defined");
    std::vector<int> seen(n_mat, 0);
    for (const auto &g : materials)
        for (int p : g.phases) {
            if (p < 0 || p >= n_mat || seen[p])
                throw std::runtime_error("MaterialManager: Invalid or duplicate phase " + std::to_string(p));
            seen[p] = 1;
        }
    for (int p = 0; p < n_mat; ++p)
        if (!seen[p])
            throw std::runtime_error("MaterialManager: Phase " + std::to_string(p) + " not assigned");
    has_ref_material = has(j, "reference_material");
    if (has_ref_material) {
        auto r = j["reference_material"].get<std::vector<std::vector<double>>>();
        if ((int) r.size() != n_str)
            throw std::invalid_argument("reference_material must be " + std::to_string(n_str) + "x" +
                                        std::to_string(n_str));
        ref_material.clear();
        for (const auto &row : r) {
            if ((int) row.size() != n_str)
                throw std::invalid_argument("reference_material must be " + std::to_string(n_str) + "x" +
                                            std::to_string(n_str));
            ref_material.insert(ref_material.end(), row.begin(), row.end());
        }
        // MaterialManager::compute_reference_stiffness (MaterialManager.h:189-192):
        // Cholesky of the lower triangle, as Eigen's LLT reads it
        std::vector<double> Lm(n_str * n_str, 0.0);
        for (int c = 0; c < n_str; ++c) {
            double d = ref_material[c * n_str + c];
            for (int k = 0; k < c; ++k)
                d -= Lm[c * n_str + k] * Lm[c * n_str + k];
            if (!(d > 0.0))
                throw std::invalid_argument("reference_material must be symmetric positive definite");
            Lm[c * n_str + c] = std::sqrt(d);
            for (int i = c + 1; i < n_str; ++i) {
                double v = ref_material[i * n_str + c];
                for (int k = 0; k < c; ++k)
                    v -= Lm[i * n_str + k] * Lm[c * n_str + k];
                Lm[i * n_str + c] = v / Lm[c * n_str + c];
            }
        }
    }
}

void Reader::ReadMS()
{
    hid_t file = h5().Fopen(ms_filename.c_str(), kH5AccRdonly, kH5Default);
    if (file < 0)
        throw std::runtime_error("cannot open microstructure file '" + ms_filename + "'");
    hid_t dset = h5().Dopen2(file, ms_datasetname.c_str(), kH5Default);
    if (dset < 0) {
        h5().Fclose(file);
        throw std::runtime_error("cannot open dataset '" + ms_datasetname + "'");
    }
    hid_t   space = h5().Dget_space(dset);
    hsize_t fd[3] = {0, 0, 0};
    if (h5().Sget_simple_extent_ndims(space) != 3) {
        h5().Sclose(space), h5().Dclose(dset), h5().Fclose(file);
        throw std::runtime_error("microstructure dataset must be 3-D");
    }
    h5().Sget_simple_extent_dims(space, fd, nullptr);
    bool zyx = false;
    if (h5().Aexists(dset, "permute_order") > 0) {
        hid_t attr = h5().Aopen(dset, "permute_order", kH5Default);
        hid_t at   = h5().Aget_type(attr);
        if (h5().Tis_variable_str(at) > 0) {
            char *s = nullptr;
            if (h5().Aread(attr, at, &s) >= 0 && s) {
                zyx = s[0] == 'z' || s[0] == 'Z';
                h5().free_memory(s);
            }
        } else {
            std::vector<char> s(h5().Tget_size(at) + 1, 0);
            if (h5().Aread(attr, at, s.data()) >= 0)
                zyx = s[0] == 'z' || s[0] == 'Z';
        }
        h5().Tclose(at);
        h5().Aclose(attr);
    }
    std::vector<uint16_t> raw(fd[0] * fd[1] * fd[2]);
    // the reference reads as native unsigned short whatever the stored type (reader.cpp:260)
    if (h5().Dread(dset, h5().t_ushort, kH5Default, kH5Default, kH5Default, raw.data()) < 0) {
        h5().Sclose(space), h5().Dclose(dset), h5().Fclose(file);
        throw std::runtime_error("reading the microstructure failed");
    }
    h5().Sclose(space);
    h5().Dclose(dset);
    h5().Fclose(file);
    dims = zyx ? std::vector<int>{(int) fd[2], (int) fd[1], (int) fd[0]}
               : std::vector<int>{(int) fd[0], (int) fd[1], (int) fd[2]};
    if (dims[2] % 2 != 0)
        throw std::invalid_argument("n_z must be even");
    l_e = {L[0] / dims[0], L[1] / dims[1], L[2] / dims[2]};
    const size_t nx = dims[0], ny = dims[1], nz = dims[2];
    ms.resize(nx * ny * nz);
    for (size_t x = 0; x < nx; ++x)
        for (size_t y = 0; y < ny; ++y)
            for (size_t z = 0; z < nz; ++z)
                ms[(x * ny + y) * nz + z] = zyx ? raw[(z * ny + y) * nx + x] : raw[(x * ny + y) * nz + z];
    const int mx = ms.empty() ? 0 : *std::max_element(ms.begin(), ms.end());
    if (mx >= n_mat)
        throw std::runtime_error("microstructure holds phase id " + std::to_string(mx) + " without a material");
}

void Reader::OpenResultsFile(const char *output_fn)
{
    results_file_id = h5().Fcreate(output_fn, kH5AccTrunc, kH5Default, kH5Default);
    if (results_file_id < 0)
        throw std::runtime_error("Failed to create results file");
}

void Reader::CloseResultsFile()
{
    if (results_file_id >= 0) {
        h5().Fclose(results_file_id);
        results_file_id = -1;
    }
}

bool Reader::wants(const char *field) const
{
    return std::find(resultsToWrite.begin(), resultsToWrite.end(), field) != resultsToWrite.end();
}

std::string Reader::path_of(const char *field, int load_idx, int time_idx) const
{
    char name[4096];
    std::snprintf(name, sizeof(name), "%s/load%i/time_step%i/%s", dataset_name.c_str(), load_idx, time_idx, field);
    return name;
}

// create every group on an absolute path except the last component
void Reader::safe_create_group(const char *name) const
{
    if (name[0] != '/')
        return;
    std::string buf(name);
    for (size_t pos = buf.find('/', 1); pos != std::string::npos; pos = buf.find('/', pos + 1)) {
        const std::string g = buf.substr(0, pos);
        if (h5().Lexists(results_file_id, g.c_str(), kH5Default) > 0)
            continue;
        hid_t gid = h5().Gcreate2(results_file_id, g.c_str(), kH5Default, kH5Default, kH5Default);
        if (gid < 0)
            throw std::runtime_error("cannot create group " + g);
        h5().Gclose(gid);
    }
}

void Reader::writeData(const char *field, int load_idx, int time_idx, const double *data, const hsize_t *dims_,
                       int ndims) const
{
    if (!wants(field))
        return;
    if (results_file_id < 0)
        throw std::runtime_error("WriteData: results file is not open");
    const std::string name = path_of(field, load_idx, time_idx);
    safe_create_group(name.c_str());
    if (h5().Lexists(results_file_id, name.c_str(), kH5Default) > 0)
        h5().Ldelete(results_file_id, name.c_str(), kH5Default);
    hid_t space = h5().Screate_simple(ndims, dims_, nullptr);
    hid_t dset  = h5().Dcreate2(results_file_id, name.c_str(), h5().t_double, space, kH5Default, kH5Default,
                             kH5Default);
    if (dset < 0 || h5().Dwrite(dset, h5().t_double, kH5Default, kH5Default, kH5Default, data) < 0) {
        h5().Sclose(space);
        throw std::runtime_error("Error writing data to dataset " + name);
    }
    h5().Dclose(dset);
    h5().Sclose(space);
}

namespace {
template <typename T>
hid_t h5_type();
template <>
hid_t h5_type<double>() { return h5().t_double; }
template <>
hid_t h5_type<int>() { return h5().t_int; }
template <>
hid_t h5_type<uint16_t>() { return h5().t_ushort; }
template <>
hid_t h5_type<float>() { return h5().t_float; }
}  // namespace

template <typename T>
void Reader::writeSlab(const char *field, int load_idx, int time_idx, const T *data,
                       const std::vector<int> &extra_dims) const
{
    if (!wants(field))
        return;
    if (results_file_id < 0)
        throw std::runtime_error("WriteSlab: results file is not open");
    const std::string name = path_of(field, load_idx, time_idx);
    safe_create_group(name.c_str());
    const size_t         nx = dims[0], ny = dims[1], nz = dims[2];
    size_t               ex = 1;
    std::vector<hsize_t> fd = {nz, ny, nx};
    for (int e : extra_dims) {
        fd.push_back((hsize_t) e);
        ex *= (size_t) e;
    }
    // logical [X][Y][Z][extra] -> file [Z][Y][X][extra]
    std::vector<T> tmp(nx * ny * nz * ex);
    for (size_t x = 0; x < nx; ++x)
        for (size_t y = 0; y < ny; ++y)
            for (size_t z = 0; z < nz; ++z)
                std::memcpy(&tmp[((z * ny + y) * nx + x) * ex], &data[((x * ny + y) * nz + z) * ex], sizeof(T) * ex);
    if (h5().Lexists(results_file_id, name.c_str(), kH5Default) > 0)
        h5().Ldelete(results_file_id, name.c_str(), kH5Default);
    hid_t space = h5().Screate_simple((int) fd.size(), fd.data(), nullptr);
    hid_t dset  = h5().Dcreate2(results_file_id, name.c_str(), h5_type<T>(), space, kH5Default, kH5Default,
                             kH5Default);
    if (dset < 0) {
        h5().Sclose(space);
        throw std::runtime_error("WriteSlab: cannot create " + name);
    }
    const char perm[] = "zyx";
    hid_t      at     = h5().Tcopy(h5().t_c_s1);
    h5().Tset_size(at, 4);
    hid_t as   = h5().Screate(kH5SScalar);
    hid_t attr = h5().Acreate2(dset, "permute_order", at, as, kH5Default, kH5Default);
    h5().Awrite(attr, at, perm);
    h5().Aclose(attr);
    h5().Sclose(as);
    h5().Tclose(at);
    const herr_t st = h5().Dwrite(dset, h5_type<T>(), kH5Default, kH5Default, kH5Default, tmp.data());
    h5().Dclose(dset);
    h5().Sclose(space);
    if (st < 0)
        throw std::runtime_error("WriteSlab: H5Dwrite failed for " + name);
}

template void Reader::writeSlab<double>(const char *, int, int, const double *, const std::vector<int> &) const;
template void Reader::writeSlab<int>(const char *, int, int, const int *, const std::vector<int> &) const;
template void Reader::writeSlab<uint16_t>(const char *, int, int, const uint16_t *, const std::vector<int> &) const;
template void Reader::writeSlab<float>(const char *, int, int, const float *, const std::vector<int> &) const;

}  // namespace fans_host
