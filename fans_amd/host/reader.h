// Host-side input/output of the drop-in CLI: the reference's JSON input format
// and HDF5 result layout for a single-rank run.
//
// Restates the Reader of include/reader.h / src/reader.cpp:
//   ReadInputFile  src/reader.cpp:63-183   (keys, defaults, error messages)
//   ReadMS         src/reader.cpp:227-411  (uint16 phase ids, permute_order "zyx")
//   writeData      include/reader.h:330-339 (<dataset>_results/<prefix>/load<i>/time_step<t>/<field>)
//   writeSlab      include/reader.h:341-351, WriteSlab :161-327 (file layout Z Y X extra..., attribute
//                  permute_order = "zyx")
// with serial HDF5 (one rank: no MPI-IO property lists).
#pragma once

#include <H5public.h>
#include <H5Ipublic.h>

#include <cstdint>
#include <string>
#include <vector>

namespace fans_host {

struct MaterialGroup {
    std::string              matmodel;
    std::vector<int>         phases;
    std::vector<std::string> keys;    // material_properties keys in model parameter order
    std::vector<double>      params;  // [n_phases][n_keys]
};

struct MixedBC {
    std::vector<int>    idx_E, idx_F;
    std::vector<double> strain_path;  // [n_steps][idx_E.size()]
    std::vector<double> stress_path;  // [n_steps][idx_F.size()]
    int                 n_steps = 0;
};

struct LoadCase {
    bool                             mixed = false;
    std::vector<std::vector<double>> g0_path;  // pure strain: [n_steps][n_str]
    MixedBC                          mbc;
    int                              n_steps = 0;
};

class Reader {
  public:
    // input (ReadInputFile)
    std::string                ms_filename, ms_datasetname, results_prefix, dataset_name;
    std::vector<double>        L;
    std::string                err_type = "absolute", err_measure = "Linfinity";
    double                     TOL  = 1e-10;
    int                        n_it = 100;
    bool                       extrapolate_displacement = true;
    int                        ls_max_iter = 5;
    double                     ls_tol      = 1e-2;
    std::string                problemType, method, strain_type = "small", FE_type = "HEX8";
    std::vector<std::string>   resultsToWrite;
    std::vector<LoadCase>      load_cases;
    std::vector<MaterialGroup> materials;
    bool                       has_ref_material = false;
    std::vector<double>        ref_material;  // n_str x n_str row major
    int                        n_str = 6, howmany = 3;

    // microstructure (ReadMS): phase ids in x-major (x, y, z) order
    std::vector<int>      dims;
    std::vector<double>   l_e;
    std::vector<uint16_t> ms;
    int                   n_mat = 0;  // number of phases (max id + 1)

    ~Reader();

    // throws std::runtime_error / std::invalid_argument with the reference's messages
    void ReadInputFile(const char *input_fn);
    void ReadMS();
    void OpenResultsFile(const char *output_fn);
    void CloseResultsFile();

    bool wants(const char *field) const;

    // Reader::writeData: small datasets (rank 1 or 2), written only if requested
    void writeData(const char *field, int load_idx, int time_idx, const double *data, const hsize_t *dims,
                   int ndims) const;
    // Reader::writeSlab: data in logical [X][Y][Z][extra...] order, written as Z Y X extra...
    template <typename T>
    void writeSlab(const char *field, int load_idx, int time_idx, const T *data,
                   const std::vector<int> &extra_dims) const;

  private:
    hid_t results_file_id = -1;
    void  safe_create_group(const char *name) const;
    std::string path_of(const char *field, int load_idx, int time_idx) const;
};

}  // namespace fans_host
