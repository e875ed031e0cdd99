// Serial HDF5 C API loaded at run time (dlopen), so the CLI does not link
// against the HDF5 installation's directory, whose older libstdc++ would
// shadow the one the ROCm runtime needs.  Library: $FANS_HDF5_LIB, else
// /opt/conda/lib/libhdf5.so.103, else libhdf5.so from the loader path.
#pragma once

#include <H5public.h>  // hid_t, herr_t, hsize_t, htri_t (types only)
#include <H5Ipublic.h>

#include <cstddef>

namespace fans_host {

struct H5Api {
    herr_t (*open)(void);
    hid_t (*Fopen)(const char *, unsigned, hid_t);
    hid_t (*Fcreate)(const char *, unsigned, hid_t, hid_t);
    herr_t (*Fclose)(hid_t);
    hid_t (*Dopen2)(hid_t, const char *, hid_t);
    hid_t (*Dcreate2)(hid_t, const char *, hid_t, hid_t, hid_t, hid_t, hid_t);
    herr_t (*Dread)(hid_t, hid_t, hid_t, hid_t, hid_t, void *);
    herr_t (*Dwrite)(hid_t, hid_t, hid_t, hid_t, hid_t, const void *);
    herr_t (*Dclose)(hid_t);
    hid_t (*Dget_space)(hid_t);
    int (*Sget_simple_extent_ndims)(hid_t);
    int (*Sget_simple_extent_dims)(hid_t, hsize_t *, hsize_t *);
    hid_t (*Screate_simple)(int, const hsize_t *, const hsize_t *);
    hid_t (*Screate)(int);
    herr_t (*Sclose)(hid_t);
    htri_t (*Aexists)(hid_t, const char *);
    hid_t (*Aopen)(hid_t, const char *, hid_t);
    hid_t (*Acreate2)(hid_t, const char *, hid_t, hid_t, hid_t, hid_t);
    hid_t (*Aget_type)(hid_t);
    herr_t (*Aread)(hid_t, hid_t, void *);
    herr_t (*Awrite)(hid_t, hid_t, const void *);
    herr_t (*Aclose)(hid_t);
    htri_t (*Tis_variable_str)(hid_t);
    size_t (*Tget_size)(hid_t);
    hid_t (*Tcopy)(hid_t);
    herr_t (*Tset_size)(hid_t, size_t);
    herr_t (*Tclose)(hid_t);
    herr_t (*free_memory)(void *);
    htri_t (*Lexists)(hid_t, const char *, hid_t);
    herr_t (*Ldelete)(hid_t, const char *, hid_t);
    hid_t (*Gcreate2)(hid_t, const char *, hid_t, hid_t, hid_t);
    herr_t (*Gclose)(hid_t);
    // predefined types (the H5T_*_g globals after H5open)
    hid_t t_double, t_float, t_int, t_ushort, t_uchar, t_c_s1;
};

// loads on first use; throws std::runtime_error when the library is missing
const H5Api &h5();

constexpr hid_t    kH5Default   = 0;  // H5P_DEFAULT, H5S_ALL
constexpr unsigned kH5AccRdonly = 0u;
constexpr unsigned kH5AccTrunc  = 2u;
constexpr int      kH5SScalar   = 0;

}  // namespace fans_host
