// Run-time binding of the serial HDF5 C API (see h5dyn.h).
#include "h5dyn.h"

#include <dlfcn.h>

#include <cstdlib>
#include <stdexcept>
#include <string>

namespace fans_host {

namespace {

template <typename F>
void bind(void *lib, F &fn, const char *name)
{
    void *p = dlsym(lib, name);
    if (!p)
        throw std::runtime_error(std::string("HDF5 library lacks ") + name);
    fn = reinterpret_cast<F>(p);
}

hid_t global_type(void *lib, const char *name)
{
    void *p = dlsym(lib, name);
    if (!p)
        throw std::runtime_error(std::string("HDF5 library lacks ") + name);
    return *static_cast<hid_t *>(p);
}

H5Api load()
{
    const char *env   = std::getenv("FANS_HDF5_LIB");
    const char *cands[] = {env, "/opt/conda/lib/libhdf5.so.103", "libhdf5.so", "libhdf5_serial.so"};
    void       *lib   = nullptr;
    for (const char *c : cands)
        if (c && (lib = dlopen(c, RTLD_NOW | RTLD_LOCAL)))
            break;
    if (!lib)
        throw std::runtime_error("serial HDF5 library not found (set FANS_HDF5_LIB)");
    H5Api a{};
    bind(lib, a.open, "H5open");
    bind(lib, a.Fopen, "H5Fopen");
    bind(lib, a.Fcreate, "H5Fcreate");
    bind(lib, a.Fclose, "H5Fclose");
    bind(lib, a.Dopen2, "H5Dopen2");
    bind(lib, a.Dcreate2, "H5Dcreate2");
    bind(lib, a.Dread, "H5Dread");
    bind(lib, a.Dwrite, "H5Dwrite");
    bind(lib, a.Dclose, "H5Dclose");
    bind(lib, a.Dget_space, "H5Dget_space");
    bind(lib, a.Sget_simple_extent_ndims, "H5Sget_simple_extent_ndims");
    bind(lib, a.Sget_simple_extent_dims, "H5Sget_simple_extent_dims");
    bind(lib, a.Screate_simple, "H5Screate_simple");
    bind(lib, a.Screate, "H5Screate");
    bind(lib, a.Sclose, "H5Sclose");
    bind(lib, a.Aexists, "H5Aexists");
    bind(lib, a.Aopen, "H5Aopen");
    bind(lib, a.Acreate2, "H5Acreate2");
    bind(lib, a.Aget_type, "H5Aget_type");
    bind(lib, a.Aread, "H5Aread");
    bind(lib, a.Awrite, "H5Awrite");
    bind(lib, a.Aclose, "H5Aclose");
    bind(lib, a.Tis_variable_str, "H5Tis_variable_str");
    bind(lib, a.Tget_size, "H5Tget_size");
    bind(lib, a.Tcopy, "H5Tcopy");
    bind(lib, a.Tset_size, "H5Tset_size");
    bind(lib, a.Tclose, "H5Tclose");
    bind(lib, a.free_memory, "H5free_memory");
    bind(lib, a.Lexists, "H5Lexists");
    bind(lib, a.Ldelete, "H5Ldelete");
    bind(lib, a.Gcreate2, "H5Gcreate2");
    bind(lib, a.Gclose, "H5Gclose");
    if (a.open() < 0)
        throw std::runtime_error("H5open failed");
    a.t_double = global_type(lib, "H5T_NATIVE_DOUBLE_g");
    a.t_float  = global_type(lib, "H5T_NATIVE_FLOAT_g");
    a.t_int    = global_type(lib, "H5T_NATIVE_INT_g");
    a.t_ushort = global_type(lib, "H5T_NATIVE_USHORT_g");
    a.t_uchar  = global_type(lib, "H5T_NATIVE_UCHAR_g");
    a.t_c_s1   = global_type(lib, "H5T_C_S1_g");
    return a;
}

}  // namespace

const H5Api &h5()
{
    static const H5Api api = load();
    return api;
}

}  // namespace fans_host
