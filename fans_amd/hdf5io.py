"""HDF5 input/output of the Python host: the microstructure reader of
src/reader.cpp:227-411 (Reader::ReadMS) and the result layout of
include/reader.h:88-351 (writeData / writeSlab), over the serial HDF5 C
library through ctypes (no h5py in this environment).

Layout conventions kept from the reference:
  * a microstructure dataset with attribute permute_order = "zyx" is stored
    [z][y][x] and read into logical [x][y][z]; otherwise [x][y][z];
  * results go to <datasetname>_results/<results_prefix>/load<i>/time_step<t>/<field>;
    slab fields are stored [z][y][x][extra...] with permute_order = "zyx".
"""
from __future__ import annotations

import ctypes as C
import ctypes.util
import os

import numpy as np

_LIB = None
hid_t = C.c_int64
herr_t = C.c_int
hsize_t = C.c_uint64

_CANDIDATES = ["/opt/conda/lib/libhdf5.so", "libhdf5.so", "libhdf5_serial.so"]


def lib():
    global _LIB
    if _LIB is not None:
        return _LIB
    names = list(_CANDIDATES)
    found = ctypes.util.find_library("hdf5")
    if found:
        names.insert(1, found)
    last = None
    for n in names:
        try:
            L = C.CDLL(n)
            break
        except OSError as e:  # try the next location
            last = e
    else:
        raise RuntimeError(f"libhdf5 not found ({last})")
    L.H5open.restype = herr_t
    L.H5open()
    for fn, res, args in [
        ("H5Fopen", hid_t, [C.c_char_p, C.c_uint, hid_t]),
        ("H5Fcreate", hid_t, [C.c_char_p, C.c_uint, hid_t, hid_t]),
        ("H5Fclose", herr_t, [hid_t]),
        ("H5Dopen2", hid_t, [hid_t, C.c_char_p, hid_t]),
        ("H5Dcreate2", hid_t, [hid_t, C.c_char_p, hid_t, hid_t, hid_t, hid_t, hid_t]),
        ("H5Dread", herr_t, [hid_t, hid_t, hid_t, hid_t, hid_t, C.c_void_p]),
        ("H5Dwrite", herr_t, [hid_t, hid_t, hid_t, hid_t, hid_t, C.c_void_p]),
        ("H5Dget_space", hid_t, [hid_t]),
        ("H5Dget_type", hid_t, [hid_t]),
        ("H5Dclose", herr_t, [hid_t]),
        ("H5Sget_simple_extent_ndims", C.c_int, [hid_t]),
        ("H5Sget_simple_extent_dims", C.c_int, [hid_t, C.POINTER(hsize_t), C.POINTER(hsize_t)]),
        ("H5Screate_simple", hid_t, [C.c_int, C.POINTER(hsize_t), C.POINTER(hsize_t)]),
        ("H5Screate", hid_t, [C.c_int]),
        ("H5Sclose", herr_t, [hid_t]),
        ("H5Aexists", C.c_int, [hid_t, C.c_char_p]),
        ("H5Aopen", hid_t, [hid_t, C.c_char_p, hid_t]),
        ("H5Acreate2", hid_t, [hid_t, C.c_char_p, hid_t, hid_t, hid_t, hid_t]),
        ("H5Aget_type", hid_t, [hid_t]),
        ("H5Aread", herr_t, [hid_t, hid_t, C.c_void_p]),
        ("H5Awrite", herr_t, [hid_t, hid_t, C.c_void_p]),
        ("H5Aclose", herr_t, [hid_t]),
        ("H5Tis_variable_str", C.c_int, [hid_t]),
        ("H5Tget_size", C.c_size_t, [hid_t]),
        ("H5Tcopy", hid_t, [hid_t]),
        ("H5Tset_size", herr_t, [hid_t, C.c_size_t]),
        ("H5Tclose", herr_t, [hid_t]),
        ("H5free_memory", herr_t, [C.c_void_p]),
        ("H5Lexists", C.c_int, [hid_t, C.c_char_p, hid_t]),
        ("H5Gcreate2", hid_t, [hid_t, C.c_char_p, hid_t, hid_t, hid_t]),
        ("H5Gclose", herr_t, [hid_t]),
    ]:
        f = getattr(L, fn)
        f.restype, f.argtypes = res, args
    _LIB = L
    return L


def _native(name):
    return hid_t.in_dll(lib(), name + "_g").value


_NP2H5 = {np.dtype(np.float64): "H5T_NATIVE_DOUBLE", np.dtype(np.float32): "H5T_NATIVE_FLOAT",
          np.dtype(np.int32): "H5T_NATIVE_INT", np.dtype(np.uint16): "H5T_NATIVE_USHORT",
          np.dtype(np.uint8): "H5T_NATIVE_UCHAR"}
H5F_ACC_RDONLY, H5F_ACC_TRUNC, H5S_SCALAR = 0, 2, 0


def _check(v, what):
    if v < 0:
        raise RuntimeError(f"HDF5: {what} failed")
    return v


def _read_str_attr(obj, name):
    L = lib()
    if L.H5Aexists(obj, name.encode()) <= 0:
        return None
    a = _check(L.H5Aopen(obj, name.encode(), 0), "H5Aopen")
    t = L.H5Aget_type(a)
    try:
        if L.H5Tis_variable_str(t) > 0:
            p = C.c_char_p()
            _check(L.H5Aread(a, t, C.byref(p)), "H5Aread")
            s = p.value.decode()
            L.H5free_memory(p)
        else:
            n = L.H5Tget_size(t)
            buf = C.create_string_buffer(n + 1)
            _check(L.H5Aread(a, t, buf), "H5Aread")
            s = buf.value.decode()
    finally:
        L.H5Tclose(t)
        L.H5Aclose(a)
    return s


def read_dataset(path, dataset, dtype=np.float64):
    """Whole dataset as numpy (converted to `dtype` by HDF5) and its
    permute_order attribute (or None)."""
    L = lib()
    f = _check(L.H5Fopen(os.fsencode(path), H5F_ACC_RDONLY, 0), f"open {path}")
    try:
        d = _check(L.H5Dopen2(f, dataset.encode(), 0), f"open dataset {dataset}")
        try:
            sp = L.H5Dget_space(d)
            nd = L.H5Sget_simple_extent_ndims(sp)
            dims = (hsize_t * max(nd, 1))()
            L.H5Sget_simple_extent_dims(sp, dims, None)
            L.H5Sclose(sp)
            shape = tuple(int(dims[i]) for i in range(nd))
            out = np.empty(shape, dtype=dtype)
            _check(L.H5Dread(d, _native(_NP2H5[np.dtype(dtype)]), 0, 0, 0, out.ctypes.data), "H5Dread")
            perm = _read_str_attr(d, "permute_order")
        finally:
            L.H5Dclose(d)
    finally:
        L.H5Fclose(f)
    return out, perm


def exists(path, name):
    L = lib()
    f = _check(L.H5Fopen(os.fsencode(path), H5F_ACC_RDONLY, 0), f"open {path}")
    try:
        parts = [p for p in name.split("/") if p]
        cur = ""
        for p in parts:
            cur += "/" + p
            if L.H5Lexists(f, cur.encode(), 0) <= 0:
                return False
        return True
    finally:
        L.H5Fclose(f)


def read_microstructure(path, dataset):
    """Reader::ReadMS (reader.cpp:227-300): phase ids as uint16 in logical
    [x][y][z] order."""
    if not dataset.startswith("/"):
        dataset = "/" + dataset
    raw, perm = read_dataset(path, dataset, np.uint16)
    if raw.ndim != 3:
        raise ValueError("microstructure dataset must be 3-D")
    if perm and perm[0] in "zZ":
        raw = raw.transpose(2, 1, 0)
    return np.ascontiguousarray(raw)


def write_microstructure(path, dataset, ms_xyz, zyx=True):
    """Write a microstructure file (uint8 phase ids), [z][y][x] with
    permute_order = "zyx" by default -- the layout of the reference's test
    microstructures."""
    arr = np.asarray(ms_xyz, dtype=np.uint8)
    with ResultsFile(path) as rf:
        rf.write(dataset, arr.transpose(2, 1, 0) if zyx else arr, permute="zyx" if zyx else None)


class ResultsFile:
    """H5Fcreate(TRUNC) + Reader::WriteData / WriteSlab for one rank."""

    def __init__(self, path):
        self.f = _check(lib().H5Fcreate(os.fsencode(path), H5F_ACC_TRUNC, 0, 0), f"create {path}")

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def close(self):
        if self.f is not None:
            lib().H5Fclose(self.f)
            self.f = None

    def _groups(self, name):
        L = lib()
        parts = [p for p in name.split("/") if p][:-1]
        cur = ""
        for p in parts:
            cur += "/" + p
            if L.H5Lexists(self.f, cur.encode(), 0) <= 0:
                g = _check(L.H5Gcreate2(self.f, cur.encode(), 0, 0, 0), f"create group {cur}")
                L.H5Gclose(g)

    def write(self, name, data, permute=None):
        L = lib()
        data = np.ascontiguousarray(data)
        self._groups(name)
        dims = (hsize_t * data.ndim)(*data.shape)
        sp = L.H5Screate_simple(data.ndim, dims, None)
        t = _native(_NP2H5[data.dtype])
        d = _check(L.H5Dcreate2(self.f, name.encode(), t, sp, 0, 0, 0), f"create dataset {name}")
        try:
            _check(L.H5Dwrite(d, t, 0, 0, 0, data.ctypes.data), f"write {name}")
            if permute:
                at = L.H5Tcopy(_native("H5T_C_S1"))
                L.H5Tset_size(at, len(permute) + 1)
                asp = L.H5Screate(H5S_SCALAR)
                a = L.H5Acreate2(d, b"permute_order", at, asp, 0, 0)
                L.H5Awrite(a, at, C.create_string_buffer(permute.encode(), len(permute) + 1))
                L.H5Aclose(a)
                L.H5Sclose(asp)
                L.H5Tclose(at)
        finally:
            L.H5Dclose(d)
            L.H5Sclose(sp)

    def write_slab(self, name, data_xyz):
        """[X][Y][Z][extra...] -> file [Z][Y][X][extra...] with permute_order."""
        a = np.asarray(data_xyz)
        self.write(name, np.swapaxes(a, 0, 2), permute="zyx")
