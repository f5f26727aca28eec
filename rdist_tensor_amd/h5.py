"""Minimal HDF5 reader / writer over the system libhdf5 (ctypes, C API).

The reference's tensor files are HDF5 (tensor_io/process_frostt_tensor.cpp
writes them with the C API, exafac/sparse_tensor_e.py reads them with h5py):
one-dimensional datasets at the file's root.  This module writes and reads
exactly that: contiguous 1-D datasets of unsigned 32/64-bit integers or
doubles.  The library is found through EXAFAC_LIBHDF5, the loader's search
path, or /opt/conda/lib; available() says whether one was found.
"""
import ctypes as C
import ctypes.util
import os

import numpy as np

_hid = C.c_int64          # hid_t (HDF5 >= 1.10)
_hsize = C.c_uint64       # hsize_t
H5F_ACC_RDONLY = 0x0000
H5F_ACC_TRUNC = 0x0002
H5P_DEFAULT = 0
H5S_ALL = 0
H5T_INTEGER = 0
H5T_FLOAT = 1
H5T_SGN_NONE = 0

_LIB = None
_TYPES = {}


def _candidates():
    env = os.environ.get("EXAFAC_LIBHDF5")
    if env:
        yield env
    found = ctypes.util.find_library("hdf5")
    if found:
        yield found
    yield "/opt/conda/lib/libhdf5.so"


def lib():
    global _LIB
    if _LIB is not None:
        return _LIB
    err = None
    for path in _candidates():
        try:
            L = C.CDLL(path)
            break
        except OSError as e:
            err = e
    else:
        raise OSError(f"no HDF5 library found ({err})")
    sig = {
        "H5open": ([], C.c_int),
        "H5Fcreate": ([C.c_char_p, C.c_uint, _hid, _hid], _hid),
        "H5Fopen": ([C.c_char_p, C.c_uint, _hid], _hid),
        "H5Fclose": ([_hid], C.c_int),
        "H5Screate_simple": ([C.c_int, C.POINTER(_hsize), C.POINTER(_hsize)], _hid),
        "H5Sclose": ([_hid], C.c_int),
        "H5Sget_simple_extent_ndims": ([_hid], C.c_int),
        "H5Sget_simple_extent_dims": ([_hid, C.POINTER(_hsize), C.POINTER(_hsize)], C.c_int),
        "H5Dcreate2": ([_hid, C.c_char_p, _hid, _hid, _hid, _hid, _hid], _hid),
        "H5Dopen2": ([_hid, C.c_char_p, _hid], _hid),
        "H5Dwrite": ([_hid, _hid, _hid, _hid, _hid, C.c_void_p], C.c_int),
        "H5Dread": ([_hid, _hid, _hid, _hid, _hid, C.c_void_p], C.c_int),
        "H5Dget_space": ([_hid], _hid),
        "H5Dget_type": ([_hid], _hid),
        "H5Dclose": ([_hid], C.c_int),
        "H5Tget_class": ([_hid], C.c_int),
        "H5Tget_size": ([_hid], C.c_size_t),
        "H5Tget_sign": ([_hid], C.c_int),
        "H5Tclose": ([_hid], C.c_int),
    }
    for name, (args, res) in sig.items():
        f = getattr(L, name)
        f.argtypes = args
        f.restype = res
    if L.H5open() < 0:
        raise OSError("H5open failed")
    for key, sym in (("u4", "H5T_NATIVE_UINT32_g"), ("u8", "H5T_NATIVE_UINT64_g"), ("f8", "H5T_NATIVE_DOUBLE_g")):
        _TYPES[key] = _hid.in_dll(L, sym).value
    _LIB = L
    return L


def available():
    try:
        lib()
        return True
    except OSError:
        return False


def _ok(v, what):
    if v < 0:
        raise OSError(f"HDF5: {what} failed")
    return v


def _mem_type(dtype):
    dtype = np.dtype(dtype)
    key = {np.dtype(np.uint32): "u4", np.dtype(np.uint64): "u8", np.dtype(np.float64): "f8"}.get(dtype)
    if key is None:
        raise ValueError(f"unsupported dataset type {dtype}")
    return _TYPES[key]


def write(path, datasets):
    """Creates `path` (truncating) with one contiguous 1-D dataset per entry of
    `datasets` (name -> 1-D uint32 / uint64 / float64 array), stored in the
    native type as the reference's converter stores them."""
    L = lib()
    fid = _ok(L.H5Fcreate(os.fsencode(path), H5F_ACC_TRUNC, H5P_DEFAULT, H5P_DEFAULT), "H5Fcreate")
    try:
        for name, arr in datasets.items():
            a = np.ascontiguousarray(arr)
            if a.ndim != 1:
                raise ValueError(f"{name}: only 1-D datasets")
            t = _mem_type(a.dtype)
            dims = (_hsize * 1)(len(a))
            sid = _ok(L.H5Screate_simple(1, dims, None), "H5Screate_simple")
            try:
                did = _ok(L.H5Dcreate2(fid, name.encode(), t, sid, H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT),
                          f"H5Dcreate2({name})")
                try:
                    if len(a):
                        _ok(L.H5Dwrite(did, t, H5S_ALL, H5S_ALL, H5P_DEFAULT, a.ctypes.data), f"H5Dwrite({name})")
                finally:
                    L.H5Dclose(did)
            finally:
                L.H5Sclose(sid)
    finally:
        L.H5Fclose(fid)


def read(path, names=None):
    """Reads 1-D datasets of `path` into a dict (unsigned integers as uint32 /
    uint64 by their stored size, floating point as float64)."""
    L = lib()
    fid = _ok(L.H5Fopen(os.fsencode(path), H5F_ACC_RDONLY, H5P_DEFAULT), f"H5Fopen({path})")
    out = {}
    try:
        for name in names:
            did = _ok(L.H5Dopen2(fid, name.encode(), H5P_DEFAULT), f"H5Dopen2({name})")
            try:
                sid = _ok(L.H5Dget_space(did), "H5Dget_space")
                tid = _ok(L.H5Dget_type(did), "H5Dget_type")
                try:
                    if L.H5Sget_simple_extent_ndims(sid) != 1:
                        raise ValueError(f"{name}: not a 1-D dataset")
                    n = (_hsize * 1)()
                    L.H5Sget_simple_extent_dims(sid, n, None)
                    cls, size = L.H5Tget_class(tid), L.H5Tget_size(tid)
                    if cls == H5T_INTEGER and L.H5Tget_sign(tid) == H5T_SGN_NONE and size in (4, 8):
                        dtype = np.uint32 if size == 4 else np.uint64
                    elif cls == H5T_FLOAT:
                        dtype = np.float64
                    else:
                        raise ValueError(f"{name}: unsupported stored type (class {cls}, {size} bytes)")
                    a = np.empty(int(n[0]), dtype)
                    if len(a):
                        _ok(L.H5Dread(did, _mem_type(dtype), H5S_ALL, H5S_ALL, H5P_DEFAULT, a.ctypes.data),
                            f"H5Dread({name})")
                    out[name] = a
                finally:
                    L.H5Tclose(tid)
                    L.H5Sclose(sid)
            finally:
                L.H5Dclose(did)
    finally:
        L.H5Fclose(fid)
    return out


def is_hdf5(path):
    """True when the file starts with the HDF5 signature."""
    try:
        with open(path, "rb") as f:
            return f.read(8) == b"\x89HDF\r\n\x1a\n"
    except OSError:
        return False
