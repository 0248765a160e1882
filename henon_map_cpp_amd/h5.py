"""HDF5 datasets through the HDF5 C library (libhdf5, loaded with ctypes), for
the parameter-scan output: the reference's run_sim.py writes each scan's
survival turns with h5py as dataset "steps" of a new file
(htcondor_jobs/run_sim.py:55-59); h5py is not importable here, but the library
it wraps is. write_dataset stores a numpy array of uint8/int32/uint32/int64/
uint64/float64 under the little-endian standard type h5py would choose for it
(contiguous layout); read_dataset reads one back.

The library is looked up as $HM_HDF5_LIB, then libhdf5.so / libhdf5.so.103 on
the loader path, then next to the running interpreter (sys.prefix/lib). Without
it, the first call raises ImportError."""
from __future__ import annotations

import ctypes
import os
import sys

import numpy as np

_hid = ctypes.c_int64
_hsize = ctypes.c_uint64
_herr = ctypes.c_int
_lib = None

# numpy dtype -> (file type, memory type) global names
_TYPES = {
    np.dtype(np.uint8): ("H5T_STD_U8LE_g", "H5T_NATIVE_UINT8_g"),
    np.dtype(np.int32): ("H5T_STD_I32LE_g", "H5T_NATIVE_INT32_g"),
    np.dtype(np.uint32): ("H5T_STD_U32LE_g", "H5T_NATIVE_UINT32_g"),
    np.dtype(np.int64): ("H5T_STD_I64LE_g", "H5T_NATIVE_INT64_g"),
    np.dtype(np.uint64): ("H5T_STD_U64LE_g", "H5T_NATIVE_UINT64_g"),
    np.dtype(np.float64): ("H5T_IEEE_F64LE_g", "H5T_NATIVE_DOUBLE_g"),
}
_H5F_ACC_RDONLY, _H5F_ACC_TRUNC = 0x0000, 0x0002
_H5P_DEFAULT = _H5S_ALL = _H5E_DEFAULT = 0
_H5T_INTEGER, _H5T_FLOAT = 0, 1


def _candidates():
    env = os.environ.get("HM_HDF5_LIB")
    if env:
        yield env
    yield "libhdf5.so"
    yield "libhdf5.so.103"
    yield os.path.join(sys.prefix, "lib", "libhdf5.so")
    yield os.path.join(sys.base_prefix, "lib", "libhdf5.so")
    yield "/opt/conda/lib/libhdf5.so"


def lib():
    """The loaded HDF5 library (initialised, automatic error printing off)."""
    global _lib
    if _lib is not None:
        return _lib
    errors = []
    for name in _candidates():
        try:
            L = ctypes.CDLL(name)
            break
        except OSError as e:
            errors.append(f"{name}: {e}")
    else:
        raise ImportError("HDF5 output needs the HDF5 C library (libhdf5); tried " + "; ".join(errors))
    sig = {
        "H5open": (_herr, []),
        "H5Eset_auto2": (_herr, [_hid, ctypes.c_void_p, ctypes.c_void_p]),
        "H5Fcreate": (_hid, [ctypes.c_char_p, ctypes.c_uint, _hid, _hid]),
        "H5Fopen": (_hid, [ctypes.c_char_p, ctypes.c_uint, _hid]),
        "H5Fclose": (_herr, [_hid]),
        "H5Screate_simple": (_hid, [ctypes.c_int, ctypes.POINTER(_hsize), ctypes.POINTER(_hsize)]),
        "H5Sget_simple_extent_ndims": (ctypes.c_int, [_hid]),
        "H5Sget_simple_extent_dims": (ctypes.c_int, [_hid, ctypes.POINTER(_hsize), ctypes.POINTER(_hsize)]),
        "H5Sclose": (_herr, [_hid]),
        "H5Dcreate2": (_hid, [_hid, ctypes.c_char_p, _hid, _hid, _hid, _hid, _hid]),
        "H5Dopen2": (_hid, [_hid, ctypes.c_char_p, _hid]),
        "H5Dget_space": (_hid, [_hid]),
        "H5Dget_type": (_hid, [_hid]),
        "H5Dwrite": (_herr, [_hid, _hid, _hid, _hid, _hid, ctypes.c_void_p]),
        "H5Dread": (_herr, [_hid, _hid, _hid, _hid, _hid, ctypes.c_void_p]),
        "H5Dclose": (_herr, [_hid]),
        "H5Tget_class": (ctypes.c_int, [_hid]),
        "H5Tget_size": (ctypes.c_size_t, [_hid]),
        "H5Tget_sign": (ctypes.c_int, [_hid]),
        "H5Tclose": (_herr, [_hid]),
    }
    for fn, (res, args) in sig.items():
        f = getattr(L, fn)
        f.restype, f.argtypes = res, args
    if L.H5open() < 0:
        raise OSError("H5open failed")
    L.H5Eset_auto2(_H5E_DEFAULT, None, None)
    _lib = L
    return L


def _type_id(name):
    return _hid.in_dll(lib(), name).value


def _ok(v, what):
    if v < 0:
        raise OSError(f"HDF5: {what} failed")
    return v


def write_dataset(path: str, name: str, data) -> None:
    """Creates (or truncates) the HDF5 file `path` with one dataset `name`
    holding `data` (any of the dtypes above, any shape)."""
    L = lib()
    a = np.ascontiguousarray(data)
    if a.dtype not in _TYPES:
        raise TypeError(f"write_dataset: unsupported dtype {a.dtype}")
    ftype, mtype = (_type_id(t) for t in _TYPES[a.dtype])
    dims = (_hsize * max(a.ndim, 1))(*(a.shape or (1,)))
    f = _ok(L.H5Fcreate(os.fsencode(path), _H5F_ACC_TRUNC, _H5P_DEFAULT, _H5P_DEFAULT), f"creating {path}")
    try:
        sp = _ok(L.H5Screate_simple(max(a.ndim, 1), dims, None), "dataspace")
        try:
            d = _ok(L.H5Dcreate2(f, name.encode(), ftype, sp, _H5P_DEFAULT, _H5P_DEFAULT, _H5P_DEFAULT),
                    f"creating dataset {name}")
            try:
                _ok(L.H5Dwrite(d, mtype, _H5S_ALL, _H5S_ALL, _H5P_DEFAULT, a.ctypes.data), f"writing {name}")
            finally:
                L.H5Dclose(d)
        finally:
            L.H5Sclose(sp)
    finally:
        _ok(L.H5Fclose(f), f"closing {path}")


def read_dataset(path: str, name: str) -> np.ndarray:
    """Dataset `name` of the HDF5 file `path` as a numpy array (integer and
    float datasets of the sizes above)."""
    L = lib()
    f = _ok(L.H5Fopen(os.fsencode(path), _H5F_ACC_RDONLY, _H5P_DEFAULT), f"opening {path}")
    try:
        d = _ok(L.H5Dopen2(f, name.encode(), _H5P_DEFAULT), f"opening dataset {name}")
        try:
            t = _ok(L.H5Dget_type(d), "dataset type")
            try:
                cls, size, sign = L.H5Tget_class(t), L.H5Tget_size(t), L.H5Tget_sign(t)
            finally:
                L.H5Tclose(t)
            if cls == _H5T_FLOAT and size == 8:
                dt = np.dtype(np.float64)
            elif cls == _H5T_INTEGER:
                dt = np.dtype(("u" if sign == 0 else "i") + str(size))
            else:
                raise TypeError(f"read_dataset: unsupported type class {cls} of size {size}")
            if dt not in _TYPES:
                raise TypeError(f"read_dataset: unsupported integer type {dt}")
            sp = _ok(L.H5Dget_space(d), "dataspace")
            try:
                nd = _ok(L.H5Sget_simple_extent_ndims(sp), "rank")
                dims = (_hsize * max(nd, 1))()
                _ok(L.H5Sget_simple_extent_dims(sp, dims, None), "extent")
            finally:
                L.H5Sclose(sp)
            out = np.empty(tuple(dims[k] for k in range(nd)), dtype=dt)
            _ok(L.H5Dread(d, _type_id(_TYPES[dt][1]), _H5S_ALL, _H5S_ALL, _H5P_DEFAULT, out.ctypes.data),
                f"reading {name}")
            return out
        finally:
            L.H5Dclose(d)
    finally:
        L.H5Fclose(f)
