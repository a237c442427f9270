"""A small HDF5 writer for FileReduction (reductions.jl:57-91), over the HDF5 C library through ctypes.

Arrays are written the way HDF5.jl writes a Julia array: HDF5 stores dimensions slowest-first, Julia
arrays are column-major, so a Julia array of size (d1, ..., dk) appears in the file with dimensions
(dk, ..., d1).  The Python mirror keeps Julia's shapes in its numpy arrays ((ndofs, natoms) positions),
so each array is written transposed; a vector of arrays is stacked on a new last Julia axis
(reshape(reduce(hcat, v), size(v[1])..., :)), i.e. a new first file axis.  Complex numbers use HDF5.jl's
compound type {r, i}.  The library is found through NQCD_HDF5_LIB, the loader's search path or the
conda prefix of the image."""
import ctypes as C
import ctypes.util
import os

import numpy as np

_H5F_ACC_RDWR, _H5F_ACC_EXCL = 0x0001, 0x0004
_H5P_DEFAULT = 0
_H5S_ALL = 0
_H5S_SCALAR = 0
_H5T_COMPOUND = 6

_lib = None


def _candidates():
    env = os.environ.get("NQCD_HDF5_LIB")
    if env:
        yield env
    found = ctypes.util.find_library("hdf5")
    if found:
        yield found
    for prefix in (os.environ.get("CONDA_PREFIX"), "/opt/conda", "/usr"):
        if prefix:
            yield os.path.join(prefix, "lib", "libhdf5.so")
    yield "/usr/lib/x86_64-linux-gnu/hdf5/serial/libhdf5.so"


def library():
    """The HDF5 C library (loaded once); raises OSError naming the paths tried when it is absent."""
    global _lib
    if _lib is not None:
        return _lib
    tried = []
    for path in _candidates():
        tried.append(path)
        try:
            lib = C.CDLL(path)
        except OSError:
            continue
        hid, herr = C.c_int64, C.c_int
        lib.H5open.restype = herr
        if lib.H5open() < 0:
            continue
        sig = {
            "H5Fcreate": (hid, [C.c_char_p, C.c_uint, hid, hid]),
            "H5Fopen": (hid, [C.c_char_p, C.c_uint, hid]),
            "H5Fclose": (herr, [hid]),
            "H5Gcreate2": (hid, [hid, C.c_char_p, hid, hid, hid]),
            "H5Gclose": (herr, [hid]),
            "H5Lexists": (C.c_int, [hid, C.c_char_p, hid]),
            "H5Screate": (hid, [C.c_int]),
            "H5Screate_simple": (hid, [C.c_int, C.POINTER(C.c_ulonglong), C.POINTER(C.c_ulonglong)]),
            "H5Sclose": (herr, [hid]),
            "H5Dcreate2": (hid, [hid, C.c_char_p, hid, hid, hid, hid, hid]),
            "H5Dwrite": (herr, [hid, hid, hid, hid, hid, C.c_void_p]),
            "H5Dclose": (herr, [hid]),
            "H5Tcreate": (hid, [C.c_int, C.c_size_t]),
            "H5Tinsert": (herr, [hid, C.c_char_p, C.c_size_t, hid]),
            "H5Tclose": (herr, [hid]),
        }
        for name, (res, args) in sig.items():
            f = getattr(lib, name)
            f.restype, f.argtypes = res, args
        lib.types = {k: C.c_int64.in_dll(lib, k + "_g").value for k in
                     ("H5T_NATIVE_DOUBLE", "H5T_NATIVE_INT64", "H5T_NATIVE_UINT8", "H5T_IEEE_F64LE", "H5T_STD_I64LE",
                      "H5T_STD_U8LE")}
        _lib = lib
        return lib
    raise OSError("FileReduction needs the HDF5 C library (tried: " + ", ".join(tried) + "); set NQCD_HDF5_LIB")


def _check(rc, what):
    if rc < 0:
        raise OSError(f"HDF5: {what} failed")
    return rc


class File:
    """An HDF5 file opened as HDF5.jl's h5open(name, "cw"): read-write, created when absent."""

    def __init__(self, filename):
        self.lib = library()
        name = os.fsencode(filename)
        if os.path.exists(filename):
            self.fid = _check(self.lib.H5Fopen(name, _H5F_ACC_RDWR, _H5P_DEFAULT), f"opening {filename}")
        else:
            self.fid = _check(self.lib.H5Fcreate(name, _H5F_ACC_EXCL, _H5P_DEFAULT, _H5P_DEFAULT), f"creating {filename}")

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def close(self):
        if self.fid is not None:
            self.lib.H5Fclose(self.fid)
            self.fid = None

    def create_group(self, name):
        """A new group (an existing name is an error, as HDF5.create_group)."""
        b = name.encode()
        if self.lib.H5Lexists(self.fid, b, _H5P_DEFAULT) > 0:
            raise ValueError(f"HDF5 group {name} already exists")
        return Group(self.lib, _check(self.lib.H5Gcreate2(self.fid, b, _H5P_DEFAULT, _H5P_DEFAULT, _H5P_DEFAULT),
                                      f"creating group {name}"))


class Group:
    def __init__(self, lib, gid):
        self.lib, self.gid = lib, gid

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.lib.H5Gclose(self.gid)

    def write(self, name, data):
        """Writes a numpy array (already in file order, C-contiguous) or a numpy scalar as dataset `name`."""
        lib, t = self.lib, self.lib.types
        a = np.asarray(data)
        scalar = a.ndim == 0
        owned = []
        if np.iscomplexobj(a):
            a = np.ascontiguousarray(a, dtype=np.complex128)
            mem = _check(lib.H5Tcreate(_H5T_COMPOUND, 16), "complex type")
            owned.append(mem)
            _check(lib.H5Tinsert(mem, b"r", 0, t["H5T_NATIVE_DOUBLE"]), "complex type")
            _check(lib.H5Tinsert(mem, b"i", 8, t["H5T_NATIVE_DOUBLE"]), "complex type")
            ftype = mem
        elif a.dtype == np.bool_:
            a = np.ascontiguousarray(a, dtype=np.uint8)
            mem, ftype = t["H5T_NATIVE_UINT8"], t["H5T_STD_U8LE"]
        elif np.issubdtype(a.dtype, np.integer):
            a = np.ascontiguousarray(a, dtype=np.int64)
            mem, ftype = t["H5T_NATIVE_INT64"], t["H5T_STD_I64LE"]
        elif np.issubdtype(a.dtype, np.floating):
            a = np.ascontiguousarray(a, dtype=np.float64)
            mem, ftype = t["H5T_NATIVE_DOUBLE"], t["H5T_IEEE_F64LE"]
        else:
            raise ValueError("Cannot convert output type to HDF5 format")
        if scalar:  # np.ascontiguousarray made it a one-element vector; the dataspace stays scalar
            space = _check(lib.H5Screate(_H5S_SCALAR), "dataspace")
        else:
            dims = (C.c_ulonglong * a.ndim)(*a.shape)
            space = _check(lib.H5Screate_simple(a.ndim, dims, None), "dataspace")
        try:
            ds = _check(lib.H5Dcreate2(self.gid, name.encode(), ftype, space, _H5P_DEFAULT, _H5P_DEFAULT, _H5P_DEFAULT),
                        f"creating dataset {name}")
            try:
                _check(lib.H5Dwrite(ds, mem, _H5S_ALL, _H5S_ALL, _H5P_DEFAULT, a.ctypes.data_as(C.c_void_p)),
                       f"writing dataset {name}")
            finally:
                lib.H5Dclose(ds)
        finally:
            lib.H5Sclose(space)
            for x in owned:
                lib.H5Tclose(x)


def file_order(value):
    """The array HDF5.jl would write for one output value (reductions.jl:77-86): numbers and arrays of
    numbers as they are, a vector of arrays stacked on a new last axis; returned in file (reversed) order.
    Anything else raises, as the reference does."""
    if isinstance(value, (list, tuple)) and len(value) > 0 and all(isinstance(x, np.ndarray) for x in value):
        first = value[0].shape
        if any(x.shape != first for x in value):
            raise ValueError("Cannot convert output type to HDF5 format")
        return np.ascontiguousarray(np.stack([x.T for x in value]))
    if isinstance(value, (list, tuple)):
        arr = np.asarray(value)
        if arr.dtype == object:
            raise ValueError("Cannot convert output type to HDF5 format")
        return np.ascontiguousarray(arr.T)
    if isinstance(value, (np.ndarray, np.generic, int, float, complex, bool)) and not isinstance(value, str):
        arr = np.asarray(value)
        if arr.dtype == object:
            raise ValueError("Cannot convert output type to HDF5 format")
        return arr if arr.ndim == 0 else np.ascontiguousarray(arr.T)
    raise ValueError("Cannot convert output type to HDF5 format")
