"""Readers for the qew tools' outputs.

The tools write the reference's HDF5 dataset keys (src/qew_io.h:24-63) into
either container:

* an HDF5 file (the default), one FP64 dataset per key, read here through the
  HDF5 C library with ctypes (h5py is not required);
* a "qew-kv 1" text archive (an output name ending in .kv), one dataset per
  line: `key<TAB>count<TAB>values` with values printed as %.17g.

Both readers return {key: value}, where a scalar dataset gives a float and a
profile (a key ending in "/h", or any 1-D dataset) a numpy array.
"""
import ctypes
import os

import numpy as np

_HDF5_ROOTS = ("/usr/local", "/opt/conda")


def read_kv(path):
    """Returns {key: numpy array (or float for single values)}."""
    out = {}
    with open(path) as f:
        header = f.readline().strip()
        if header != "# qew-kv 1":
            raise ValueError(f"{path}: not a qew-kv archive (header {header!r})")
        for line in f:
            line = line.rstrip("\n")
            if not line:
                continue
            key, count, values = line.split("\t")
            arr = np.array([float(v) for v in values.split()]) if values else np.empty(0)
            if arr.size != int(count):
                raise ValueError(f"{path}: {key} has {arr.size} values, header says {count}")
            out[key] = float(arr[0]) if arr.size == 1 and not key.endswith("/h") else arr
    return out


class _Hdf5:
    """The few HDF5 C entry points the reader needs (hid_t is 64-bit since 1.10)."""

    H5I_DATASET = 5
    H5_INDEX_NAME = 0
    H5_ITER_NATIVE = 2

    def __init__(self, path):
        lib = ctypes.CDLL(path)
        hid, herr = ctypes.c_int64, ctypes.c_int
        self.lib = lib
        lib.H5open.restype = herr
        if lib.H5open() < 0:
            raise OSError(f"{path}: H5open failed")
        lib.H5Eset_auto2.argtypes = [hid, ctypes.c_void_p, ctypes.c_void_p]
        lib.H5Eset_auto2(0, None, None)
        lib.H5Fopen.restype, lib.H5Fopen.argtypes = hid, [ctypes.c_char_p, ctypes.c_uint, hid]
        lib.H5Fclose.argtypes = [hid]
        lib.H5Oopen.restype, lib.H5Oopen.argtypes = hid, [hid, ctypes.c_char_p, hid]
        lib.H5Oclose.argtypes = [hid]
        lib.H5Iget_type.restype, lib.H5Iget_type.argtypes = ctypes.c_int, [hid]
        lib.H5Dget_space.restype, lib.H5Dget_space.argtypes = hid, [hid]
        lib.H5Dget_type.restype, lib.H5Dget_type.argtypes = hid, [hid]
        lib.H5Tget_class.restype, lib.H5Tget_class.argtypes = ctypes.c_int, [hid]
        lib.H5Tclose.argtypes = [hid]
        lib.H5Sget_simple_extent_ndims.restype = ctypes.c_int
        lib.H5Sget_simple_extent_ndims.argtypes = [hid]
        lib.H5Sget_simple_extent_dims.argtypes = [hid, ctypes.c_void_p, ctypes.c_void_p]
        lib.H5Sclose.argtypes = [hid]
        lib.H5Dread.restype = herr
        lib.H5Dread.argtypes = [hid, hid, hid, hid, hid, ctypes.c_void_p]
        self.native_double = hid.in_dll(lib, "H5T_NATIVE_DOUBLE_g").value
        # H5Lvisit's callback: (group, name, link info, user data) -> herr_t
        self.visit_cb = ctypes.CFUNCTYPE(herr, hid, ctypes.c_char_p, ctypes.c_void_p,
                                         ctypes.c_void_p)
        self.visit = getattr(lib, "H5Lvisit2", None) or lib.H5Lvisit
        self.visit.restype = herr
        self.visit.argtypes = [hid, ctypes.c_int, ctypes.c_int, self.visit_cb, ctypes.c_void_p]


_hdf5 = None


def _hdf5_library():
    global _hdf5
    if _hdf5 is None:
        cands = [os.environ.get("QEW_HDF5_LIBRARY", "")]
        cands += [os.path.join(r, "lib", "libhdf5.so") for r in _HDF5_ROOTS]
        errors = []
        for c in cands:
            if c and os.path.exists(c):
                try:
                    _hdf5 = _Hdf5(c)
                    break
                except OSError as e:
                    errors.append(str(e))
        if _hdf5 is None:
            raise OSError("no loadable HDF5 C library (set QEW_HDF5_LIBRARY) " + "; ".join(errors))
    return _hdf5


def read_h5(path):
    """Returns {key: float | numpy array} for every FP64 dataset in an HDF5 file,
    keyed by its path without the leading '/'."""
    h5 = _hdf5_library()
    lib = h5.lib
    f = lib.H5Fopen(os.fsencode(path), 0, 0)  # H5F_ACC_RDONLY, H5P_DEFAULT
    if f < 0:
        raise OSError(f"{path}: not an HDF5 file")
    names = []

    def collect(_group, name, _info, _data):
        names.append(name)
        return 0

    out = {}
    try:
        if h5.visit(f, h5.H5_INDEX_NAME, h5.H5_ITER_NATIVE, h5.visit_cb(collect), None) < 0:
            raise OSError(f"{path}: cannot list its objects")
        for name in names:
            obj = lib.H5Oopen(f, name, 0)
            if obj < 0:
                raise OSError(f"{path}: cannot open {name!r}")
            try:
                if lib.H5Iget_type(obj) != h5.H5I_DATASET:
                    continue
                key = name.decode()
                t = lib.H5Dget_type(obj)
                cls = lib.H5Tget_class(t)
                lib.H5Tclose(t)
                if cls != 1:  # H5T_FLOAT
                    raise ValueError(f"{path}: dataset {key!r} is not floating point")
                space = lib.H5Dget_space(obj)
                nd = lib.H5Sget_simple_extent_ndims(space)
                dims = (ctypes.c_uint64 * max(nd, 1))()
                if nd > 0:
                    lib.H5Sget_simple_extent_dims(space, dims, None)
                lib.H5Sclose(space)
                shape = tuple(dims[i] for i in range(nd))
                buf = np.empty(shape, dtype=np.float64)
                if lib.H5Dread(obj, h5.native_double, 0, 0, 0, buf.ctypes.data) < 0:
                    raise OSError(f"{path}: cannot read {key!r}")
                out[key] = float(buf) if nd == 0 else buf
            finally:
                lib.H5Oclose(obj)
    finally:
        lib.H5Fclose(f)
    return out


def read_output(path):
    """Reads a tool's output file: a qew-kv archive for a .kv name, else HDF5."""
    return read_kv(path) if str(path).endswith(".kv") else read_h5(path)
