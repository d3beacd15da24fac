"""Minimal serial HDF5 access through the C library (ctypes), for the
diagnostics and restart files of pic2d.f90 (fields%05d.h5, particles.h5,
energy_tot.h5, spectra.h5; :1359-2115).  The reference writes them with
parallel HDF5 from Fortran; a file written here has the same dataset names,
types and shapes as seen from C (Fortran dims reversed), so readers of the
reference's files read these alike.

The library is looked up in $PBPIC_HDF5_LIB, then /opt/conda/lib, then the
loader's default path; HDF5 >= 1.10 (64-bit hid_t).  Only what the
diagnostics need: files, groups, whole-dataset write/read of contiguous
numpy arrays of float32/float64/int32/int64.
"""
import ctypes
import os

import numpy as np

_H5F_ACC_RDONLY, _H5F_ACC_RDWR, _H5F_ACC_TRUNC = 0x0, 0x1, 0x2
_H5P_DEFAULT = 0
_H5S_ALL = 0

_lib = None
_types = {}


class H5Error(RuntimeError):
    pass


def _load():
    global _lib
    if _lib is not None:
        return _lib
    cands = [os.environ.get("PBPIC_HDF5_LIB"), "/opt/conda/lib/libhdf5.so", "libhdf5.so",
             "libhdf5_serial.so"]
    err = None
    for c in cands:
        if not c:
            continue
        try:
            L = ctypes.CDLL(c)
            break
        except OSError as e:
            err = e
    else:
        raise H5Error(f"no HDF5 C library found (set PBPIC_HDF5_LIB): {err}")
    hid, herr, hs = ctypes.c_int64, ctypes.c_int, ctypes.c_uint64
    cs, vp, u = ctypes.c_char_p, ctypes.c_void_p, ctypes.c_uint
    P = ctypes.POINTER
    sig = {
        "H5open": (herr, []),
        "H5Fcreate": (hid, [cs, u, hid, hid]),
        "H5Fopen": (hid, [cs, u, hid]),
        "H5Fclose": (herr, [hid]),
        "H5Gcreate2": (hid, [hid, cs, hid, hid, hid]),
        "H5Gopen2": (hid, [hid, cs, hid]),
        "H5Gclose": (herr, [hid]),
        "H5Screate_simple": (hid, [ctypes.c_int, P(hs), P(hs)]),
        "H5Sclose": (herr, [hid]),
        "H5Sget_simple_extent_ndims": (ctypes.c_int, [hid]),
        "H5Sget_simple_extent_dims": (ctypes.c_int, [hid, P(hs), P(hs)]),
        "H5Dcreate2": (hid, [hid, cs, hid, hid, hid, hid, hid]),
        "H5Dopen2": (hid, [hid, cs, hid]),
        "H5Dclose": (herr, [hid]),
        "H5Dwrite": (herr, [hid, hid, hid, hid, hid, vp]),
        "H5Dread": (herr, [hid, hid, hid, hid, hid, vp]),
        "H5Dget_space": (hid, [hid]),
        "H5Dget_type": (hid, [hid]),
        "H5Tget_class": (ctypes.c_int, [hid]),
        "H5Tget_size": (ctypes.c_size_t, [hid]),
        "H5Tclose": (herr, [hid]),
        "H5Lexists": (ctypes.c_int, [hid, cs, hid]),
    }
    for name, (res, args) in sig.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    if L.H5open() < 0:
        raise H5Error("H5open failed")
    for key, sym in (("float32", "H5T_NATIVE_FLOAT_g"), ("float64", "H5T_NATIVE_DOUBLE_g"),
                     ("int32", "H5T_NATIVE_INT_g"), ("int64", "H5T_NATIVE_LLONG_g")):
        _types[key] = ctypes.c_int64.in_dll(L, sym).value
    _lib = L
    return L


def available():
    try:
        _load()
        return True
    except H5Error:
        return False


def _chk(v, what):
    if v < 0:
        raise H5Error(f"HDF5 call failed: {what}")
    return v


class File:
    """An HDF5 file: mode "w" (truncate), "a" (read/write, created if
    missing) or "r"."""

    def __init__(self, path, mode="r"):
        L = _load()
        self._L = L
        p = os.fsencode(path)
        if mode == "w" or (mode == "a" and not os.path.exists(path)):
            self.id = _chk(L.H5Fcreate(p, _H5F_ACC_TRUNC, _H5P_DEFAULT, _H5P_DEFAULT), path)
        elif mode in ("a", "r"):
            flag = _H5F_ACC_RDWR if mode == "a" else _H5F_ACC_RDONLY
            self.id = _chk(L.H5Fopen(p, flag, _H5P_DEFAULT), path)
        else:
            raise ValueError(mode)

    def close(self):
        if self.id is not None:
            self._L.H5Fclose(self.id)
            self.id = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def _loc(self, name):
        """Opens the group holding `name` ("a/b/c" -> group a/b), creating it
        in write modes; returns (loc id, leaf name, opened group ids)."""
        parts = [x for x in name.split("/") if x]
        loc, opened = self.id, []
        for g in parts[:-1]:
            gb = g.encode()
            if self._L.H5Lexists(loc, gb, _H5P_DEFAULT) > 0:
                gid = _chk(self._L.H5Gopen2(loc, gb, _H5P_DEFAULT), g)
            else:
                gid = _chk(self._L.H5Gcreate2(loc, gb, _H5P_DEFAULT, _H5P_DEFAULT, _H5P_DEFAULT), g)
            opened.append(gid)
            loc = gid
        return loc, parts[-1].encode(), opened

    def __contains__(self, name):
        parts = [x for x in name.split("/") if x]
        loc, opened = self.id, []
        try:
            for i, g in enumerate(parts):
                if self._L.H5Lexists(loc, g.encode(), _H5P_DEFAULT) <= 0:
                    return False
                if i < len(parts) - 1:
                    loc = _chk(self._L.H5Gopen2(loc, g.encode(), _H5P_DEFAULT), g)
                    opened.append(loc)
            return True
        finally:
            for g in reversed(opened):
                self._L.H5Gclose(g)

    def write(self, name, array):
        """Creates dataset `name` (or overwrites one of the same shape and
        type) with the whole array."""
        a = np.ascontiguousarray(array)
        key = a.dtype.name
        if key not in _types:
            raise TypeError(f"unsupported dtype {a.dtype}")
        loc, leaf, opened = self._loc(name)
        try:
            if self._L.H5Lexists(loc, leaf, _H5P_DEFAULT) > 0:
                d = _chk(self._L.H5Dopen2(loc, leaf, _H5P_DEFAULT), name)
            else:
                dims = (ctypes.c_uint64 * max(a.ndim, 1))(*(a.shape or (1,)))
                sp = _chk(self._L.H5Screate_simple(max(a.ndim, 1), dims, None), name)
                d = self._L.H5Dcreate2(loc, leaf, _types[key], sp, _H5P_DEFAULT, _H5P_DEFAULT,
                                       _H5P_DEFAULT)
                self._L.H5Sclose(sp)
                _chk(d, name)
            try:
                _chk(self._L.H5Dwrite(d, _types[key], _H5S_ALL, _H5S_ALL, _H5P_DEFAULT,
                                      a.ctypes.data_as(ctypes.c_void_p)), name)
            finally:
                self._L.H5Dclose(d)
        finally:
            for g in reversed(opened):
                self._L.H5Gclose(g)

    def read(self, name):
        parts = [x for x in name.split("/") if x]
        loc, opened = self.id, []
        try:
            for g in parts[:-1]:
                loc = _chk(self._L.H5Gopen2(loc, g.encode(), _H5P_DEFAULT), g)
                opened.append(loc)
            d = _chk(self._L.H5Dopen2(loc, parts[-1].encode(), _H5P_DEFAULT), name)
            try:
                sp = self._L.H5Dget_space(d)
                nd = self._L.H5Sget_simple_extent_ndims(sp)
                dims = (ctypes.c_uint64 * max(nd, 1))()
                self._L.H5Sget_simple_extent_dims(sp, dims, None)
                self._L.H5Sclose(sp)
                t = self._L.H5Dget_type(d)
                cls, size = self._L.H5Tget_class(t), self._L.H5Tget_size(t)
                self._L.H5Tclose(t)
                # H5T_INTEGER = 0, H5T_FLOAT = 1
                dt = {(0, 4): np.int32, (0, 8): np.int64, (1, 4): np.float32,
                      (1, 8): np.float64}.get((cls, size))
                if dt is None:
                    raise H5Error(f"{name}: unsupported type class {cls} size {size}")
                out = np.empty(tuple(dims[:nd]), dt)
                _chk(self._L.H5Dread(d, _types[np.dtype(dt).name], _H5S_ALL, _H5S_ALL,
                                     _H5P_DEFAULT, out.ctypes.data_as(ctypes.c_void_p)), name)
                return out
            finally:
                self._L.H5Dclose(d)
        finally:
            for g in reversed(opened):
                self._L.H5Gclose(g)
