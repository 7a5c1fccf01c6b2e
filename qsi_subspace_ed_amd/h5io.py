"""HDF5 files in the reference's layouts, through the HDF5 C library (ctypes).

  basis file   dataset `basis` (or a sector name), (D, 2) native uint64, row =
               [lo, hi] (components/common/include/basis_io_h5.hpp:30-70,
               read back as in :90-141)
  .eigs.h5     `eigenvalues` (k,) float64; `eigenvectors` with dims (D, k)
               holding the column-major buffer (vector j at offset j*D), as
               diag_DOQSI_ham writes Eigen's buffer (src/diag_DOQSI_ham.cpp:
               149-173); `latfile_json` and `dset_name` variable-length
               strings (src/expectation_eval.cpp:268-284)
  .out.h5      `latfile_json`, `eigenvalues`, `ring` (n, nv, nv), `ring_ring`
               (n, n, nv, nv), `partial_vol_sl0..3` (src/eval_observables.cpp:
               120-206, src/expectation_eval.cpp:210-249)

The library is found through QSI_HDF5_LIB, the conda prefix of this image or
the loader path; without it these functions raise.  No counterpart is needed
on the device: the files are written from host arrays.
"""
import ctypes as C
import ctypes.util
import os

import numpy as np

_H5 = None
_hid = C.c_int64
_H5P_DEFAULT = 0
_H5S_ALL = 0
_H5F_ACC_RDONLY = 0x0000
_H5F_ACC_TRUNC = 0x0002
_H5S_SCALAR = 0
_H5T_VARIABLE = C.c_size_t(-1).value


def _candidates():
    env = os.environ.get("QSI_HDF5_LIB")
    if env:
        yield env
    yield "/opt/conda/lib/libhdf5.so.103"
    yield "/opt/conda/lib/libhdf5.so"
    found = ctypes.util.find_library("hdf5")
    if found:
        yield found


def _lib():
    global _H5
    if _H5 is not None:
        return _H5
    err = []
    for path in _candidates():
        try:
            h = C.CDLL(path)
            break
        except OSError as e:
            err.append(f"{path}: {e}")
    else:
        raise RuntimeError("HDF5 C library not found (set QSI_HDF5_LIB): " + "; ".join(err))
    h.H5open.restype = C.c_int
    if h.H5open() < 0:
        raise RuntimeError("H5open failed")
    for name, args, res in (
            ("H5Fcreate", [C.c_char_p, C.c_uint, _hid, _hid], _hid),
            ("H5Fopen", [C.c_char_p, C.c_uint, _hid], _hid),
            ("H5Fclose", [_hid], C.c_int),
            ("H5Screate_simple", [C.c_int, C.POINTER(C.c_uint64), C.c_void_p], _hid),
            ("H5Screate", [C.c_int], _hid),
            ("H5Sclose", [_hid], C.c_int),
            ("H5Sget_simple_extent_ndims", [_hid], C.c_int),
            ("H5Sget_simple_extent_dims", [_hid, C.POINTER(C.c_uint64), C.c_void_p], C.c_int),
            ("H5Dcreate2", [_hid, C.c_char_p, _hid, _hid, _hid, _hid, _hid], _hid),
            ("H5Dopen2", [_hid, C.c_char_p, _hid], _hid),
            ("H5Dwrite", [_hid, _hid, _hid, _hid, _hid, C.c_void_p], C.c_int),
            ("H5Dread", [_hid, _hid, _hid, _hid, _hid, C.c_void_p], C.c_int),
            ("H5Dget_space", [_hid], _hid),
            ("H5Dget_type", [_hid], _hid),
            ("H5Dvlen_reclaim", [_hid, _hid, _hid, C.c_void_p], C.c_int),
            ("H5Dclose", [_hid], C.c_int),
            ("H5Tcopy", [_hid], _hid),
            ("H5Tset_size", [_hid, C.c_size_t], C.c_int),
            ("H5Tclose", [_hid], C.c_int)):
        f = getattr(h, name)
        f.argtypes = args
        f.restype = res
    h.T_DOUBLE = _hid.in_dll(h, "H5T_NATIVE_DOUBLE_g").value
    h.T_UINT64 = _hid.in_dll(h, "H5T_NATIVE_UINT64_g").value
    h.T_C_S1 = _hid.in_dll(h, "H5T_C_S1_g").value
    _H5 = h
    return h


def _ok(v, what):
    if v < 0:
        raise RuntimeError(f"HDF5: {what} failed")
    return v


def _write(fid, name, arr, dims, htype):
    h = _lib()
    d = (C.c_uint64 * len(dims))(*dims)
    sid = _ok(h.H5Screate_simple(len(dims), d, None), "H5Screate_simple")
    try:
        did = _ok(h.H5Dcreate2(fid, name.encode(), htype, sid, _H5P_DEFAULT, _H5P_DEFAULT,
                               _H5P_DEFAULT), f"H5Dcreate2({name})")
        try:
            _ok(h.H5Dwrite(did, htype, _H5S_ALL, _H5S_ALL, _H5P_DEFAULT,
                           arr.ctypes.data_as(C.c_void_p)), f"H5Dwrite({name})")
        finally:
            h.H5Dclose(did)
    finally:
        h.H5Sclose(sid)


def _write_string(fid, name, value):
    """Variable-length C string in a scalar dataspace (write_string_to_hdf5)."""
    h = _lib()
    sid = _ok(h.H5Screate(_H5S_SCALAR), "H5Screate")
    tid = _ok(h.H5Tcopy(h.T_C_S1), "H5Tcopy")
    try:
        _ok(h.H5Tset_size(tid, _H5T_VARIABLE), "H5Tset_size")
        did = _ok(h.H5Dcreate2(fid, name.encode(), tid, sid, _H5P_DEFAULT, _H5P_DEFAULT,
                               _H5P_DEFAULT), f"H5Dcreate2({name})")
        try:
            buf = C.c_char_p(value.encode())
            _ok(h.H5Dwrite(did, tid, _H5S_ALL, _H5S_ALL, _H5P_DEFAULT, C.byref(buf)),
                f"H5Dwrite({name})")
        finally:
            h.H5Dclose(did)
    finally:
        h.H5Tclose(tid)
        h.H5Sclose(sid)


def _read(fid, name, htype, dtype):
    h = _lib()
    did = _ok(h.H5Dopen2(fid, name.encode(), _H5P_DEFAULT), f"H5Dopen2({name})")
    try:
        sid = _ok(h.H5Dget_space(did), "H5Dget_space")
        try:
            nd = _ok(h.H5Sget_simple_extent_ndims(sid), "ndims")
            dims = (C.c_uint64 * max(nd, 1))()
            _ok(h.H5Sget_simple_extent_dims(sid, dims, None), "dims")
            shape = tuple(int(dims[i]) for i in range(nd))
        finally:
            h.H5Sclose(sid)
        out = np.empty(shape, dtype=dtype)
        _ok(h.H5Dread(did, htype, _H5S_ALL, _H5S_ALL, _H5P_DEFAULT,
                      out.ctypes.data_as(C.c_void_p)), f"H5Dread({name})")
        return out
    finally:
        h.H5Dclose(did)


def _read_string(fid, name):
    h = _lib()
    did = _ok(h.H5Dopen2(fid, name.encode(), _H5P_DEFAULT), f"H5Dopen2({name})")
    try:
        tid = _ok(h.H5Dget_type(did), "H5Dget_type")
        sid = _ok(h.H5Dget_space(did), "H5Dget_space")
        try:
            buf = C.c_char_p()
            _ok(h.H5Dread(did, tid, _H5S_ALL, _H5S_ALL, _H5P_DEFAULT, C.byref(buf)),
                f"H5Dread({name})")
            value = buf.value.decode()
            h.H5Dvlen_reclaim(tid, sid, _H5P_DEFAULT, C.byref(buf))
            return value
        finally:
            h.H5Sclose(sid)
            h.H5Tclose(tid)
    finally:
        h.H5Dclose(did)


class _File:
    def __init__(self, path, write):
        h = _lib()
        if write:
            self.fid = _ok(h.H5Fcreate(str(path).encode(), _H5F_ACC_TRUNC, _H5P_DEFAULT,
                                       _H5P_DEFAULT), f"H5Fcreate({path})")
        else:
            self.fid = _ok(h.H5Fopen(str(path).encode(), _H5F_ACC_RDONLY, _H5P_DEFAULT),
                           f"H5Fopen({path})")

    def __enter__(self):
        return self.fid

    def __exit__(self, *exc):
        _lib().H5Fclose(self.fid)


def write_basis(path, states, dset="basis"):
    """(D, 2) uint64 [lo, hi] rows as write_basis_hdf5 writes them."""
    a = np.ascontiguousarray(states, dtype=np.uint64)
    if a.ndim != 2 or a.shape[1] != 2:
        raise ValueError("states must be (D, 2) uint64")
    with _File(path, True) as fid:
        _write(fid, dset, a, a.shape, _lib().T_UINT64)


def read_basis(path, dset="basis"):
    """The (D, 2) uint64 basis dataset of a basis file (read_basis_hdf5)."""
    with _File(path, False) as fid:
        a = _read(fid, dset, _lib().T_UINT64, np.uint64)
    if a.ndim != 2 or a.shape[1] != 2:
        raise ValueError(f"{path}:{dset} is not (D, 2)")
    return a


def write_eigs(path, eigenvalues, eigenvectors, latfile, dset_name="basis"):
    """The .eigs.h5 of diag_DOQSI_ham: eigenvectors (D, k) are stored as the
    column-major buffer under dims (D, k)."""
    ev = np.ascontiguousarray(eigenvalues, dtype=np.float64).reshape(-1)
    V = np.asarray(eigenvectors, dtype=np.float64)
    if V.ndim == 1:
        V = V.reshape(-1, 1)
    buf = np.ascontiguousarray(V.T)  # vector j contiguous at offset j*D
    with _File(path, True) as fid:
        _write(fid, "eigenvalues", ev, ev.shape, _lib().T_DOUBLE)
        _write(fid, "eigenvectors", buf, (V.shape[0], V.shape[1]), _lib().T_DOUBLE)
        _write_string(fid, "latfile_json", os.path.basename(str(latfile)))
        _write_string(fid, "dset_name", dset_name)


def read_eigs(path):
    """(eigenvalues, eigenvectors (D, k), latfile_json, dset_name) of an
    .eigs.h5 written by diag_DOQSI_ham or write_eigs."""
    with _File(path, False) as fid:
        ev = _read(fid, "eigenvalues", _lib().T_DOUBLE, np.float64)
        raw = _read(fid, "eigenvectors", _lib().T_DOUBLE, np.float64)
        lat = _read_string(fid, "latfile_json")
        dn = _read_string(fid, "dset_name")
    D, k = raw.shape
    return ev, raw.reshape(-1).reshape(k, D).T.copy(), lat, dn


def write_observables(path, eigenvalues, datasets, latfile):
    """The .out.h5 of eval_observables: latfile_json, eigenvalues and each
    array of `datasets` (e.g. observables.ring_observables) under its name."""
    ev = np.ascontiguousarray(eigenvalues, dtype=np.float64).reshape(-1)
    with _File(path, True) as fid:
        _write_string(fid, "latfile_json", os.path.basename(str(latfile)))
        _write(fid, "eigenvalues", ev, ev.shape, _lib().T_DOUBLE)
        for name, arr in datasets.items():
            a = np.ascontiguousarray(arr, dtype=np.float64)
            _write(fid, name, a, a.shape, _lib().T_DOUBLE)


def read_dataset(path, name):
    """One float64 dataset of a file, with its shape."""
    with _File(path, False) as fid:
        return _read(fid, name, _lib().T_DOUBLE, np.float64)
