"""Reads an .hdf5 file back through the HDF5 library that libweldform_io.so
links (ctypes over the C and high-level APIs): the names of the root
group's datasets in name order, and each dataset's values."""
import ctypes
import os

import numpy as np

from weldform_amd import build

H5F_ACC_RDONLY = 0
H5P_DEFAULT = 0
H5T_INTEGER, H5T_FLOAT = 0, 1


class _GInfo(ctypes.Structure):
    _fields_ = [("storage_type", ctypes.c_int), ("nlinks", ctypes.c_uint64), ("max_corder", ctypes.c_int64),
                ("mounted", ctypes.c_bool)]


def available():
    return build.hdf5_prefix()[0] is not None


def _lib():
    _, libdir = build.hdf5_prefix()
    L = ctypes.CDLL(os.path.join(libdir, "libhdf5_hl.so"))
    hid = ctypes.c_int64
    L.H5open.restype = ctypes.c_int
    L.H5Fopen.argtypes, L.H5Fopen.restype = [ctypes.c_char_p, ctypes.c_uint, hid], hid
    L.H5Fclose.argtypes, L.H5Fclose.restype = [hid], ctypes.c_int
    L.H5Gget_info.argtypes, L.H5Gget_info.restype = [hid, ctypes.POINTER(_GInfo)], ctypes.c_int
    L.H5Lget_name_by_idx.argtypes = [hid, ctypes.c_char_p, ctypes.c_int, ctypes.c_int, ctypes.c_uint64,
                                     ctypes.c_char_p, ctypes.c_size_t, hid]
    L.H5Lget_name_by_idx.restype = ctypes.c_ssize_t
    L.H5LTget_dataset_info.argtypes = [hid, ctypes.c_char_p, ctypes.POINTER(ctypes.c_uint64),
                                       ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_size_t)]
    L.H5LTget_dataset_info.restype = ctypes.c_int
    L.H5LTget_dataset_ndims.argtypes = [hid, ctypes.c_char_p, ctypes.POINTER(ctypes.c_int)]
    L.H5LTget_dataset_ndims.restype = ctypes.c_int
    L.H5LTread_dataset_float.argtypes = [hid, ctypes.c_char_p, ctypes.c_void_p]
    L.H5LTread_dataset_float.restype = ctypes.c_int
    L.H5LTread_dataset_int.argtypes = [hid, ctypes.c_char_p, ctypes.c_void_p]
    L.H5LTread_dataset_int.restype = ctypes.c_int
    L.H5open()
    return L


def read(path):
    """{name: (dims tuple, numpy array)} of every dataset under /."""
    L = _lib()
    f = L.H5Fopen(str(path).encode(), H5F_ACC_RDONLY, H5P_DEFAULT)
    if f < 0:
        raise OSError(f"cannot open {path}")
    try:
        info = _GInfo()
        assert L.H5Gget_info(f, ctypes.byref(info)) >= 0
        out = {}
        for k in range(info.nlinks):
            buf = ctypes.create_string_buffer(256)
            assert L.H5Lget_name_by_idx(f, b".", 0, 0, k, buf, 256, H5P_DEFAULT) > 0
            name = buf.value
            nd = ctypes.c_int(0)
            assert L.H5LTget_dataset_ndims(f, name, ctypes.byref(nd)) >= 0
            dims = (ctypes.c_uint64 * max(nd.value, 1))()
            cls, size = ctypes.c_int(0), ctypes.c_size_t(0)
            assert L.H5LTget_dataset_info(f, name, dims, ctypes.byref(cls), ctypes.byref(size)) >= 0
            shape = tuple(dims[i] for i in range(nd.value))
            if cls.value == H5T_FLOAT:
                a = np.empty(int(np.prod(shape)), dtype=np.float32)
                assert size.value == 4 and L.H5LTread_dataset_float(f, name, a.ctypes.data) >= 0
            elif cls.value == H5T_INTEGER:
                a = np.empty(int(np.prod(shape)), dtype=np.int32)
                assert size.value == 4 and L.H5LTread_dataset_int(f, name, a.ctypes.data) >= 0
            else:
                raise AssertionError(f"unexpected class {cls.value} of {name}")
            out[name.decode()] = (shape, a)
        return out
    finally:
        L.H5Fclose(f)
