"""Output files of the reference's solvers from a device context:
Domain::WriteXDMF (FileKey.hdf5 + FileKey.xmf, Source/Output.cpp:75-518) and
Domain::WriteCSV (FileKey.csv, Source/Output.cpp:521-554).

The particle tables are packed on the device (Solver.pack_xdmf /
Solver.pack_csv, one device-to-host copy each); libweldform_io.so writes the
files.  `table_from_fields` packs the same tables on the host from FP64
particle fields in the caller's order (as the reference's writers pack them
from its Particle objects)."""
import ctypes
import os

import numpy as np

from . import capi

IO_LIB_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "lib", "libweldform_io.so")

# float datasets of the .hdf5: name -> (offset, width); offsets are the
# WF_XDMF_* constants of weldform_hip.h (dataset D of width w occupies
# f[D*n : (D+w)*n])
XDMF_FLOAT = {
    "Position": (0, 3), "Velocity": (3, 3), "Acceleration": (6, 3), "Pressure": (9, 1), "Density": (10, 1),
    "Property1": (11, 1), "Property2": (12, 1), "Property3": (13, 1), "Mass": (14, 1), "h": (15, 1),
    "Sigma": (16, 6), "ShearS": (22, 6), "Strain": (28, 6), "StrainRate": (34, 6), "Strain_pl": (40, 6),
    "gradcorrmat": (46, 6), "Sigma_eq": (52, 1), "Temperature": (53, 1), "Pl_Strain": (54, 1),
    "q_friction": (55, 1), "c_shearabs": (56, 1), "ps_en": (57, 1), "Displacement": (58, 3),
    "Contact Force": (61, 3),
}
XDMF_INT = {"Tag": 0, "Neighbors": 1, "ContNeib": 2}
CSV_HEADER = "X, Y, Z, ID, Sigma_eq, Pl_Strain, vx, vy, vz, ax, ay, az, CFx, CFy, CZFz, p, Temp"

_io = None


class OutputError(RuntimeError):
    pass


def io_lib():
    """Load libweldform_io.so (raises if it has not been built)."""
    global _io
    if _io is not None:
        return _io
    if not os.path.exists(IO_LIB_PATH):
        raise RuntimeError(f"{IO_LIB_PATH} is missing: build it with `python -m weldform_amd.build`")
    L = ctypes.CDLL(IO_LIB_PATH)
    L.wfio_has_hdf5.argtypes, L.wfio_has_hdf5.restype = [], ctypes.c_int
    L.wfio_last_error.argtypes, L.wfio_last_error.restype = [], ctypes.c_char_p
    L.wfio_write_xdmf.argtypes = [ctypes.c_char_p, ctypes.c_int64, ctypes.POINTER(ctypes.c_float),
                                  ctypes.POINTER(ctypes.c_int32), ctypes.POINTER(ctypes.c_char_p)]
    L.wfio_write_xdmf.restype = ctypes.c_int
    L.wfio_write_csv.argtypes = [ctypes.c_char_p, ctypes.c_int64, ctypes.POINTER(ctypes.c_double), ctypes.c_int32]
    L.wfio_write_csv.restype = ctypes.c_int
    _io = L
    return L


def has_hdf5():
    return bool(io_lib().wfio_has_hdf5())


def _check(rc):
    if rc != 0:
        raise OutputError(io_lib().wfio_last_error().decode())


def sigma_eq(shear):
    """Particle::CalculateEquivalentStress (Source/Particle.cpp:623-633) of
    ShearStress tensors (n, 3, 3), in the reference's order of operations."""
    S = np.asarray(shear, dtype=np.float64)
    j2 = 0.5 * (S[:, 0, 0] * S[:, 0, 0] + 2.0 * S[:, 0, 1] * S[:, 1, 0] + 2.0 * S[:, 0, 2] * S[:, 2, 0]
                + S[:, 1, 1] * S[:, 1, 1] + 2.0 * S[:, 1, 2] * S[:, 2, 1] + S[:, 2, 2] * S[:, 2, 2])
    return np.sqrt(3.0 * j2)


def _t6(t):
    t = np.asarray(t, dtype=np.float64)
    return np.stack([t[:, 0, 0], t[:, 1, 1], t[:, 2, 2], t[:, 0, 1], t[:, 1, 2], t[:, 0, 2]], axis=1)


def table_from_fields(fl, props=None):
    """Host packing of both tables from FP64 fields in the caller's order.
    fl: dict with x, v, a, disp, pressure, density, mass, h (per particle),
    sigma, shear, strain, strain_rate, strain_pl (n,3,3), T, pl_strain,
    ps_energy, contforce (n,3), id, nb, contnb (int).  Returns (f, iv, rows)
    laid out as Solver.pack_xdmf / Solver.pack_csv return them."""
    n = len(fl["x"])
    f = np.zeros(capi.XDMF_NF * n, dtype=np.float32)

    def put(name, vals):
        off, w = XDMF_FLOAT[name]
        f[off * n:(off + w) * n] = np.asarray(vals, dtype=np.float64).reshape(-1).astype(np.float32)

    put("Position", fl["x"])
    put("Velocity", fl["v"])
    put("Acceleration", fl["a"])
    put("Pressure", fl["pressure"])
    put("Density", fl["density"])
    for k in range(3):
        put(f"Property{k + 1}", np.zeros(n) if props is None else props[k])
    put("Mass", fl["mass"])
    put("h", fl["h"])
    put("Sigma", _t6(fl["sigma"]))
    put("ShearS", _t6(fl["sigma"]))  # the reference writes Sigma under ShearS
    put("Strain", _t6(fl["strain"]))
    put("StrainRate", _t6(fl["strain_rate"]))
    put("Strain_pl", _t6(fl["strain_pl"]))
    put("gradcorrmat", np.zeros((n, 6)))
    seq = sigma_eq(fl["shear"])
    put("Sigma_eq", seq)
    put("Temperature", fl["T"])
    put("Pl_Strain", fl["pl_strain"])
    put("q_friction", np.zeros(n))
    put("c_shearabs", np.zeros(n))
    put("ps_en", fl["ps_energy"])
    put("Displacement", fl["disp"])
    put("Contact Force", fl["contforce"])
    iv = np.concatenate([np.asarray(fl["id"], dtype=np.int32), np.asarray(fl["nb"], dtype=np.int32),
                         np.asarray(fl["contnb"], dtype=np.int32)])
    rows = np.column_stack([fl["x"], np.asarray(fl["id"], dtype=np.float64), seq, fl["pl_strain"], fl["v"], fl["a"],
                            fl["contforce"], fl["pressure"], fl["T"]])
    return f, iv, rows


def datasets(f, iv, n):
    """The .hdf5 datasets of a packed table: name -> array (1-D, n*width)."""
    out = {name: f[off * n:(off + w) * n] for name, (off, w) in XDMF_FLOAT.items()}
    out.update({name: iv[k * n:(k + 1) * n] for name, k in XDMF_INT.items()})
    return out


def write_xdmf_table(key, n, f, iv, prop_names=None):
    f = np.ascontiguousarray(f, dtype=np.float32)
    iv = np.ascontiguousarray(iv, dtype=np.int32)
    names = None
    if prop_names is not None:
        names = (ctypes.c_char_p * 3)(*[p.encode() for p in prop_names])
    _check(io_lib().wfio_write_xdmf(str(key).encode(), n, f.ctypes.data_as(ctypes.POINTER(ctypes.c_float)),
                                    iv.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)), names))


def write_csv_rows(key, rows, threads=8):
    rows = np.ascontiguousarray(rows, dtype=np.float64)
    _check(io_lib().wfio_write_csv(str(key).encode(), len(rows), rows.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                                   int(threads)))


def write_xdmf(solver, key, prop_names=None):
    """Domain::WriteXDMF of the solver's current state: key.hdf5 + key.xmf."""
    f, iv = solver.pack_xdmf()
    write_xdmf_table(key, solver.n, f, iv, prop_names)


def write_csv(solver, key, threads=8):
    """Domain::WriteCSV of the solver's current state: key.csv."""
    write_csv_rows(key, solver.pack_csv(), threads)
