"""ctypes wrapper of the CPU oracle (TEST INFRASTRUCTURE).

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import
this module; the product package never does.  See wf_oracle.h for the list
of reference files the oracle restates.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")

MAX_NB = 100
BILINEAR, HOLLOMON, JOHNSON_COOK = 0, 1, 2

# double fields and their widths, in wf_oracle.h enum order
F64 = ["X", "V", "A", "XPREV", "DISP", "RHO", "DRHO", "MASS", "H", "REFRHO", "P",
       "CS", "P0", "G", "ALPHA", "BETA", "SIGMA", "SHEAR", "SR", "RR", "STRAIN",
       "EPL", "EPLINC", "PLSTRAIN", "DPLSTRAIN", "SIGMAY", "EFFSR", "EP", "ET",
       "T", "DTDT", "KT", "CPT", "THEXP", "QCONV", "QSRC", "QPL", "QCONT",
       "HCONV", "TINF", "PSENERGY", "NORMAL", "CONTFORCE", "TI", "TIN", "TIDIST", "TIR", "GCM",
       "QFRIC", "CSHEAR", "FHFL", "DAMD", "ZWAB", "VXSPH", "SUMDEN"]
F64_DIM = dict(X=3, V=3, A=3, XPREV=3, DISP=3, SIGMA=9, SHEAR=9, SR=9, RR=9,
               STRAIN=9, EPL=9, EPLINC=9, NORMAL=3, CONTFORCE=3, TIR=9, GCM=9, VXSPH=3)
I32 = ["ID", "MATERIAL", "MATIDX", "MODEL", "FREE", "THBC", "NB", "ELEMENT", "MESHIDX", "IDORIG"]


class Material(ctypes.Structure):
    _fields_ = [("model", ctypes.c_int), ("E", ctypes.c_double), ("nu", ctypes.c_double),
                ("K", ctypes.c_double), ("m", ctypes.c_double), ("sy0", ctypes.c_double),
                ("A", ctypes.c_double), ("B", ctypes.c_double), ("n", ctypes.c_double),
                ("C", ctypes.c_double), ("eps_0", ctypes.c_double), ("jm", ctypes.c_double),
                ("Tm", ctypes.c_double), ("Tt", ctypes.c_double)]


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("oracle library not built: run `make -C oracle`")
        L = ctypes.CDLL(LIB_PATH)
        vp, i32, i64, d = ctypes.c_void_p, ctypes.c_int, ctypes.c_int64, ctypes.c_double
        dp = ctypes.POINTER(ctypes.c_double)
        L.wfo_create.restype = vp
        L.wfo_create.argtypes = [i32]
        L.wfo_destroy.argtypes = [vp]
        L.wfo_f64.restype = dp
        L.wfo_f64.argtypes = [vp, i32]
        L.wfo_i32.restype = ctypes.POINTER(ctypes.c_int)
        L.wfo_i32.argtypes = [vp, i32]
        L.wfo_set_param.argtypes = [vp, ctypes.c_char_p, d]
        L.wfo_get_param.restype = d
        L.wfo_get_param.argtypes = [vp, ctypes.c_char_p]
        L.wfo_set_material.argtypes = [vp, i32, ctypes.POINTER(Material)]
        L.wfo_set_material_friction.argtypes = [vp, i32, d, d]
        L.wfo_set_material_damage.argtypes = [vp, i32, i32, dp]
        L.wfo_calc_force2233.argtypes = [vp, d, i32]
        L.wfo_set_threads.argtypes = [i32]
        L.wfo_last_error.restype = ctypes.c_char_p
        L.wfo_add_velocity_bc.argtypes = [vp, i32, i32, dp, i32, d]
        L.wfo_add_temperature_bc.argtypes = [vp, i32, d]
        L.wfo_set_bc_angular.argtypes = [vp, i32, dp]
        L.wfo_add_amplitude.argtypes = [vp, i32, i32, dp, dp]
        L.wfo_assign_zone.argtypes = [vp, dp, dp, i32]
        L.wfo_gen_cylinder.restype = i64
        L.wfo_gen_cylinder.argtypes = [dp, d, d, d, dp, i64]
        for name in ["wfo_solve_init"]:
            getattr(L, name).argtypes = [vp, d]
        for name in ["wfo_step_leapfrog", "wfo_step_kickdrift", "wfo_neighbour_search",
                     "wfo_start_acceleration", "wfo_calc_accel", "wfo_calc_dens",
                     "wfo_calc_rates", "wfo_general_after", "wfo_energies", "wfo_calc_grad_corr"]:
            getattr(L, name).argtypes = [vp]
        L.wfo_stress_all.argtypes = [vp, d]
        L.wfo_thermal_calcs.argtypes = [vp, d]
        L.wfo_plastic_work_heat.argtypes = [vp, d]
        L.wfo_npairs.restype = i64
        L.wfo_npairs.argtypes = [vp]
        L.wfo_pairs.restype = ctypes.POINTER(ctypes.c_int)
        L.wfo_pairs.argtypes = [vp]
        L.wfo_reaction.argtypes = [vp, i32, dp]
        L.wfo_add_plane_mesh.argtypes = [vp, i32, i32, i32, i32, dp, dp, i32, d]
        L.wfo_add_mesh.argtypes = [vp, i32, i32, i32, dp, i32, ctypes.POINTER(ctypes.c_int), dp, d]
        L.wfo_contact_force.argtypes = [vp, i32, dp]
        L.wfo_group_contact_force.argtypes = [vp, i32, i32, dp]
        L.wfo_ncontact_pairs.restype = i64
        L.wfo_ncontact_pairs.argtypes = [vp]
        L.wfo_contact_pairs.restype = ctypes.POINTER(ctypes.c_int)
        L.wfo_contact_pairs.argtypes = [vp]
        L.wfo_kernel.restype = d
        L.wfo_kernel.argtypes = [d, d]
        L.wfo_grad_kernel.restype = d
        L.wfo_grad_kernel.argtypes = [d, d]
        L.wfo_kernel_dim.restype = d
        L.wfo_kernel_dim.argtypes = [d, d, ctypes.c_int]
        L.wfo_grad_kernel_dim.restype = d
        L.wfo_grad_kernel_dim.argtypes = [d, d, ctypes.c_int]
        L.wfo_sym3_eig.restype = None
        L.wfo_sym3_eig.argtypes = [dp, dp, dp]
        _lib = L
    return _lib


def _dp(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def set_threads(n):
    lib().wfo_set_threads(int(n))


def kernel(q, h, dim=3):
    return lib().wfo_kernel_dim(float(q), float(h), int(dim))


def grad_kernel(q, h, dim=3):
    return lib().wfo_grad_kernel_dim(float(q), float(h), int(dim))


def sym3_eig(A):
    """(w, V) with A = V diag(w) V^T for a symmetric 3x3 A."""
    A = np.ascontiguousarray(A, dtype=np.float64).reshape(9)
    w = np.zeros(3)
    V = np.zeros(9)
    lib().wfo_sym3_eig(_dp(A), _dp(w), _dp(V))
    return w, V.reshape(3, 3)


def gen_cylinder(V, R, L, r):
    """AddCylinderLength lattice positions, shape (n, 3)."""
    V = np.ascontiguousarray(V, dtype=np.float64)
    n = lib().wfo_gen_cylinder(_dp(V), R, L, r, None, 0)
    out = np.zeros((n, 3))
    lib().wfo_gen_cylinder(_dp(V), R, L, r, _dp(out), n)
    return out


class Domain:
    """Oracle domain: numpy views onto the C arrays (original particle order)."""

    def __init__(self, n):
        self.n = int(n)
        self.h = lib().wfo_create(self.n)

    def __del__(self):
        if getattr(self, "h", None):
            lib().wfo_destroy(self.h)
            self.h = None

    def _check(self, rc):
        if rc != 0:
            raise RuntimeError("oracle: " + lib().wfo_last_error().decode())

    def f(self, name):
        k = F64.index(name)
        dim = F64_DIM.get(name, 1)
        p = lib().wfo_f64(self.h, k)
        a = np.ctypeslib.as_array(p, shape=(self.n * dim,))
        return a.reshape(self.n, 3, 3) if dim == 9 else (a.reshape(self.n, 3) if dim == 3 else a)

    def i(self, name):
        p = lib().wfo_i32(self.h, I32.index(name))
        return np.ctypeslib.as_array(p, shape=(self.n,))

    def set_param(self, name, v):
        self._check(lib().wfo_set_param(self.h, name.encode(), float(v)))

    def param(self, name):
        return lib().wfo_get_param(self.h, name.encode())

    def set_material(self, idx, **kw):
        m = Material(**kw)
        self._check(lib().wfo_set_material(self.h, int(idx), ctypes.byref(m)))

    def set_material_damage(self, idx, model, D=(0.0, 0.0, 0.0, 0.0, 0.0)):
        Dv = np.ascontiguousarray(D, dtype=np.float64)
        self._check(lib().wfo_set_material_damage(self.h, int(idx), int(model), _dp(Dv)))

    def set_material_friction(self, idx, sta, dyn):
        self._check(lib().wfo_set_material_friction(self.h, int(idx), float(sta), float(dyn)))

    def add_velocity_bc(self, zone, value, value_type=0, amp_id=0, amp_factor=1.0):
        v = np.ascontiguousarray(value, dtype=np.float64)
        self._check(lib().wfo_add_velocity_bc(self.h, zone, value_type, _dp(v), amp_id, amp_factor))

    def set_bc_angular(self, zone, w):
        a = np.ascontiguousarray(w, dtype=np.float64)
        self._check(lib().wfo_set_bc_angular(self.h, int(zone), _dp(a)))

    def add_temperature_bc(self, zone, T):
        self._check(lib().wfo_add_temperature_bc(self.h, zone, float(T)))

    def add_amplitude(self, aid, t, v):
        t = np.ascontiguousarray(t, dtype=np.float64)
        v = np.ascontiguousarray(v, dtype=np.float64)
        self._check(lib().wfo_add_amplitude(self.h, aid, len(t), _dp(t), _dp(v)))

    def assign_zone(self, start, end, zid):
        s = np.ascontiguousarray(start, dtype=np.float64)
        e = np.ascontiguousarray(end, dtype=np.float64)
        return lib().wfo_assign_zone(self.h, _dp(s), _dp(e), zid)

    def solve_init(self, dt):
        self._check(lib().wfo_solve_init(self.h, float(dt)))

    def step(self, scheme="leapfrog"):
        fn = lib().wfo_step_leapfrog if scheme == "leapfrog" else lib().wfo_step_kickdrift
        self._check(fn(self.h))

    def neighbour_search(self):
        self._check(lib().wfo_neighbour_search(self.h))

    def pairs(self):
        P = lib().wfo_npairs(self.h)
        if P == 0:
            return np.zeros((0, 2), dtype=np.int32)
        return np.ctypeslib.as_array(lib().wfo_pairs(self.h), shape=(2 * P,)).reshape(P, 2).copy()

    def start_acceleration(self):
        self._check(lib().wfo_start_acceleration(self.h))

    def calc_accel(self):
        self._check(lib().wfo_calc_accel(self.h))

    def calc_grad_corr(self):
        self._check(lib().wfo_calc_grad_corr(self.h))

    def calc_dens(self):
        self._check(lib().wfo_calc_dens(self.h))

    def calc_rates(self):
        self._check(lib().wfo_calc_rates(self.h))

    def calc_force2233(self, xsph=0.0, shepard=False):
        self._check(lib().wfo_calc_force2233(self.h, float(xsph), int(bool(shepard))))

    def stress_all(self, dt):
        self._check(lib().wfo_stress_all(self.h, float(dt)))

    def thermal_calcs(self, dt):
        self._check(lib().wfo_thermal_calcs(self.h, float(dt)))

    def plastic_work_heat(self, dt):
        self._check(lib().wfo_plastic_work_heat(self.h, float(dt)))

    def energies(self):
        self._check(lib().wfo_energies(self.h))

    def general_after(self):
        self._check(lib().wfo_general_after(self.h))

    def add_plane_mesh(self, first, zone, axis, flip, p1, p2, dens, hfac=1.1):
        """AxisPlaneMesh + CalcSpheres + AddTrimeshParticles; the mesh's
        2 dens^2 particles are written from index `first` on."""
        a = np.ascontiguousarray(p1, dtype=np.float64)
        b = np.ascontiguousarray(p2, dtype=np.float64)
        rc = lib().wfo_add_plane_mesh(self.h, int(first), int(zone), int(axis), int(bool(flip)), _dp(a), _dp(b),
                                      int(dens), float(hfac))
        if rc < 0:
            self._check(rc)
        return rc

    def add_mesh(self, first, zone, nodes, elems, hfac=1.1, normals=None):
        """A triangle mesh from explicit nodes (nn, 3), 0-based elements
        (ne, 3) and optional initial normals (ne, 3) + CalcSpheres +
        AddTrimeshParticles from index `first`."""
        nd = np.ascontiguousarray(nodes, dtype=np.float64).reshape(-1, 3)
        el = np.ascontiguousarray(elems, dtype=np.int32).reshape(-1, 3)
        nr = None if normals is None else np.ascontiguousarray(normals, dtype=np.float64).reshape(-1, 3)
        rc = lib().wfo_add_mesh(self.h, int(first), int(zone), len(nd), _dp(nd), len(el),
                                el.ctypes.data_as(ctypes.POINTER(ctypes.c_int)), None if nr is None else _dp(nr),
                                float(hfac))
        if rc < 0:
            self._check(rc)
        return rc

    def contact_force(self, mesh):
        out = np.zeros(3)
        self._check(lib().wfo_contact_force(self.h, mesh, _dp(out)))
        return out

    def group_contact_force(self, group, mesh):
        out = np.zeros(3)
        self._check(lib().wfo_group_contact_force(self.h, group, mesh, _dp(out)))
        return out

    def contact_pairs(self):
        P = lib().wfo_ncontact_pairs(self.h)
        if P == 0:
            return np.zeros((0, 2), dtype=np.int32)
        return np.ctypeslib.as_array(lib().wfo_contact_pairs(self.h), shape=(2 * P,)).reshape(P, 2).copy()

    def reaction(self, zone):
        out = np.zeros(3)
        self._check(lib().wfo_reaction(self.h, zone, _dp(out)))
        return out
