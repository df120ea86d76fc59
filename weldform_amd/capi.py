"""ctypes binding of the C-ABI in include/weldform_hip.h.

Mirrors the reference's SPH::Domain operator interface at the level the
Python tests and the benchmark need: particle fields in and out, boundary
conditions, SolveDiffUpdateLeapFrog / SolveDiffUpdateKickDrift steps.
The library is the compiled HIP path; there is no CPU fallback: if
libweldform_hip.so is missing, loading fails loudly.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "lib", "libweldform_hip.so")

OK, EINVAL, EHIP, ESTATE, ENBOVERFLOW, ENAN, EUNSUPPORTED, ENOMEM = 0, -1, -2, -3, -4, -5, -6, -7
FP64, MIXED = 0, 1
LEAPFROG, KICKDRIFT = 0, 1
PART_SEARCH, PART_ACCEL, PART_KICK1, PART_KICK2, PART_REST = range(5)
BILINEAR, HOLLOMON, JOHNSON_COOK = 0, 1, 2

FIELDS = {
    "x": (0, 3), "v": (1, 3), "a": (2, 3), "disp": (3, 3), "density": (4, 1),
    "mass": (5, 1), "h": (6, 1), "sigma": (7, 9), "shear": (8, 9),
    "strain_rate": (9, 9), "rot_rate": (10, 9), "strain": (11, 9),
    "strain_pl": (12, 9), "pl_strain": (13, 1), "dpl_strain": (14, 1),
    "sigmay": (15, 1), "eff_sr": (16, 1), "pressure": (17, 1), "T": (18, 1),
    "dTdt": (19, 1), "ddensity": (20, 1), "q_plheat": (21, 1),
    "strain_pl_incr": (22, 9), "q_conv": (23, 1), "h_conv": (24, 1),
    "T_inf": (25, 1), "q_source": (26, 1), "ps_energy": (27, 1), "contforce": (28, 3), "gradcorr": (29, 9),
    "dam_D": (30, 1), "zwab": (31, 1), "vxsph": (32, 3), "sumden": (33, 1),
}
RIGID_FIELDS = {"x": 0, "v": 1, "normal": 2}
INT_FIELDS = {"id": 0, "material": 1, "free": 2, "thermal_bc": 3, "nb": 4, "contnb": 5}
EVAL = {"accel": 0, "density": 1, "rates": 2, "stress": 3}
KERNELS = {"accel": 0, "density": 1, "rates": 2, "neighbours": 3, "step": 4, "surface": 5}


class WFError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"weldform_hip error {code}: {msg}")
        self.code = code


class Config(ctypes.Structure):
    _fields_ = [("precision", ctypes.c_int32), ("ts_nb_inc", ctypes.c_int32),
                ("max_nb", ctypes.c_int32), ("thermal_solver", ctypes.c_int32),
                ("pl_heating", ctypes.c_int32), ("device", ctypes.c_int32),
                ("no_graphs", ctypes.c_int32), ("stage_max", ctypes.c_int32), ("lanes", ctypes.c_int32),
                ("auto_ts", ctypes.c_int32), ("cellfac", ctypes.c_double),
                ("pl_work_heat_frac", ctypes.c_double), ("dom_max", ctypes.c_double * 3),
                ("dom_min", ctypes.c_double * 3), ("cfl", ctypes.c_double), ("sqrt_h_a", ctypes.c_double),
                ("min_ts", ctypes.c_double), ("grad_corr", ctypes.c_int32), ("locking", ctypes.c_int32)]


class Material(ctypes.Structure):
    _fields_ = [("model", ctypes.c_int32), ("reserved", ctypes.c_int32)] + [
        (n, ctypes.c_double) for n in
        ["E", "nu", "G", "Cs", "ref_density", "P0", "alpha", "beta", "Ep",
         "hol_K", "hol_m", "hol_sy0", "jc_A", "jc_B", "jc_n", "jc_C", "jc_eps0",
         "jc_m", "jc_Tm", "jc_Tt", "k_T", "cp_T", "th_exp", "ti", "ti_n", "ti_dist"]]


class Scalars(ctypes.Structure):
    _fields_ = [("time", ctypes.c_double), ("dt", ctypes.c_double),
                ("steps", ctypes.c_int64), ("nb_searches", ctypes.c_int64),
                ("npairs", ctypes.c_int64), ("max_nb_count", ctypes.c_int32),
                ("reserved", ctypes.c_int32)] + [
        (n, ctypes.c_double) for n in
        ["kin_energy_sum", "int_energy_sum", "plastic_work", "max_pl_strain", "max_T", "min_T",
         "contact_force_sum", "contact_reaction_sum", "ext_forces_work"]] + [("contact_pairs", ctypes.c_int64)] + [
        (n, ctypes.c_double) for n in ["contact_friction_work", "accum_cont_heat_cond", "min_force_ts"]]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_ if k != "reserved"}


class AdjConfig(ctypes.Structure):
    _fields_ = [("steps", ctypes.c_int64), ("checkpoint_every", ctypes.c_int32), ("loss_kind", ctypes.c_int32),
                ("loss_index", ctypes.c_int32), ("loss_component", ctypes.c_int32), ("material", ctypes.c_int32),
                ("reserved", ctypes.c_int32)]


ADJ_PLANE, ADJ_ZONE = 0, 1
ADJ_PARAMS = ("jc_A", "jc_B", "jc_n", "jc_C", "jc_m", "friction_sta", "friction_dyn")
ADJ_FIELDS = {"x": (0, 3), "v": (1, 3), "density": (4, 1), "shear": (8, 9), "sigmay": (15, 1), "pl_strain": (13, 1)}

EXPORTS = [
    "wf_abi_version", "wf_last_error", "wf_device_count", "wf_create", "wf_destroy",
    "wf_set_material", "wf_set_field", "wf_get_field", "wf_set_int_field",
    "wf_get_int_field", "wf_add_amplitude", "wf_add_velocity_bc",
    "wf_add_temperature_bc", "wf_init", "wf_step", "wf_sync", "wf_get_scalars",
    "wf_get_reaction_history", "wf_rebuild_neighbours", "wf_get_neighbour_pairs",
    "wf_eval", "wf_set_timing", "wf_get_kernel_time", "wf_get_neighbour_stats", "wf_get_stage_stats",
    "wf_gen_cylinder", "wf_add_rigid_plane", "wf_set_contact", "wf_get_contact_force_history",
    "wf_get_rigid_field", "wf_adj_forward", "wf_adj_backward", "wf_adj_get_field", "wf_adj_restore",
    "wf_pack_xdmf", "wf_pack_csv", "wf_add_rigid_mesh", "wf_set_bc_angular_velocity",
    "wf_set_contact_heat", "wf_set_rigid_temperature", "wf_prepare", "wf_step_part",
    "wf_pack_xdmf_noalloc", "wf_pack_xdmf_noalloc_max", "wf_get_cell_grids",
    "wf_set_material_friction", "wf_set_group_sums", "wf_get_group_history",
    "wf_set_contact_law", "wf_set_friction_function", "wf_set_material_damage", "wf_calc_force2233",
    "wf_set_dimension",
]

# contact laws (Domain::contact_alg, Source/Domain.h:64)
FRASER, WANG, SEO, LSDYNA = 0, 1, 2, 4

# output tables (wf_pack_xdmf / wf_pack_csv): floats and ints per particle,
# CSV columns
XDMF_NF, XDMF_NI, CSV_NCOL = 64, 3, 17
XDMF_Q_FRICTION, XDMF_C_SHEARABS = 55, 56

_lib = None


def lib():
    """Load libweldform_hip.so (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"{LIB_PATH} is missing: build it with `python -m weldform_amd.build`")
    L = ctypes.CDLL(LIB_PATH)
    vp, i32, i64, d = ctypes.c_void_p, ctypes.c_int32, ctypes.c_int64, ctypes.c_double
    dp = ctypes.POINTER(ctypes.c_double)
    ip = ctypes.POINTER(ctypes.c_int32)
    I64P = ctypes.POINTER(ctypes.c_int64)
    sig = {
        "wf_abi_version": ([], ctypes.c_int),
        "wf_last_error": ([vp], ctypes.c_char_p),
        "wf_device_count": ([ip], ctypes.c_int),
        "wf_create": ([ctypes.POINTER(Config), i64, ctypes.POINTER(vp)], ctypes.c_int),
        "wf_destroy": ([vp], ctypes.c_int),
        "wf_set_material": ([vp, i32, ctypes.POINTER(Material)], ctypes.c_int),
        "wf_set_field": ([vp, i32, dp], ctypes.c_int),
        "wf_get_field": ([vp, i32, dp], ctypes.c_int),
        "wf_set_int_field": ([vp, i32, ip], ctypes.c_int),
        "wf_get_int_field": ([vp, i32, ip], ctypes.c_int),
        "wf_add_amplitude": ([vp, i32, i32, dp, dp], ctypes.c_int),
        "wf_add_velocity_bc": ([vp, i32, i32, dp, i32, d], ctypes.c_int),
        "wf_add_temperature_bc": ([vp, i32, d], ctypes.c_int),
        "wf_init": ([vp, d], ctypes.c_int),
        "wf_step": ([vp, i32, i64], ctypes.c_int),
        "wf_prepare": ([vp, i32], ctypes.c_int),
        "wf_step_part": ([vp, i32, i32], ctypes.c_int),
        "wf_sync": ([vp], ctypes.c_int),
        "wf_get_scalars": ([vp, ctypes.POINTER(Scalars)], ctypes.c_int),
        "wf_get_reaction_history": ([vp, i32, dp, i64, I64P], ctypes.c_int),
        "wf_rebuild_neighbours": ([vp], ctypes.c_int),
        "wf_get_neighbour_pairs": ([vp, ip, i64, I64P], ctypes.c_int),
        "wf_eval": ([vp, i32, d], ctypes.c_int),
        "wf_set_timing": ([vp, i32], ctypes.c_int),
        "wf_get_kernel_time": ([vp, i32, dp, I64P], ctypes.c_int),
        "wf_get_neighbour_stats": ([vp, I64P, ip], ctypes.c_int),
        "wf_get_stage_stats": ([vp, I64P, I64P, I64P], ctypes.c_int),
        "wf_get_cell_grids": ([vp, ip, ip], ctypes.c_int),
        "wf_set_material_friction": ([vp, i32, d, d], ctypes.c_int),
        "wf_set_contact_law": ([vp, i32, d, i32], ctypes.c_int),
        "wf_set_friction_function": ([vp, i32, d, d], ctypes.c_int),
        "wf_set_dimension": ([vp, i32], ctypes.c_int),
        "wf_set_material_damage": ([vp, i32, i32, dp], ctypes.c_int),
        "wf_calc_force2233": ([vp, d, i32], ctypes.c_int),
        "wf_set_group_sums": ([vp, i32, i64], ctypes.c_int),
        "wf_get_group_history": ([vp, i32, i32, i32, dp, i64, I64P], ctypes.c_int),
        "wf_gen_cylinder": ([dp, d, d, d, dp, i64], i64),
        "wf_add_rigid_plane": ([vp, i32, i32, i32, dp, dp, i32, d], ctypes.c_int),
        "wf_set_contact": ([vp, d, d, d], ctypes.c_int),
        "wf_set_contact_heat": ([vp, ctypes.c_int32, ctypes.c_int32, d], ctypes.c_int),
        "wf_set_rigid_temperature": ([vp, ctypes.c_int64, ctypes.POINTER(ctypes.c_double)], ctypes.c_int),
        "wf_get_contact_force_history": ([vp, i32, dp, i64, I64P], ctypes.c_int),
        "wf_get_rigid_field": ([vp, i32, dp], ctypes.c_int),
        "wf_adj_forward": ([vp, ctypes.POINTER(AdjConfig), dp, dp, dp], ctypes.c_int),
        "wf_adj_backward": ([vp, dp], ctypes.c_int),
        "wf_adj_get_field": ([vp, i32, dp], ctypes.c_int),
        "wf_adj_restore": ([vp], ctypes.c_int),
        "wf_pack_xdmf": ([vp, ctypes.POINTER(ctypes.c_float), ip], ctypes.c_int),
        "wf_pack_xdmf_noalloc": ([vp, ctypes.POINTER(ctypes.c_float), ip], ctypes.c_int),
        "wf_pack_xdmf_noalloc_max": ([], ctypes.c_int64),
        "wf_pack_csv": ([vp, dp], ctypes.c_int),
        "wf_add_rigid_mesh": ([vp, i32, i64, dp, i64, ip, dp, d], ctypes.c_int),
        "wf_set_bc_angular_velocity": ([vp, i32, dp], ctypes.c_int),
    }
    for name, (args, res) in sig.items():
        fn = getattr(L, name)
        fn.argtypes = args
        fn.restype = res
    _lib = L
    return L


def _dp(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def _ip(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_int32))


def gen_cylinder(V, R, L, r):
    """Domain::AddCylinderLength lattice positions, shape (n, 3)."""
    V = np.ascontiguousarray(V, dtype=np.float64)
    n = lib().wf_gen_cylinder(_dp(V), R, L, r, None, 0)
    if n < 0:
        raise WFError(n, "bad cylinder parameters")
    out = np.zeros((n, 3))
    lib().wf_gen_cylinder(_dp(V), R, L, r, _dp(out), n)
    return out


def device_count():
    c = ctypes.c_int32(0)
    rc = lib().wf_device_count(ctypes.byref(c))
    return c.value if rc == 0 else 0


class Solver:
    """One device context (one GPU, one HIP stream)."""

    def __init__(self, n, precision=FP64, ts_nb_inc=5, max_nb=128, thermal_solver=False,
                 pl_heating=False, device=0, cellfac=2.0, pl_work_heat_frac=0.9,
                 dom_max=(-1e11, -1e11, -1e11), dom_min=(1e11, 1e11, 1e11), graphs=True, stage_max=0,
                 lanes=0, auto_ts=0, cfl=0.7, sqrt_h_a=0.0025, min_ts=0.0, grad_corr=False, locking=False):
        cfg = Config(precision=precision, ts_nb_inc=ts_nb_inc, max_nb=max_nb,
                     thermal_solver=int(thermal_solver), pl_heating=int(pl_heating),
                     device=device, no_graphs=int(not graphs), stage_max=stage_max, lanes=lanes, cellfac=cellfac,
                     pl_work_heat_frac=pl_work_heat_frac, auto_ts=auto_ts, cfl=cfl, sqrt_h_a=sqrt_h_a,
                     min_ts=min_ts, grad_corr=int(grad_corr), locking=int(locking))
        cfg.dom_max[:] = list(dom_max)
        cfg.dom_min[:] = list(dom_min)
        self.n = int(n)
        self.ctx = ctypes.c_void_p()
        self._check(lib().wf_create(ctypes.byref(cfg), self.n, ctypes.byref(self.ctx)), glob=True)

    def close(self):
        if getattr(self, "ctx", None) and self.ctx.value:
            lib().wf_destroy(self.ctx)
            self.ctx = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, rc, glob=False):
        if rc != 0:
            msg = lib().wf_last_error(None if glob else self.ctx)
            raise WFError(rc, msg.decode() if msg else "")

    # fields -------------------------------------------------------------
    def set(self, name, arr):
        fid, dim = FIELDS[name]
        a = np.ascontiguousarray(arr, dtype=np.float64).reshape(-1)
        if a.size != dim * self.n:
            raise ValueError(f"{name}: expected {dim * self.n} values, got {a.size}")
        self._check(lib().wf_set_field(self.ctx, fid, _dp(a)))

    def get(self, name):
        fid, dim = FIELDS[name]
        a = np.zeros(dim * self.n)
        self._check(lib().wf_get_field(self.ctx, fid, _dp(a)))
        return a.reshape(self.n, 3, 3) if dim == 9 else (a.reshape(self.n, 3) if dim == 3 else a)

    def set_int(self, name, arr):
        a = np.ascontiguousarray(arr, dtype=np.int32).reshape(-1)
        if a.size != self.n:
            raise ValueError(f"{name}: expected {self.n} values")
        self._check(lib().wf_set_int_field(self.ctx, INT_FIELDS[name], _ip(a)))

    def get_int(self, name):
        a = np.zeros(self.n, dtype=np.int32)
        self._check(lib().wf_get_int_field(self.ctx, INT_FIELDS[name], _ip(a)))
        return a

    def set_material(self, idx, **kw):
        m = Material(**kw)
        self._check(lib().wf_set_material(self.ctx, int(idx), ctypes.byref(m)))

    # boundary conditions ---------------------------------------------------
    def add_amplitude(self, aid, t, v):
        t = np.ascontiguousarray(t, dtype=np.float64)
        v = np.ascontiguousarray(v, dtype=np.float64)
        self._check(lib().wf_add_amplitude(self.ctx, aid, len(t), _dp(t), _dp(v)))

    def add_velocity_bc(self, zone, value, value_type=0, amp_id=0, amp_factor=1.0):
        v = np.ascontiguousarray(value, dtype=np.float64)
        self._check(lib().wf_add_velocity_bc(self.ctx, zone, value_type, _dp(v), amp_id, amp_factor))

    def set_bc_angular_velocity(self, zone, w):
        a = np.ascontiguousarray(w, dtype=np.float64)
        self._check(lib().wf_set_bc_angular_velocity(self.ctx, int(zone), _dp(a)))

    def add_temperature_bc(self, zone, T):
        self._check(lib().wf_add_temperature_bc(self.ctx, zone, float(T)))

    # solve -----------------------------------------------------------------
    def init(self, dt):
        self._check(lib().wf_init(self.ctx, float(dt)))

    def step(self, nsteps=1, scheme=LEAPFROG):
        self._check(lib().wf_step(self.ctx, scheme, int(nsteps)))

    def step_part(self, part, scheme=LEAPFROG):
        """One part of a step (PART_SEARCH .. PART_REST), for host hooks between parts."""
        self._check(lib().wf_step_part(self.ctx, scheme, int(part)))

    def prepare(self, scheme=LEAPFROG):
        """Capture the step's HIP graphs now rather than in the first steps."""
        self._check(lib().wf_prepare(self.ctx, scheme))

    def sync(self):
        self._check(lib().wf_sync(self.ctx))

    def scalars(self):
        s = Scalars()
        self._check(lib().wf_get_scalars(self.ctx, ctypes.byref(s)))
        return s.as_dict()

    def reaction_history(self, zone, cap=1 << 16):
        out = np.zeros(3 * cap)
        cnt = ctypes.c_int64(0)
        self._check(lib().wf_get_reaction_history(self.ctx, zone, _dp(out), cap, ctypes.byref(cnt)))
        return out[:3 * cnt.value].reshape(-1, 3)

    def add_rigid_plane(self, zone, axis, flip, p1, p2, dens, hfac=1.1):
        a = np.ascontiguousarray(p1, dtype=np.float64)
        b = np.ascontiguousarray(p2, dtype=np.float64)
        rc = lib().wf_add_rigid_plane(self.ctx, int(zone), int(axis), int(bool(flip)), _dp(a), _dp(b), int(dens),
                                      float(hfac))
        if rc < 0:
            self._check(rc)
        return rc

    def add_rigid_mesh(self, zone, nodes, elems, hfac=1.1, normals=None):
        """A rigid triangle mesh: nodes (nn, 3), elems (ne, 3) 0-based, and
        optionally the initial element normals (ne, 3)."""
        nd = np.ascontiguousarray(nodes, dtype=np.float64).reshape(-1, 3)
        el = np.ascontiguousarray(elems, dtype=np.int32).reshape(-1, 3)
        nr = None if normals is None else np.ascontiguousarray(normals, dtype=np.float64).reshape(-1, 3)
        rc = lib().wf_add_rigid_mesh(self.ctx, int(zone), len(nd), _dp(nd), len(el), _ip(el),
                                     None if nr is None else _dp(nr), float(hfac))
        if rc < 0:
            self._check(rc)
        return rc

    def set_contact(self, friction_sta, friction_dyn, damping=0.0):
        self._check(lib().wf_set_contact(self.ctx, friction_sta, friction_dyn, damping))

    def set_contact_heat(self, conduction=False, friction=False, hc=0.0):
        """Contact heat conduction with conductance hc and friction
        heating (Domain::cont_heat_cond / contact_hc / cont_heat_fric)."""
        self._check(lib().wf_set_contact_heat(self.ctx, int(bool(conduction)), int(bool(friction)), float(hc)))

    def set_rigid_temperature(self, T):
        """Temperatures of the rigid particles, one each, in mesh order."""
        t = np.ascontiguousarray(T, dtype=np.float64).ravel()
        self._check(lib().wf_set_rigid_temperature(self.ctx, len(t), _dp(t)))

    def contact_force_history(self, mesh, cap=1 << 16):
        out = np.zeros(3 * cap)
        cnt = ctypes.c_int64(0)
        self._check(lib().wf_get_contact_force_history(self.ctx, mesh, _dp(out), cap, ctypes.byref(cnt)))
        return out[:3 * cnt.value].reshape(-1, 3)

    def rigid(self, name, nrig):
        out = np.zeros((nrig, 3))
        self._check(lib().wf_get_rigid_field(self.ctx, RIGID_FIELDS[name], _dp(out)))
        return out

    # adjoint window (FP64 Leapfrog): loss = sum_n (F_n - target_n)^2 over the
    # plane contact force or zone reaction component F_n of each step
    def adj_forward(self, target, checkpoint_every=20, kind=ADJ_PLANE, index=0, component=2, material=0):
        t = np.ascontiguousarray(target, dtype=np.float64)
        cfg = AdjConfig(steps=len(t), checkpoint_every=int(checkpoint_every), loss_kind=int(kind),
                        loss_index=int(index), loss_component=int(component), material=int(material))
        loss = ctypes.c_double(0)
        hist = np.zeros(len(t))
        self._check(lib().wf_adj_forward(self.ctx, ctypes.byref(cfg), _dp(t), ctypes.byref(loss), _dp(hist)))
        return loss.value, hist

    def adj_backward(self):
        g = np.zeros(len(ADJ_PARAMS))
        self._check(lib().wf_adj_backward(self.ctx, _dp(g)))
        return g

    def adj_field(self, name):
        f, w = ADJ_FIELDS[name]
        out = np.zeros((self.n, w)) if w > 1 else np.zeros(self.n)
        self._check(lib().wf_adj_get_field(self.ctx, f, _dp(out)))
        return out

    def adj_restore(self):
        self._check(lib().wf_adj_restore(self.ctx))

    # output tables ----------------------------------------------------------
    def pack_xdmf(self):
        """Domain::WriteXDMF's datasets packed on the device: (f, iv), float32
        [XDMF_NF * n] and int32 [XDMF_NI * n], dataset after dataset."""
        f = np.empty(XDMF_NF * self.n, dtype=np.float32)
        iv = np.empty(XDMF_NI * self.n, dtype=np.int32)
        self._check(lib().wf_pack_xdmf(self.ctx, f.ctypes.data_as(ctypes.POINTER(ctypes.c_float)), _ip(iv)))
        return f, iv

    def pack_xdmf_noalloc(self):
        """pack_xdmf staged in the library's static device scratch (no device
        allocation; contexts of at most pack_xdmf_noalloc_max() particles)."""
        f = np.empty(XDMF_NF * self.n, dtype=np.float32)
        iv = np.empty(XDMF_NI * self.n, dtype=np.int32)
        self._check(lib().wf_pack_xdmf_noalloc(self.ctx, f.ctypes.data_as(ctypes.POINTER(ctypes.c_float)), _ip(iv)))
        return f, iv

    def pack_csv(self):
        """Domain::WriteCSV's rows packed on the device, shape (n, CSV_NCOL)."""
        rows = np.empty(CSV_NCOL * self.n)
        self._check(lib().wf_pack_csv(self.ctx, _dp(rows)))
        return rows.reshape(self.n, CSV_NCOL)

    def rebuild_neighbours(self):
        self._check(lib().wf_rebuild_neighbours(self.ctx))

    def neighbour_pairs(self):
        cnt = ctypes.c_int64(0)
        self._check(lib().wf_get_neighbour_pairs(self.ctx, None, 0, ctypes.byref(cnt)))
        out = np.zeros(2 * max(cnt.value, 1), dtype=np.int32)
        self._check(lib().wf_get_neighbour_pairs(self.ctx, _ip(out), cnt.value, ctypes.byref(cnt)))
        return out[:2 * cnt.value].reshape(-1, 2)

    def eval(self, what, dt=0.0):
        self._check(lib().wf_eval(self.ctx, EVAL[what], float(dt)))

    def set_timing(self, on):
        self._check(lib().wf_set_timing(self.ctx, int(bool(on))))

    def kernel_time(self, kernel):
        ms = ctypes.c_double(0)
        cnt = ctypes.c_int64(0)
        self._check(lib().wf_get_kernel_time(self.ctx, KERNELS[kernel], ctypes.byref(ms), ctypes.byref(cnt)))
        return ms.value, cnt.value

    def set_material_friction(self, index, friction_sta, friction_dyn):
        self._check(lib().wf_set_material_friction(self.ctx, index, friction_sta, friction_dyn))

    def set_contact_law(self, law, pfac=0.5, fric_bound=False):
        """law: 0 Fraser, 1 Wang, 2 Seo, 4 LSDyna (Domain::contact_alg)."""
        self._check(lib().wf_set_contact_law(self.ctx, law, pfac, int(bool(fric_bound))))

    def set_material_damage(self, index, model, D=(0.0, 0.0, 0.0, 0.0, 0.0)):
        """model 1: Johnson-Cook damage with D = (D1..D5); 0: none."""
        Dv = np.ascontiguousarray(D, dtype=np.float64)
        self._check(lib().wf_set_material_damage(self.ctx, index, model, _dp(Dv)))

    def calc_force2233(self, xsph=0.0, shepard=False):
        """CalcForce2233 over every pair at the current state (no time advance)."""
        self._check(lib().wf_calc_force2233(self.ctx, float(xsph), int(bool(shepard))))

    def set_dimension(self, dim):
        """Domain::Dimension: 3, or 2 for plane strain (before init)."""
        self._check(lib().wf_set_dimension(self.ctx, int(dim)))

    def set_friction_function(self, linear, m=0.0, b=0.0):
        self._check(lib().wf_set_friction_function(self.ctx, int(bool(linear)), m, b))

    def set_group_sums(self, zone=-1, cap=4096):
        self._check(lib().wf_set_group_sums(self.ctx, zone, cap))

    def group_history(self, what, group, mesh=0, cap=1 << 16):
        """Per-group sums, oldest first: what 0 the zone reaction, 1 the
        contact force of rigid mesh `mesh` on the group; shape (steps, 3)."""
        out = np.zeros(3 * cap)
        cnt = ctypes.c_int64(0)
        self._check(lib().wf_get_group_history(self.ctx, what, group, mesh, _dp(out), cap, ctypes.byref(cnt)))
        return out[:3 * cnt.value].reshape(-1, 3)

    def cell_grids(self):
        """(reference grid cells per axis, SPH search grid cells per axis)."""
        r = np.zeros(3, dtype=np.int32)
        g = np.zeros(3, dtype=np.int32)
        self._check(lib().wf_get_cell_grids(self.ctx, _ip(r), _ip(g)))
        return r, g

    def stage_stats(self):
        """(tiles, tiles staged in LDS, total size of the staged sets) after
        the last neighbour build; FP64 contexts stage nothing."""
        t, st, sv = ctypes.c_int64(0), ctypes.c_int64(0), ctypes.c_int64(0)
        self._check(lib().wf_get_stage_stats(self.ctx, ctypes.byref(t), ctypes.byref(st), ctypes.byref(sv)))
        return t.value, st.value, sv.value

    def neighbour_stats(self):
        p = ctypes.c_int64(0)
        m = ctypes.c_int32(0)
        self._check(lib().wf_get_neighbour_stats(self.ctx, ctypes.byref(p), ctypes.byref(m)))
        return p.value, m.value
