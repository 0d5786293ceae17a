"""Python mirror of the C ABI in include/mpflip.h (ctypes).

`Simulator` plays the role of the reference's `class cspray`
(Multiphase/Multiphase/spray.h:18-389): it owns the device state through an
`mpf_ctx` and exposes one method per host step method of the reference
(hashAndSortParticles, addexternalforce, computeLevelset+sweepLSAndMardGrid,
mapvelp2g_bubble, sweepU/setWaterBoundaryU, project_CG_bubble,
mapvelg2p_bubble, CollisionSolid, updateTemperature, MeltSolid/Freezing/
GenerateGasParticle, runMC_smooth) and `step()` for bubblesim
(spray.cpp:979-1233).  Host buffers are numpy arrays; device buffers are
anything with .data_ptr() that the library's HIP runtime allocated
(hipmem.DeviceArray).

The compute path is the HIP library `libmpflip.so` built next to this file;
importing fails loudly when it is missing — there is no CPU fallback.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libmpflip.so")

# ---- constants (include/mpflip.h) ----
TYPEFLUID, TYPEAIR, TYPEBOUNDARY, TYPEVACUUM, TYPEAIRSOLO, TYPESOLID = range(6)
CELL_UNDEF = 0xFFFFFFFF
HOST, DEVICE = 0, 1
FP32, FP64, MIXED = 0, 1, 2
SOLVER_REFERENCE_CG, SOLVER_JACOBI_PCG = 0, 1
OPERATOR_UNIT, OPERATOR_VARDENS = 0, 1
ADVECT_EULER_COLLIDE, ADVECT_RK2, ADVECT_FREEZING = 0, 1, 2
HEAT_CG, HEAT_JACOBI = 0, 1
MC_LEVELSET, MC_LIQUID_GAS, MC_GAS, MC_SOLID = 0, 1, 2, 3

STATUS = {0: "OK", 1: "INVALID_ARG", 2: "NONFINITE", 3: "NO_CONVERGENCE", 4: "CAPACITY",
          5: "HIP_ERROR", 6: "IO"}
OK, INVALID_ARG, NONFINITE, NO_CONVERGENCE, CAPACITY, HIP_ERROR, IO = range(7)

FIELDS = [
    "MARK", "LSFLUID", "LSAIR", "LSMERGE", "PHIGRAD_X", "PHIGRAD_Y", "PHIGRAD_Z",
    "PHIFLUID", "PHIAIR", "PHISOLID", "DIV", "PRESSURE", "TEMPERATURE", "TEMPERATURE_CHANGE",
    "CELL_START", "CELL_END",
    "WATER_UX", "WATER_UY", "WATER_UZ", "AIR_UX", "AIR_UY", "AIR_UZ",
    "WATER_UX_OLD", "WATER_UY_OLD", "WATER_UZ_OLD", "AIR_UX_OLD", "AIR_UY_OLD", "AIR_UZ_OLD",
    "SOLID_UX", "SOLID_UY", "SOLID_UZ", "GRID_HASH", "GRID_INDEX", "SURFACE_TENSION",
]
FIELD_ID = {n: i for i, n in enumerate(FIELDS)}

ABI_FUNCTIONS = [
    "mpf_abi_version", "mpf_config_default", "mpf_create", "mpf_destroy", "mpf_last_error",
    "mpf_stream", "mpf_synchronize", "mpf_set_particles", "mpf_get_particles",
    "mpf_particle_count", "mpf_field_info", "mpf_get_field", "mpf_set_field",
    "mpf_set_seed_cells", "mpf_hash_sort", "mpf_apply_forces", "mpf_levelset_update", "mpf_p2g",
    "mpf_extrapolate_bc", "mpf_project", "mpf_g2p", "mpf_advect", "mpf_heat", "mpf_phase_change",
    "mpf_marching_cubes", "mpf_get_mesh", "mpf_get_mc_cases", "mpf_step",
    "mpf_set_kernel_timing", "mpf_kernel_times", "mpf_reset_kernel_times",
    "mpf_project_adjoint", "mpf_extrapolate_adjoint", "mpf_p2g_adjoint", "mpf_g2p_adjoint",
    "mpf_advect_adjoint", "mpf_hash_sort_noalloc", "mpf_update_seed_cells", "mpf_set_pour_source",
    "mpf_delete_particles", "mpf_pour", "mpf_mc_hash_sort", "mpf_get_mc_field", "mpf_get_mc_cells",
    "mpf_mark_solo_air", "mpf_surface_tension", "mpf_drag_force", "mpf_update_solo_air",
    "mpf_correct_positions", "mpf_correct_bubble_positions", "mpf_solid_motion",
    "mpf_get_mesh_normals", "mpf_write_mesh", "mpf_read_point_cloud",
]


class Config(ctypes.Structure):
    """mpf_config (field order and types as in include/mpflip.h)."""
    _fields_ = [
        ("nx", ctypes.c_int32), ("ny", ctypes.c_int32), ("nz", ctypes.c_int32),
        ("dx", ctypes.c_float), ("dt", ctypes.c_float),
        ("gravity", ctypes.c_float * 3), ("gmin", ctypes.c_float * 3),
        ("samplespace", ctypes.c_float), ("m0", ctypes.c_float), ("airm0", ctypes.c_float),
        ("particle_capacity", ctypes.c_int64),
        ("buoyancy_kick", ctypes.c_int32), ("bubble_max_vel", ctypes.c_float),
        ("buoyancy_rate_air", ctypes.c_float), ("buoyancy_rate_solo", ctypes.c_float),
        ("precision", ctypes.c_int32), ("solver", ctypes.c_int32), ("op", ctypes.c_int32),
        ("rho_liquid", ctypes.c_float), ("rho_air", ctypes.c_float),
        ("tol", ctypes.c_double), ("max_iters", ctypes.c_int32), ("ref_max_iters", ctypes.c_int32),
        ("refine_every", ctypes.c_int32), ("check_every", ctypes.c_int32),
        ("solid_constraint", ctypes.c_int32), ("flags_from_counts", ctypes.c_int32),
        ("flip_alpha", ctypes.c_float), ("integrator", ctypes.c_int32), ("fix_solids", ctypes.c_int32),
        ("bounce_vel", ctypes.c_float), ("bounce_pos", ctypes.c_float),
        ("heat_enabled", ctypes.c_int32), ("heat_solver", ctypes.c_int32),
        ("heat_jacobi_iters", ctypes.c_int32), ("heat_max_iters", ctypes.c_int32),
        ("heat_alpha", ctypes.c_float * 6), ("default_temperature", ctypes.c_float),
        ("heat_add_bottom", ctypes.c_int32), ("heat_increase_bottom", ctypes.c_float),
        ("heat_bottom_clamp", ctypes.c_float), ("heat_extend_boundary", ctypes.c_int32),
        ("heat_vacuum_neumann", ctypes.c_int32), ("heat_melt_and_boil", ctypes.c_int32),
        ("alpha_temp_trans", ctypes.c_float),
        ("melt_enabled", ctypes.c_int32), ("freeze_enabled", ctypes.c_int32),
        ("gas_enabled", ctypes.c_int32), ("gas_every", ctypes.c_int32),
        ("melting_point", ctypes.c_float), ("boiling_point", ctypes.c_float),
        ("liquid_heat_th", ctypes.c_float), ("temperature0", ctypes.c_float),
        ("dissolve_gas_rate", ctypes.c_float), ("vapor_gen_rate", ctypes.c_float),
        ("mc_enabled", ctypes.c_int32), ("mc_max_verts", ctypes.c_int64), ("mc_max_tris", ctypes.c_int64),
        ("seed", ctypes.c_uint64),
        ("adjoint_enabled", ctypes.c_int32),
        ("delete_enabled", ctypes.c_int32), ("delete_vacuum_sweeps", ctypes.c_int32),
        ("pour_every", ctypes.c_int32), ("pour_pos_rand", ctypes.c_float), ("pour_vel_rand", ctypes.c_float),
        ("pour_temperature", ctypes.c_float), ("seed_update_every", ctypes.c_int32),
        ("mc_source", ctypes.c_int32), ("mc_gas_airsolo", ctypes.c_int32), ("mc_every", ctypes.c_int32),
        ("mc_smooth_iters", ctypes.c_int32),
        ("solo_air_enabled", ctypes.c_int32), ("solo_tol_in", ctypes.c_float), ("solo_tol_out", ctypes.c_float),
        ("sph_viscosity", ctypes.c_float), ("surface_tension_sigma", ctypes.c_float),
        ("drag_enabled", ctypes.c_int32), ("drag_solo", ctypes.c_float), ("drag_grid", ctypes.c_float),
        ("drag_particle_update", ctypes.c_int32), ("correct_position", ctypes.c_int32),
        ("correction_spring", ctypes.c_float), ("correction_radius", ctypes.c_float),
        ("correct_bubble_air_ls", ctypes.c_int32),
        ("solid_bounce", ctypes.c_float), ("rigid_lock_y", ctypes.c_int32),
        ("pcg_recompute_q", ctypes.c_int32),
    ]


class SolveInfo(ctypes.Structure):
    _fields_ = [("iterations", ctypes.c_int32), ("refinements", ctypes.c_int32),
                ("converged", ctypes.c_int32), ("pad_", ctypes.c_int32),
                ("rel_residual", ctypes.c_double), ("abs_rr", ctypes.c_double)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_ if k != "pad_"}


class PhaseInfo(ctypes.Structure):
    _fields_ = [("melted", ctypes.c_int32), ("frozen", ctypes.c_int32), ("gas_added", ctypes.c_int32),
                ("deleted", ctypes.c_int32), ("particles", ctypes.c_int64)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_ if k != "pad_"}


class McInfo(ctypes.Structure):
    _fields_ = [("active_voxels", ctypes.c_int64), ("verts", ctypes.c_int64), ("tris", ctypes.c_int64)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class StepInfo(ctypes.Structure):
    _fields_ = [("project", SolveInfo), ("heat", SolveInfo), ("phase", PhaseInfo), ("mc", McInfo),
                ("particles", ctypes.c_int64), ("poured", ctypes.c_int64)]


_lib_handle = None


def lib():
    """Load libmpflip.so (built by `make -C 2017multiphase_amd`); raises if absent."""
    global _lib_handle
    if _lib_handle is not None:
        return _lib_handle
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(
            "libmpflip.so is not built (%s). Run __graft_entry__.build() or "
            "`make -C 2017multiphase_amd`; there is no CPU fallback." % LIB_PATH)
    L = ctypes.CDLL(LIB_PATH)
    vp, i32, i64, c = ctypes.c_void_p, ctypes.c_int32, ctypes.c_int64, ctypes
    L.mpf_abi_version.restype = i32
    L.mpf_config_default.argtypes = [c.POINTER(Config), i32, i32, i32]
    L.mpf_create.argtypes = [c.POINTER(Config), c.POINTER(vp)]
    L.mpf_destroy.argtypes = [vp]
    L.mpf_last_error.argtypes = [vp]
    L.mpf_last_error.restype = c.c_char_p
    L.mpf_stream.argtypes = [vp]
    L.mpf_stream.restype = vp
    L.mpf_synchronize.argtypes = [vp]
    L.mpf_set_particles.argtypes = [vp, i64] + [vp] * 8 + [i32]
    L.mpf_get_particles.argtypes = [vp, c.POINTER(i64)] + [vp] * 8 + [i32]
    L.mpf_particle_count.argtypes = [vp]
    L.mpf_particle_count.restype = i64
    L.mpf_field_info.argtypes = [vp, i32, c.POINTER(i64), c.POINTER(i32)]
    L.mpf_get_field.argtypes = [vp, i32, vp, i32]
    L.mpf_set_field.argtypes = [vp, i32, vp, i32]
    L.mpf_set_seed_cells.argtypes = [vp, vp, i32]
    for n in ("mpf_hash_sort", "mpf_apply_forces", "mpf_levelset_update", "mpf_p2g", "mpf_g2p", "mpf_advect",
              "mpf_mark_solo_air", "mpf_surface_tension", "mpf_drag_force", "mpf_update_solo_air",
              "mpf_correct_positions", "mpf_correct_bubble_positions", "mpf_solid_motion"):
        getattr(L, n).argtypes = [vp]
    L.mpf_extrapolate_bc.argtypes = [vp, i32]
    L.mpf_project.argtypes = [vp, c.POINTER(SolveInfo)]
    L.mpf_heat.argtypes = [vp, c.POINTER(SolveInfo)]
    L.mpf_phase_change.argtypes = [vp, i32, c.POINTER(PhaseInfo)]
    L.mpf_marching_cubes.argtypes = [vp, c.POINTER(McInfo)]
    L.mpf_get_mesh.argtypes = [vp, vp, vp, i32]
    L.mpf_get_mc_cases.argtypes = [vp, vp, i32]
    L.mpf_get_mesh_normals.argtypes = [vp, vp, i32]
    L.mpf_write_mesh.argtypes = [vp, c.c_char_p, i32, c.c_char_p]
    L.mpf_read_point_cloud.argtypes = [c.c_char_p, i64, vp, c.POINTER(i64)]
    L.mpf_mc_hash_sort.argtypes = [vp]
    L.mpf_get_mc_field.argtypes = [vp, vp, i32]
    L.mpf_get_mc_cells.argtypes = [vp, vp, vp, i32]
    L.mpf_step.argtypes = [vp, i32, c.POINTER(StepInfo)]
    L.mpf_set_kernel_timing.argtypes = [vp, i32]
    L.mpf_kernel_times.argtypes = [vp, vp, vp, vp, i32]
    L.mpf_kernel_times.restype = i32
    L.mpf_reset_kernel_times.argtypes = [vp]
    L.mpf_project_adjoint.argtypes = [vp, vp, vp, i32, c.POINTER(SolveInfo)]
    L.mpf_extrapolate_adjoint.argtypes = [vp, vp, vp, i32]
    L.mpf_p2g_adjoint.argtypes = [vp, vp, vp, vp, i32]
    L.mpf_g2p_adjoint.argtypes = [vp, vp, vp, vp, vp, vp, i32]
    L.mpf_advect_adjoint.argtypes = [vp, vp, vp, vp, vp, i32, c.POINTER(i64)]
    L.mpf_hash_sort_noalloc.argtypes = [c.POINTER(Config), i64, vp, vp, vp, vp, vp, vp]
    L.mpf_update_seed_cells.argtypes = [vp, i32]
    L.mpf_set_pour_source.argtypes = [vp, i64, vp, vp, c.c_int8]
    L.mpf_delete_particles.argtypes = [vp, c.POINTER(i64)]
    L.mpf_pour.argtypes = [vp, i32, c.POINTER(i64)]
    _lib_handle = L
    return L


def default_config(nx, ny, nz, **overrides):
    """Reference defaults (initparam, spray.cpp:23-79, 573-618) plus overrides."""
    cfg = Config()
    st = lib().mpf_config_default(ctypes.byref(cfg), nx, ny, nz)
    if st != OK:
        raise ValueError("mpf_config_default failed: %s" % STATUS.get(st, st))
    for k, v in overrides.items():
        cur = getattr(cfg, k)
        if isinstance(cur, ctypes.Array):
            for i, x in enumerate(v):
                cur[i] = x
        else:
            setattr(cfg, k, v)
    return cfg


MESH_POVRAY, MESH_OBJ, MESH_BOBJ_GZ = 0, 1, 2


def read_point_cloud(path, cap=None):
    """readdata_solid (LoadSolidData.cpp:12-59): the points of an .asc file, float64[n, 3]."""
    L = lib()
    n = ctypes.c_int64(0)
    st = L.mpf_read_point_cloud(str(path).encode(), 0, None, ctypes.byref(n))
    if st != OK:
        raise MpfError(st, L.mpf_last_error(None).decode())
    out = np.empty((n.value if cap is None else min(cap, n.value), 3))
    st = L.mpf_read_point_cloud(str(path).encode(), len(out), _ptr(out), ctypes.byref(n))
    if st != OK:
        raise MpfError(st, L.mpf_last_error(None).decode())
    return out


def mc_voxels(cfg):
    """Voxels of the marching-cubes lattice of cfg.mc_source."""
    if cfg.mc_source == MC_LEVELSET:
        return (cfg.nx - 1) * (cfg.ny - 1) * (cfg.nz - 1)
    return 8 * cfg.nx * cfg.ny * cfg.nz


def hash_sort_noalloc(cfg, n, pos, grid_hash, grid_index, cell_start, cell_end, stream=None):
    """mpf_hash_sort_noalloc on device buffers (hipmem.DeviceArray or raw
    device addresses as int): no allocation, asynchronous on `stream` (a raw
    hipStream_t as int, or None for the null stream)."""
    def addr(a):
        if a is None:
            return None
        return ctypes.c_void_p(a.data_ptr() if hasattr(a, "data_ptr") else int(a))
    # device buffers: hipmem.DeviceArray (the library's own HIP runtime) or raw addresses
    L = lib()
    st = L.mpf_hash_sort_noalloc(ctypes.byref(cfg), int(n), addr(pos), addr(grid_hash), addr(grid_index),
                                 addr(cell_start), addr(cell_end), ctypes.c_void_p(stream) if stream else None)
    if st != OK:
        raise MpfError(st, L.mpf_last_error(None).decode())


class MpfError(RuntimeError):
    def __init__(self, status, msg):
        super().__init__("%s: %s" % (STATUS.get(status, status), msg))
        self.status = status


def _ptr(a):
    if a is None:
        return None
    if hasattr(a, "data_ptr"):
        return ctypes.c_void_p(a.data_ptr())
    return a.ctypes.data_as(ctypes.c_void_p)


_PART_DTYPES = [("pos", np.float32, 3), ("vel", np.float32, 3), ("flag", np.int8, 1),
                ("mass", np.float32, 1), ("T", np.float32, 1), ("lheat", np.float32, 1),
                ("solubility", np.float32, 1), ("gascontain", np.float32, 1)]


class Simulator:
    """One simulation context on the current HIP device (see module docstring)."""

    def __init__(self, cfg):
        self.L = lib()
        self.cfg = cfg
        h = ctypes.c_void_p()
        st = self.L.mpf_create(ctypes.byref(cfg), ctypes.byref(h))
        if st != OK:
            raise MpfError(st, self.L.mpf_last_error(None).decode())
        self.h = h
        self.G = cfg.nx * cfg.ny * cfg.nz

    # -- lifecycle --
    def close(self):
        if self.h:
            self.L.mpf_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, st, allow=(OK,)):
        if st not in allow:
            raise MpfError(st, self.L.mpf_last_error(self.h).decode())
        return st

    @property
    def stream(self):
        return self.L.mpf_stream(self.h)

    def synchronize(self):
        self._check(self.L.mpf_synchronize(self.h))

    # -- data movement --
    def set_particles(self, n, where=HOST, **arrays):
        ptrs = []
        keep = []
        for name, dt, w in _PART_DTYPES:
            a = arrays.get(name)
            if a is not None and where == HOST:
                a = np.ascontiguousarray(a, dtype=dt)
                assert a.size == n * w, (name, a.size, n * w)
                keep.append(a)
            ptrs.append(_ptr(a))
        self._check(self.L.mpf_set_particles(self.h, n, *ptrs, where))

    def get_particles(self):
        n = self.particle_count()
        out = {}
        ptrs = []
        for name, dt, w in _PART_DTYPES:
            a = np.empty(n * w if w == 1 else (n, w), dtype=dt)
            out[name] = a
            ptrs.append(_ptr(a))
        nn = ctypes.c_int64()
        self._check(self.L.mpf_get_particles(self.h, ctypes.byref(nn), *ptrs, HOST))
        return out

    def particle_count(self):
        return int(self.L.mpf_particle_count(self.h))

    def field_info(self, name):
        n = ctypes.c_int64()
        es = ctypes.c_int32()
        self._check(self.L.mpf_field_info(self.h, FIELD_ID[name], ctypes.byref(n), ctypes.byref(es)))
        return n.value, es.value

    def get_field(self, name):
        n, es = self.field_info(name)
        if name == "MARK":
            dt = np.int8
        elif name in ("CELL_START", "CELL_END", "GRID_HASH", "GRID_INDEX"):
            dt = np.uint32
        else:
            dt = np.float64 if es == 8 else np.float32
        a = np.empty(n, dtype=dt)
        if n:
            self._check(self.L.mpf_get_field(self.h, FIELD_ID[name], _ptr(a), HOST))
        return a

    def set_field(self, name, values):
        n, es = self.field_info(name)
        if name == "MARK":
            dt = np.int8
        elif name in ("CELL_START", "CELL_END"):
            dt = np.uint32
        else:
            dt = np.float64 if es == 8 else np.float32
        a = np.ascontiguousarray(values, dtype=dt).ravel()
        assert a.size == n, (name, a.size, n)
        self._check(self.L.mpf_set_field(self.h, FIELD_ID[name], _ptr(a), HOST))

    def set_seed_cells(self, cells):
        a = np.ascontiguousarray(cells, dtype=np.int32)
        self._check(self.L.mpf_set_seed_cells(self.h, _ptr(a), a.size))

    # -- step entry points --
    def hash_sort(self):
        self._check(self.L.mpf_hash_sort(self.h))

    def apply_forces(self):
        self._check(self.L.mpf_apply_forces(self.h))

    def levelset_update(self):
        self._check(self.L.mpf_levelset_update(self.h))

    def p2g(self):
        self._check(self.L.mpf_p2g(self.h))

    def extrapolate_bc(self, recompute_phi=True):
        self._check(self.L.mpf_extrapolate_bc(self.h, 1 if recompute_phi else 0))

    def project(self, allow_no_convergence=False):
        info = SolveInfo()
        allow = (OK, NO_CONVERGENCE) if allow_no_convergence else (OK,)
        self._check(self.L.mpf_project(self.h, ctypes.byref(info)), allow)
        return info.as_dict()

    def g2p(self):
        self._check(self.L.mpf_g2p(self.h))

    def advect(self):
        self._check(self.L.mpf_advect(self.h))

    def solid_motion(self):
        self._check(self.L.mpf_solid_motion(self.h))

    def mark_solo_air(self):
        self._check(self.L.mpf_mark_solo_air(self.h))

    def surface_tension(self):
        self._check(self.L.mpf_surface_tension(self.h))

    def drag_force(self):
        self._check(self.L.mpf_drag_force(self.h))

    def update_solo_air(self):
        self._check(self.L.mpf_update_solo_air(self.h))

    def correct_positions(self):
        self._check(self.L.mpf_correct_positions(self.h))

    def correct_bubble_positions(self):
        self._check(self.L.mpf_correct_bubble_positions(self.h))

    def heat(self):
        info = SolveInfo()
        self._check(self.L.mpf_heat(self.h, ctypes.byref(info)))
        return info.as_dict()

    def phase_change(self, frame):
        info = PhaseInfo()
        self._check(self.L.mpf_phase_change(self.h, frame, ctypes.byref(info)))
        return info.as_dict()

    def marching_cubes(self):
        info = McInfo()
        self._check(self.L.mpf_marching_cubes(self.h, ctypes.byref(info)))
        return info.as_dict()

    def get_mesh(self, nverts, ntris):
        v = np.empty((nverts, 3), dtype=np.float32)
        t = np.empty((ntris, 3), dtype=np.uint32)
        self._check(self.L.mpf_get_mesh(self.h, _ptr(v), _ptr(t), HOST))
        return v, t

    def get_mesh_normals(self, nverts):
        a = np.empty((nverts, 3), dtype=np.float32)
        self._check(self.L.mpf_get_mesh_normals(self.h, _ptr(a), HOST))
        return a

    def write_mesh(self, path, fmt, object_name="water"):
        """outputPovRaywater / outputOBJwater / outputBOBJwater of the last mesh."""
        self._check(self.L.mpf_write_mesh(self.h, str(path).encode(), fmt, object_name.encode()))

    def get_mc_cases(self):
        n = mc_voxels(self.cfg)
        a = np.empty(n, dtype=np.uint8)
        self._check(self.L.mpf_get_mc_cases(self.h, _ptr(a), HOST))
        return a

    def mc_hash_sort(self):
        """preMC (spray.cu:759-765): hash, stable sort and reorder at 2N."""
        self._check(self.L.mpf_mc_hash_sort(self.h))

    def get_mc_field(self):
        c = self.cfg
        a = np.empty((2 * c.nx + 1) * (2 * c.ny + 1) * (2 * c.nz + 1), dtype=np.float32)
        self._check(self.L.mpf_get_mc_field(self.h, _ptr(a), HOST))
        return a

    def get_mc_cells(self):
        c = self.cfg
        n = 8 * c.nx * c.ny * c.nz
        s0 = np.empty(n, dtype=np.uint32)
        s1 = np.empty(n, dtype=np.uint32)
        self._check(self.L.mpf_get_mc_cells(self.h, _ptr(s0), _ptr(s1), HOST))
        return s0, s1

    def update_seed_cells(self, frame):
        self._check(self.L.mpf_update_seed_cells(self.h, frame))

    def set_pour_source(self, pos, vel, flag=TYPEFLUID):
        pos = np.ascontiguousarray(pos, dtype=np.float32).reshape(-1, 3)
        vel = np.ascontiguousarray(vel, dtype=np.float32).reshape(-1, 3)
        assert pos.shape == vel.shape
        self._check(self.L.mpf_set_pour_source(self.h, len(pos), _ptr(pos), _ptr(vel), flag))

    def delete_particles(self):
        n = ctypes.c_int64()
        self._check(self.L.mpf_delete_particles(self.h, ctypes.byref(n)))
        return int(n.value)

    def pour(self, frame):
        n = ctypes.c_int64()
        self._check(self.L.mpf_pour(self.h, frame, ctypes.byref(n)))
        return int(n.value)

    def step(self, frame, allow_no_convergence=False):
        info = StepInfo()
        allow = (OK, NO_CONVERGENCE) if allow_no_convergence else (OK,)
        self._check(self.L.mpf_step(self.h, frame, ctypes.byref(info)), allow)
        return {"project": info.project.as_dict(), "heat": info.heat.as_dict(),
                "phase": info.phase.as_dict(), "mc": info.mc.as_dict(), "particles": info.particles,
                "poured": info.poured}

    # -- adjoints (FP64 cotangents in host memory) --
    def face_sizes(self):
        nx, ny, nz = self.cfg.nx, self.cfg.ny, self.cfg.nz
        f = [(nx + 1) * ny * nz, nx * (ny + 1) * nz, nx * ny * (nz + 1)]
        return f + f

    def _faces_in(self, faces):
        """6 FP64 face arrays (or None = zero) -> (keep-alive list, void*[6] or None)."""
        if faces is None:
            return [], None
        arrs = [None if a is None else np.ascontiguousarray(a, dtype=np.float64).ravel() for a in faces]
        for a, n in zip(arrs, self.face_sizes()):
            assert a is None or a.size == n, (a.size, n)
        tab = (ctypes.c_void_p * 6)(*[None if a is None else a.ctypes.data for a in arrs])
        return arrs, tab

    def _faces_out(self):
        arrs = [np.empty(n, dtype=np.float64) for n in self.face_sizes()]
        return arrs, (ctypes.c_void_p * 6)(*[a.ctypes.data for a in arrs])

    @staticmethod
    def _vec(a, n):
        if a is None:
            return None
        a = np.ascontiguousarray(a, dtype=np.float64).ravel()
        assert a.size == n, (a.size, n)
        return a

    def project_adjoint(self, g_out, allow_no_convergence=False):
        """g_out: 6 face cotangents of mpf_project's output -> 6 of its input."""
        keep, tin = self._faces_in(g_out)
        res, tout = self._faces_out()
        info = SolveInfo()
        allow = (OK, NO_CONVERGENCE) if allow_no_convergence else (OK,)
        self._check(self.L.mpf_project_adjoint(self.h, ctypes.cast(tin, ctypes.c_void_p) if tin else None,
                                               ctypes.cast(tout, ctypes.c_void_p), HOST, ctypes.byref(info)), allow)
        return res, info.as_dict()

    def extrapolate_adjoint(self, g_out):
        keep, tin = self._faces_in(g_out)
        res, tout = self._faces_out()
        self._check(self.L.mpf_extrapolate_adjoint(self.h, ctypes.cast(tin, ctypes.c_void_p) if tin else None,
                                                   ctypes.cast(tout, ctypes.c_void_p), HOST))
        return res

    def p2g_adjoint(self, g_faces):
        """-> (g_vel[P,3], g_pos[P,3]) from the cotangents of the P2G output faces."""
        P = self.particle_count()
        keep, tin = self._faces_in(g_faces)
        gv = np.empty((P, 3), dtype=np.float64)
        gx = np.empty((P, 3), dtype=np.float64)
        self._check(self.L.mpf_p2g_adjoint(self.h, ctypes.cast(tin, ctypes.c_void_p) if tin else None,
                                           _ptr(gv), _ptr(gx), HOST))
        return gv, gx

    def g2p_adjoint(self, g_vel_out):
        """-> dict(vel, u[6], u_old[6], pos) cotangents of G2P's inputs."""
        P = self.particle_count()
        gvo = self._vec(g_vel_out, 3 * P)
        gv = np.empty((P, 3), dtype=np.float64)
        gx = np.empty((P, 3), dtype=np.float64)
        gu, tu = self._faces_out()
        guo, tuo = self._faces_out()
        self._check(self.L.mpf_g2p_adjoint(self.h, _ptr(gvo), _ptr(gv), ctypes.cast(tu, ctypes.c_void_p),
                                           ctypes.cast(tuo, ctypes.c_void_p), _ptr(gx), HOST))
        return {"vel": gv, "u": gu, "u_old": guo, "pos": gx}

    def advect_adjoint(self, g_pos_out, g_vel_out):
        """-> (g_pos_in[P,3], g_vel_in[P,3], particles inside the solid band)."""
        P = self.particle_count()
        gxo, gvo = self._vec(g_pos_out, 3 * P), self._vec(g_vel_out, 3 * P)
        gx = np.empty((P, 3), dtype=np.float64)
        gv = np.empty((P, 3), dtype=np.float64)
        band = ctypes.c_int64()
        self._check(self.L.mpf_advect_adjoint(self.h, _ptr(gxo), _ptr(gvo), _ptr(gx), _ptr(gv), HOST,
                                              ctypes.byref(band)))
        return gx, gv, int(band.value)

    # -- timing --
    def set_kernel_timing(self, on):
        self._check(self.L.mpf_set_kernel_timing(self.h, 1 if on else 0))

    def kernel_times(self):
        n = self.L.mpf_kernel_times(self.h, None, None, None, 0)
        if n <= 0:
            return {}
        names = ctypes.create_string_buffer(48 * n)
        counts = (ctypes.c_int64 * n)()
        ms = (ctypes.c_double * n)()
        self.L.mpf_kernel_times(self.h, names, counts, ms, n)
        out = {}
        for i in range(n):
            nm = names.raw[48 * i:48 * (i + 1)].split(b"\0", 1)[0].decode()
            out[nm] = (int(counts[i]), float(ms[i]))
        return out

    def reset_kernel_times(self):
        self._check(self.L.mpf_reset_kernel_times(self.h))
