"""ctypes binding of the CPU oracle (oracle/mpf_oracle.cpp) — TEST INFRASTRUCTURE.

Only tests/, __graft_entry__.smoke() (as the checker) and bench.py's CPU
baseline may import this module.  `Oracle` mirrors the product's Simulator
interface on host memory so a test drives both sides with the same calls.
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "liboracle.so")
_L = None


def build(force=False):
    cmd = ["make", "-C", _HERE] + (["-B"] if force else [])
    subprocess.run(cmd, check=True, stdout=subprocess.DEVNULL)


def lib():
    global _L
    if _L is not None:
        return _L
    if not os.path.exists(LIB_PATH):
        build()
    import importlib
    pkg = importlib.import_module("2017multiphase_amd")
    L = ctypes.CDLL(LIB_PATH)
    vp, i32, i64, c = ctypes.c_void_p, ctypes.c_int32, ctypes.c_int64, ctypes
    L.orc_create.argtypes = [c.POINTER(pkg.Config), c.c_int]
    L.orc_create.restype = vp
    L.orc_destroy.argtypes = [vp]
    L.orc_set_threads.argtypes = [c.c_int]
    L.orc_set_particles.argtypes = [vp, i64] + [vp] * 8
    L.orc_get_particles.argtypes = [vp, c.POINTER(i64)] + [vp] * 8
    L.orc_particle_count.argtypes = [vp]
    L.orc_set_kinematics64.argtypes = [vp, vp, vp]
    L.orc_particle_count.restype = i64
    L.orc_field_count.argtypes = [vp, c.c_int]
    L.orc_field_count.restype = i64
    L.orc_get_field.argtypes = [vp, c.c_int, vp]
    L.orc_set_field.argtypes = [vp, c.c_int, vp]
    L.orc_set_seed_cells.argtypes = [vp, vp, i32]
    L.orc_get_seed_cells.argtypes = [vp, vp]
    L.orc_update_seed_cells.argtypes = [vp, c.c_int]
    L.orc_set_pour_source.argtypes = [vp, i64, vp, vp, c.c_int8]
    L.orc_pour.argtypes = [vp, c.c_int]
    L.orc_pour.restype = i64
    L.orc_delete_particles.argtypes = [vp, c.c_int]
    L.orc_delete_particles.restype = i64
    for n in ("orc_hash_sort", "orc_apply_forces", "orc_levelset_update", "orc_p2g", "orc_g2p", "orc_advect"):
        getattr(L, n).argtypes = [vp]
    L.orc_extrapolate_bc.argtypes = [vp, c.c_int]
    L.orc_project.argtypes = [vp, c.POINTER(pkg.SolveInfo)]
    L.orc_heat.argtypes = [vp, c.POINTER(pkg.SolveInfo)]
    L.orc_heat_grid.argtypes = [vp, c.POINTER(pkg.SolveInfo)]
    L.orc_heat_g2p.argtypes = [vp]
    L.orc_set_solve_float.argtypes = [vp, c.c_int]
    for n in ("orc_phigrad", "orc_mark_from_counts", "orc_mark_solid_cells"):
        getattr(L, n).argtypes = [vp]
    L.orc_phase_change.argtypes = [vp, c.c_int, c.POINTER(pkg.PhaseInfo)]
    L.orc_marching_cubes.argtypes = [vp, c.POINTER(pkg.McInfo)]
    L.orc_get_mesh.argtypes = [vp, vp, vp]
    L.orc_get_mc_cases.argtypes = [vp, vp]
    L.orc_mc_hash_sort.argtypes = [vp]
    L.orc_get_mesh_normals.argtypes = [vp, vp]
    for n in ("mark_solo_air", "surface_tension", "drag_force", "update_solo_air", "correct_positions",
              "correct_bubble_positions", "solid_motion"):
        getattr(L, "orc_" + n).argtypes = [vp]
    L.orc_get_mc_field.argtypes = [vp, vp]
    L.orc_get_mc_cells.argtypes = [vp, vp, vp]
    L.orc_step.argtypes = [vp, c.c_int, c.POINTER(pkg.StepInfo)]
    L.orc_hash_key.argtypes = [c.POINTER(pkg.Config), c.c_float, c.c_float, c.c_float]
    L.orc_hash_key.restype = c.c_uint32
    L.orc_stable_sort_pairs.argtypes = [i64, vp, vp]
    L.orc_apply_operator.argtypes = [vp, vp, vp]
    L.orc_apply_heat_operator.argtypes = [vp, vp, vp]
    L.orc_uniform.argtypes = [c.c_uint64] * 4
    L.orc_uniform.restype = c.c_double
    L.orc_mc_tables.argtypes = [vp, vp, vp]
    _L = L
    return L


def _p(a):
    return None if a is None else a.ctypes.data_as(ctypes.c_void_p)


class Oracle:
    def __init__(self, cfg, use_double=False):
        import importlib
        self.pkg = importlib.import_module("2017multiphase_amd")
        self.L = lib()
        self.cfg = cfg
        self.h = self.L.orc_create(ctypes.byref(cfg), 1 if use_double else 0)
        if not self.h:
            raise ValueError("orc_create failed")
        self.G = cfg.nx * cfg.ny * cfg.nz

    def close(self):
        if self.h:
            self.L.orc_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _ok(self, st, allow=(0,)):
        if st not in allow:
            raise RuntimeError("oracle status %d" % st)
        return st

    def set_particles(self, n, **a):
        names = ["pos", "vel", "flag", "mass", "T", "lheat", "solubility", "gascontain"]
        arrs = []
        for nm in names:
            x = a.get(nm)
            if x is not None:
                x = np.ascontiguousarray(x, dtype=np.int8 if nm == "flag" else np.float32)
            arrs.append(x)
        self._ok(self.L.orc_set_particles(self.h, n, *[_p(x) for x in arrs]))

    def set_kinematics64(self, pos=None, vel=None):
        """Overwrite positions/velocities (FP64, current order) keeping all discrete state."""
        n = self.particle_count()
        pos = None if pos is None else np.ascontiguousarray(pos, dtype=np.float64).reshape(-1)
        vel = None if vel is None else np.ascontiguousarray(vel, dtype=np.float64).reshape(-1)
        assert pos is None or pos.size == 3 * n
        assert vel is None or vel.size == 3 * n
        self._ok(self.L.orc_set_kinematics64(self.h, _p(pos), _p(vel)))

    def get_particles(self):
        n = int(self.L.orc_particle_count(self.h))
        out = {"pos": np.empty((n, 3)), "vel": np.empty((n, 3)), "flag": np.empty(n, np.int8), "mass": np.empty(n),
               "T": np.empty(n), "lheat": np.empty(n), "solubility": np.empty(n), "gascontain": np.empty(n)}
        nn = ctypes.c_int64()
        self._ok(self.L.orc_get_particles(self.h, ctypes.byref(nn), *[_p(out[k]) for k in
                                                                      ["pos", "vel", "flag", "mass", "T", "lheat",
                                                                       "solubility", "gascontain"]]))
        return out

    def particle_count(self):
        return int(self.L.orc_particle_count(self.h))

    def get_field(self, name):
        f = self.pkg.FIELD_ID[name]
        n = int(self.L.orc_field_count(self.h, f))
        if name == "MARK":
            a = np.empty(n, np.int8)
        elif name in ("CELL_START", "CELL_END", "GRID_HASH", "GRID_INDEX"):
            a = np.empty(n, np.uint32)
        else:
            a = np.empty(n, np.float64)
        self._ok(self.L.orc_get_field(self.h, f, _p(a)))
        return a

    def set_field(self, name, values):
        f = self.pkg.FIELD_ID[name]
        n = int(self.L.orc_field_count(self.h, f))
        if name == "MARK":
            a = np.ascontiguousarray(values, np.int8).ravel()
        elif name in ("CELL_START", "CELL_END"):
            a = np.ascontiguousarray(values, np.uint32).ravel()
        else:
            a = np.ascontiguousarray(values, np.float64).ravel()
        assert a.size == n
        self._ok(self.L.orc_set_field(self.h, f, _p(a)))

    def set_seed_cells(self, cells):
        a = np.ascontiguousarray(cells, np.int32)
        self._ok(self.L.orc_set_seed_cells(self.h, _p(a), a.size))

    def get_seed_cells(self, cap=4096):
        a = np.empty(cap, np.int32)
        n = self.L.orc_get_seed_cells(self.h, _p(a))
        return a[:n].copy()

    def update_seed_cells(self, frame):
        self._ok(self.L.orc_update_seed_cells(self.h, frame))

    def set_pour_source(self, pos, vel, flag=0):
        pos = np.ascontiguousarray(pos, np.float32).reshape(-1)
        vel = np.ascontiguousarray(vel, np.float32).reshape(-1)
        self._ok(self.L.orc_set_pour_source(self.h, pos.size // 3, _p(pos), _p(vel), flag))

    def pour(self, frame):
        return int(self.L.orc_pour(self.h, frame))

    def delete_particles(self, keep_lists=True):
        return int(self.L.orc_delete_particles(self.h, 1 if keep_lists else 0))

    def hash_sort(self):
        self._ok(self.L.orc_hash_sort(self.h))

    def apply_forces(self):
        self._ok(self.L.orc_apply_forces(self.h))

    def levelset_update(self):
        self._ok(self.L.orc_levelset_update(self.h))

    def p2g(self):
        self._ok(self.L.orc_p2g(self.h))

    def extrapolate_bc(self, recompute_phi=True):
        self._ok(self.L.orc_extrapolate_bc(self.h, 1 if recompute_phi else 0))

    def project(self, allow_no_convergence=False):
        info = self.pkg.SolveInfo()
        self._ok(self.L.orc_project(self.h, ctypes.byref(info)), (0, 3) if allow_no_convergence else (0,))
        return info.as_dict()

    def g2p(self):
        self._ok(self.L.orc_g2p(self.h))

    def advect(self):
        self._ok(self.L.orc_advect(self.h))

    def solid_motion(self):
        self._ok(self.L.orc_solid_motion(self.h))

    def mark_solo_air(self):
        self._ok(self.L.orc_mark_solo_air(self.h))

    def surface_tension(self):
        self._ok(self.L.orc_surface_tension(self.h))

    def drag_force(self):
        self._ok(self.L.orc_drag_force(self.h))

    def update_solo_air(self):
        self._ok(self.L.orc_update_solo_air(self.h))

    def correct_positions(self):
        self._ok(self.L.orc_correct_positions(self.h))

    def correct_bubble_positions(self):
        self._ok(self.L.orc_correct_bubble_positions(self.h))

    def heat(self):
        info = self.pkg.SolveInfo()
        self._ok(self.L.orc_heat(self.h, ctypes.byref(info)))
        return info.as_dict()

    def heat_grid(self):
        """Grid step of the heat update on the current TEMPERATURE field."""
        info = self.pkg.SolveInfo()
        self._ok(self.L.orc_heat_grid(self.h, ctypes.byref(info)))
        return info.as_dict()

    def heat_g2p(self):
        """Grid-to-particle temperature map and latent-heat rules."""
        self._ok(self.L.orc_heat_g2p(self.h))

    def set_solve_float(self, on=True):
        self._ok(self.L.orc_set_solve_float(self.h, 1 if on else 0))

    def phigrad(self):
        self._ok(self.L.orc_phigrad(self.h))

    def mark_from_counts(self):
        self._ok(self.L.orc_mark_from_counts(self.h))

    def mark_solid_cells(self):
        self._ok(self.L.orc_mark_solid_cells(self.h))

    def phase_change(self, frame):
        info = self.pkg.PhaseInfo()
        self._ok(self.L.orc_phase_change(self.h, frame, ctypes.byref(info)))
        return info.as_dict()

    def marching_cubes(self):
        info = self.pkg.McInfo()
        self._ok(self.L.orc_marching_cubes(self.h, ctypes.byref(info)))
        return info.as_dict()

    def get_mesh(self, nverts, ntris):
        v = np.empty((nverts, 3))
        t = np.empty((ntris, 3), np.uint32)
        self._ok(self.L.orc_get_mesh(self.h, _p(v), _p(t)))
        return v, t

    def get_mesh_normals(self, nverts):
        a = np.empty((nverts, 3))
        self._ok(self.L.orc_get_mesh_normals(self.h, _p(a)))
        return a

    def get_mc_cases(self):
        a = np.empty(self.pkg.mc_voxels(self.cfg), np.uint8)
        self._ok(self.L.orc_get_mc_cases(self.h, _p(a)))
        return a

    def mc_hash_sort(self):
        self._ok(self.L.orc_mc_hash_sort(self.h))

    def get_mc_field(self):
        c = self.cfg
        a = np.empty((2 * c.nx + 1) * (2 * c.ny + 1) * (2 * c.nz + 1))
        self._ok(self.L.orc_get_mc_field(self.h, _p(a)))
        return a

    def get_mc_cells(self):
        c = self.cfg
        n = 8 * c.nx * c.ny * c.nz
        s0 = np.empty(n, np.uint32)
        s1 = np.empty(n, np.uint32)
        self._ok(self.L.orc_get_mc_cells(self.h, _p(s0), _p(s1)))
        return s0, s1

    def step(self, frame, allow_no_convergence=False):
        info = self.pkg.StepInfo()
        self._ok(self.L.orc_step(self.h, frame, ctypes.byref(info)), (0, 3) if allow_no_convergence else (0,))
        return {"project": info.project.as_dict(), "heat": info.heat.as_dict(), "phase": info.phase.as_dict(),
                "mc": info.mc.as_dict(), "particles": info.particles, "poured": info.poured}

    def apply_operator(self, x):
        x = np.ascontiguousarray(x, np.float64)
        y = np.empty_like(x)
        self._ok(self.L.orc_apply_operator(self.h, _p(x), _p(y)))
        return y

    def apply_heat_operator(self, x):
        x = np.ascontiguousarray(x, np.float64)
        y = np.empty_like(x)
        self._ok(self.L.orc_apply_heat_operator(self.h, _p(x), _p(y)))
        return y


def hash_key(cfg, x, y, z):
    return int(lib().orc_hash_key(ctypes.byref(cfg), x, y, z))


def stable_sort_pairs(keys, vals):
    k = np.ascontiguousarray(keys, np.uint32).copy()
    v = np.ascontiguousarray(vals, np.uint32).copy()
    lib().orc_stable_sort_pairs(k.size, _p(k), _p(v))
    return k, v


def uniform(seed, a, b, c):
    return lib().orc_uniform(seed, a, b, c)


def mc_tables():
    et = np.empty(256, np.int32)
    tt = np.empty(256 * 16, np.int32)
    nv = np.empty(256, np.int32)
    lib().orc_mc_tables(_p(et), _p(tt), _p(nv))
    return et, tt.reshape(256, 16), nv


def set_threads(n):
    return lib().orc_set_threads(n)
