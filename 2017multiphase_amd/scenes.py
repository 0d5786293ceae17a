"""Seeded synthetic scenes for the five workloads (host-side set-up only).

The reference builds its scenes in cspray::initparam / initparticle_*
(spray.cpp:83-547, spray.cu:395-484) from hard-coded tables and a point-cloud
file; those scenes are not reproduced.  These generators build the workloads
this library is measured on, each with its own seed (default = workload index):

  bubble_rise    64^3, lower 60% water, air sphere r=8 cells at (32,32,16)
  boiling        128^3, 70% liquid, heated floor, 64 nucleation sites
  melt_freeze    256^3, four 40^3 solid blocks (T=0.8) in liquid (T=1.1)
  diff_slab      96^3, water slab of 40 cells + 6 air bubbles (r 4..10)
  bubble_column  384^3, 85% liquid, 512 bubbles at the floor (r 3..12)

Every scene takes `n` to rescale the grid (lengths scale with n / n_default),
so tests run the same physics at small sizes.  Particles: 8 per occupied
cell, uniform in the cell; velocities start at zero.
"""
import numpy as np

from . import (ADVECT_EULER_COLLIDE, FP32, HEAT_JACOBI, MIXED, OPERATOR_VARDENS, SOLVER_JACOBI_PCG,
               TYPEAIR, TYPEFLUID, TYPESOLID, default_config)

DEFAULT_N = {"bubble_rise": 64, "boiling": 128, "melt_freeze": 256, "diff_slab": 96, "bubble_column": 384}
SEED = {"bubble_rise": 0, "boiling": 1, "melt_freeze": 2, "diff_slab": 3, "bubble_column": 4}
PER_CELL = 8


def _fill(rng, cells, n, dx):
    """8 particles per listed cell (flat cell ids, x-major), uniform in the cell."""
    cells = np.asarray(cells, dtype=np.int64)
    i = cells // (n * n)
    j = (cells // n) % n
    k = cells % n
    base = np.stack([i, j, k], axis=1).astype(np.float32)
    base = np.repeat(base, PER_CELL, axis=0)
    jit = rng.random((base.shape[0], 3), dtype=np.float32)
    return ((base + jit) * np.float32(dx)).astype(np.float32)


def _grid(n):
    c = np.arange(n, dtype=np.int64)
    I, J, K = np.meshgrid(c, c, c, indexing="ij")
    return I.ravel(), J.ravel(), K.ravel()


def _interior(I, J, K, n):
    return (I >= 1) & (I <= n - 2) & (J >= 1) & (J <= n - 2) & (K >= 1) & (K <= n - 2)


def _spheres_mask(I, J, K, centres, radii):
    m = np.zeros(I.shape, dtype=bool)
    for (cx, cy, cz), r in zip(centres, radii):
        d2 = (I + 0.5 - cx) ** 2 + (J + 0.5 - cy) ** 2 + (K + 0.5 - cz) ** 2
        m |= d2 < r * r
    return m


def _assemble(rng, n, dx, groups):
    """groups: list of (cell_mask, flag, temperature, lheat)."""
    pos, flag, T, LH = [], [], [], []
    for mask, fl, t, lh in groups:
        cells = np.flatnonzero(mask)
        p = _fill(rng, cells, n, dx)
        pos.append(p)
        flag.append(np.full(len(p), fl, dtype=np.int8))
        T.append(t(len(p)) if callable(t) else np.full(len(p), t, dtype=np.float32))
        LH.append(np.full(len(p), lh, dtype=np.float32))
    pos = np.concatenate(pos)
    return pos, np.concatenate(flag), np.concatenate(T).astype(np.float32), np.concatenate(LH)


def make(name, n=None, seed=None, precision=None, capacity_factor=1.15):
    """Returns dict(cfg, n_particles, particles (dict of arrays), seeds, name, n, steps, dt)."""
    n0 = DEFAULT_N[name]
    n = n or n0
    s = SEED[name] if seed is None else seed
    rng = np.random.default_rng(s)
    sc = n / float(n0)
    I, J, K = _grid(n)
    inner = _interior(I, J, K, n)
    seeds = np.zeros(0, dtype=np.int32)
    extra = {}
    if name == "bubble_rise":
        liquid = inner & (K < int(0.6 * n))
        air = liquid & _spheres_mask(I, J, K, [(32 * sc, 32 * sc, 16 * sc)], [8 * sc])
        liquid &= ~air
        pos, flag, T, LH = _assemble(rng, n, 1.0 / 64, [(liquid, TYPEFLUID, 293.15, 10.0), (air, TYPEAIR, 293.15, 0.0)])
        cfg = default_config(n, n, n, dt=2e-3, flip_alpha=0.95, precision=FP32 if precision is None else precision,
                             solver=SOLVER_JACOBI_PCG, op=OPERATOR_VARDENS, rho_liquid=1000.0, rho_air=1.0,
                             tol=1e-6, heat_enabled=0, gas_enabled=0, mc_enabled=1)
        steps, dt = 100, 2e-3
    elif name == "boiling":
        liquid = inner & (K < int(0.7 * n))
        tnoise = lambda m: (1.0 + rng.uniform(-0.01, 0.01, m)).astype(np.float32)
        pos, flag, T, LH = _assemble(rng, n, 1.0 / 64, [(liquid, TYPEFLUID, tnoise, 10.0)])
        nseed = max(1, int(round(64 * sc * sc)))
        xs = rng.integers(3, n - 3, size=(4 * nseed, 2))
        ids = np.unique(xs[:, 0] * n * n + xs[:, 1] * n + 2)[:nseed]
        seeds = ids.astype(np.int32)
        cfg = default_config(n, n, n, dt=1e-3, precision=MIXED if precision is None else precision,
                             solver=SOLVER_JACOBI_PCG, op=OPERATOR_VARDENS, rho_liquid=1000.0, rho_air=5.0,
                             tol=1e-6, refine_every=20, heat_enabled=1, heat_solver=HEAT_JACOBI,
                             heat_jacobi_iters=10, default_temperature=1.0, heat_add_bottom=1,
                             heat_increase_bottom=0.2, heat_bottom_clamp=1.2, heat_extend_boundary=1,
                             heat_vacuum_neumann=1, melting_point=0.0, boiling_point=1.05, temperature0=1.0,
                             liquid_heat_th=10.0, dissolve_gas_rate=0.95, gas_enabled=1, gas_every=2,
                             heat_alpha=[0.18, 0.08, 0.0, 0.0001, 0.08, 0.2], mc_enabled=1)
        extra["solub0"] = 0.95
        steps, dt = 200, 1e-3
    elif name == "melt_freeze":
        top = int(0.68 * n)
        liquid = inner & (K < top)
        b = max(2, int(round(40 * sc)))
        blocks = []
        tries = 0
        while len(blocks) < 4 and tries < 10000:
            tries += 1
            lo = rng.integers(2, [n - 2 - b, n - 2 - b, max(3, top - 1 - b)], size=3)
            if any(np.all(np.abs(lo - q) < b + 1) for q in blocks):
                continue
            blocks.append(lo)
        solid = np.zeros(I.shape, dtype=bool)
        for lo in blocks:
            solid |= ((I >= lo[0]) & (I < lo[0] + b) & (J >= lo[1]) & (J < lo[1] + b) & (K >= lo[2]) & (K < lo[2] + b))
        solid &= liquid
        liquid &= ~solid
        pos, flag, T, LH = _assemble(rng, n, 1.0 / 64, [(liquid, TYPEFLUID, 1.1, 0.1), (solid, TYPESOLID, 0.8, 0.0)])
        cfg = default_config(n, n, n, dt=1e-3, precision=MIXED if precision is None else precision,
                             solver=SOLVER_JACOBI_PCG, op=OPERATOR_VARDENS, rho_liquid=1000.0, rho_air=1.0,
                             tol=1e-6, refine_every=20, solid_constraint=1, fix_solids=1, heat_enabled=1,
                             heat_solver=HEAT_JACOBI, heat_jacobi_iters=10, default_temperature=1.1,
                             heat_add_bottom=0, melting_point=1.0, boiling_point=10.0, liquid_heat_th=0.1,
                             temperature0=1.0, melt_enabled=1, freeze_enabled=1, gas_enabled=0, mc_enabled=1)
        extra["blocks"] = np.array(blocks)
        steps, dt = 100, 1e-3
    elif name == "diff_slab":
        liquid = inner & (K < 1 + int(round(40 * sc)))
        cen, rad = [], []
        for _ in range(6):
            r = rng.uniform(4, 10) * sc
            cen.append((rng.uniform(r + 2, n - r - 2), rng.uniform(r + 2, n - r - 2),
                        rng.uniform(r + 2, max(r + 3, 40 * sc - r))))
            rad.append(r)
        air = liquid & _spheres_mask(I, J, K, cen, rad)
        liquid &= ~air
        pos, flag, T, LH = _assemble(rng, n, 1.0 / 64, [(liquid, TYPEFLUID, 293.15, 10.0), (air, TYPEAIR, 293.15, 0.0)])
        cfg = default_config(n, n, n, dt=2e-3, precision=FP32 if precision is None else precision,
                             solver=SOLVER_JACOBI_PCG, op=OPERATOR_VARDENS, rho_liquid=1000.0, rho_air=1.0,
                             tol=1e-6, heat_enabled=0, gas_enabled=0, mc_enabled=0)
        steps, dt = 16, 2e-3
    elif name == "bubble_column":
        liquid = inner & (K < int(0.85 * n))
        nb = max(1, int(round(512 * sc * sc)))
        cen, rad = [], []
        for _ in range(nb):
            r = rng.uniform(3, 12) * sc
            cen.append((rng.uniform(r + 1, n - r - 1), rng.uniform(r + 1, n - r - 1), rng.uniform(r + 1, 3 * r + 2)))
            rad.append(r)
        air = liquid & _spheres_mask(I, J, K, cen, rad)
        liquid &= ~air
        pos, flag, T, LH = _assemble(rng, n, 1.0 / 64, [(liquid, TYPEFLUID, 293.15, 10.0), (air, TYPEAIR, 293.15, 0.0)])
        cfg = default_config(n, n, n, dt=1e-3, precision=MIXED if precision is None else precision,
                             solver=SOLVER_JACOBI_PCG, op=OPERATOR_VARDENS, rho_liquid=1000.0, rho_air=1.0,
                             tol=1e-8, refine_every=20, heat_enabled=0, gas_enabled=0, mc_enabled=1)
        steps, dt = 50, 1e-3
    else:
        raise KeyError(name)
    P = len(flag)
    m0 = np.float32(cfg.m0)
    mass = np.full(P, m0, dtype=np.float32)
    sol = np.zeros(P, dtype=np.float32)
    gas = np.zeros(P, dtype=np.float32)
    if "solub0" in extra:
        sol[flag == TYPEFLUID] = np.float32(extra["solub0"] * cfg.airm0)
        gas[flag == TYPEFLUID] = np.float32(0.88) * sol[flag == TYPEFLUID]
    cfg.particle_capacity = int(P * capacity_factor) + 1024
    cfg.seed = s
    parts = {"pos": pos, "vel": np.zeros_like(pos), "flag": flag, "mass": mass, "T": T, "lheat": LH,
             "solubility": sol, "gascontain": gas}
    return {"name": name, "n": n, "cfg": cfg, "n_particles": P, "particles": parts, "seeds": seeds,
            "steps": steps, "dt": dt, "extra": extra}
