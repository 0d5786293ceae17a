"""Pins of the oracle's remaining bubblesim stages (no GPU).

Each stage is fixed by a configuration whose answer follows from the
reference's rule or a published method, not from the oracle's loop:
verifySoloAirParticle on hand-placed particles (each branch of its flag
rule, spray_k.cu:4223-4261); compsurfacetension_k on the analytic normals of
a sphere, whose divergence is 2/r (spray_k.cu:3543-3637), and the
particle-side force direction (3639-3680); calDragForce on one solo air
particle in a uniform liquid field, the face shares being the particle's
fractional position (5127-5165); the SPH density, pressure and forces of
calcDensPress_Air / enforceForceSoloAirP (4264-4373) against Mueller, Charypar
and Gross, "Particle-Based Fluid Simulation for Interactive Applications"
(SCA 2003): poly6 density, spiky pressure gradient and viscosity laplacian;
correctparticlepos's spring, its 0.3-cell cap and wall clamp (1330-1411); and
correctbubblepos / correctbubblepos_air at a planar interface, where the
trilinear level set is exact (3884-3991).
"""
import numpy as np
import pytest

H = 1.0 / 64
F32 = np.float32


def cfg_of(pkg, n=12, **kw):
    kw.setdefault("particle_capacity", 10000)
    kw.setdefault("mc_enabled", 0)
    return pkg.default_config(n, n, n, **kw)


def cell_centres(n):
    I, J, K = np.meshgrid(*[np.arange(n)] * 3, indexing="ij")
    return (np.stack([I, J, K], -1).reshape(-1, 3) + 0.5) * H


# --------------------------------------------------------------- solo air
def test_mark_solo_air_branches(pkg, orc):
    """tol1 = -1.45 h, tol2 = -0.5 h (defaults).  A: a cluster of 12 within h
    of each other where ls = -2 h: its AIR particles turn AIRSOLO (ls < tol2 h)
    and take the air grid velocity, its AIRSOLO particle turns AIR (count
    >= 10).  Where ls = -0.3 h: B, an isolated AIR particle (count 1) turns
    AIRSOLO; C, an AIRSOLO particle in the x/y wall band stays; D, an AIRSOLO
    particle with one AIR neighbour (count 2 < 3) stays and so does the AIR
    one; E, an AIRSOLO particle with two AIR neighbours (count 3) turns AIR."""
    n = 16
    cfg = cfg_of(pkg, n)
    U = np.array([0.3, -0.2, 0.7])
    c1 = np.array([8.5, 8.5, 8.5]) * H
    rng = np.random.default_rng(3)
    pts = list(c1 + rng.uniform(-0.2, 0.2, (12, 3)) * H)
    flags = [pkg.TYPEAIR] * 11 + [pkg.TYPEAIRSOLO]                                     # A: 0..11
    pts.append(np.array([4.5, 12.5, 4.5]) * H); flags.append(pkg.TYPEAIR)             # B: 12
    pts.append(np.array([1.5, 8.5, 8.5]) * H); flags.append(pkg.TYPEAIRSOLO)          # C: 13
    pts += [np.array([12.5, 4.5, 12.5]) * H + np.array(d) * H for d in ([0, 0, 0], [0.3, 0, 0], [0, 0.3, 0])]
    flags += [pkg.TYPEAIRSOLO, pkg.TYPEFLUID, pkg.TYPEAIR]                             # D: 14..16
    pts += [np.array([12.5, 12.5, 12.5]) * H + np.array(d) * H for d in ([0, 0, 0], [0.3, 0, 0], [0, 0.3, 0])]
    flags += [pkg.TYPEAIRSOLO, pkg.TYPEAIR, pkg.TYPEAIR]                               # E: 17..19
    P = len(pts)
    pos = np.array(pts, F32)
    o = orc.Oracle(cfg, use_double=True)
    o.set_particles(P, pos=pos, flag=np.array(flags, np.int8))
    o.hash_sort()
    order = o.get_field("GRID_INDEX")
    x = cell_centres(n)
    ls = np.full(n ** 3, -0.3 * H)
    ls[np.max(np.abs(x - c1), axis=1) < 2.1 * H] = -2.0 * H
    o.set_field("LSMERGE", ls)
    for a, nm in enumerate(("AIR_UX", "AIR_UY", "AIR_UZ")):
        o.set_field(nm, np.full(o.get_field(nm).size, U[a]))
    o.mark_solo_air()
    got = o.get_particles()
    fl = np.empty(P, np.int8)
    fl[order] = got["flag"]
    vel = np.empty((P, 3))
    vel[order] = got["vel"]
    assert np.all(fl[:11] == pkg.TYPEAIRSOLO) and np.allclose(vel[:11], U, rtol=0, atol=1e-15)
    assert fl[11] == pkg.TYPEAIR and np.all(vel[11] == 0)
    assert fl[12] == pkg.TYPEAIRSOLO and np.allclose(vel[12], U, rtol=0, atol=1e-15)
    assert fl[13] == pkg.TYPEAIRSOLO and np.all(vel[13] == 0)
    assert list(fl[14:17]) == [pkg.TYPEAIRSOLO, pkg.TYPEFLUID, pkg.TYPEAIR]
    assert list(fl[17:20]) == [pkg.TYPEAIR] * 3
    assert np.all(vel[14:20] == 0)


# --------------------------------------------------------------- surface tension
def test_surface_tension_sphere_curvature(pkg, orc):
    """Analytic unit normals of a sphere of centre c: the cell force is
    sigma * div(n) = sigma * 2/|x - c| to the accuracy of the central
    difference; BOUNDARY cells are 0; an AIR particle near the interface with
    AIR cells around gets v -= dt * sf(x) n(x), other types are untouched."""
    n = 24
    sigma = 0.7
    cfg = cfg_of(pkg, n, surface_tension_sigma=sigma)
    o = orc.Oracle(cfg, use_double=True)
    x = cell_centres(n)
    cen = np.array([12.2, 11.9, 12.1]) * H
    d = x - cen
    r = np.linalg.norm(d, axis=1)
    for a, nm in enumerate(("PHIGRAD_X", "PHIGRAD_Y", "PHIGRAD_Z")):
        o.set_field(nm, d[:, a] / r * 3.0)  # any positive scale: the stage normalises
    ls = r - 6 * H
    o.set_field("LSMERGE", ls)
    mark = np.where(ls > 0, pkg.TYPEAIR, pkg.TYPEFLUID).astype(np.int8).reshape(n, n, n)
    mark[[0, -1], :, :] = pkg.TYPEBOUNDARY
    o.set_field("MARK", mark.ravel())
    pp = np.array([cen + np.array([6.3, 0.2, -0.1]) * H, cen + np.array([0.1, -6.2, 0.3]) * H,
                   cen + np.array([-6.4, 0.1, 0.2]) * H], F32)
    fl = np.array([pkg.TYPEAIR, pkg.TYPEAIR, pkg.TYPEFLUID], np.int8)
    v0 = np.array([[0.1, 0.2, 0.3]] * 3, F32)
    o.set_particles(3, pos=pp, vel=v0, flag=fl)
    o.surface_tension()
    sf = o.get_field("SURFACE_TENSION").reshape(n, n, n)
    assert np.all(sf[[0, -1], :, :] == 0)
    inner = (np.abs(r - 6 * H) < 2 * H).reshape(n, n, n)
    inner[[0, 1, -2, -1], :, :] = False
    ratio = sf[inner] / (sigma * 2 / r.reshape(n, n, n)[inner])
    assert np.all(np.abs(ratio - 1) < 0.05), (ratio.min(), ratio.max())
    v = o.get_particles()["vel"]
    dt = cfg.dt
    for q in range(2):
        dv = (v[q] - v0[q]) / dt
        nrm = (pp[q] - cen) / np.linalg.norm(pp[q] - cen)
        mag = np.linalg.norm(dv)
        assert mag > 0 and np.allclose(dv / mag, -nrm, atol=0.05)  # sf > 0: pushed inward
        assert abs(mag / (sigma * 2 / (6.3 * H)) - 1) < 0.1
    assert np.allclose(v[2], v0[2])


# --------------------------------------------------------------- drag
def test_drag_uniform_liquid(pkg, orc):
    """One AIRSOLO particle with velocity v in a uniform liquid field U: the
    drag b = solo |U - v| (U - v); the liquid faces of its cell get
    -b grid dt (1 - w) on the low face and -b grid dt w on the high face per
    axis, w the fractional position; two particles add; other types and
    other faces are untouched; the particle itself changes only with
    drag_particle_update (v += b dt)."""
    n = 10
    for upd in (0, 1):
        cfg = cfg_of(pkg, n, drag_particle_update=upd)
        o = orc.Oracle(cfg, use_double=True)
        U = np.array([0.4, -0.3, 0.2])
        fr = np.array([[0.25, 0.5, 0.75], [0.6, 0.1, 0.3]])
        pos = ((np.array([4, 5, 6]) + fr) * H).astype(F32)
        pos = np.concatenate([pos, [(np.array([2.5, 2.5, 2.5]) * H).astype(F32)]])
        vel = np.array([[0.0, 0.1, -0.2], [0.3, 0.0, 0.1], [1.0, 1.0, 1.0]], F32)
        flag = np.array([pkg.TYPEAIRSOLO, pkg.TYPEAIRSOLO, pkg.TYPEAIR], np.int8)
        o.set_particles(3, pos=pos, vel=vel, flag=flag)
        o.hash_sort()
        names = ("WATER_UX", "WATER_UY", "WATER_UZ")
        for a, nm in enumerate(names):
            o.set_field(nm, np.full(o.get_field(nm).size, U[a]))
        o.drag_force()
        dims = {0: (n + 1, n, n), 1: (n, n + 1, n), 2: (n, n, n + 1)}
        exp = {a: np.full(dims[a], U[a]) for a in range(3)}
        bs = []
        for q in range(2):
            dvq = U - vel[q].astype(float)
            b = cfg.drag_solo * np.linalg.norm(dvq) * dvq
            bs.append(b)
            ch = -b * cfg.drag_grid * cfg.dt
            w = pos[q].astype(float) / H - np.array([4, 5, 6])
            for a in range(3):
                lo = [4, 5, 6]
                hi = list(lo)
                hi[a] += 1
                exp[a][tuple(lo)] += ch[a] * (1 - w[a])
                exp[a][tuple(hi)] += ch[a] * w[a]
        for a, nm in enumerate(names):
            assert np.allclose(o.get_field(nm).reshape(dims[a]), exp[a], rtol=0, atol=1e-14)
        got = o.get_particles()
        order = o.get_field("GRID_INDEX")
        v = np.empty((3, 3))
        v[order] = got["vel"]
        for q in range(2):
            want = vel[q] + (bs[q] * cfg.dt if upd else 0)
            assert np.allclose(v[q], want, rtol=0, atol=1e-12)
        assert np.allclose(v[2], vel[2])


# --------------------------------------------------------------- SPH solo air
def mueller_sph(pos, vel, air, h, m, mu, rho0, dt):
    """Mueller et al. 2003: rho_i = sum_j m W_poly6, p_i = 1.5 (rho_i - rho0/2),
    a_i = (1/rho_i) sum_j m [-(p_i + p_j) / (2 rho_j) grad W_spiky
    + mu (v_j - v_i) / rho_j lap W_visc]."""
    P = len(pos)
    rho = np.ones(P)
    for i in range(P):
        if not air[i]:
            continue
        s = 0.0
        for j in range(P):
            r2 = np.sum((pos[j] - pos[i]) ** 2)
            if air[j] and r2 < h * h:
                s += m * 315 / (64 * np.pi * h ** 9) * (h * h - r2) ** 3
        rho[i] = s if s != 0 else 1.0
    p = 1.5 * (rho - rho0 * 0.5)
    out = vel.copy()
    for i in range(P):
        if not air[i]:
            continue
        a = np.zeros(3)
        for j in range(P):
            d = pos[i] - pos[j]
            r = np.linalg.norm(d)
            if air[j] and 0 < r < h:
                grad = -45 / (np.pi * h ** 6) * (h - r) ** 2 * d / r
                lap = 45 / (np.pi * h ** 6) * (h - r)
                a += m * (-(p[i] + p[j]) / (2 * rho[j]) * grad + mu * (vel[j] - vel[i]) / rho[j] * lap)
        out[i] = vel[i] + a / rho[i] * dt
    return out


def test_sph_solo_air_against_mueller(pkg, orc):
    n = 10
    cfg = cfg_of(pkg, n)
    rng = np.random.default_rng(5)
    P = 40
    pos = (np.array([5.0, 5.0, 5.0]) * H + rng.uniform(-1.2, 1.2, (P, 3)) * H).astype(F32)
    vel = rng.normal(0, 0.3, (P, 3)).astype(F32)
    flag = rng.choice([pkg.TYPEAIR, pkg.TYPEAIRSOLO, pkg.TYPEFLUID], P).astype(np.int8)
    o = orc.Oracle(cfg, use_double=True)
    o.set_particles(P, pos=pos, vel=vel, flag=flag)
    o.hash_sort()
    order = o.get_field("GRID_INDEX")
    o.update_solo_air()
    v = np.empty((P, 3))
    v[order] = o.get_particles()["vel"]
    air = np.isin(flag, [pkg.TYPEAIR, pkg.TYPEAIRSOLO])
    want = mueller_sph(pos.astype(float), vel.astype(float), air, H, float(F32(cfg.airm0)), float(F32(cfg.sph_viscosity)),
                       float(F32(cfg.rho_liquid)), float(F32(cfg.dt)))
    # the reference's constants are float (pi as 3.141592f): relative 1e-6
    dv_got, dv_want = v - vel, want - vel
    assert np.max(np.abs(dv_got - dv_want)) <= 1e-5 * np.max(np.abs(dv_want)) + 1e-12
    assert np.max(np.abs(dv_want[air])) > 0
    assert np.array_equal(v[~air], vel[~air].astype(float))


# --------------------------------------------------------------- position correction
def test_correct_positions_spring_cap_and_clamp(pkg, orc):
    """Two FLUID particles d apart (0.1 re < d < re, re = 0.5 h): each moves
    away from the other by dt k re m (1 - d^2/re^2) when that is below 0.3 h,
    else by exactly 0.3 h; closer than 0.1 re nothing moves; SOLID and
    AIRSOLO particles push but do not move; positions clamp to
    [gmin + h + ss/2, gmax - h - ss/2]."""
    n = 10
    re = 0.5 * H
    for spring, capped in ((5.0, False), (500000.0, True)):
        cfg = cfg_of(pkg, n, correction_spring=spring)
        m = cfg.m0
        a = np.array([5.3, 5.5, 5.5]) * H
        b = a + np.array([0.15, 0.0, 0.0]) * H  # d = 0.15 h = 0.3 re
        c = np.array([3.5, 3.5, 3.5]) * H
        c2 = c + np.array([0.02, 0, 0]) * H    # 0.04 re: too close to push
        s0 = np.array([7.5, 7.5, 7.5]) * H
        s1 = s0 + np.array([0, 0.3, 0]) * H
        pos = np.array([a, b, c, c2, s0, s1], F32)
        flag = np.array([0, 0, 0, 0, pkg.TYPESOLID, 0], np.int8)
        mass = np.full(6, m, F32)
        o = orc.Oracle(cfg, use_double=True)
        o.set_particles(6, pos=pos, flag=flag, mass=mass)
        o.hash_sort()
        order = o.get_field("GRID_INDEX")
        o.correct_positions()
        got = np.empty((6, 3))
        got[order] = o.get_particles()["pos"]
        p64 = pos.astype(float)
        d = np.linalg.norm(p64[1] - p64[0])
        mag = cfg.dt * spring * re * float(F32(m)) * (1 - d * d / (re * re))
        cap = float(F32(0.3)) * H  # the reference's 0.3f
        step = cap if capped else mag
        assert capped == (mag > cap)
        u = (p64[0] - p64[1]) / d
        assert np.allclose(got[0], p64[0] + step * u, rtol=0, atol=1e-12)
        assert np.allclose(got[1], p64[1] - step * u, rtol=0, atol=1e-12)
        assert np.allclose(got[2:5], p64[2:5])
        assert got[5][1] > p64[5][1]   # pushed by the solid particle, which stays
    # wall clamp
    cfg = cfg_of(pkg, n)
    lo = cfg.dx + 0.5 * cfg.samplespace
    pos = np.array([[0.2 * H, 5.5 * H, 5.5 * H], [5.5 * H, 9.9 * H, 5.5 * H]], F32)
    o = orc.Oracle(cfg, use_double=True)
    o.set_particles(2, pos=pos, flag=np.zeros(2, np.int8), mass=np.full(2, cfg.m0, F32))
    o.hash_sort()
    o.correct_positions()
    got = o.get_particles()["pos"]
    assert np.isclose(got[:, 0].min(), lo) and np.isclose(got[:, 1].max(), n * H - lo)


def planar(pkg, orc, z0, zs, flags, air_ls=0, lsair=None):
    n = 12
    cfg = cfg_of(pkg, n, correct_bubble_air_ls=air_ls)
    o = orc.Oracle(cfg, use_double=True)
    x = cell_centres(n)
    o.set_field("LSMERGE", 2.0 * (x[:, 2] - z0))  # gradient 2: the stage divides by it
    o.set_field("PHIGRAD_X", np.zeros(n ** 3))
    o.set_field("PHIGRAD_Y", np.zeros(n ** 3))
    o.set_field("PHIGRAD_Z", np.full(n ** 3, 2.0))
    if lsair is not None:
        o.set_field("LSAIR", lsair(x))
    pos = np.array([[6.1 * H, 5.7 * H, z] for z in zs], F32)
    o.set_particles(len(zs), pos=pos, flag=np.array(flags, np.int8))
    o.correct_bubble_positions()
    return o.get_particles()["pos"][:, 2], pos[:, 2].astype(float)


def test_correct_bubble_planar_interface(pkg, orc):
    """lsmerge = z - z0 (liquid below): a FLUID particle 0.2 h above z0 is
    taken back to z0 and then rs = h/2 into the liquid; an AIR particle
    0.4 h below goes to z0 (|d| > 0.3 h), one 0.2 h below stays; a FLUID
    particle 0.1 h below is pushed to rs below z0, one 0.3 h below stays
    (s d >= rs/2); AIRSOLO and SOLID never move."""
    z0 = 6.0 * H
    zs = [z0 + 0.2 * H, z0 - 0.4 * H, z0 - 0.2 * H, z0 - 0.1 * H, z0 - 0.3 * H, z0 + 0.2 * H, z0 + 0.2 * H]
    fl = [pkg.TYPEFLUID, pkg.TYPEAIR, pkg.TYPEAIR, pkg.TYPEFLUID, pkg.TYPEFLUID, pkg.TYPEAIRSOLO, pkg.TYPESOLID]
    z, zin = planar(pkg, orc, z0, zs, fl)
    want = [z0 - 0.5 * H, z0, zin[2], z0 - 0.5 * H, zin[4], zin[5], zin[6]]
    assert np.allclose(z, want, rtol=0, atol=1e-12)


def test_correct_bubble_air_levelset_variant(pkg, orc):
    """correctbubblepos_air: a FLUID particle with lsair d < -1.3 rs is moved
    by -(d - rs) along the air normal, to rs outside the air; one with
    d > -1.3 rs stays; the merged-level-set AIR rule still applies."""
    z0 = 6.0 * H
    za = 3.0 * H  # air level set zero plane, air below: lsair = za - z, gradient -z
    zs = [2.0 * H, 2.8 * H, z0 - 0.4 * H]
    fl = [pkg.TYPEFLUID, pkg.TYPEFLUID, pkg.TYPEAIR]
    n = 12
    cfg = cfg_of(pkg, n, correct_bubble_air_ls=1)
    o = orc.Oracle(cfg, use_double=True)
    x = cell_centres(n)
    o.set_field("LSMERGE", 2.0 * (x[:, 2] - z0))
    o.set_field("PHIGRAD_X", np.zeros(n ** 3))
    o.set_field("PHIGRAD_Y", np.zeros(n ** 3))
    o.set_field("PHIGRAD_Z", np.full(n ** 3, 2.0))
    o.set_field("LSAIR", 2.0 * (x[:, 2] - za))  # its gradient (2) is taken inside the stage
    pos = np.array([[6.1 * H, 5.7 * H, z] for z in zs], F32)
    o.set_particles(3, pos=pos, flag=np.array(fl, np.int8))
    o.correct_bubble_positions()
    z = o.get_particles()["pos"][:, 2]
    rs = 0.5 * H
    d0 = float(pos[0, 2]) - za  # lsair / |grad lsair|
    assert d0 < -1.3 * rs
    assert np.isclose(z[0], float(pos[0, 2]) - (d0 - rs), rtol=0, atol=1e-12)  # = za + rs
    assert np.isclose(z[0], za + rs, rtol=0, atol=1e-12)
    assert np.isclose(z[1], float(pos[1, 2]), rtol=0, atol=1e-15)
    assert np.isclose(z[2], z0, rtol=0, atol=1e-12)


# --------------------------------------------------------------- rigid solid motion
def solid_oracle(pkg, orc, pos, vel, flag, **kw):
    n = 16
    cfg = cfg_of(pkg, n, fix_solids=0, **kw)
    o = orc.Oracle(cfg, use_double=True)
    o.set_particles(len(pos), pos=np.array(pos, F32), vel=np.array(vel, F32), flag=np.array(flag, np.int8))
    o.solid_motion()
    p = o.get_particles()
    return cfg, p["pos"], p["vel"]


def test_solid_motion_translation(pkg, orc):
    """A solid whose particles all move with T0 (away from the walls):
    the centre moves by dt T0, the rotation estimate is 0 (sum r = 0), so
    every particle moves by dt T0 and keeps T0; other particles are
    untouched."""
    rng = np.random.default_rng(8)
    pos = (np.array([8.0, 8.0, 8.0]) + rng.uniform(-2, 2, (40, 3))) * H
    T0 = np.array([0.3, -0.2, 0.1])
    vel = np.tile(T0, (40, 1))
    flag = [pkg.TYPESOLID] * 39 + [pkg.TYPEFLUID]
    cfg, p, v = solid_oracle(pkg, orc, pos, vel, flag)
    x0 = np.array(pos, F32).astype(float)
    t0 = np.float32(T0).astype(float)
    assert np.allclose(p[:39], x0[:39] + cfg.dt * t0, rtol=0, atol=1e-7)
    assert np.allclose(v[:39], t0, rtol=0, atol=1e-7)
    assert np.array_equal(p[39], x0[39]) and np.array_equal(v[39], t0)


def test_solid_motion_ring_rotation(pkg, orc):
    """Particles on a ring of radius a about the z axis through c, moving
    with v = w z x r: sum r x v = n a^2 w and the mean |r| is a, so the
    rotation estimate is exactly w; positions turn by w dt about c and
    velocities stay the rigid ones."""
    a, w = 3.0 * H, 2.5
    c = np.array([8.0, 8.0, 8.0]) * H
    ph = np.linspace(0, 2 * np.pi, 24, endpoint=False)
    r = np.stack([a * np.cos(ph), a * np.sin(ph), np.zeros_like(ph)], 1)
    pos = c + r
    vel = np.stack([-w * r[:, 1], w * r[:, 0], np.zeros_like(ph)], 1)
    cfg, p, v = solid_oracle(pkg, orc, pos, vel, [pkg.TYPESOLID] * 24)
    x0 = np.array(pos, F32).astype(float)
    cen = x0.mean(0)
    th = w * float(F32(cfg.dt))
    rot = np.array([[np.cos(th), -np.sin(th), 0], [np.sin(th), np.cos(th), 0], [0, 0, 1]])
    want = (x0 - cen) @ rot.T + cen
    assert np.max(np.abs(p - want)) < 2e-6 * a  # float host arithmetic of the reference
    v0 = np.array(vel, F32).astype(float)
    assert np.max(np.abs(v - v0)) < 1e-5 * w * a


def test_solid_motion_wall_push_and_fixed(pkg, orc):
    """A solid particle below gmin + h + 0.3 ss takes (tmin - x) bounce n
    on that velocity component before the sums (n solid particles), which
    moves the body's translation by that over n; with fix_solids the solid
    velocities are zeroed and nothing moves; rigid_lock_y drops T.y."""
    n_in = 4
    base = (np.array([8.0, 8.0, 8.0]) + np.array([[0, 0, 0], [0.5, 0, 0], [0, 0.5, 0], [0, 0, 0.5]])) * H
    low = np.array([[8.0 * H, 8.0 * H, 0.5 * H]])
    pos = np.concatenate([base, low])
    vel = np.zeros((5, 3))
    cfg, p, v = solid_oracle(pkg, orc, pos, vel, [pkg.TYPESOLID] * 5)
    tmin = float(F32(cfg.dx + 0.3 * cfg.samplespace))
    push = (tmin - float(F32(0.5 * H))) * float(F32(cfg.solid_bounce)) * 5
    Tz = push / 5
    assert np.isclose(p[:, 2].mean() - np.array(pos, F32)[:, 2].astype(float).mean(), cfg.dt * Tz, rtol=1e-4)
    cfg2, p2, v2 = solid_oracle(pkg, orc, pos, np.ones((5, 3)), [pkg.TYPESOLID] * 4 + [pkg.TYPEFLUID], rigid_lock_y=1)
    assert np.allclose(p2[:4, 1], np.array(pos, F32)[:4, 1], rtol=0, atol=1e-7)  # no y translation
    n = 16
    cfg3 = cfg_of(pkg, n, fix_solids=1)
    o = orc.Oracle(cfg3, use_double=True)
    o.set_particles(5, pos=np.array(pos, F32), vel=np.ones((5, 3), F32), flag=np.array([pkg.TYPESOLID] * 5, np.int8))
    o.solid_motion()
    q = o.get_particles()
    assert np.all(q["vel"] == 0) and np.array_equal(q["pos"], np.array(pos, F32).astype(float))


# --------------------------------------------------------------- freezing-scene collision
def test_collide_freezing_closed_forms(pkg, orc):
    """CollisionWithSolid_Freezing next to a flat solid slab (cells k <= 2):
    the solid distance grows along +z only, so the normal is (0, 0, 1).  Two
    fluid particles 0.2 h apart at z = 3.3 h (phisolid <= 1) overlap by
    0.5 h - 0.2 h: each rises by 0.3 h and loses its normal velocity; a
    particle whose nearest neighbour is 0.6 h away (adhesion) loses
    0.1 |v| along the normal; one far from the solid only moves by v dt; a
    particle past the wall margin bounces with -0.5 v; then x += v dt."""
    n = 12
    cfg = cfg_of(pkg, n, integrator=pkg.ADVECT_FREEZING, fix_solids=1)
    sol = np.array([[i + 0.5, j + 0.5, k + 0.5] for i in range(n) for j in range(n) for k in range(3)]) * H
    v = np.array([0.1, -0.2, 0.3])
    a = np.array([5.3, 5.5, 3.3]) * H
    b = a + np.array([0.2, 0, 0]) * H
    c0 = np.array([8.3, 8.5, 3.3]) * H
    c1 = c0 + np.array([0.6, 0, 0]) * H
    far = np.array([5.5, 5.5, 8.5]) * H
    wall = np.array([5.5, 0.9 * 1.0, 6.5]) * H  # y below gmin + h + 0.3 ss
    fl_pts = np.array([a, b, c0, c1, far, wall])
    pos = np.concatenate([sol, fl_pts]).astype(F32)
    vel = np.concatenate([np.zeros_like(sol), np.tile(v, (6, 1))]).astype(F32)
    flag = np.array([pkg.TYPESOLID] * len(sol) + [pkg.TYPEFLUID] * 6, np.int8)
    o = orc.Oracle(cfg, use_double=True)
    o.set_particles(len(pos), pos=pos, vel=vel, flag=flag)
    o.hash_sort()
    order = o.get_field("GRID_INDEX")
    o.advect()
    got = o.get_particles()
    P = np.empty((len(pos), 3)); P[order] = got["pos"]
    V = np.empty((len(pos), 3)); V[order] = got["vel"]
    q = len(sol)
    x0 = pos.astype(float)
    v0 = np.float32(v).astype(float)
    dt = float(F32(cfg.dt))
    for s in (q, q + 1):  # collision: pushed up by the overlap, normal velocity removed
        vn = np.array([v0[0], v0[1], 0.0])
        want = x0[s] + np.array([0, 0, 0.3 * H]) + vn * dt
        assert np.allclose(V[s], vn, atol=1e-12) and np.allclose(P[s], want, atol=1e-9)
    for s in (q + 2, q + 3):  # adhesion
        vn = v0 - np.array([0, 0, 0.1 * np.linalg.norm(v0)])
        assert np.allclose(V[s], vn, atol=1e-7) and np.allclose(P[s], x0[s] + vn * dt, atol=1e-9)
    assert np.allclose(V[q + 4], v0) and np.allclose(P[q + 4], x0[q + 4] + v0 * dt, atol=1e-12)
    lo = float(F32(cfg.dx + 0.3 * cfg.samplespace))
    vw = v0 * np.array([1, -0.5, 1])
    assert np.allclose(V[q + 5], vw) and np.isclose(P[q + 5][1], lo + vw[1] * dt, atol=1e-9)
    assert np.array_equal(P[:q], x0[:q])  # solids stay
