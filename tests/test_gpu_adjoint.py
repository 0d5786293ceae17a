"""Adjoint entry points (mpf_*_adjoint) against the forward stages.

Each adjoint is the transposed Jacobian of one forward stage at the state of
that stage's call, with its discrete decisions held fixed.

* Linear maps and the linear parts of the others (projection, extrapolation,
  P2G in the particle velocities, G2P in velocities and grid fields) are
  checked by the transposition identity <g, J d> = <J^T g, d> with random d
  and g, J applied by the GPU forward itself.  The identity part of a stage
  is subtracted on both sides (<g, (J - I) d> = <(J^T - I) g, d>) so a wrong
  non-identity term cannot hide behind it.
* The nonlinear dependence on particle positions (P2G weights and their
  normalisation, G2P interpolation, advection clamp) is checked against
  central finite differences of the FP64 oracle, started from the GPU's own
  state (sorted particles, cell lists) with every discrete decision held,
  at a step of 1e-7 cell sizes so that no particle crosses a kink of the
  piecewise-smooth maps.

Bar: relative error <= 1e-4 (the adjoint-vs-FD bar), tighter where the
forward's own arithmetic allows it (stated per test).
"""
import numpy as np
import pytest

from helpers import load_scene, say

pytestmark = pytest.mark.gpu

FACE = ["WATER_UX", "WATER_UY", "WATER_UZ", "AIR_UX", "AIR_UY", "AIR_UZ"]
FACE_OLD = [f + "_OLD" for f in FACE]
FD_EPS = 1e-7  # in cell sizes


def _make(pkg, scenes, name, n, **over):
    s = scenes.make(name, n=n)
    s["cfg"].adjoint_enabled = 1
    for k, v in over.items():
        setattr(s["cfg"], k, v)
    sim = pkg.Simulator(s["cfg"])
    load_scene(sim, s)
    return s, sim


def _stir(sim, seed=13, scale=0.2):
    """Random particle velocities (keeps order and positions)."""
    p = sim.get_particles()
    rng = np.random.default_rng(seed)
    vel = (p["vel"] + rng.standard_normal(p["vel"].shape) * scale).astype(np.float32)
    sim.set_particles(len(p["flag"]), **dict(p, vel=vel))


def _to_p2g(pkg, scenes, name, n, **over):
    s, sim = _make(pkg, scenes, name, n, **over)
    sim.hash_sort()
    _stir(sim)
    sim.hash_sort()
    sim.apply_forces()
    sim.levelset_update()
    return s, sim


def _rel(a, b):
    return abs(a - b) / max(abs(a), abs(b), 1e-300)


def _dot(xs, ys):
    return float(sum(np.dot(np.asarray(x, np.float64).ravel(), np.asarray(y, np.float64).ravel())
                     for x, y in zip(xs, ys)))


def _rand_faces(sim, rng, dtype=np.float64):
    return [rng.standard_normal(n).astype(dtype) for n in sim.face_sizes()]


PROJ = [
    ("diff_slab", 24, dict(precision=1, tol=1e-12), 1e-6),
    ("diff_slab", 24, dict(precision=2, tol=1e-10), 1e-5),
    ("diff_slab", 24, dict(precision=0, tol=1e-6), 1e-4),
    ("melt_freeze", 24, dict(precision=1, tol=1e-12), 1e-6),
    ("bubble_rise", 16, dict(precision=1, tol=1e-12, op=0), 1e-6),
]


@pytest.mark.parametrize("name,n,over,tol", PROJ)
def test_project_adjoint(pkg, scenes, name, n, over, tol):
    say("project_adjoint", name, n, over)
    s, sim = _to_p2g(pkg, scenes, name, n, **over)
    sim.p2g()
    sim.extrapolate_bc(True)
    rng = np.random.default_rng(5)
    d = _rand_faces(sim, rng, np.float32)
    for nm, a in zip(FACE, d):
        sim.set_field(nm, a)
    sim.project()
    out = [sim.get_field(nm) for nm in FACE]
    g = _rand_faces(sim, rng)
    gi, info = sim.project_adjoint(g)
    say("  adjoint solve", info)
    lhs = _dot(g, [o.astype(np.float64) - a for o, a in zip(out, d)])
    rhs = _dot([a - b for a, b in zip(gi, g)], d)
    e = _rel(lhs, rhs)
    say("  <g,(P-I)d> = %.12e  <(P^T-I)g,d> = %.12e  rel %.2e" % (lhs, rhs, e))
    assert info["converged"]
    assert e <= tol


def test_extrapolate_adjoint(pkg, scenes):
    say("extrapolate_adjoint")
    s, sim = _to_p2g(pkg, scenes, "diff_slab", 24)
    sim.p2g()
    rng = np.random.default_rng(6)
    for trial in range(2):
        d = _rand_faces(sim, rng, np.float32)
        for nm, a in zip(FACE, d):
            sim.set_field(nm, a)
        sim.extrapolate_bc(True)
        out = [sim.get_field(nm) for nm in FACE]
        g = _rand_faces(sim, rng)
        gi = sim.extrapolate_adjoint(g)
        lhs = _dot(g, [o.astype(np.float64) - a for o, a in zip(out, d)])
        rhs = _dot([a - b for a, b in zip(gi, g)], d)
        e = _rel(lhs, rhs)
        say("  trial %d: %.12e vs %.12e rel %.2e" % (trial, lhs, rhs, e))
        assert e <= 1e-6


def test_p2g_adjoint_velocity(pkg, scenes):
    say("p2g_adjoint velocity")
    s, sim = _to_p2g(pkg, scenes, "diff_slab", 24)
    p = sim.get_particles()
    rng = np.random.default_rng(7)
    dv = rng.standard_normal(p["vel"].shape).astype(np.float32)
    sim.set_particles(len(p["flag"]), **dict(p, vel=dv))
    sim.hash_sort()  # same positions: identity permutation, same cell lists
    sim.p2g()
    out = [sim.get_field(nm) for nm in FACE]
    g = _rand_faces(sim, rng)
    gv, gx = sim.p2g_adjoint(g)
    lhs = _dot(g, out)
    rhs = _dot([gv], [dv])
    e = _rel(lhs, rhs)
    say("  %.12e vs %.12e rel %.2e" % (lhs, rhs, e))
    assert e <= 1e-5


def _oracle_from(sim, orc, cfg, fields):
    o = orc.Oracle(cfg, use_double=True)
    p = sim.get_particles()
    o.set_particles(len(p["flag"]), **p)
    for nm in fields:
        o.set_field(nm, sim.get_field(nm))
    return o, p


def test_p2g_adjoint_position_fd(pkg, scenes, orc):
    say("p2g_adjoint position FD")
    s, sim = _to_p2g(pkg, scenes, "diff_slab", 24)
    o, p = _oracle_from(sim, orc, s["cfg"], ["CELL_START", "CELL_END", "MARK"])
    rng = np.random.default_rng(8)
    g = _rand_faces(sim, rng)
    gv, gx = sim.p2g_adjoint(g)
    h = float(s["cfg"].dx)
    x0 = p["pos"].astype(np.float64)
    v0 = p["vel"].astype(np.float64)
    for trial in range(2):
        d = rng.standard_normal(x0.shape) * h
        f = []
        for sgn in (1.0, -1.0):
            o.set_kinematics64(pos=x0 + sgn * FD_EPS * d, vel=v0)
            o.p2g()
            f.append(_dot(g, [o.get_field(nm) for nm in FACE]))
        fd = (f[0] - f[1]) / (2 * FD_EPS)
        ad = _dot([gx], [d])
        e = _rel(fd, ad)
        say("  trial %d: FD %.10e adjoint %.10e rel %.2e" % (trial, fd, ad, e))
        assert e <= 1e-4


@pytest.mark.parametrize("alpha", [1.0, 0.95])
def test_g2p_adjoint_linear(pkg, scenes, alpha):
    say("g2p_adjoint linear", alpha)
    s, sim = _to_p2g(pkg, scenes, "diff_slab", 24, flip_alpha=alpha)
    p = sim.get_particles()
    rng = np.random.default_rng(9)
    dv = rng.standard_normal(p["vel"].shape).astype(np.float32)
    du = _rand_faces(sim, rng, np.float32)
    duo = _rand_faces(sim, rng, np.float32)
    sim.set_particles(len(p["flag"]), **dict(p, vel=dv))
    for nm, a in zip(FACE, du):
        sim.set_field(nm, a)
    for nm, a in zip(FACE_OLD, duo):
        sim.set_field(nm, a)
    sim.g2p()
    vout = sim.get_particles()["vel"]
    g = rng.standard_normal(vout.shape)
    r = sim.g2p_adjoint(g)
    lhs = _dot([g], [vout])
    rhs = _dot([r["vel"]] + r["u"] + r["u_old"], [dv] + du + duo)
    e = _rel(lhs, rhs)
    say("  %.12e vs %.12e rel %.2e" % (lhs, rhs, e))
    assert e <= 1e-5


@pytest.mark.parametrize("alpha", [1.0, 0.95])
def test_g2p_adjoint_position_fd(pkg, scenes, orc, alpha):
    say("g2p_adjoint position FD", alpha)
    s, sim = _to_p2g(pkg, scenes, "diff_slab", 24, flip_alpha=alpha, precision=1, tol=1e-10)
    sim.p2g()
    sim.extrapolate_bc(True)
    sim.project()
    sim.extrapolate_bc(False)
    old = {nm: sim.get_field(nm) for nm in FACE_OLD}
    o, p = _oracle_from(sim, orc, s["cfg"], FACE + FACE_OLD)
    sim.g2p()
    rng = np.random.default_rng(10)
    g = rng.standard_normal(p["vel"].shape)
    r = sim.g2p_adjoint(g)
    h = float(s["cfg"].dx)
    x0 = p["pos"].astype(np.float64)
    v0 = p["vel"].astype(np.float64)
    for trial in range(2):
        d = rng.standard_normal(x0.shape) * h
        f = []
        for sgn in (1.0, -1.0):
            for nm in FACE_OLD:
                o.set_field(nm, old[nm])
            o.set_kinematics64(pos=x0 + sgn * FD_EPS * d, vel=v0)
            o.g2p()
            f.append(_dot([g], [o.get_particles()["vel"]]))
        fd = (f[0] - f[1]) / (2 * FD_EPS)
        ad = _dot([r["pos"]], [d])
        e = _rel(fd, ad)
        say("  trial %d: FD %.10e adjoint %.10e rel %.2e" % (trial, fd, ad, e))
        assert e <= 1e-4


def test_advect_adjoint_fd(pkg, scenes, orc):
    say("advect_adjoint FD")
    s, sim = _to_p2g(pkg, scenes, "diff_slab", 24, precision=1, tol=1e-10)
    sim.p2g()
    sim.extrapolate_bc(True)
    sim.project()
    sim.extrapolate_bc(False)
    sim.g2p()
    o, p = _oracle_from(sim, orc, s["cfg"], ["CELL_START", "CELL_END", "MARK"])
    rng = np.random.default_rng(11)
    gxo = rng.standard_normal(p["pos"].shape)
    gvo = rng.standard_normal(p["vel"].shape)
    gx, gv, band = sim.advect_adjoint(gxo, gvo)
    say("  particles in the solid band:", band)
    assert band == 0
    h = float(s["cfg"].dx)
    x0 = p["pos"].astype(np.float64)
    v0 = p["vel"].astype(np.float64)
    for trial in range(2):
        dx = rng.standard_normal(x0.shape) * h
        dv = rng.standard_normal(v0.shape) * h
        f = []
        for sgn in (1.0, -1.0):
            o.set_kinematics64(pos=x0 + sgn * FD_EPS * dx, vel=v0 + sgn * FD_EPS * dv)
            o.advect()
            q = o.get_particles()
            f.append(_dot([gxo, gvo], [q["pos"], q["vel"]]))
        fd = (f[0] - f[1]) / (2 * FD_EPS)
        ad = _dot([gx, gv], [dx, dv])
        e = _rel(fd, ad)
        say("  trial %d: FD %.10e adjoint %.10e rel %.2e" % (trial, fd, ad, e))
        assert e <= 1e-4
