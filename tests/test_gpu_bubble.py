"""GPU parity of the remaining bubblesim stages against the float oracle.

Each stage runs on the GPU and on the oracle from identical inputs (the
GPU's sorted particles, level sets, flags and face velocities; random face
velocities so the interpolations see non-trivial fields) and its outputs are
compared bit-exact: the per-element expressions are restated in the same
operation order, and the two reductions that the reference leaves to
atomics or races (drag onto the liquid faces, SPH neighbour velocities)
have one defined order shared by both sides.
"""
import numpy as np
import pytest

from helpers import load_scene, say

pytestmark = pytest.mark.gpu

SCENES = [("bubble_rise", 16), ("boiling", 16), ("melt_freeze", 24)]
AIR_SCENES = [("bubble_rise", 16), ("bubble_rise", 24)]  # scenes that start with air particles
FACES = ["WATER_UX", "WATER_UY", "WATER_UZ", "AIR_UX", "AIR_UY", "AIR_UZ"]
CELLS = ["LSMERGE", "LSAIR", "LSFLUID", "PHIGRAD_X", "PHIGRAD_Y", "PHIGRAD_Z", "MARK"]


def bitexact(a, b, what):
    a = np.asarray(a)
    b = np.asarray(b).astype(a.dtype)
    bad = np.flatnonzero(a.ravel() != b.ravel())
    assert bad.size == 0, "%s: %d mismatches, first at %d: gpu=%r oracle=%r" % (
        what, bad.size, bad[0], a.ravel()[bad[0]], b.ravel()[bad[0]])


def prepared(pkg, scenes, orc, name, n, seed=0, solo_every=0, **over):
    s = scenes.make(name, n=n)
    for k, v in over.items():
        setattr(s["cfg"], k, v)
    if solo_every:  # turn every k-th AIR particle into AIRSOLO
        fl = s["particles"]["flag"]
        air = np.flatnonzero(fl == pkg.TYPEAIR)
        fl[air[::solo_every]] = pkg.TYPEAIRSOLO
    sim = pkg.Simulator(s["cfg"])
    load_scene(sim, s)
    sim.hash_sort()
    sim.levelset_update()
    rng = np.random.default_rng(seed)
    for nm in FACES:
        m = sim.get_field(nm).size
        sim.set_field(nm, rng.normal(0, 0.2, m).astype(np.float32))
    o = orc.Oracle(s["cfg"])
    sync(sim, o)
    return s, sim, o


def sync(sim, o):
    p = sim.get_particles()
    o.set_particles(len(p["flag"]), **p)
    o.hash_sort()  # the particles are in cell order: the identity, same lists
    bitexact(o.get_field("GRID_INDEX"), np.arange(len(p["flag"])), "identity order")
    for nm in FACES + CELLS:
        o.set_field(nm, sim.get_field(nm))


def same_particles(sim, o, what):
    pg, po = sim.get_particles(), o.get_particles()
    for k in ("pos", "vel", "flag"):
        bitexact(pg[k], po[k].astype(pg[k].dtype), "%s: particle %s" % (what, k))


@pytest.mark.parametrize("name,n", SCENES)
def test_mark_solo_air(pkg, scenes, orc, name, n):
    say("mark_solo_air", name, n)
    s, sim, o = prepared(pkg, scenes, orc, name, n)
    f0 = sim.get_particles()["flag"]
    sim.mark_solo_air()
    o.mark_solo_air()
    same_particles(sim, o, "mark_solo_air")
    f1 = sim.get_particles()["flag"]
    say("  flags changed:", int(np.count_nonzero(f0 != f1)))


@pytest.mark.parametrize("name,n", SCENES)
def test_surface_tension(pkg, scenes, orc, name, n):
    say("surface_tension", name, n)
    s, sim, o = prepared(pkg, scenes, orc, name, n)
    v0 = sim.get_particles()["vel"]
    sim.surface_tension()
    o.surface_tension()
    bitexact(sim.get_field("SURFACE_TENSION"), o.get_field("SURFACE_TENSION").astype(np.float32), "sf")
    same_particles(sim, o, "surface_tension")
    say("  particles changed:", int(np.count_nonzero(np.any(sim.get_particles()["vel"] != v0, axis=1))))


@pytest.mark.parametrize("upd", [0, 1])
@pytest.mark.parametrize("name,n", AIR_SCENES)
def test_drag_force(pkg, scenes, orc, name, n, upd):
    say("drag_force", name, n, upd)
    s, sim, o = prepared(pkg, scenes, orc, name, n, solo_every=3, drag_particle_update=upd)
    nsolo = int(np.count_nonzero(sim.get_particles()["flag"] == pkg.TYPEAIRSOLO))
    say("  solo air particles:", nsolo)
    assert nsolo > 0
    sim.drag_force()
    o.drag_force()
    for nm in FACES[:3]:
        bitexact(sim.get_field(nm), o.get_field(nm).astype(np.float32), nm)
    same_particles(sim, o, "drag")


@pytest.mark.parametrize("name,n", AIR_SCENES)
def test_update_solo_air(pkg, scenes, orc, name, n):
    say("update_solo_air", name, n)
    s, sim, o = prepared(pkg, scenes, orc, name, n, solo_every=3)
    sim.update_solo_air()
    o.update_solo_air()
    same_particles(sim, o, "sph")


@pytest.mark.parametrize("name,n", SCENES)
def test_correct_positions(pkg, scenes, orc, name, n):
    say("correct_positions", name, n)
    s, sim, o = prepared(pkg, scenes, orc, name, n)
    sim.correct_positions()
    o.correct_positions()
    same_particles(sim, o, "correct_positions")
    with pytest.raises(RuntimeError):  # positions moved: the lists are stale
        sim.correct_positions()


@pytest.mark.parametrize("air_ls", [0, 1])
@pytest.mark.parametrize("name,n", SCENES)
def test_correct_bubble_positions(pkg, scenes, orc, name, n, air_ls):
    say("correct_bubble_positions", name, n, air_ls)
    s, sim, o = prepared(pkg, scenes, orc, name, n, correct_bubble_air_ls=air_ls)
    p0 = sim.get_particles()["pos"]
    sim.correct_bubble_positions()
    o.correct_bubble_positions()
    same_particles(sim, o, "correct_bubble_positions")
    say("  particles moved:", int(np.count_nonzero(np.any(sim.get_particles()["pos"] != p0, axis=1))))


@pytest.mark.parametrize("lock_y", [0, 1])
def test_solid_motion(pkg, scenes, orc, lock_y):
    """Rigid motion of the solid blocks (fix_solids = 0) with random solid
    velocities: the GPU's double sums are added in another order than the
    oracle's, so positions and velocities agree to float rounding (1e-6 of
    their scale) rather than bit for bit; other particles bit-exact."""
    say("solid_motion", lock_y)
    s = scenes.make("melt_freeze", n=24)
    cfg = s["cfg"]
    cfg.fix_solids = 0
    cfg.rigid_lock_y = lock_y
    fl = s["particles"]["flag"]
    rng = np.random.default_rng(4)
    v = s["particles"]["vel"].copy()
    v[fl == pkg.TYPESOLID] = rng.normal(0, 0.2, (int((fl == pkg.TYPESOLID).sum()), 3)).astype(np.float32)
    s["particles"]["vel"] = v
    sim = pkg.Simulator(cfg)
    load_scene(sim, s)
    o = orc.Oracle(cfg)
    load_scene(o, s)
    sim.hash_sort()
    o.hash_sort()
    sim.solid_motion()
    o.solid_motion()
    pg, po = sim.get_particles(), o.get_particles()
    sol = pg["flag"] == pkg.TYPESOLID
    assert sol.sum() > 100
    bitexact(pg["pos"][~sol], po["pos"][~sol].astype(np.float32), "non-solid pos")
    bitexact(pg["vel"][~sol], po["vel"][~sol].astype(np.float32), "non-solid vel")
    dp = np.abs(pg["pos"][sol] - po["pos"][sol]).max()
    dv = np.abs(pg["vel"][sol] - po["vel"][sol]).max()
    say("  max |dpos| %.2e  max |dvel| %.2e" % (dp, dv))
    assert dp <= 1e-6 and dv <= 1e-6
    with pytest.raises(RuntimeError):  # the solid moved: the cell lists are stale
        sim.mark_solo_air()


def test_step_with_moving_solid(pkg, scenes, orc):
    """Two steps of the melt/freeze scene with the solid blocks free to move:
    particle counts equal, solid positions within 1e-5 of the float oracle
    (float rounding of the rigid sums), all positions within 1e-4 rel L2."""
    say("step with moving solid")
    s = scenes.make("melt_freeze", n=24)
    cfg = s["cfg"]
    cfg.fix_solids = 0
    cfg.tol = 1e-9
    sim = pkg.Simulator(cfg)
    load_scene(sim, s)
    o = orc.Oracle(cfg)
    load_scene(o, s)
    for f in range(2):
        gi = sim.step(f)
        oi = o.step(f, allow_no_convergence=True)
        assert gi["particles"] == oi["particles"]
    pg, po = sim.get_particles(), o.get_particles()
    sol = (pg["flag"] == pkg.TYPESOLID) & (po["flag"] == pkg.TYPESOLID)
    d = np.abs(pg["pos"][sol] - po["pos"][sol]).max()
    from helpers import rel_l2
    e = rel_l2(pg["pos"], po["pos"])
    say("  solid max |dpos| %.2e, positions rel L2 %.2e" % (d, e))
    assert d <= 1e-5 and e <= 1e-4


def test_collide_freezing(pkg, scenes, orc):
    """CollisionWithSolid_Freezing (MPF_ADVECT_FREEZING) on the melt/freeze
    scene: positions and velocities bit-exact with the float oracle (the
    solid distance field is bit-exact and neighbour positions are the pass's
    input on both sides)."""
    say("collide freezing")
    s = scenes.make("melt_freeze", n=24)
    cfg = s["cfg"]
    cfg.integrator = pkg.ADVECT_FREEZING
    rng = np.random.default_rng(9)
    s["particles"]["vel"] = rng.normal(0, 0.3, s["particles"]["vel"].shape).astype(np.float32)
    sim = pkg.Simulator(cfg)
    load_scene(sim, s)
    sim.hash_sort()
    o = orc.Oracle(cfg)
    sync(sim, o)
    sim.advect()
    o.advect()
    same_particles(sim, o, "collide_freezing")
    moved = np.abs(sim.get_particles()["pos"] - o.get_particles()["pos"]).max()
    assert moved == 0
