"""Parity at the benchmark's full size: the melt/freeze scene at 256^3
(89.3M particles, 16.8M cells), where the GPU path runs thousands of blocks
per kernel, multi-run PCG tiles and the XCD block order.

* hash and stable sort: keys, permutation and cell ranges bit-exact against
  the oracle's counting sort, keys non-decreasing, the permutation a
  permutation, equal keys in index order, ranges partitioning the particles;
  sorting the sorted particles again gives the identity (the full sort path,
  not the library's shortcut);
* one projection (mixed form to 1e-10) against the FP64 oracle (1e-12):
  pressure and liquid face velocities within 1e-5 relative L2;
* one full mpf_step: every solve converged on its true residual, particles
  left in cell order, and the step's marching-cubes mesh (case indices,
  triangle indices, smoothed vertices, normals) bit-exact against the
  oracle's meshing of the GPU's own level set.
"""
import numpy as np
import pytest

from helpers import load_scene, rel_l2, say

pytestmark = pytest.mark.gpu

N = 256
FACE_W = ["WATER_UX", "WATER_UY", "WATER_UZ"]


def bitexact(a, b, what):
    a = np.asarray(a)
    b = np.asarray(b).astype(a.dtype)
    bad = np.flatnonzero(a != b)
    assert bad.size == 0, "%s: %d mismatches, first at %d: gpu=%r oracle=%r" % (
        what, bad.size, bad[0], a[bad[0]], b[bad[0]])


@pytest.fixture(scope="module")
def scene(scenes):
    say("melt_freeze %d^3: generating the scene" % N)
    s = scenes.make("melt_freeze", n=N)
    say("  %d particles" % s["n_particles"])
    return s


def test_fullsize_hash_sort(pkg, orc, scene):
    say("full-size hash sort")
    sim = pkg.Simulator(scene["cfg"])
    load_scene(sim, scene)
    sim.hash_sort()
    keys = sim.get_field("GRID_HASH")
    perm = sim.get_field("GRID_INDEX")
    cs, ce = sim.get_field("CELL_START"), sim.get_field("CELL_END")
    P = scene["n_particles"]
    G = N ** 3
    assert keys.size == P and perm.size == P
    assert np.all(keys[1:] >= keys[:-1])
    assert np.all(np.bincount(perm, minlength=P) == 1)
    same = keys[1:] == keys[:-1]
    assert np.all(perm[1:][same] > perm[:-1][same])
    counts = np.bincount(keys, minlength=G)
    occ = counts > 0
    assert np.all(cs[~occ] == pkg.CELL_UNDEF)
    assert np.array_equal(ce[occ].astype(np.int64) - cs[occ].astype(np.int64), counts[occ])
    assert np.array_equal(cs[occ].astype(np.int64), np.concatenate([[0], np.cumsum(counts)[:-1]])[occ])
    say("  invariants hold; oracle sort")
    o = orc.Oracle(scene["cfg"])
    load_scene(o, scene)
    o.hash_sort()
    bitexact(keys, o.get_field("GRID_HASH"), "GRID_HASH")
    bitexact(perm, o.get_field("GRID_INDEX"), "GRID_INDEX")
    bitexact(cs, o.get_field("CELL_START"), "CELL_START")
    bitexact(ce[occ], o.get_field("CELL_END")[occ], "CELL_END")
    o.close()
    say("  bit-exact; sorting the sorted particles again")
    p = sim.get_particles()
    sim.set_particles(P, **p)
    del p
    sim.hash_sort()
    bitexact(sim.get_field("GRID_INDEX"), np.arange(P, dtype=np.uint32), "second sort is the identity")
    bitexact(sim.get_field("GRID_HASH"), keys, "second sort keys")
    sim.close()


def test_fullsize_projection_vs_fp64(pkg, orc, scene):
    say("full-size projection, mixed form against the FP64 oracle")
    cfg = pkg.Config.from_buffer_copy(scene["cfg"])
    cfg.tol = 1e-10
    cfg.max_iters = 20000
    sim = pkg.Simulator(cfg)
    load_scene(sim, scene)
    sim.hash_sort()
    sim.apply_forces()
    sim.levelset_update()
    sim.p2g()
    sim.extrapolate_bc(True)
    tight = pkg.Config.from_buffer_copy(cfg)
    tight.tol = 1e-12
    o64 = orc.Oracle(tight, use_double=True)
    for f in ["MARK", "LSMERGE"] + FACE_W + ["AIR_UX", "AIR_UY", "AIR_UZ"]:
        o64.set_field(f, sim.get_field(f))
    gi = sim.project()
    say("  gpu", gi)
    oi = o64.project()
    say("  fp64 oracle", oi)
    assert gi["converged"] == 1 and gi["rel_residual"] <= 1e-9
    for f in ["PRESSURE"] + FACE_W:
        e = rel_l2(sim.get_field(f), o64.get_field(f))
        say("  %s rel L2 vs FP64 = %.2e" % (f, e))
        assert e <= 1e-5, f
    o64.close()
    sim.close()


def test_fullsize_step_and_mesh(pkg, orc, scene):
    say("full-size step and its mesh")
    sim = pkg.Simulator(scene["cfg"])
    load_scene(sim, scene)
    info = sim.step(0)
    say("  step", info["project"], info["mc"])
    assert info["project"]["converged"] == 1 and info["project"]["rel_residual"] <= scene["cfg"].tol
    assert info["particles"] == sim.particle_count()
    keys = sim.get_field("GRID_HASH")
    assert keys.size == sim.particle_count() and np.all(keys[1:] >= keys[:-1])
    del keys
    o = orc.Oracle(scene["cfg"])
    o.set_field("LSFLUID", sim.get_field("LSFLUID"))
    oi = o.marching_cubes()
    say("  oracle mesh", oi)
    assert info["mc"] == oi and oi["tris"] > 0
    bitexact(sim.get_mc_cases(), o.get_mc_cases(), "case index")
    vg, tg = sim.get_mesh(oi["verts"], oi["tris"])
    vo, to = o.get_mesh(oi["verts"], oi["tris"])
    bitexact(tg, to, "triangle indices")
    bitexact(vg, vo.astype(np.float32), "smoothed vertices")
    bitexact(sim.get_mesh_normals(oi["verts"]), o.get_mesh_normals(oi["verts"]).astype(np.float32), "normals")
    o.close()
    sim.close()
