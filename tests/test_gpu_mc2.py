"""GPU parity of the particle-source marching cubes (preMC + runMC_*).

mpf_mc_hash_sort against the oracle's 2N sort: keys, permutation, 2N cell
lists and the reordered particles bit-exact.  mpf_marching_cubes on each
particle source (liquid and gas, gas, solid): the node field bit-exact with
the float oracle (its expressions are restated operation by operation, in
the reference's summation order), case indices, triangle indices and vertex
positions bit-exact.  A full mpf_step with a particle source ends with the
particles in 2N order and the mesh the oracle builds from them.
"""
import numpy as np
import pytest

from helpers import load_scene, say

pytestmark = pytest.mark.gpu

SCENES = [("bubble_rise", 16), ("boiling", 16), ("melt_freeze", 24)]


def bitexact(a, b, what):
    a = np.asarray(a)
    b = np.asarray(b).astype(a.dtype)
    bad = np.flatnonzero(a != b)
    assert bad.size == 0, "%s: %d mismatches, first at %d: gpu=%r oracle=%r" % (
        what, bad.size, bad[0], a[bad[0]], b[bad[0]])


def make(pkg, scenes, name, n, src):
    s = scenes.make(name, n=n)
    s["cfg"].mc_enabled = 1
    s["cfg"].mc_source = src
    sim = pkg.Simulator(s["cfg"])
    load_scene(sim, s)
    return s, sim


@pytest.mark.parametrize("name,n", SCENES)
def test_mc_hash_sort_2n(pkg, scenes, orc, name, n):
    say("mc_hash_sort", name, n)
    s, sim = make(pkg, scenes, name, n, pkg.MC_LIQUID_GAS)
    o = orc.Oracle(s["cfg"])
    load_scene(o, s)
    sim.mc_hash_sort()
    o.mc_hash_sort()
    for f in ("GRID_HASH", "GRID_INDEX"):
        bitexact(sim.get_field(f), o.get_field(f), f)
    gs, ge = sim.get_mc_cells()
    os_, oe = o.get_mc_cells()
    bitexact(gs, os_, "2N cell start")
    ne = gs != pkg.CELL_UNDEF
    bitexact(ge[ne], oe[ne], "2N cell end")
    assert int((ge[ne] - gs[ne]).astype(np.int64).sum()) == s["n_particles"]
    pg, po = sim.get_particles(), o.get_particles()
    for k in pg:
        bitexact(pg[k], po[k].astype(pg[k].dtype), "particle " + k)


@pytest.mark.parametrize("src", [1, 2, 3])
@pytest.mark.parametrize("name,n", SCENES)
def test_mc_particle_source(pkg, scenes, orc, name, n, src):
    say("mc particle source", src, name, n)
    s, sim = make(pkg, scenes, name, n, src)
    o = orc.Oracle(s["cfg"])
    load_scene(o, s)
    sim.mc_hash_sort()
    o.mc_hash_sort()
    gi = sim.marching_cubes()
    oi = o.marching_cubes()
    say("  gpu", gi, "oracle", oi)
    bitexact(sim.get_mc_field(), o.get_mc_field().astype(np.float32), "node field")
    bitexact(sim.get_mc_cases(), o.get_mc_cases(), "case index")
    assert gi == oi
    vg, tg = sim.get_mesh(gi["verts"], gi["tris"])
    vo, to = o.get_mesh(oi["verts"], oi["tris"])
    bitexact(tg, to, "triangle indices")
    bitexact(vg, vo.astype(np.float32), "vertices")
    if src == 1:
        assert gi["tris"] > 0


def test_marching_cubes_needs_2n_order(pkg, scenes):
    """A particle source refuses to mesh until the particles are in 2N order,
    and again after anything reorders them."""
    say("mc order guard")
    s, sim = make(pkg, scenes, "bubble_rise", 16, pkg.MC_LIQUID_GAS)
    with pytest.raises(RuntimeError):
        sim.marching_cubes()
    sim.mc_hash_sort()
    sim.marching_cubes()
    sim.hash_sort()
    with pytest.raises(RuntimeError):
        sim.marching_cubes()


@pytest.mark.parametrize("name,n", [("bubble_rise", 16), ("melt_freeze", 24)])
def test_step_with_particle_source(pkg, scenes, orc, name, n):
    say("step with particle mc", name, n)
    s, sim = make(pkg, scenes, name, n, pkg.MC_LIQUID_GAS)
    info = sim.step(0)
    keys = sim.get_field("GRID_HASH")
    assert keys.size == sim.particle_count() and np.all(np.diff(keys.astype(np.int64)) >= 0)
    p = sim.get_particles()
    o = orc.Oracle(s["cfg"])
    o.set_particles(len(p["flag"]), **p)
    o.mc_hash_sort()
    bitexact(o.get_field("GRID_INDEX"), np.arange(len(p["flag"])), "2N order after the step")
    oi = o.marching_cubes()
    assert info["mc"] == oi
    bitexact(sim.get_mc_cases(), o.get_mc_cases(), "case index")


@pytest.mark.parametrize("src", [0, 1])
def test_mesh_smoothing_and_normals(pkg, scenes, orc, src):
    """smoothMesh (5 iterations) and calNormals on the GPU against the float
    oracle: vertices and normals bit-exact (each vertex sums its triangles in
    the order of the triangle list on both sides)."""
    say("mesh smoothing and normals", src)
    s, sim = make(pkg, scenes, "bubble_rise", 16, src)
    assert s["cfg"].mc_smooth_iters == 5
    o = orc.Oracle(s["cfg"])
    load_scene(o, s)
    if src == 0:
        sim.hash_sort()
        sim.levelset_update()
        o.set_field("LSFLUID", sim.get_field("LSFLUID"))
    else:
        sim.mc_hash_sort()
        o.mc_hash_sort()
    gi = sim.marching_cubes()
    oi = o.marching_cubes()
    assert gi == oi and gi["tris"] > 0
    vg, tg = sim.get_mesh(gi["verts"], gi["tris"])
    vo, to = o.get_mesh(oi["verts"], oi["tris"])
    bitexact(tg, to, "triangles")
    bitexact(vg, vo.astype(np.float32), "smoothed vertices")
    bitexact(sim.get_mesh_normals(gi["verts"]), o.get_mesh_normals(oi["verts"]).astype(np.float32), "normals")


def test_mesh_files(pkg, scenes, tmp_path):
    """outputPovRaywater / outputOBJwater / outputBOBJwater layouts, read
    back: counts, vertices and normals (%f text: 1e-6; bobj: bit for bit),
    face indices (OBJ from 1)."""
    import gzip
    import re
    say("mesh files")
    s, sim = make(pkg, scenes, "bubble_rise", 16, 0)
    sim.hash_sort()
    sim.levelset_update()
    info = sim.marching_cubes()
    nv, nt = info["verts"], info["tris"]
    v, t = sim.get_mesh(nv, nt)
    nrm = sim.get_mesh_normals(nv)
    sim.write_mesh(tmp_path / "w.pov", pkg.MESH_POVRAY, "water")
    sim.write_mesh(tmp_path / "w.obj", pkg.MESH_OBJ)
    sim.write_mesh(tmp_path / "w.bobj.gz", pkg.MESH_BOBJ_GZ)
    txt = open(tmp_path / "w.obj").read().splitlines()
    V = np.array([[float(x) for x in l.split()[1:]] for l in txt if l.startswith("v ")])
    N = np.array([[float(x) for x in l.split()[1:]] for l in txt if l.startswith("vn ")])
    F = np.array([[int(x) for x in l.split()[1:]] for l in txt if l.startswith("f ")])
    assert V.shape == (nv, 3) and N.shape == (nv, 3) and F.shape == (nt, 3)
    assert np.max(np.abs(V - v)) <= 1e-6 and np.max(np.abs(N - nrm)) <= 1e-6
    assert np.array_equal(F - 1, t.astype(np.int64))
    pov = open(tmp_path / "w.pov").read()
    assert pov.startswith("#declare watermesh=mesh2{") and "material{water_material}" in pov
    blocks = re.findall(r"(vertex_vectors|normal_vectors|face_indices)\{\n(\d+),\n(.*?)\}", pov, re.S)
    assert [b[0] for b in blocks] == ["vertex_vectors", "normal_vectors", "face_indices"]
    assert [int(b[1]) for b in blocks] == [nv, nv, nt]
    PV = np.array([[float(x) for x in m] for m in re.findall(r"<\s*([-\d.e]+),\s*([-\d.e]+),\s*([-\d.e]+)\s*>", blocks[0][2])])
    PF = np.array([[int(x) for x in m] for m in re.findall(r"<(\d+), (\d+), (\d+)>", blocks[2][2])])
    assert np.max(np.abs(PV - v)) <= 1e-6 and np.array_equal(PF, t.astype(np.int64))
    raw = gzip.open(tmp_path / "w.bobj.gz").read()
    n1 = np.frombuffer(raw[:4], np.int32)[0]
    BV = np.frombuffer(raw[4:4 + 12 * n1], np.float32).reshape(-1, 3)
    o = 4 + 12 * n1
    n2 = np.frombuffer(raw[o:o + 4], np.int32)[0]
    BN = np.frombuffer(raw[o + 4:o + 4 + 12 * n2], np.float32).reshape(-1, 3)
    o += 4 + 12 * n2
    ni = np.frombuffer(raw[o:o + 4], np.int32)[0]
    BI = np.frombuffer(raw[o + 4:o + 4 + 4 * ni], np.int32).reshape(-1, 3)
    assert n1 == nv and n2 == nv and ni == 3 * nt and len(raw) == o + 4 + 4 * ni
    assert np.array_equal(BV, v) and np.array_equal(BN, nrm) and np.array_equal(BI, t.astype(np.int32))
