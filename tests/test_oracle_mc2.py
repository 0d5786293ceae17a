"""Pins of the oracle's particle-source marching cubes (no GPU).

preMC re-hashes the particles on the 2N grid (calcHashD_MC, spray_k.cu:
1030-1048; hashAndSortParticles_MC, spray.cu:639-662) and runMC_smooth
builds a node field on the (2N+1)^3 lattice from those lists
(genWaterDensfield_liquidAndGas / _Gas / _GY, spray_k.cu:1528-1772) before
classifying it.  The pins: 2N keys of hand-placed points and the stable
order; the field of one particle (centre = the particle, phi = r - d inside
the support, -r outside, lattice faces fixed), the particle types each source
counts, the weighted centre of a random cloud against a brute force over all
particles with the published kernel form (1 - (d/R)^2)^3 (the reference's
(R^2 - d^2)^3 up to the factor R^6, which cancels), the GY falloff f =
1 - (1 - gamma)^3 on both sides of tlow, and the mesh numbering of
GetEdgeID (spray_k.cu:1920-1957) read from the reference source as data.
"""
import os
import re

import numpy as np
import pytest

H = 1.0 / 64
HH = H / 2  # lattice spacing of the particle sources
F32 = np.float32


def cfg_of(pkg, n, src, **kw):
    kw.setdefault("particle_capacity", 100000)
    return pkg.default_config(n, n, n, mc_enabled=1, mc_source=src, **kw)


def nodes(n):
    m = 2 * n + 1
    I, J, K = np.meshgrid(*[np.arange(m)] * 3, indexing="ij")
    return np.stack([I, J, K], -1).reshape(-1, 3), m


def face_mask(ijk, m):
    return np.any(ijk == 0, axis=1) | np.any(ijk == m - 1, axis=1)


# --------------------------------------------------------------- 2N hash
def test_mc_hash_known_answers_and_stable_order(pkg, orc):
    """Points at 2N cell centres hash to their cell; points on a 2N cell's
    low face go to that cell; components outside [0, 2n) clamp to 0; the
    order is stable by (key, index) and the lists cover every particle."""
    n = 8
    m = 2 * n
    cfg = cfg_of(pkg, n, pkg.MC_LIQUID_GAS)
    rng = np.random.default_rng(4)
    cells = rng.integers(0, m, (300, 3))
    pos = ((cells + 0.5) * HH).astype(F32)
    face = rng.integers(0, m, (40, 3))
    pos = np.concatenate([pos, (face * HH).astype(F32)])
    cells = np.concatenate([cells, face])
    out = np.array([[-0.01, 3.5 * HH, 2.5 * HH], [5.5 * HH, 1.2, 0.5 * HH], [2.5 * HH, 7.5 * HH, -1e-3]], F32)
    pos = np.concatenate([pos, out])
    cells = np.concatenate([cells, [[0, 3, 2], [5, 0, 0], [2, 7, 0]]])
    dup = pos[:20].copy()
    pos = np.concatenate([pos, dup])
    cells = np.concatenate([cells, cells[:20]])
    P = len(pos)
    key = (cells[:, 0] * m + cells[:, 1]) * m + cells[:, 2]
    o = orc.Oracle(cfg)
    o.set_particles(P, pos=pos)
    o.mc_hash_sort()
    order = np.lexsort((np.arange(P), key))
    assert np.array_equal(o.get_field("GRID_INDEX"), order)
    assert np.array_equal(o.get_field("GRID_HASH"), key[order])
    cs, ce = o.get_mc_cells()
    sk = key[order]
    for c in np.unique(key):
        assert cs[c] == np.searchsorted(sk, c, "left") and ce[c] == np.searchsorted(sk, c, "right")
    empty = np.setdiff1d(np.arange(m ** 3), key)
    assert np.all(cs[empty] == pkg.CELL_UNDEF)
    got = o.get_particles()["pos"].reshape(-1, 3)
    assert np.array_equal(got.astype(F32), pos[order])


# --------------------------------------------------------------- node fields
def one_particle_field(pkg, orc, src, x0, fl, **kw):
    n = 8
    cfg = cfg_of(pkg, n, src, **kw)
    o = orc.Oracle(cfg, use_double=True)
    o.set_particles(1, pos=np.array([x0], F32), flag=np.array([fl], np.int8))
    o.mc_hash_sort()
    o.marching_cubes()
    return o.get_mc_field(), n


@pytest.mark.parametrize("src,rfac", [(1, 0.5), (2, 0.8)])
def test_single_particle_field(pkg, orc, src, rfac):
    """One counted particle: the centre is the particle, so phi = r - |x - p|
    within the support |x - p| < R = 2 * (h/2) and -r beyond it, r = 0.5 h/2
    (liquid and gas) or 0.8 h/2 (gas); lattice faces are -1000."""
    x0 = np.array([8.1, 7.9, 8.05]) * HH
    fl = pkg.TYPEFLUID if src == 1 else pkg.TYPEAIR
    f, n = one_particle_field(pkg, orc, src, x0, fl)
    ijk, m = nodes(n)
    x = ijk * HH
    d = np.linalg.norm(x - np.float32(x0).astype(float), axis=1)
    r = float(np.float32(rfac)) * HH  # the reference's float literals
    exp = np.where(d < 2 * HH, r - d, -r)
    exp[face_mask(ijk, m)] = -1000.0
    assert np.max(np.abs(f - exp)) < 1e-12
    assert np.count_nonzero(f > 0) > 0  # the particle's own neighbourhood is inside


@pytest.mark.parametrize("src,counted", [
    (1, {0: True, 1: True, 4: True, 5: False, 2: False, 3: False}),
    (2, {0: False, 1: True, 4: True, 5: False, 2: False, 3: False}),
    (3, {0: False, 1: False, 4: False, 5: True, 2: False, 3: False}),
])
def test_source_particle_types(pkg, orc, src, counted):
    """sumcelldens_liquidAndGas counts FLUID, AIR and AIRSOLO; sumcelldens_Gas
    AIR and AIRSOLO (not in SCENE_INTERACTION); sumcelldens3 under runMC_solid
    only SOLID.  A particle of another type leaves the empty-field value."""
    x0 = np.array([8.3, 7.6, 8.9]) * HH
    for fl, yes in counted.items():
        f, n = one_particle_field(pkg, orc, src, x0, fl)
        ijk, m = nodes(n)
        inner = ~face_mask(ijk, m)
        rfac = float(np.float32({1: 0.5, 2: 0.8, 3: 0.75}[src]))
        empty = np.allclose(f[inner], -rfac * HH, rtol=0, atol=1e-15)
        assert empty != yes, (src, fl)
    # MPF_MC_GAS without AIRSOLO (the reference's SCENE_INTERACTION branch)
    f, n = one_particle_field(pkg, orc, 2, x0, pkg.TYPEAIRSOLO, mc_gas_airsolo=0)
    ijk, m = nodes(n)
    assert np.allclose(f[~face_mask(ijk, m)], -float(np.float32(0.8)) * HH, rtol=0, atol=1e-15)


def test_cloud_centre_bruteforce(pkg, orc):
    """Random FLUID/AIR/SOLID cloud, liquid-and-gas source: every interior
    node's phi against a brute force over ALL particles with the published
    kernel (1 - (d/R)^2)^3, R = h: only the counted types within R enter,
    whatever cells they sit in."""
    n = 8
    cfg = cfg_of(pkg, n, 1)
    rng = np.random.default_rng(11)
    P = 900
    pos = rng.uniform(0.5 * H, (n - 0.5) * H, (P, 3)).astype(F32)
    flag = rng.choice([0, 1, 4, 5], P).astype(np.int8)
    o = orc.Oracle(cfg, use_double=True)
    o.set_particles(P, pos=pos, flag=flag)
    o.mc_hash_sort()
    o.marching_cubes()
    f = o.get_mc_field()
    ijk, m = nodes(n)
    x = ijk * HH
    pp = pos.astype(float)
    use = np.isin(flag, [0, 1, 4])
    R = 2 * HH
    exp = np.empty(len(x))
    for q in range(len(x)):
        d = np.linalg.norm(pp - x[q], axis=1)
        sel = use & (d < R)
        if not sel.any():
            exp[q] = -0.5 * HH
            continue
        w = (1 - (d[sel] / R) ** 2) ** 3
        c = (pp[sel] * w[:, None]).sum(0) / w.sum()
        exp[q] = 0.5 * HH - np.linalg.norm(x[q] - c)
    fm = face_mask(ijk, m)
    exp[fm] = -1000.0
    assert np.max(np.abs(f - exp)) < 1e-11
    assert np.count_nonzero((f > 0) & ~fm) > 100 and np.count_nonzero((f < 0) & ~fm) > 100


def test_gy_solid_falloff(pkg, orc):
    """genWaterDensfield_GY with one SOLID particle (centre = particle,
    support R = h/2, kernel max((1 - (d/(4 h/2))^2)^3, 0) > 0 there):
    Ev = max(centre)/(4 h/2); below tlow = 0.49 phi = d - 0.75 h/2, above it
    phi = d - 0.75 h/2 (1 - (1 - gamma)^3), gamma = (0.51 - Ev)/0.02;
    nodes outside the support -0.75 h/2; lattice faces min(phi, -10)."""
    for x0 in (np.array([1.2, 1.1, 1.3]) * HH, np.array([8.3, 7.6, 8.9]) * HH):
        f, n = one_particle_field(pkg, orc, 3, x0, pkg.TYPESOLID)
        ijk, m = nodes(n)
        x = ijk * HH
        c = np.float32(x0).astype(float)
        d = np.linalg.norm(x - c, axis=1)
        Ev = c.max() / (4 * HH)
        fr = 1.0 if Ev < np.float32(0.49) else 1 - (1 - (np.float32(0.51) - Ev) / (np.float32(0.51) - np.float32(0.49))) ** 3
        r = np.float32(0.75) * HH
        exp = np.where(d < HH, d - r * fr, -r)
        fm = face_mask(ijk, m)
        exp[fm] = np.minimum(exp[fm], -10.0)
        inside = (d < HH)
        assert inside.any()
        assert np.allclose(f, exp, rtol=1e-12, atol=1e-12 * max(1.0, abs(fr) * HH))


# --------------------------------------------------------------- mesh numbering
REF_K = "/root/reference/Multiphase/Multiphase/spray_k.cu"


def ref_edge_table():
    """(di, dj, dk, slot) of the 12 cube edges from GetEdgeID's switch."""
    src = open(REF_K, encoding="latin-1").read()
    i = src.find("int GetEdgeID(")
    body = src[i:src.find("default:", i)]
    tab = {}
    for e, args, plus in re.findall(r"case (\d+):\s*return GetVertexID\(([^)]*)\)\s*(?:\+\s*(\d))?;", body):
        d = [1 if "+ 1" in a or "+1" in a else 0 for a in args.split(",")]
        tab[int(e)] = (d[0], d[1], d[2], int(plus) if plus else 0)
    assert sorted(tab) == list(range(12))
    return tab


@pytest.mark.skipif(not os.path.exists(REF_K), reason="reference source not mounted")
def test_mesh_numbering_follows_getedgeid(pkg, orc):
    """Vertex k of the mesh is the k-th active edge in the id order
    3 * node + slot of GetEdgeID, each triangle names the edges the
    reference's table gives for its voxel's case, and every vertex is the
    zero of the linear interpolant along its edge."""
    tab = ref_edge_table()
    n = 6
    cfg = cfg_of(pkg, n, 1, mc_smooth_iters=0)  # vertices on their edges
    rng = np.random.default_rng(2)
    P = 400
    pos = rng.uniform(1.0 * H, (n - 1.0) * H, (P, 3)).astype(F32)
    o = orc.Oracle(cfg, use_double=True)
    o.set_particles(P, pos=pos, flag=np.zeros(P, np.int8))
    o.mc_hash_sort()
    info = o.marching_cubes()
    assert info["tris"] > 50
    v, t = o.get_mesh(info["verts"], info["tris"])
    f = o.get_mc_field()
    m = 2 * n + 1
    F = f.reshape(m, m, m)
    cases = o.get_mc_cases().reshape(m - 1, m - 1, m - 1)
    _, tt, nv = orc.mc_tables()
    # active edges in id order
    ids = []
    for (i, j, k), ci in np.ndenumerate(cases):
        for q in range(nv[ci]):
            di, dj, dk, slot = tab[int(tt[ci][q])]
            ids.append(3 * (((i + di) * m + (j + dj)) * m + (k + dk)) + slot)
    uniq = np.unique(ids)
    assert len(uniq) == len(v)
    vid = {e: q for q, e in enumerate(uniq)}
    w = 0
    for (i, j, k), ci in np.ndenumerate(cases):
        for q in range(nv[ci]):
            di, dj, dk, slot = tab[int(tt[ci][q])]
            e = 3 * (((i + di) * m + (j + dj)) * m + (k + dk)) + slot
            assert t.ravel()[w] == vid[e]
            w += 1
    for e, q in vid.items():
        node, slot = divmod(int(e), 3)
        a = {0: 1, 1: 0, 2: 2}[slot]
        i0 = np.array(np.unravel_index(node, (m, m, m)))
        i1 = i0.copy()
        i1[a] += 1
        f0, f1 = F[tuple(i0)], F[tuple(i1)]
        s = (v[q][a] - i0[a] * HH) / HH
        assert 0 <= s <= 1
        assert abs(f0 + s * (f1 - f0)) < 1e-9 * max(abs(f0), abs(f1))
        other = [x for x in range(3) if x != a]
        assert np.allclose(v[q][other], i0[other] * HH, rtol=0, atol=1e-15)


def test_step_leaves_particles_in_2n_order(pkg, scenes, orc):
    """mpf_step order with a particle source: preMC runs after the phase
    change, so the step ends with the particles sorted by 2N key."""
    s = scenes.make("bubble_rise", n=16)
    cfg = s["cfg"]
    cfg.mc_enabled = 1
    cfg.mc_source = pkg.MC_LIQUID_GAS
    o = orc.Oracle(cfg)
    o.set_particles(s["n_particles"], **s["particles"])
    info = o.step(0)
    assert info["mc"]["tris"] > 0
    keys = o.get_field("GRID_HASH")
    assert keys.size == o.particle_count() and np.all(np.diff(keys.astype(np.int64)) >= 0)
    pos = o.get_particles()["pos"].reshape(-1, 3)
    k2 = np.floor(pos.astype(F32) / F32(HH)).astype(np.int64).clip(0, 31)
    assert np.array_equal((k2[:, 0] * 32 + k2[:, 1]) * 32 + k2[:, 2], keys)


# --------------------------------------------------------------- smoothing, normals, point clouds
def _mesh(pkg, orc, iters, sdf_fn, n=16):
    cfg = pkg.default_config(n, n, n, mc_enabled=1, particle_capacity=10, mc_smooth_iters=iters)
    o = orc.Oracle(cfg, use_double=True)
    c = (np.stack(np.meshgrid(*[np.arange(n)] * 3, indexing="ij"), -1).reshape(-1, 3) + 0.5) * H
    o.set_field("LSFLUID", sdf_fn(c))
    info = o.marching_cubes()
    v, t = o.get_mesh(info["verts"], info["tris"])
    return v, t, o.get_mesh_normals(info["verts"])


def test_mesh_smoothing_and_normals_against_numpy(pkg, orc):
    """smoothMesh: each iteration moves every vertex by 0.5 x the mean of the
    edge vectors to the other corners of its triangles (the mu pass is a
    no-op, its sums being cleared); calNormals: the normalised sum of
    cross(p2 - p1, p3 - p1) over a vertex's triangles.  Both recomputed here
    with numpy from the unsmoothed mesh of the same level set."""
    cen = np.array([8.1, 7.7, 8.3]) * H
    sdf = lambda c: np.linalg.norm(c - cen, axis=1) - 4.3 * H
    v0, t0, _ = _mesh(pkg, orc, 0, sdf)
    v2, t2, n2 = _mesh(pkg, orc, 2, sdf)
    assert np.array_equal(t0, t2)
    v = v0.copy()
    for _ in range(2):
        d = np.zeros_like(v)
        w = np.zeros(len(v))
        for a, b, c in ((0, 1, 2), (1, 0, 2), (2, 0, 1)):
            np.add.at(d, t0[:, a], v[t0[:, b]] - v[t0[:, a]])
            np.add.at(d, t0[:, a], v[t0[:, c]] - v[t0[:, a]])
            np.add.at(w, t0[:, a], 2)
        v = v + 0.5 * d / w[:, None]
    assert np.max(np.abs(v - v2)) < 1e-13
    fn = np.cross(v[t0[:, 1]] - v[t0[:, 0]], v[t0[:, 2]] - v[t0[:, 0]])
    nn = np.zeros_like(v)
    for a in range(3):
        np.add.at(nn, t0[:, a], fn)
    nn /= np.linalg.norm(nn, axis=1)[:, None]
    assert np.max(np.abs(nn - n2)) < 1e-12
    r0 = np.linalg.norm(v0 - cen, axis=1).mean()
    r2 = np.linalg.norm(v2 - cen, axis=1).mean()
    assert r2 < r0  # Laplacian smoothing shrinks a sphere
    assert np.all(((v2 - cen) * n2).sum(1) > 0)  # normals point out of the liquid


def test_mesh_normals_of_a_plane(pkg, orc):
    """A planar level set z = z0: every vertex lies on the plane, every
    normal is exactly (0, 0, 1) (out of the liquid below), and smoothing
    keeps z (x and y of the normals exactly 0, z 1 to rounding)."""
    z0 = 7.3 * H
    v, t, nrm = _mesh(pkg, orc, 5, lambda c: c[:, 2] - z0)
    assert len(t) > 100
    assert np.allclose(v[:, 2], z0, rtol=0, atol=1e-15)
    assert np.all(nrm[:, :2] == 0) and np.allclose(nrm[:, 2], 1.0, rtol=0, atol=1e-15)


def test_read_point_cloud(pkg, tmp_path):
    """readdata_solid's .asc layout: the count, then 'index x y z' lines;
    separators of one character after each number; a short file fails."""
    rng = np.random.default_rng(2)
    pts = rng.normal(0, 1, (7, 3))
    p = tmp_path / "cloud.asc"
    with open(p, "w") as f:
        f.write("7\n")
        for i, q in enumerate(pts):
            f.write("%d %.17g %.17g %.17g\n" % (i, q[0], q[1], q[2]))
    got = pkg.read_point_cloud(p)
    assert got.shape == (7, 3) and np.array_equal(got, pts)
    assert np.array_equal(pkg.read_point_cloud(p, cap=3), pts[:3])
    with open(p, "w") as f:
        f.write("9\n0 1 2 3\n")
    with pytest.raises(RuntimeError):
        pkg.read_point_cloud(p)
    with pytest.raises(RuntimeError):
        pkg.read_point_cloud(tmp_path / "missing.asc")
