"""Pins of the oracle functions that the whole-stage tests leave open (no GPU).

Each test fixes one reference rule by a closed form, a known answer or a
brute-force count that is derived independently of the oracle's own loop:
the extrapolation weights of sweepu_k_bubble on a linear distance field and
on its tie case, the wall rules of setWaterBoundaryU_k, the two-phase blend
of cptdivergence_bubble2 at a planar interface, the midpoint step and wall
clamp of advectparticle_RK2_bubble, the response of CollisionWithSolid_k at a
planar solid, heat conservation, the backward-Euler decay of a cosine mode,
the heater, the boundary copy and the grid-to-particle blend, nucleation on
hand-computed cells, the count-based flags of markfluid_dense, the level-set
gradient of computePhigra and the solid-cell marking.  A dropped term, a
wrong sign or a swapped index changes at least one asserted value.
"""
import numpy as np
import pytest

H = 1.0 / 64
F32 = np.float32


def small_cfg(pkg, n=8, **kw):
    kw.setdefault("particle_capacity", 200000)
    kw.setdefault("mc_enabled", 0)
    return pkg.default_config(n, n, n, **kw)


def face_dims(n):
    return {0: (n + 1, n, n), 1: (n, n + 1, n), 2: (n, n, n + 1)}


def face_grid(dims):
    return np.meshgrid(*[np.arange(d) for d in dims], indexing="ij")


def boxed_marks(n, inner):
    m = np.full((n, n, n), inner, np.int8)
    m[[0, -1], :, :] = 2
    m[:, [0, -1], :] = 2
    m[:, :, [0, -1]] = 2
    return m


def in_cell_order(o, arrays):
    """Particle arrays of the oracle back in the caller's original order."""
    order = o.get_field("GRID_INDEX")
    out = {}
    for k, v in arrays.items():
        w = np.empty_like(v)
        w[order] = v
        out[k] = w
    return out


# ------------------------------------------------------------- sweepu_k_bubble
NAMES = ("WATER_UX", "WATER_UY", "WATER_UZ")


def test_sweepu_linear_distance_field(pkg, orc):
    """sweepu_k_bubble (spray_k.cu:3779-3880) with a distance field
    phi = a i + b j + c k (a, b, c > 0) and no phase cell: only the octant
    (-1,-1,-1) passes (every other octant has a negative weight), its weights
    are the gradient components (a, b, c) normalised, each paired with the
    upwind face along its own axis.  For a linear u with slopes (sx, sy, sz)
    per index, one pass gives u - d with d = (a sx + b sy + c sz)/(a+b+c), two
    passes u - 2d where the upwind faces were updated too.  Faces the index
    ranges exclude keep their value."""
    n = 12
    cfg = small_cfg(pkg, n)
    o = orc.Oracle(cfg, use_double=True)
    o.set_field("MARK", np.full(n ** 3, 3, np.int8))
    I, J, K = np.meshgrid(*[np.arange(n)] * 3, indexing="ij")
    a, b, c = 0.5, 0.25, 1.0
    phi = (a * I + b * J + c * K).astype(float).ravel()
    o.set_field("PHIFLUID", phi)
    o.set_field("PHIAIR", phi)
    slopes = {0: (0.3, -0.7, 1.1), 1: (-0.4, 0.9, 0.2), 2: (0.6, 0.5, -0.8)}
    dims = face_dims(n)
    u0 = {}
    for ax, name in enumerate(NAMES):
        ii, jj, kk = face_grid(dims[ax])
        sx, sy, sz = slopes[ax]
        u0[ax] = 2.0 + sx * ii + sy * jj + sz * kk
        o.set_field(name, u0[ax])
        o.set_field(name.replace("WATER", "AIR"), u0[ax])
    o.extrapolate_bc(False)
    for ax, name in enumerate(NAMES):
        got = o.get_field(name).reshape(dims[ax])
        idx = face_grid(dims[ax])
        lim = n
        upd = (idx[ax] > 1) & (idx[ax] < lim - 1)
        for q in range(3):
            if q != ax:
                upd &= idx[q] >= 1
        deep = upd.copy()
        deep &= idx[ax] >= 3
        for q in range(3):
            if q != ax:
                deep &= idx[q] >= 2
        sx, sy, sz = slopes[ax]
        d = (a * sx + b * sy + c * sz) / (a + b + c)
        assert deep.sum() > 100
        np.testing.assert_allclose(got[deep], u0[ax][deep] - 2 * d, rtol=0, atol=2e-6, err_msg=name)
        # faces no pass may touch and the wall rules leave alone
        walls = np.zeros_like(upd)
        if ax == 0:
            walls = (idx[0] <= 1) | (idx[0] >= n - 1)
        elif ax == 1:
            walls = (idx[1] <= 1) | (idx[1] >= n - 1)
        else:
            walls = (idx[2] <= 1) | (idx[2] >= n - 1)
        keep = ~upd & ~walls
        assert keep.sum() > 10
        assert np.array_equal(got[keep], u0[ax][keep]), name


def test_sweepu_planar_front_copies_upwind_face(pkg, orc):
    """sweepPhi + sweepu_k_bubble on a liquid slab below k = 5 (spray.cu:570-596):
    faces with no liquid cell on either side take the value of the face
    below them (the distance grows only along z), so after two Jacobi passes
    the two layers above the slab hold the last wet value."""
    n = 12
    k0 = 5
    cfg = small_cfg(pkg, n)
    o = orc.Oracle(cfg, use_double=True)
    mark = np.full((n, n, n), 3, np.int8)
    mark[:, :, :k0] = 0
    o.set_field("MARK", mark)
    rng = np.random.default_rng(21)
    dims = face_dims(n)
    u0 = {}
    for ax, name in enumerate(NAMES):
        u0[ax] = rng.standard_normal(dims[ax])
        o.set_field(name, u0[ax])
        o.set_field(name.replace("WATER", "AIR"), u0[ax])
    o.extrapolate_bc(True)
    phi = o.get_field("PHIFLUID").reshape(n, n, n)
    assert np.all(phi[:, :, :k0] == -0.5) and np.all(phi[:, :, k0] == 0.5)
    ux = o.get_field("WATER_UX").reshape(dims[0])
    uy = o.get_field("WATER_UY").reshape(dims[1])
    uz = o.get_field("WATER_UZ").reshape(dims[2])
    sl = slice(2, n - 1)  # faces the wall rules leave alone
    for kk in (k0, k0 + 1):
        np.testing.assert_allclose(ux[sl, 1:-1, kk], u0[0][sl, 1:-1, k0 - 1], rtol=1e-15)
        np.testing.assert_allclose(uy[1:-1, sl, kk], u0[1][1:-1, sl, k0 - 1], rtol=1e-15)
    for kk in (k0 + 1, k0 + 2):
        np.testing.assert_allclose(uz[1:-1, 1:-1, kk], u0[2][1:-1, 1:-1, k0], rtol=1e-15)
    # wet faces are not extrapolated
    assert np.array_equal(ux[sl, :, :k0], u0[0][sl, :, :k0])
    assert np.array_equal(uz[:, :, 2:k0 + 1], u0[2][:, :, 2:k0 + 1])


def test_sweepu_tie_takes_one_third_of_last_octant(pkg, orc):
    """With no liquid the distance field is 3N everywhere, every weight is 0
    and every in-range octant passes with wsum == 0 (spray_k.cu:3808-3811):
    the last one in loop order, (+1,+1,+1), sets the face to a third of the
    sum of its +x, +y and +z neighbours along the face's own array."""
    n = 10
    cfg = small_cfg(pkg, n)
    o = orc.Oracle(cfg, use_double=True)
    o.set_field("MARK", np.full(n ** 3, 3, np.int8))
    rng = np.random.default_rng(22)
    dims = face_dims(n)
    u0 = {}
    for ax, name in enumerate(NAMES):
        u0[ax] = rng.standard_normal(dims[ax])
        o.set_field(name, u0[ax])
        o.set_field(name.replace("WATER", "AIR"), u0[ax])
    o.extrapolate_bc(True)
    assert np.all(o.get_field("PHIFLUID") == 3 * n)
    w = float(F32(1.0) / F32(3.0))

    def one_pass(v, ax):
        out = v.copy()
        idx = face_grid(dims[ax])
        full = (idx[ax] > 1) & (idx[ax] < n - 1)
        for q in range(3):
            if q != ax:
                full &= idx[q] + 1 <= n - 1
        sh = [np.roll(v, -1, axis=q) for q in range(3)]
        out[full] = w * sh[0][full] + w * sh[1][full] + w * sh[2][full]
        return out, full

    for ax, name in enumerate(NAMES):
        u1, full = one_pass(u0[ax], ax)
        u2, _ = one_pass(u1, ax)
        # faces whose +1 neighbours took the (+1,+1,+1) octant in the first pass
        chk = full.copy()
        for q in range(3):
            chk &= np.roll(full, -1, axis=q)
        # the wall rule of the z array zeroes k >= NZ-1 (ceiling handled below)
        got = o.get_field(name).reshape(dims[ax])
        assert chk.sum() > 50
        np.testing.assert_allclose(got[chk], u2[chk], rtol=1e-12, atol=1e-14, err_msg=name)


# ------------------------------------------------------------- setWaterBoundaryU_k
def test_wall_rules_known_answers(pkg, orc):
    """setWaterBoundaryU_k (spray_k.cu:956-995): ux = 0 for i <= 1 or i >= NX-1,
    uy likewise in j, uz = 0 for k <= 1 or k >= NZ; faces between a BOUNDARY
    and a non-BOUNDARY cell are 0; the ceiling face k = NZ-1 keeps uz(k-1)
    when it is negative, else 0.  The liquid fills the interior, so the
    extrapolation before the rules touches no face next to a liquid cell."""
    n = 10
    cfg = small_cfg(pkg, n)
    o = orc.Oracle(cfg, use_double=True)
    mark = boxed_marks(n, 0)
    mark[4, 4, 4] = 2  # an interior BOUNDARY cell
    o.set_field("MARK", mark)
    rng = np.random.default_rng(23)
    dims = face_dims(n)
    u0 = {}
    for ax, name in enumerate(NAMES):
        u0[ax] = rng.standard_normal(dims[ax]) + 0.1
        o.set_field(name, u0[ax])
        o.set_field(name.replace("WATER", "AIR"), np.zeros(dims[ax]))
    o.extrapolate_bc(True)
    ux = o.get_field("WATER_UX").reshape(dims[0])
    uy = o.get_field("WATER_UY").reshape(dims[1])
    uz = o.get_field("WATER_UZ").reshape(dims[2])
    j, k = 3, 6
    for i in (0, 1, n - 1, n):
        assert ux[i, j, k] == 0.0
    for i in range(2, n - 1):
        assert ux[i, j, k] == u0[0][i, j, k]
    assert ux[4, 4, 4] == 0.0 and ux[5, 4, 4] == 0.0
    assert ux[3, 4, 4] == u0[0][3, 4, 4] and ux[6, 4, 4] == u0[0][6, 4, 4]
    i = 3
    for jj in (0, 1, n - 1, n):
        assert uy[i, jj, k] == 0.0
    assert uy[4, 4, 4] == 0.0 and uy[4, 5, 4] == 0.0 and uy[4, 6, 4] == u0[1][4, 6, 4]
    for kk in (0, 1, n):
        assert np.all(uz[1:-1, 1:-1, kk] == 0.0)
    assert uz[4, 4, 4] == 0.0 and uz[4, 4, 5] == 0.0 and uz[4, 4, 3] == u0[2][4, 4, 3]
    below = u0[2][1:-1, 1:-1, n - 2]
    ceil = uz[1:-1, 1:-1, n - 1]
    assert np.any(below < 0) and np.any(below > 0)
    assert np.array_equal(ceil, np.where(below < 0, below, 0.0))


# ------------------------------------------------------------- cptdivergence_bubble2
@pytest.mark.parametrize("axis", [2, 0])
def test_two_phase_divergence_planar_interface(pkg, orc, axis):
    """cptdivergence_bubble2 (spray_k.cu:3138-3263) at a planar liquid/air
    interface at 5.3 cells along `axis`, with lsmerge the signed distance and
    linear fields u_w = (aw x, bw y, cw z), u_a = (aa x, ba y, ca z): on the
    interface face the liquid cell reads theta u_w + (1-theta) u_a with theta
    = 0.8 (its own share of the centre-to-centre segment), the air cell reads
    theta u_a + (1-theta) u_w with theta = 0.2; elsewhere each phase reads its
    own field, so div = aw+bw+cw in the liquid and aa+ba+ca in the air."""
    n = 10
    cfg = small_cfg(pkg, n, op=0, solver=0, ref_max_iters=1)
    o = orc.Oracle(cfg, use_double=True)
    zI = 5.3
    coord = np.meshgrid(*[np.arange(n)] * 3, indexing="ij")[axis]
    mark = np.where(coord < 5, 0, 1).astype(np.int8)
    mark = np.where(boxed_marks(n, 0) == 2, 2, mark).astype(np.int8)
    o.set_field("MARK", mark)
    o.set_field("LSMERGE", ((coord + 0.5 - zI) * H).ravel())
    sw = np.array([0.5, -2.0, 1.25])
    sa = np.array([-1.5, 0.75, 3.0])
    dims = face_dims(n)
    for ax, name in enumerate(NAMES):
        x = face_grid(dims[ax])[ax] * H
        o.set_field(name, sw[ax] * x)
        o.set_field(name.replace("WATER", "AIR"), sa[ax] * x)
    o.project(allow_no_convergence=True)
    d = o.get_field("DIV").reshape(n, n, n)
    inner = mark != 2
    other = sw.sum() - sw[axis]
    other_a = sa.sum() - sa[axis]
    for kk in range(1, n - 1):
        sl = [slice(1, -1)] * 3
        sl[axis] = kk
        got = d[tuple(sl)]
        if kk < 4:
            exp = sw.sum()
        elif kk == 4:   # liquid cell under the interface face at 5
            exp = other + (0.8 * sw[axis] * 5 + 0.2 * sa[axis] * 5 - sw[axis] * 4)
        elif kk == 5:   # air cell above it
            exp = other_a + (sa[axis] * 6 - (0.2 * sa[axis] * 5 + 0.8 * sw[axis] * 5))
        else:
            exp = sa.sum()
        np.testing.assert_allclose(got, exp, rtol=1e-12, atol=1e-12, err_msg="layer %d" % kk)
    assert np.all(d[~inner] == 0)


# ------------------------------------------------------------- advectparticle_RK2_bubble
def test_rk2_midpoint_step_and_clamp(pkg, orc):
    """advectparticle_RK2_bubble (spray_k.cu:3463-3523): with a linear velocity
    field u(x) = u0 + M x on the water grid (trilinear interpolation is exact
    for it), x' = x + dt u(x + dt/2 u(x)); AIR particles use the air grid
    (zero here); SOLID and AIRSOLO particles are untouched; a position past
    gmin + 1.5h or gmax - 1.5h is clamped there and that velocity component
    zeroed."""
    n = 12
    dt = 0.01
    cfg = small_cfg(pkg, n, dt=dt, integrator=1, fix_solids=0)
    o = orc.Oracle(cfg, use_double=True)
    M = np.array([[0.5, -1.0, 0.25], [0.75, -0.25, 0.5], [-0.5, 1.5, 0.125]])
    u0 = np.array([-0.6, -0.1, 0.3])
    dims = face_dims(n)
    for ax, name in enumerate(NAMES):
        idx = face_grid(dims[ax])
        xf = [(idx[q] + (0.0 if q == ax else 0.5)) * H for q in range(3)]
        val = u0[ax] + M[ax, 0] * xf[0] + M[ax, 1] * xf[1] + M[ax, 2] * xf[2]
        o.set_field(name, val)
        o.set_field(name.replace("WATER", "AIR"), np.zeros(dims[ax]))
    pos = np.array([[5.3, 6.1, 4.7], [3.2, 8.4, 6.6], [6.0, 6.0, 6.0], [6.0, 6.0, 6.0], [6.0, 6.0, 6.0],
                    [1.52, 6.0, 6.0]], np.float32) * F32(H)
    vel = np.array([[1, 2, 3], [-1, 0.5, 2], [4, 5, 6], [7, 8, 9], [1, 1, 1], [-3, 0.5, 0.25]], np.float32)
    flag = np.array([0, 0, 1, 5, 4, 0], np.int8)
    o.set_particles(len(flag), pos=pos, vel=vel, flag=flag)
    o.advect()
    got = o.get_particles()
    x = pos.astype(np.float64)
    dt = float(F32(dt))  # the configuration holds dt as a float

    def u(p):
        return u0 + p @ M.T

    for p in (0, 1):
        mid = x[p] + 0.5 * dt * u(x[p])
        exp = x[p] + dt * u(mid)
        np.testing.assert_allclose(got["pos"][p], exp, rtol=0, atol=1e-12)
        np.testing.assert_array_equal(got["vel"][p], vel[p])
    np.testing.assert_allclose(got["pos"][2], x[2], atol=1e-15)  # air grid is zero
    np.testing.assert_array_equal(got["vel"][2], vel[2])
    for p in (3, 4):  # SOLID and AIRSOLO untouched
        np.testing.assert_array_equal(got["pos"][p], x[p])
        np.testing.assert_array_equal(got["vel"][p], vel[p])
    # particle 5 starts at 1.52h and the field carries it below 1.5h: clamp
    mid = x[5] + 0.5 * dt * u(x[5])
    exp = x[5] + dt * u(mid)
    assert exp[0] < 1.5 * H
    assert got["pos"][5][0] == pytest.approx(1.5 * H, abs=1e-15) and got["vel"][5][0] == 0
    np.testing.assert_allclose(got["pos"][5][1:], exp[1:], atol=1e-12)
    assert got["vel"][5][1] == vel[5][1] and got["vel"][5][2] == vel[5][2]


# ------------------------------------------------------------- CollisionWithSolid_k
def _trilinear_face(u, a, p, n):
    """Trilinear interpolation of face array `a` at physical position p
    (interior points only: no clamping needed)."""
    q = np.array(p, np.float64) / H - np.where(np.arange(3) == a, 0.0, 0.5)
    b = np.floor(q).astype(int)
    f = q - b
    out = 0.0
    for di in (0, 1):
        for dj in (0, 1):
            for dk in (0, 1):
                w = (f[0] if di else 1 - f[0]) * (f[1] if dj else 1 - f[1]) * (f[2] if dk else 1 - f[2])
                out += w * u[b[0] + di, b[1] + dj, b[2] + dk]
    return out


def _slab_scene(pkg, orc, n, solid_vel, fix_solids, test_pos, test_vel, bounce_pos=0.5):
    dt = 0.01
    cfg = small_cfg(pkg, n, dt=dt, integrator=0, fix_solids=fix_solids, bounce_vel=1.0, bounce_pos=bounce_pos)
    o = orc.Oracle(cfg, use_double=True)
    rng = np.random.default_rng(24)
    cells = [(i, j, k) for i in range(1, n - 1) for j in range(1, n - 1) for k in (1, 2)]
    base = np.repeat(np.array(cells, np.float32), 8, axis=0)
    spos = ((base + rng.random(base.shape, dtype=np.float32)) * F32(H)).astype(np.float32)
    pos = np.concatenate([spos, np.array(test_pos, np.float32) * F32(H)])
    vel = np.concatenate([np.tile(np.array(solid_vel, np.float32), (len(spos), 1)), np.array(test_vel, np.float32)])
    flag = np.concatenate([np.full(len(spos), 5, np.int8), np.zeros(len(test_pos), np.int8)])
    P = len(flag)
    o.set_particles(P, pos=pos, vel=vel, flag=flag, mass=np.ones(P, np.float32))
    o.hash_sort()
    o.advect()
    got = in_cell_order(o, o.get_particles())
    nt = len(test_pos)
    dims = face_dims(n)
    su = [o.get_field(nm).reshape(dims[a]) for a, nm in enumerate(("SOLID_UX", "SOLID_UY", "SOLID_UZ"))]
    x0 = pos[-nt:].astype(np.float64)
    # the solid velocity the response reads: the fluid+solid P2G on the solid
    # grid (mapvelp2g_k_fluidSolid, spray.cu:2972), interpolated at the particle
    sv = np.array([[_trilinear_face(su[a], a, x, n) for a in range(3)] for x in x0])
    return o, float(F32(dt)), got["pos"][-nt:], got["vel"][-nt:], x0, sv


def _response(v, sv, x, phi, dt, bounce_pos):
    """Closed form above a +z normal: remove the normal relative velocity when
    approaching, push out by bounce_pos (0.5 - phi) h along +z, then move."""
    v = v.copy()
    rv = v[2] - sv[2]
    if rv < 0:
        v[2] -= rv
    xn = x.copy()
    xn[2] += bounce_pos * (0.5 - phi) * H
    return v, xn + v * dt


def test_collision_planar_solid_fixed(pkg, orc):
    """CollisionWithSolid_k (spray_k.cu:5276-5345) above a solid slab filling
    cells k = 1, 2: the solid distance (initSolidPhi + 3 sweeps) is z/h - 3
    from the slab to two cells above it, so at z = 3.2h phi = 0.2 and the
    normal is +z.  An approaching particle loses its normal velocity relative
    to the solid, a receding one keeps it, both are pushed out by
    bounce_pos (0.5 - phi) h, then x += v dt.  Beyond phi = 0.5 (z = 4.1h)
    only the move happens."""
    tp = [[6.3, 5.7, 3.2], [6.5, 5.5, 3.2], [5.5, 6.5, 4.1]]
    tv = [[0.1, -0.2, -0.5], [0.05, 0.1, 0.3], [0.2, 0.1, -0.4]]
    o, dt, gp, gv, x0, sv = _slab_scene(pkg, orc, 12, (0, 0, 0), 1, tp, tv)
    phis = o.get_field("PHISOLID").reshape(12, 12, 12)
    assert np.all(phis[3:9, 3:9, 2] == -0.5) and np.all(phis[3:9, 3:9, 3] == 0.5) and np.all(phis[3:9, 3:9, 4] == 1.5)
    v = np.array(tv, np.float32).astype(np.float64)
    for p in (0, 1):
        ev, ex = _response(v[p], sv[p], x0[p], 0.2, dt, 0.5)
        np.testing.assert_allclose(gv[p], ev, atol=1e-12)
        np.testing.assert_allclose(gp[p], ex, atol=1e-12)
    assert gv[0][2] != v[0][2] and gv[1][2] == v[1][2]
    np.testing.assert_array_equal(gv[2], v[2])
    np.testing.assert_allclose(gp[2], x0[2] + v[2] * dt, atol=1e-15)


def test_collision_planar_solid_moving(pkg, orc):
    """Moving solid (fix_solids off, slab velocity 0.2 z-hat): the normal
    relative velocity against the interpolated solid-grid velocity is removed,
    so an approaching particle leaves with the solid grid's normal velocity;
    the tangential components are kept."""
    tp = [[6.3, 5.7, 3.2], [6.5, 5.5, 3.3]]
    tv = [[0.1, -0.2, -0.5], [0.0, 0.3, -0.05]]
    o, dt, gp, gv, x0, sv = _slab_scene(pkg, orc, 12, (0, 0, 0.2), 0, tp, tv, bounce_pos=0.0)
    v = np.array(tv, np.float32).astype(np.float64)
    assert np.all(np.abs(sv[:, 2]) > 0.01)
    for p in (0, 1):
        ev, ex = _response(v[p], sv[p], x0[p], 0.0, dt, 0.0)
        np.testing.assert_allclose(gv[p], ev, atol=1e-12)
        np.testing.assert_allclose(gp[p], ex, atol=1e-12)
        assert gv[p][2] == pytest.approx(sv[p][2], abs=1e-12)


# ------------------------------------------------------------- heat
def test_heat_operator_conserves_heat(pkg, orc):
    """computeAx_heat (spray_k.cu:4526-4567) with Neumann BOUNDARY neighbours:
    the neighbour terms cancel in the sum over cells, so sum(A x) =
    sum(h^2/(alpha dt) x) over the non-boundary cells for any x and any mix
    of cell types (alpha per type) -- the implicit step conserves heat."""
    n = 10
    cfg = small_cfg(pkg, n, dt=0.003)
    o = orc.Oracle(cfg, use_double=True)
    rng = np.random.default_rng(25)
    mark = rng.choice(np.array([0, 1, 3, 4, 5], np.int8), (n, n, n))
    mark[[0, -1], :, :] = 2
    mark[:, [0, -1], :] = 2
    mark[:, :, [0, -1]] = 2
    mark[4:6, 4:6, 4:6] = 2
    o.set_field("MARK", mark)
    al = np.array(cfg.heat_alpha[:6], np.float64)
    x = rng.standard_normal(n ** 3) * 50 + 300
    y = o.apply_heat_operator(x)
    m = mark.ravel()
    nb = m != 2
    coef = H * H / al[m[nb]] / float(np.float32(0.003))
    assert np.all(y[~nb] == 0)
    assert y.sum() == pytest.approx((coef * x[nb]).sum(), rel=1e-12)


@pytest.mark.parametrize("mode", [1, 3])
def test_heat_backward_euler_cosine_decay(pkg, orc, mode):
    """One implicit step on a cosine mode cos(pi m (i - 1/2)/M) along x over
    the M = n-2 interior cells (Neumann walls): the discrete Laplacian's
    eigenvalue is 4 sin^2(pi m/(2M)), so the amplitude is divided by
    1 + lambda alpha dt / h^2, the mean is kept (Jacobi solve, converged)."""
    n = 16
    dt = 0.003
    cfg = small_cfg(pkg, n, dt=dt, heat_solver=1, heat_jacobi_iters=400, heat_add_bottom=0,
                    heat_extend_boundary=0)
    o = orc.Oracle(cfg, use_double=True)
    o.set_field("MARK", boxed_marks(n, 0))
    Mi = n - 2
    i = np.arange(n)
    prof = np.cos(np.pi * mode * (i - 0.5) / Mi)
    T0 = 300.0 + 5.0 * np.broadcast_to(prof[:, None, None], (n, n, n))
    o.set_field("TEMPERATURE", T0.ravel())
    o.heat_grid()
    T = o.get_field("TEMPERATURE").reshape(n, n, n)
    alpha = float(np.float32(cfg.heat_alpha[0]))
    lam = 4 * np.sin(np.pi * mode / (2 * Mi)) ** 2
    hh = float(np.float32(H)) ** 2
    exp = 300.0 + 5.0 * prof / (1 + lam * alpha * float(np.float32(dt)) / hh)
    inner = np.s_[1:-1, 1:-1, 1:-1]
    np.testing.assert_allclose(T[inner], np.broadcast_to(exp[1:-1, None, None], (Mi, Mi, Mi)), rtol=1e-12)
    dT = o.get_field("TEMPERATURE_CHANGE").reshape(n, n, n)
    np.testing.assert_allclose(dT[inner], (T - T0)[inner], rtol=1e-12, atol=1e-12)
    assert np.all(dT[0] == 0)


def test_heater_clamp_and_boundary_copy(pkg, orc):
    """addHeatAtBottom (spray_k.cu:4492-4504): cells with 1 <= i,j <= N-2 and
    k <= 3 gain heat_increase_bottom, every cell is clamped to
    heat_bottom_clamp.  setBoundaryHeat (4585-4605): a wall cell copies its
    inward neighbour by the first matching rule (i = N-1, i = 0, j = N-1,
    j = 0, k = N-1, k = 0), reading the pass's input (wall cells of the
    solve are 0, so edge cells copy 0).  No diffusion (0 Jacobi sweeps)."""
    n = 8
    cfg = small_cfg(pkg, n, heat_solver=1, heat_jacobi_iters=0, heat_add_bottom=1, heat_increase_bottom=1.5,
                    heat_bottom_clamp=378.0, heat_extend_boundary=1)
    o = orc.Oracle(cfg, use_double=True)
    o.set_field("MARK", boxed_marks(n, 0))
    rng = np.random.default_rng(26)
    T0 = 300.0 + rng.random((n, n, n)) * 10
    T0[3, 3, 2] = 377.0
    T0[4, 4, 6] = 390.0
    o.set_field("TEMPERATURE", T0.ravel())
    o.heat_grid()
    T = o.get_field("TEMPERATURE").reshape(n, n, n)
    heated = T0.copy()
    heated[1:n - 1, 1:n - 1, :4] += 1.5
    heated = np.minimum(heated, 378.0)
    inner = np.s_[1:-1, 1:-1, 1:-1]
    np.testing.assert_allclose(T[inner], heated[inner], rtol=1e-15)
    assert T[3, 3, 2] == 378.0 and T[4, 4, 6] == 378.0 and T[3, 3, 4] == pytest.approx(T0[3, 3, 4])
    assert T[0, 3, 4] == pytest.approx(heated[1, 3, 4]) and T[n - 1, 3, 4] == pytest.approx(heated[n - 2, 3, 4])
    assert T[3, 0, 4] == pytest.approx(heated[3, 1, 4]) and T[3, 3, n - 1] == pytest.approx(heated[3, 3, n - 2])
    assert T[3, 3, 0] == pytest.approx(heated[3, 3, 1])
    assert T[0, 0, 4] == 0.0 and T[n - 1, n - 1, n - 1] == 0.0


@pytest.mark.parametrize("melt_and_boil", [0, 1])
def test_heat_g2p_blend_and_latent_heat(pkg, orc, melt_and_boil):
    """mapHeatg2p(_MeltAndBoil) (spray_k.cu:4655-4679) on uniform grid fields
    Tg and dT: T' = at (T + dT) + (1 - at) Tg (SOLID: 0.95 T + 0.05 T' in the
    melt-and-boil form); updateLatentHeat_k (5546-5573): SOLID above the
    melting point stores the excess as latent heat at the melting point,
    FLUID below it (above the boiling point) likewise, FLUID in between resets
    its latent heat to the threshold; AIR is untouched."""
    n = 8
    at, mp, bp, th = 0.3, 273.15, 373.15, 10.0
    cfg = small_cfg(pkg, n, alpha_temp_trans=at, melting_point=mp, boiling_point=bp, liquid_heat_th=th,
                    heat_melt_and_boil=melt_and_boil)
    o = orc.Oracle(cfg, use_double=True)
    Tg, dT = 300.0, -4.0
    o.set_field("TEMPERATURE", np.full(n ** 3, Tg))
    o.set_field("TEMPERATURE_CHANGE", np.full(n ** 3, dT))
    pos = np.full((6, 3), 4.2 * H, np.float32)
    flag = np.array([5, 5, 0, 0, 0, 1], np.int8)
    T = np.array([100.0, 400.0, 150.0, 600.0, 300.0, 321.0], np.float32)
    lh = np.array([1.0, 2.0, 3.0, 4.0, 5.0, 6.0], np.float32)
    o.set_particles(6, pos=pos, flag=flag, T=T, lheat=lh)
    o.heat_g2p()
    got = o.get_particles()
    at32 = float(np.float32(at))
    mp32, bp32, th32 = (float(np.float32(v)) for v in (mp, bp, th))
    nt = at32 * (T.astype(float) + dT) + (1 - at32) * Tg
    newT = nt.copy()
    if melt_and_boil:
        newT[flag == 5] = float(np.float32(0.95)) * T[flag == 5].astype(float) + float(np.float32(0.05)) * nt[flag == 5]
    expT, expL = newT.copy(), lh.astype(float).copy()
    for p in range(6):
        if flag[p] == 5 and newT[p] > mp32:
            expL[p] += newT[p] - mp32
            expT[p] = mp32
        if flag[p] == 0:
            if newT[p] < mp32:
                expL[p] -= mp32 - newT[p]
                expT[p] = mp32
            elif newT[p] > bp32:
                expL[p] += newT[p] - bp32
                expT[p] = bp32
            else:
                expL[p] = th32
    np.testing.assert_allclose(got["T"], expT, rtol=1e-12)
    np.testing.assert_allclose(got["lheat"], expL, rtol=1e-12)


# ------------------------------------------------------------- nucleation
def test_nucleation_hand_computed_cells(pkg, orc):
    """updatesolubility + GenerateGasParticle_k (spray_k.cu:4719-4817) on three
    cells: A holds liquid with gas - solubility + vapour > airm0 and an AIR
    particle, so it emits an AIRSOLO particle at the gas centre + (rp - 1/2)
    (max(1/2, rn) h/4) 2; B holds the same liquid but no gas and is no seed,
    so it does not emit; C is a seed cell without gas and emits at its corner
    + h rp.  Emitting cells cap their liquid's gas content at its solubility
    and its latent heat at the threshold.  rp, rn are the library's counter
    hash of (seed, frame, cell, 0..3)."""
    n = 10
    T0 = 273.15
    cfg = small_cfg(pkg, n, gas_enabled=1, gas_every=1, melt_enabled=0, freeze_enabled=0, temperature0=T0,
                    dissolve_gas_rate=1.0, liquid_heat_th=10.0, vapor_gen_rate=0.01, seed=5)
    o = orc.Oracle(cfg, use_double=True)
    m0 = float(np.float32(cfg.airm0))
    cellA, cellB, cellC = (4, 4, 4), (5, 4, 6), (4, 5, 3)
    pos, flag, gas, lh, T = [], [], [], [], []
    for cell, with_gas in ((cellA, True), (cellB, False)):
        for q in range(3):
            pos.append((np.array(cell) + [0.2 + 0.2 * q, 0.5, 0.5]) * H)
            flag.append(0)
            gas.append(1.5 * m0)
            lh.append(12.0)
            T.append(T0)
        if with_gas:
            pos.append((np.array(cell) + [0.7, 0.3, 0.8]) * H)
            pos.append((np.array(cell) + [0.3, 0.6, 0.2]) * H)
            flag += [1, 4]
            gas += [0.0, 0.0]
            lh += [0.0, 0.0]
            T += [T0, T0]
    for q in range(2):
        pos.append((np.array(cellC) + [0.5, 0.25 + 0.5 * q, 0.5]) * H)
        flag.append(0)
        gas.append(1.5 * m0)
        lh.append(12.0)
        T.append(T0)
    P = len(flag)
    pos = np.array(pos, np.float32)
    mass = np.full(P, np.float32(m0), np.float32)
    o.set_particles(P, pos=pos, flag=np.array(flag, np.int8), gascontain=np.array(gas, np.float32),
                    lheat=np.array(lh, np.float32), T=np.array(T, np.float32), mass=mass)
    o.set_seed_cells([4 * 100 + 5 * 10 + 3])
    o.hash_sort()
    o.set_field("TEMPERATURE", np.full(n ** 3, 280.0))
    info = o.phase_change(0)
    assert info["gas_added"] == 2 and info["particles"] == P + 2
    got = o.get_particles()
    # solubility at T = T0 is dissolve_gas_rate * airm0 exactly
    liq = got["flag"][:P] == 0
    np.testing.assert_allclose(got["solubility"][:P][liq], m0, rtol=1e-12)
    new = {"pos": got["pos"][P:], "flag": got["flag"][P:], "T": got["T"][P:], "mass": got["mass"][P:]}
    assert list(new["flag"]) == [4, 4] and np.all(new["T"] == 280.0) and np.allclose(new["mass"], m0)

    def rnd(cell, q):
        return orc.uniform(5, 0, cell[0] * 100 + cell[1] * 10 + cell[2], q)

    # slots follow the cell index: A (444) before C (453)
    rpA = np.array([rnd(cellA, q) for q in range(3)])
    rnA = rnd(cellA, 3)
    gasc = ((np.array(cellA) + [0.7, 0.3, 0.8]) + (np.array(cellA) + [0.3, 0.6, 0.2])) / 2 * H
    expA = gasc + (rpA - 0.5) * (max(0.5, rnA) * 0.25 * H) * 2
    np.testing.assert_allclose(new["pos"][0], expA, atol=1e-7 * H)
    rpC = np.array([rnd(cellC, q) for q in range(3)])
    np.testing.assert_allclose(new["pos"][1], np.array(cellC) * H + H * rpC, atol=1e-7 * H)
    # liquid of the emitting cells capped, of cell B untouched
    cells = np.floor(got["pos"][:P] / H + 1e-9).astype(int)
    for cell, emits in ((cellA, True), (cellB, False), (cellC, True)):
        sel = liq & np.all(cells == cell, axis=1)
        assert sel.sum() >= 2
        if emits:
            np.testing.assert_allclose(got["gascontain"][:P][sel], m0, rtol=1e-12)
            np.testing.assert_allclose(got["lheat"][:P][sel], 10.0, rtol=1e-12)
        else:
            np.testing.assert_allclose(got["gascontain"][:P][sel], 1.5 * m0, rtol=1e-6)
            np.testing.assert_allclose(got["lheat"][:P][sel], 12.0, rtol=1e-12)


def test_solubility_temperature_law(pkg, orc):
    """updatesolubility: s = rate airm0 exp(1018.9 (1/T - 1/T0)) for FLUID only."""
    n = 8
    T0 = 273.15
    cfg = small_cfg(pkg, n, gas_enabled=1, gas_every=1, temperature0=T0, dissolve_gas_rate=0.8)
    o = orc.Oracle(cfg, use_double=True)
    T = np.array([280.0, 300.0, 350.0, 320.0], np.float32)
    flag = np.array([0, 0, 0, 5], np.int8)
    o.set_particles(4, pos=np.full((4, 3), 4.5 * H, np.float32), flag=flag, T=T,
                    solubility=np.full(4, 7.0, np.float32))
    o.hash_sort()
    o.phase_change(0)
    got = o.get_particles()
    f = lambda v: float(np.float32(v))
    exp = f(0.8) * f(cfg.airm0) * np.exp(f(1018.9) * (1 / T.astype(float) - 1 / f(T0)))
    np.testing.assert_allclose(got["solubility"][:3], exp[:3], rtol=1e-12)
    assert got["solubility"][3] == 7.0


# ------------------------------------------------------------- flags and gradient
def test_count_flags_bruteforce(pkg, orc):
    """markfluid_dense + markBoundaryCell (spray_k.cu:506-545): per cell,
    cfs = #FLUID + #SOLID, ca = #AIR (AIRSOLO counts for neither); VACUUM when
    both are 0, FLUID when cfs > ca, else AIR; the outer layer BOUNDARY."""
    n = 8
    cfg = small_cfg(pkg, n)
    o = orc.Oracle(cfg)
    rng = np.random.default_rng(27)
    P = 1500
    pos = (rng.random((P, 3)) * n * H).astype(np.float32)
    flag = rng.choice(np.array([0, 1, 4, 5], np.int8), P, p=[0.35, 0.35, 0.2, 0.1])
    o.set_particles(P, pos=pos, flag=flag)
    o.hash_sort()
    o.mark_from_counts()
    cell = np.floor(pos.astype(np.float64) / H).astype(int).clip(0, n - 1)
    cid = (cell[:, 0] * n + cell[:, 1]) * n + cell[:, 2]
    cfs = np.bincount(cid[np.isin(flag, [0, 5])], minlength=n ** 3)
    ca = np.bincount(cid[flag == 1], minlength=n ** 3)
    exp = np.where((cfs == 0) & (ca == 0), 3, np.where(cfs > ca, 0, 1)).reshape(n, n, n)
    exp = np.where(boxed_marks(n, 0) == 2, 2, exp)
    assert np.array_equal(o.get_field("MARK").reshape(n, n, n), exp)
    assert (exp == 1).any() and (exp == 0).any() and (exp == 3).any()


def test_solid_cell_marking_bruteforce(pkg, orc):
    """Solid constraint: a non-boundary cell with particles of which at least
    half (and at least one) are SOLID becomes SOLID; other flags stay."""
    n = 8
    cfg = small_cfg(pkg, n, solid_constraint=1)
    o = orc.Oracle(cfg)
    rng = np.random.default_rng(28)
    P = 900
    pos = (rng.random((P, 3)) * n * H).astype(np.float32)
    flag = rng.choice(np.array([0, 1, 5], np.int8), P, p=[0.4, 0.2, 0.4])
    o.set_particles(P, pos=pos, flag=flag)
    o.hash_sort()
    m0 = boxed_marks(n, 0)
    m0[2, :, :] = 1
    o.set_field("MARK", m0)
    o.mark_solid_cells()
    cell = np.floor(pos.astype(np.float64) / H).astype(int).clip(0, n - 1)
    cid = (cell[:, 0] * n + cell[:, 1]) * n + cell[:, 2]
    cs = np.bincount(cid[flag == 5], minlength=n ** 3).reshape(n, n, n)
    co = np.bincount(cid[flag != 5], minlength=n ** 3).reshape(n, n, n)
    exp = np.where((m0 != 2) & (cs > 0) & (cs >= co), 5, m0)
    assert np.array_equal(o.get_field("MARK").reshape(n, n, n), exp)
    assert (exp == 5).sum() > 5 and ((m0 != 2) & (cs > 0) & (cs < co)).any()


def test_phigrad_linear_field(pkg, orc):
    """computePhigra (spray_k.cu:3994-4019): central differences (x1 - x0)/2h,
    a missing neighbour replaced by the cell itself; a linear level set
    gives its slope inside and half of it on the outer layer."""
    n = 8
    cfg = small_cfg(pkg, n)
    o = orc.Oracle(cfg, use_double=True)
    I, J, K = np.meshgrid(*[np.arange(n)] * 3, indexing="ij")
    a, b, c = 0.75, -1.5, 2.25
    o.set_field("LSMERGE", ((a * (I + 0.5) + b * (J + 0.5) + c * (K + 0.5)) * H).ravel())
    o.phigrad()
    gx = o.get_field("PHIGRAD_X").reshape(n, n, n)
    gy = o.get_field("PHIGRAD_Y").reshape(n, n, n)
    gz = o.get_field("PHIGRAD_Z").reshape(n, n, n)
    np.testing.assert_allclose(gx[1:-1], a, rtol=1e-12)
    np.testing.assert_allclose(gx[[0, -1]], a / 2, rtol=1e-12)
    np.testing.assert_allclose(gy[:, 1:-1], b, rtol=1e-12)
    np.testing.assert_allclose(gy[:, [0, -1]], b / 2, rtol=1e-12)
    np.testing.assert_allclose(gz[:, :, 1:-1], c, rtol=1e-12)
    np.testing.assert_allclose(gz[:, :, [0, -1]], c / 2, rtol=1e-12)


# ------------------------------------------------------------- deletion, pouring, seeds
def test_sweep_vacuum_diagonal_closed_form(pkg, orc):
    """sweepVacuum (spray_k.cu:4117-4132) as two-buffer passes: an AIR cell
    turns VACUUM when one of its eight corner-diagonal neighbours is VACUUM.
    From one VACUUM cell in AIR, T passes reach exactly the offsets whose
    three components share a parity and whose largest magnitude is <= T
    (each diagonal step flips every component's parity); BOUNDARY cells
    neither change nor spread."""
    n = 14
    for T in (1, 2, 3):
        cfg = small_cfg(pkg, n, delete_vacuum_sweeps=T)
        o = orc.Oracle(cfg)
        mark = boxed_marks(n, 1)
        c0 = (7, 6, 7)
        mark[c0] = 3
        o.set_field("MARK", mark)
        o.set_particles(0)
        assert o.delete_particles() == 0
        got = o.get_field("MARK").reshape(n, n, n)
        I, J, K = np.meshgrid(*[np.arange(n)] * 3, indexing="ij")
        a, b, c = I - c0[0], J - c0[1], K - c0[2]
        reach = ((a % 2) == (b % 2)) & ((b % 2) == (c % 2)) & (np.maximum(np.maximum(abs(a), abs(b)), abs(c)) <= T)
        exp = np.where(mark == 2, 2, np.where(reach, 3, 1))
        assert np.array_equal(got, exp), T


def test_delete_particle_rules_and_compaction(pkg, orc):
    """markDeleteAirParticle (spray_k.cu:4134-4169) + deleteparticles
    (4172-4194): SOLID always stays; FLUID goes where the air level set is
    below -h; AIR and AIRSOLO go with a VACUUM cell in their 3x3x3
    neighbourhood or with mass <= 1e-6.  Survivors keep their order and the
    cell lists are remapped to the compacted array (equal to a fresh sort)."""
    n = 12
    cfg = small_cfg(pkg, n, delete_vacuum_sweeps=0)
    o = orc.Oracle(cfg)
    rng = np.random.default_rng(29)
    P = 4000
    pos = ((rng.random((P, 3)) * (n - 2) + 1) * H).astype(np.float32)
    flag = rng.choice(np.array([0, 1, 4, 5], np.int8), P)
    mass = np.full(P, 1e-3, np.float32)
    mass[rng.random(P) < 0.05] = 1e-7
    o.set_particles(P, pos=pos, flag=flag, mass=mass)
    o.hash_sort()
    sp = o.get_particles()
    mark = boxed_marks(n, 1)
    vac = [(3, 3, 3), (8, 4, 6), (5, 9, 2)]
    for v in vac:
        mark[v] = 3
    o.set_field("MARK", mark)
    # lsair = (x - 6h) along x: below -h for x < 5h
    I = np.meshgrid(*[np.arange(n)] * 3, indexing="ij")[0]
    o.set_field("LSAIR", ((I + 0.5 - 6.0) * H).ravel())
    deleted = o.delete_particles(keep_lists=True)
    x = sp["pos"] / H
    cell = np.floor(x).astype(int)
    near = np.zeros(len(x), bool)
    for v in vac:
        near |= np.all(np.abs(cell - np.array(v)) <= 1, axis=1)
    keep = np.ones(len(x), bool)
    fl = sp["flag"]
    keep[(fl == 0) & (x[:, 0] - 6.0 < -1.0)] = False
    air = (fl == 1) | (fl == 4)
    keep[air & (near | (sp["mass"] <= 1e-6))] = False
    assert deleted == (~keep).sum() and 0 < deleted < P
    assert (keep[fl == 5]).all()
    got = o.get_particles()
    for k in ("pos", "flag", "mass"):
        assert np.array_equal(got[k], sp[k][keep]), k
    cs, ce = o.get_field("CELL_START"), o.get_field("CELL_END")
    ref = orc.Oracle(cfg)
    ref.set_particles(int(keep.sum()), pos=sp["pos"][keep].astype(np.float32), flag=sp["flag"][keep])
    ref.hash_sort()
    assert np.array_equal(ref.get_field("GRID_INDEX"), np.arange(keep.sum()))  # order already sorted
    assert np.array_equal(cs, ref.get_field("CELL_START"))
    ne = cs != pkg.CELL_UNDEF
    assert np.array_equal(ce[ne], ref.get_field("CELL_END")[ne])


def test_pour_known_answers(pkg, orc):
    """pouringwater (spray_k.cu:5575-5596): every pour_every frames the source
    particles are appended at src + (2u - 1) pos_rand samplespace with
    velocity src_v + (2u - 1) vel_rand, mass m0, the source's type, the pour
    temperature and latent heat = threshold; u from the counter hash of
    (seed, pour tag, source index, component 3..5 / 0..2); no pour when it
    would exceed the capacity."""
    n = 10
    cfg = small_cfg(pkg, n, pour_every=4, pour_pos_rand=0.1, pour_vel_rand=0.05, pour_temperature=300.0,
                    liquid_heat_th=10.0, seed=11, particle_capacity=30)
    o = orc.Oracle(cfg, use_double=True)
    src = np.array([[0.3, 0.4, 0.5], [0.31, 0.41, 0.5], [0.32, 0.42, 0.5]], np.float32) * 0.2
    sv = np.array([[0.0, 0.0, -1.0], [0.1, 0.0, -1.0], [0.0, 0.2, -1.0]], np.float32)
    o.set_pour_source(src, sv, flag=4)
    o.set_particles(2, pos=np.full((2, 3), 0.05, np.float32))
    assert o.pour(1) == 0 and o.particle_count() == 2
    assert o.pour(8) == 3 and o.particle_count() == 5
    got = o.get_particles()
    tag = 0x706f7572 << 32
    f = lambda v: float(np.float32(v))
    for i in range(3):
        u = [orc.uniform(11, tag, i, q) for q in range(6)]
        rv = np.array(u[:3]) * 2 - 1
        rp = np.array(u[3:]) * 2 - 1
        np.testing.assert_allclose(got["pos"][2 + i], src[i].astype(float) + rp * f(f(0.1) * f(cfg.samplespace)),
                                   rtol=0, atol=1e-12)
        np.testing.assert_allclose(got["vel"][2 + i], sv[i].astype(float) + rv * f(0.05), rtol=0, atol=1e-12)
    assert list(got["flag"][2:]) == [4, 4, 4]
    assert np.allclose(got["mass"][2:], f(cfg.m0)) and np.all(got["T"][2:] == 300.0)
    assert np.all(got["lheat"][2:] == 10.0)
    assert o.pour(12) == 3 and o.particle_count() == 8
    for _ in range(7):
        o.pour(16)
    assert o.particle_count() == 29  # 29 + 3 > 30: the next pour is skipped
    assert o.pour(20) == 0 and o.particle_count() == 29


def test_seed_refresh_ring(pkg, orc):
    """updateSeedCell (spray.cu:2933-2950): on frames that are multiples of
    seed_update_every, four seed slots (a ring over the list) get cells at
    k = 2 with 3 <= i, j < N - 3; other frames change nothing."""
    n = 16
    cfg = small_cfg(pkg, n, seed_update_every=5, seed=3)
    o = orc.Oracle(cfg)
    seeds0 = np.array([1, 2, 3, 4, 5, 6], np.int32)
    o.set_seed_cells(seeds0)
    o.update_seed_cells(3)
    assert np.array_equal(o.get_seed_cells(), seeds0)
    o.update_seed_cells(5)
    s1 = o.get_seed_cells()
    assert np.array_equal(s1[4:], seeds0[4:]) and not np.array_equal(s1[:4], seeds0[:4])
    o.update_seed_cells(10)
    s2 = o.get_seed_cells()
    assert np.array_equal(s2[2:4], s1[2:4])  # slots 4, 5, 0, 1 replaced
    for c in np.concatenate([s1[:4], s2[[4, 5, 0, 1]]]):
        i, j, k = c // (n * n), (c // n) % n, c % n
        assert k == 2 and 3 <= i < n - 3 and 3 <= j < n - 3
    o2 = orc.Oracle(cfg)
    o2.set_seed_cells(seeds0)
    o2.update_seed_cells(5)
    assert np.array_equal(o2.get_seed_cells(), s1)
