// The remaining stages of bubblesim's step (spray.cpp:1059-1150): solo-air
// classification, surface tension, drag between solo air and the liquid
// grid, the SPH step of air particles and the two position corrections.
//
// Every kernel is a gather: a particle reads its own 3x3x3 cell neighbourhood
// through the sorted cell lists (or interpolates grid fields), a face reads
// the particles of its two cells.  Two of the reference kernels are not
// deterministic and are restated so that they are:
//  * calDragForce (spray_k.cu:5127-5165) scatters float atomics onto the
//    liquid faces; here a particle pass writes each solo air particle's two
//    face shares and a face pass adds them in particle-index order (the
//    order of a sequential loop over the particles);
//  * enforceForceSoloAirP (spray_k.cu:4335-4373) reads neighbour velocities
//    while other threads overwrite them; here it reads the velocities of the
//    pass's input and writes a second buffer.
// The cube (h2 - d2)^3 of sumdensity (spray_k.cu:4264-4282, pow(x, 3.0f))
// is evaluated as x*x*x, and normalize() as v * (1 / length(v)).
#include "mpf_internal.h"
#include "interp.h"

namespace mpf {

namespace {

__device__ __forceinline__ f3 mk3(float x, float y, float z) { return f3{x, y, z}; }
__device__ __forceinline__ f3 add3(f3 a, f3 b) { return mk3(a.x + b.x, a.y + b.y, a.z + b.z); }
__device__ __forceinline__ f3 sub3(f3 a, f3 b) { return mk3(a.x - b.x, a.y - b.y, a.z - b.z); }
__device__ __forceinline__ f3 mul3(f3 a, float s) { return mk3(a.x * s, a.y * s, a.z * s); }
__device__ __forceinline__ f3 smul3(float s, f3 a) { return mk3(s * a.x, s * a.y, s * a.z); }
__device__ __forceinline__ f3 div3(f3 a, float s) { return mk3(a.x / s, a.y / s, a.z / s); }
__device__ __forceinline__ float dot3(f3 a, f3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ __forceinline__ float len3(f3 a) { return sqrtf(dot3(a, a)); }

// getijkfrompos, spray_k.cu:51-57
__device__ __forceinline__ void cell_of(const Grid& g, f3 x, int& i, int& j, int& k) {
  const float qx = (x.x - g.gx0) / g.h, qy = (x.y - g.gy0) / g.h, qz = (x.z - g.gz0) / g.h;
  i = (qx >= 0 && qx < (float)g.nx) ? (int)qx : 0;
  j = (qy >= 0 && qy < (float)g.ny) ? (int)qy : 0;
  k = (qz >= 0 && qz < (float)g.nz) ? (int)qz : 0;
}

// getVectorFromGrid, spray_k.cu:124-137
__device__ __forceinline__ f3 vec_from_grid(const Grid& g, f3 p, const float* gx, const float* gy, const float* gz) {
  return mk3(scalar_from_grid(g, p, gx), scalar_from_grid(g, p, gy), scalar_from_grid(g, p, gz));
}

// ---- markSoloAirParticle: verifySoloAirParticle, spray_k.cu:4195-4261 ----
// isInBoundaryCell (spray_k.cu:4214-4221): x and y within 2 cells of a wall
__device__ __forceinline__ bool in_boundary_band(const Grid& g, int x, int y) {
  const int level = 2;
  return x <= level || x >= g.nx - 1 - level || y <= level || y >= g.ny - 1 - level;
}

// Flags change only between AIR and AIRSOLO, and the neighbour count takes
// both, so the in-place update reads the same counts in any order.
__global__ void k_solo_air(Grid g, int64_t P, const f3* __restrict__ pos, f3* __restrict__ vel, int8_t* __restrict__ flag,
                           const float* __restrict__ lsmerge, const float* __restrict__ aux,
                           const float* __restrict__ auy, const float* __restrict__ auz,
                           const uint32_t* __restrict__ cstart, const uint32_t* __restrict__ cend, float tol1,
                           float tol2) {
  const int64_t p = xcd_block() * blockDim.x + threadIdx.x;
  if (p >= P) return;
  const int8_t fl = flag[p];
  if (fl == FLUID || fl == SOLID) return;
  const f3 x = pos[p];
  const float ls = scalar_from_grid(g, x, lsmerge);
  const float h = g.h;
  int i, j, k;
  cell_of(g, x, i, j, k);
  int cnt = 0;
  for (int di = -1; di <= 1; ++di)
    for (int dj = -1; dj <= 1; ++dj)
      for (int dk = -1; dk <= 1; ++dk) {
        if (!valid(g, i + di, j + dj, k + dk)) continue;
        const int64_t c = cid(g, i + di, j + dj, k + dk);
        const uint32_t s = cstart[c];
        if (s == MPF_CELL_UNDEF) continue;
        const uint32_t e = cend[c];
        for (uint32_t q = s; q < e; ++q) {  // cntairparticle, spray_k.cu:4195-4212
          const float d = len3(sub3(pos[q], x));
          const int8_t f = flag[q];
          if (d < h && (f == AIR || f == AIRSOLO)) ++cnt;
        }
      }
  const bool band = in_boundary_band(g, i, j);
  if ((cnt >= 10 || ls > tol1 * h) && fl == AIRSOLO && !band) {
    if (cnt >= 3) flag[p] = AIR;
  } else if (fl == AIR && (band || ls < tol2 * h || cnt <= 1)) {
    vel[p] = vel_from_grid(g, x, aux, auy, auz);
    flag[p] = AIRSOLO;
  }
}

// ---- computesurfacetension: compsurfacetension_k, spray_k.cu:3543-3637 ----
__global__ void k_st_cell(Grid g, const int8_t* __restrict__ mark, const float* __restrict__ gx,
                          const float* __restrict__ gy, const float* __restrict__ gz, float* __restrict__ sf,
                          float sigma) {
  const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= g.G) return;
  if (mark[idx] == BOUNDARY) {
    sf[idx] = 0.0f;
    return;
  }
  int i, j, k;
  ijk_of(g, idx, i, j, k);
  const float h = g.h;
  f3 cen = mk3(gx[idx], gy[idx], gz[idx]);
  float len = len3(cen);
  float res;
  if (len == 0.0f) {
    res = 0.0f;
  } else {
    cen = div3(cen, len);
    // the normalised component a of the gradient at a neighbour, or the
    // centre's when the neighbour is outside or has no gradient
    auto comp = [&](int a, int b, int q, int axis, float own) {
      if (!valid(g, a, b, q)) return own;
      const int64_t n = cid(g, a, b, q);
      const float l = len3(mk3(gx[n], gy[n], gz[n]));
      if (l == 0.0f) return own;
      const float v = axis == 0 ? gx[n] : (axis == 1 ? gy[n] : gz[n]);
      return v / l;
    };
    const float x1 = comp(i + 1, j, k, 0, cen.x), x0 = comp(i - 1, j, k, 0, cen.x);
    const float y1 = comp(i, j + 1, k, 1, cen.y), y0 = comp(i, j - 1, k, 1, cen.y);
    const float z1 = comp(i, j, k + 1, 2, cen.z), z0 = comp(i, j, k - 1, 2, cen.z);
    res = (x1 - x0 + y1 - y0 + z1 - z0) / h * 0.5f;
  }
  sf[idx] = res * sigma;
}

// enforcesurfacetension_p, spray_k.cu:3639-3680.  Its scene test leaves only
// AIR particles (AIRSOLO returns in every scene).
__global__ void k_st_particles(Grid g, int64_t P, const f3* __restrict__ pos, f3* __restrict__ vel,
                               const int8_t* __restrict__ flag, const float* __restrict__ lsmerge,
                               const float* __restrict__ sf, const float* __restrict__ gx,
                               const float* __restrict__ gy, const float* __restrict__ gz,
                               const int8_t* __restrict__ mark, float dt) {
  const int64_t p = xcd_block() * blockDim.x + threadIdx.x;
  if (p >= P) return;
  const int8_t fl = flag[p];
  if (fl == SOLID || fl == FLUID || fl == AIRSOLO) return;
  const f3 x = pos[p];
  float ils = scalar_from_grid(g, x, lsmerge);
  const float isf = scalar_from_grid(g, x, sf);
  f3 dir = vec_from_grid(g, x, gx, gy, gz);
  const float lendir = len3(dir);
  if (lendir == 0.0f) return;
  dir = div3(dir, lendir);
  ils = ils / lendir;
  int i, j, k;
  cell_of(g, x, i, j, k);
  int cnt = (mark[cid(g, i, j, k)] == AIR) ? 1 : 0;
  for (int di = -1; di <= 1; di += 2)
    for (int dj = -1; dj <= 1; dj += 2)
      for (int dk = -1; dk <= 1; dk += 2)
        if (valid(g, i + di, j + dj, k + dk) && mark[cid(g, i + di, j + dj, k + dk)] == AIR) cnt++;
  if (cnt == 0) return;
  if (fabsf(ils) < g.h) {
    const f3 f = smul3(-isf, dir);
    vel[p] = add3(vel[p], mul3(f, dt));
  }
}

// ---- enforceDragForce: calDragForce, spray_k.cu:5127-5165 ----
// Particle pass: each solo air particle's shares of the velocity change of
// its cell's low faces (s0) and high faces (s1), per axis.
__global__ void k_drag_particles(Grid g, int64_t P, const f3* __restrict__ pos, f3* __restrict__ vel,
                                 const int8_t* __restrict__ flag, const float* __restrict__ ux,
                                 const float* __restrict__ uy, const float* __restrict__ uz, float solo, float grid,
                                 float dt, int update_particle, f3* __restrict__ s0, f3* __restrict__ s1) {
  const int64_t p = xcd_block() * blockDim.x + threadIdx.x;
  if (p >= P || flag[p] != AIRSOLO) return;
  const f3 x = pos[p], v = vel[p];
  int i, j, k;
  cell_of(g, x, i, j, k);
  const f3 gv = vel_from_grid(g, x, ux, uy, uz);
  const f3 dv = sub3(gv, v);
  const f3 fb = smul3(solo * len3(dv), dv);
  const f3 ch = mul3(mul3(mk3(-fb.x, -fb.y, -fb.z), grid), dt);
  const f3 w = sub3(div3(x, g.h), mk3((float)i, (float)j, (float)k));
  s0[p] = mk3(ch.x * (1.0f - w.x), ch.y * (1.0f - w.y), ch.z * (1.0f - w.z));
  s1[p] = mk3(ch.x * w.x, ch.y * w.y, ch.z * w.z);
  if (update_particle) vel[p] = add3(v, mul3(fb, dt));
}

// Face pass, axis A: the face between cells c-1 and c (along A) takes the s1
// shares of the solo air particles of c-1, then the s0 shares of those of c
// (particle-index order).
template <int A>
__global__ void k_drag_faces(Grid g, int64_t NF, float* __restrict__ u, const int8_t* __restrict__ flag,
                             const uint32_t* __restrict__ cstart, const uint32_t* __restrict__ cend,
                             const uint8_t* __restrict__ cmask, const f3* __restrict__ s0,
                             const f3* __restrict__ s1) {
  const int64_t f = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (f >= NF) return;
  int i, j, k;
  ijk_face<A>(g, f, i, j, k);
  const int lim = A == 0 ? g.nx : (A == 1 ? g.ny : g.nz);
  const int at = A == 0 ? i : (A == 1 ? j : k);
  const uint8_t air = M_AIR | M_AIRSOLO;  // flags flip between the two after the sort
  bool any = false;
  float v = 0.0f;
  for (int side = 0; side < 2; ++side) {
    const int c = at - 1 + side;  // side 0: the cell below the face, side 1: above
    if (c < 0 || c > lim - 1) continue;
    const int64_t cell = cid(g, A == 0 ? c : i, A == 1 ? c : j, A == 2 ? c : k);
    if (!(cmask[cell] & air)) continue;
    const uint32_t s = cstart[cell];
    if (s == MPF_CELL_UNDEF) continue;
    const uint32_t e = cend[cell];
    for (uint32_t q = s; q < e; ++q) {
      if (flag[q] != AIRSOLO) continue;
      const f3 sh = side == 0 ? s1[q] : s0[q];
      const float d = A == 0 ? sh.x : (A == 1 ? sh.y : sh.z);
      if (!any) {
        v = u[f];
        any = true;
      }
      v = v + d;
    }
  }
  if (any) u[f] = v;
}

// ---- updateSoloAirParticle: calcDensPress_Air + enforceForceSoloAirP ----
// spray_k.cu:4264-4333: density of air particles (poly6 over the 27 cells)
__global__ void k_sph_density(Grid g, int64_t P, const f3* __restrict__ pos, const int8_t* __restrict__ flag,
                              const uint32_t* __restrict__ cstart, const uint32_t* __restrict__ cend,
                              float* __restrict__ pdens, float* __restrict__ ppress, float m0poly6, float rho) {
  const int64_t p = xcd_block() * blockDim.x + threadIdx.x;
  if (p >= P) return;
  const int8_t fl = flag[p];
  if (fl != AIR && fl != AIRSOLO) return;
  const f3 x = pos[p];
  const float h = g.h, h2 = h * h;
  int i, j, k;
  cell_of(g, x, i, j, k);
  float dens = 0.0f;
  for (int di = -1; di <= 1; ++di)
    for (int dj = -1; dj <= 1; ++dj)
      for (int dk = -1; dk <= 1; ++dk) {
        if (!valid(g, i + di, j + dj, k + dk)) continue;
        const int64_t c = cid(g, i + di, j + dj, k + dk);
        const uint32_t s = cstart[c];
        float res = 0.0f;
        if (s != MPF_CELL_UNDEF) {
          const uint32_t e = cend[c];
          for (uint32_t q = s; q < e; ++q) {
            const int8_t f = flag[q];
            if (f != AIR && f != AIRSOLO) continue;
            const f3 d = sub3(pos[q], x);
            const float d2 = dot3(d, d);
            if (d2 < h2) {
              const float t = h2 - d2;
              res = res + t * t * t;
            }
          }
        }
        dens = dens + res;
      }
  dens = dens * m0poly6;
  if (dens == 0.0f) dens = 1.0f;
  pdens[p] = 1.0f / dens;
  ppress[p] = 1.5f * (dens - rho * 0.5f);
}

// spray_k.cu:4305-4373: pressure (spiky) and viscosity (laplacian) forces on
// air particles; velocities of the input buffer, result in the output
// buffer (other particles copied).
__global__ void k_sph_force(Grid g, int64_t P, const f3* __restrict__ pos, const f3* __restrict__ vin,
                            f3* __restrict__ vout, const int8_t* __restrict__ flag, const uint32_t* __restrict__ cstart,
                            const uint32_t* __restrict__ cend, const float* __restrict__ pdens,
                            const float* __restrict__ ppress, float vterm, float spiky, float m0, float dt) {
  const int64_t p = xcd_block() * blockDim.x + threadIdx.x;
  if (p >= P) return;
  const int8_t fl = flag[p];
  f3 iv = vin[p];
  if (fl != AIRSOLO && fl != AIR) {
    vout[p] = iv;
    return;
  }
  const f3 x = pos[p];
  const float ip = ppress[p], id = pdens[p];
  const float h = g.h;
  int i, j, k;
  cell_of(g, x, i, j, k);
  f3 force = mk3(0.0f, 0.0f, 0.0f);
  for (int di = -1; di <= 1; ++di)
    for (int dj = -1; dj <= 1; ++dj)
      for (int dk = -1; dk <= 1; ++dk) {
        if (!valid(g, i + di, j + dj, k + dk)) continue;
        const int64_t c = cid(g, i + di, j + dj, k + dk);
        const uint32_t s = cstart[c];
        f3 res = mk3(0.0f, 0.0f, 0.0f);
        if (s != MPF_CELL_UNDEF) {
          const uint32_t e = cend[c];
          for (uint32_t q = s; q < e; ++q) {
            const f3 dir = sub3(x, pos[q]);
            const float dis = len3(dir);
            const int8_t f = flag[q];
            if (dis > 0.0f && dis < h && (f == AIRSOLO || f == AIR)) {
              const float cc = h - dis;
              const float pterm = -0.5f * cc * spiky * (ip + ppress[q]) / dis;
              const float dterm = cc * id * pdens[q];
              const f3 vq = vin[q];
              const f3 t = add3(smul3(pterm, dir), smul3(vterm, sub3(vq, iv)));
              res = add3(res, mul3(t, dterm));
            }
          }
        }
        force = add3(force, res);
      }
  force = mul3(force, m0);
  iv = add3(iv, mul3(force, dt));
  vout[p] = iv;
}

// ---- correctpos: correctparticlepos + sumcellspring, spray_k.cu:1330-1411 ----
// The spring weight m (1 - d^2/re^2) is 0 from d >= re on (re = radius h,
// 0.5 h by default), and a zero-weight pair adds +-0 to the spring, which
// changes no position: cells farther than re from the particle are skipped
// (the bound keeps a margin against the rounding of the cell boundaries), and
// a pair is evaluated only when its squared distance is below re^2 (1 + 2^-19),
// a bound under which the rounded quotient d^2/re^2 is always >= 1.
constexpr int CP_LIST = 16;  // in-range neighbours listed per thread before a drain
__global__ void __launch_bounds__(256) k_correct_pos(Grid g, int64_t P, const f3* __restrict__ pin, f3* __restrict__ pout,
                              const float* __restrict__ mass, const int8_t* __restrict__ flag,
                              const uint32_t* __restrict__ cstart, const uint32_t* __restrict__ cend, float spring_k,
                              float radius, float ss, float dt) {
  __shared__ uint32_t lds_hits[CP_LIST][256];  // blockDim.x == 256
  const int64_t p = xcd_block() * blockDim.x + threadIdx.x;
  if (p >= P) return;
  const int8_t fl = flag[p];
  f3 x = pin[p];
  if (fl == SOLID || fl == AIRSOLO) {
    pout[p] = x;
    return;
  }
  int i, j, k;
  cell_of(g, x, i, j, k);
  f3 spring = mk3(0.0f, 0.0f, 0.0f);
  const float hh = g.h;
  const float m = hh + 0.5f * ss;
  const f3 tmin = mk3(g.gx0 + m, g.gy0 + m, g.gz0 + m);
  const f3 tmax = mk3(g.gx1 - m, g.gy1 - m, g.gz1 - m);
  const float re = radius * hh;
  const float re2 = re * re;
  const float cut = re2 * (1.0f + 1.0f / 524288.0f);
  // distance from the particle to the low and high faces of its cell, per
  // axis, less a margin: a neighbour cell on that side is reachable only
  // when that gap is below re
  const float mg = 1e-3f * hh;
  const float lx = (x.x - g.gx0) - (float)i * hh, ly = (x.y - g.gy0) - (float)j * hh, lz = (x.z - g.gz0) - (float)k * hh;
  const float gl[3] = {lx - mg, ly - mg, lz - mg};
  const float gh[3] = {hh - lx - mg, hh - ly - mg, hh - lz - mg};
  auto gap = [&](int a, int d) -> float {
    const float v = d < 0 ? gl[a] : (d > 0 ? gh[a] : 0.0f);
    return v > 0.0f ? v : 0.0f;
  };
  // the reachable cells, as bits of the 27 offsets in the reference's order
  uint32_t cand = 0;
  for (int o = 0; o < 27; ++o) {
    const int di = o / 9 - 1, dj = (o / 3) % 3 - 1, dk = o % 3 - 1;
    if (!valid(g, i + di, j + dj, k + dk)) continue;
    const float gx = gap(0, di), gy = gap(1, dj), gz = gap(2, dk);
    if (gx * gx + gy * gy + gz * gz <= cut) cand |= 1u << o;
  }
  // Two passes.  The first reads distances only: the in-range neighbours of
  // the reachable cells (cells by offset, particles by index, the reference's
  // order) go to this thread's list in LDS, the first of each cell marked.
  // The second evaluates the spring terms of the list, adding the running
  // cell subtotal to the spring at each mark (sumcellspring returns one
  // subtotal per cell; cells without an in-range neighbour add +-0).  A wave
  // then pays the spring arithmetic for its lane with the most in-range
  // neighbours, not for every candidate any lane has in range.  A full list
  // is drained and refilled.  The particle itself (d = 0) adds nothing and
  // is not listed.
  uint32_t(*hits)[256] = lds_hits;
  const int tid = threadIdx.x;
  int nh = 0;
  f3 sc = mk3(0.0f, 0.0f, 0.0f);
  auto drain = [&]() {
    if (nh == 0) return;
    uint32_t e = hits[0][tid];
    f3 y = pin[e & 0x7fffffffu];
    float m = mass[e & 0x7fffffffu];
    for (int h = 0; h < nh; ++h) {
      const uint32_t ec = e;
      const f3 yc = y;
      const float mc = m;
      if (h + 1 < nh) {  // next entry's loads in flight during this one
        e = hits[h + 1][tid];
        y = pin[e & 0x7fffffffu];
        m = mass[e & 0x7fffffffu];
      }
      if (ec >> 31) {
        spring = add3(spring, sc);
        sc = mk3(0.0f, 0.0f, 0.0f);
      }
      const f3 dd = sub3(yc, x);
      const float d2 = dot3(dd, dd);
      const float dist = sqrtf(d2);
      const float w = mc * fmaxf(1.0f - (dist * dist) / (re * re), 0.0f);
      if (dist > 0.1f * re) sc = add3(sc, div3(smul3(w, sub3(x, yc)), dist));
    }
    nh = 0;
  };
  // up to 8 cells at a time: their ranges read together, then each cell's
  // particles four at a time (positions in flight together)
  constexpr int NCB = 8;
  while (cand) {
    uint32_t cs[NCB], ce[NCB];
    int nc = 0;
#pragma unroll
    for (int t = 0; t < NCB; ++t) {
      int64_t c = 0;
      if (cand) {
        const int o = __ffs(cand) - 1;
        cand &= cand - 1u;
        c = cid(g, i + o / 9 - 1, j + (o / 3) % 3 - 1, k + o % 3 - 1);
        nc = t + 1;
      }
      cs[t] = cstart[c];
      ce[t] = cend[c];
    }
#pragma unroll
    for (int t = 0; t < NCB; ++t) {
      if (t >= nc) break;
      const uint32_t s0 = cs[t];
      if (s0 == MPF_CELL_UNDEF) continue;
      const uint32_t e0 = ce[t];
      uint32_t mark = 0x80000000u;
      for (uint32_t q0 = s0; q0 < e0; q0 += 4) {
        f3 y4[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) y4[u] = pin[min(q0 + (uint32_t)u, e0 - 1u)];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
          const uint32_t q = q0 + (uint32_t)u;
          if (q >= e0) break;
          const f3 dd = sub3(y4[u], x);
          const float d2 = dot3(dd, dd);
          if (!(d2 < cut) || (int64_t)q == p) continue;
          if (nh == CP_LIST) drain();
          hits[nh][tid] = q | mark;
          mark = 0;
          ++nh;
        }
      }
    }
  }
  drain();
  spring = add3(spring, sc);
  spring = mul3(spring, spring_k * re);
  if (len3(smul3(dt, spring)) > 0.3f * hh)
    x = add3(x, div3(smul3(hh * 0.3f, spring), len3(spring)));
  else
    x = add3(x, smul3(dt, spring));
  x.x = fmaxf(tmin.x, fminf(tmax.x, x.x));
  x.y = fmaxf(tmin.y, fminf(tmax.y, x.y));
  x.z = fmaxf(tmin.z, fminf(tmax.z, x.z));
  pout[p] = x;
}

// ---- correctpos_bubble: correctbubblepos / correctbubblepos_air ----
// spray_k.cu:3884-3991 (AIRLS: the _air variant with the air level set)
__device__ __forceinline__ bool unit_dir(const Grid& g, f3 x, const float* gx, const float* gy, const float* gz,
                                         f3& dir, float& dirlen) {
  dir = vec_from_grid(g, x, gx, gy, gz);
  dirlen = len3(dir);
  if (dirlen == 0.0f) return false;
  dir = mul3(dir, 1.0f / dirlen);
  return true;
}

template <bool AIRLS>
__global__ void k_correct_bubble(Grid g, int64_t P, f3* __restrict__ pos, const int8_t* __restrict__ flag,
                                 const float* __restrict__ ls, const float* __restrict__ gx,
                                 const float* __restrict__ gy, const float* __restrict__ gz,
                                 const float* __restrict__ lsair, const float* __restrict__ ax,
                                 const float* __restrict__ ay, const float* __restrict__ az) {
  const int64_t p = xcd_block() * blockDim.x + threadIdx.x;
  if (p >= P) return;
  const int8_t fl = flag[p];
  if (fl == AIRSOLO || fl == SOLID) return;
  f3 x = pos[p];
  const float s = (fl == FLUID) ? -1.0f : 1.0f;
  const float h = g.h, rs = 0.5f * h;
  f3 dir;
  float dirlen;
  if (!unit_dir(g, x, gx, gy, gz, dir, dirlen)) return;
  float d = scalar_from_grid(g, x, ls) / dirlen;
  if (!AIRLS) {
    if (s * d < 0.0f && fabsf(d) < 0.5f * h) {
      if (fl == AIR && fabsf(d) > 0.3f * h) {
        x = sub3(x, smul3(d, dir));
      } else if (fl == FLUID) {
        x = sub3(x, smul3(d, dir));
        if (!unit_dir(g, x, gx, gy, gz, dir, dirlen)) return;
        d = scalar_from_grid(g, x, ls) / dirlen;
        x = add3(x, smul3(s * (rs - s * d), dir));
      }
    } else if (fl == FLUID && s * d < rs * 0.5f && s * d >= 0.0f) {
      x = add3(x, smul3(s * (rs - s * d), dir));
    }
  } else {
    if (s * d < 0.0f && fabsf(d) < 0.5f * h) {
      if (fl == AIR && fabsf(d) > 0.3f * h) x = sub3(x, smul3(d, dir));
    }
    if (fl == FLUID) {
      if (!unit_dir(g, x, ax, ay, az, dir, dirlen)) return;
      d = scalar_from_grid(g, x, lsair) / dirlen;
      if (d < -1.3f * rs) x = sub3(x, smul3(d - rs, dir));
    }
  }
  pos[p] = x;
}

// computePhigra, spray_k.cu:3994-4019 (for the air level set)
__global__ void k_phigrad_of(Grid g, const float* __restrict__ ls, float* __restrict__ gx, float* __restrict__ gy,
                             float* __restrict__ gz) {
  const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= g.G) return;
  int i, j, k;
  ijk_of(g, idx, i, j, k);
  const float h = g.h, lc = ls[idx];
  const float x1 = valid(g, i + 1, j, k) ? ls[cid(g, i + 1, j, k)] : lc;
  const float x0 = valid(g, i - 1, j, k) ? ls[cid(g, i - 1, j, k)] : lc;
  const float y1 = valid(g, i, j + 1, k) ? ls[cid(g, i, j + 1, k)] : lc;
  const float y0 = valid(g, i, j - 1, k) ? ls[cid(g, i, j - 1, k)] : lc;
  const float z1 = valid(g, i, j, k + 1) ? ls[cid(g, i, j, k + 1)] : lc;
  const float z0 = valid(g, i, j, k - 1) ? ls[cid(g, i, j, k - 1)] : lc;
  gx[idx] = ((x1 - x0) * 0.5f) / h;
  gy[idx] = ((y1 - y0) * 0.5f) / h;
  gz[idx] = ((z1 - z0) * 0.5f) / h;
}

}  // namespace

// ================= host stage drivers =================
void run_mark_solo_air(Ctx& c) {
  const int cur = c.cur;
  launch(c, "solo_air", dim3(nblocks(c.P, 256)), dim3(256), 0, k_solo_air, c.g, c.P, (const f3*)c.pos[cur],
         c.vel[cur], c.flag[cur], (const float*)c.lsmerge, (const float*)c.au[0], (const float*)c.au[1],
         (const float*)c.au[2], (const uint32_t*)c.cstart, (const uint32_t*)c.cend, c.cfg.solo_tol_in,
         c.cfg.solo_tol_out);
  // flags moved between AIR and AIRSOLO: the occupancy masks the gathers skip
  // cells by (the air P2G takes AIR only) are rebuilt
  run_cell_mask(c);
}

void run_surface_tension(Ctx& c) {
  const int cur = c.cur;
  launch(c, "st_cell", dim3(nblocks(c.g.G, 256)), dim3(256), 0, k_st_cell, c.g, (const int8_t*)c.mark,
         (const float*)c.phig[0], (const float*)c.phig[1], (const float*)c.phig[2], c.sf,
         c.cfg.surface_tension_sigma);
  launch(c, "st_particles", dim3(nblocks(c.P, 256)), dim3(256), 0, k_st_particles, c.g, c.P, (const f3*)c.pos[cur],
         c.vel[cur], (const int8_t*)c.flag[cur], (const float*)c.lsmerge, (const float*)c.sf,
         (const float*)c.phig[0], (const float*)c.phig[1], (const float*)c.phig[2], (const int8_t*)c.mark,
         c.cfg.dt);
}

void run_drag(Ctx& c) {
  const int cur = c.cur, nxt = 1 - c.cur;
  f3* s0 = c.pos[nxt];  // scratch until the next sort
  f3* s1 = c.vel[nxt];
  launch(c, "drag_particles", dim3(nblocks(c.P, 256)), dim3(256), 0, k_drag_particles, c.g, c.P,
         (const f3*)c.pos[cur], c.vel[cur], (const int8_t*)c.flag[cur], (const float*)c.wu[0],
         (const float*)c.wu[1], (const float*)c.wu[2], c.cfg.drag_solo, c.cfg.drag_grid, c.cfg.dt,
         c.cfg.drag_particle_update, s0, s1);
  const int8_t* fl = c.flag[cur];
  launch(c, "drag_faces_x", dim3(nblocks(c.g.FX, 256)), dim3(256), 0, k_drag_faces<0>, c.g, c.g.FX, c.wu[0], fl,
         (const uint32_t*)c.cstart, (const uint32_t*)c.cend, (const uint8_t*)c.cmask, (const f3*)s0, (const f3*)s1);
  launch(c, "drag_faces_y", dim3(nblocks(c.g.FY, 256)), dim3(256), 0, k_drag_faces<1>, c.g, c.g.FY, c.wu[1], fl,
         (const uint32_t*)c.cstart, (const uint32_t*)c.cend, (const uint8_t*)c.cmask, (const f3*)s0, (const f3*)s1);
  launch(c, "drag_faces_z", dim3(nblocks(c.g.FZ, 256)), dim3(256), 0, k_drag_faces<2>, c.g, c.g.FZ, c.wu[2], fl,
         (const uint32_t*)c.cstart, (const uint32_t*)c.cend, (const uint8_t*)c.cmask, (const f3*)s0, (const f3*)s1);
}

// SPH kernel constants of initparam (spray.cpp:584-586), from the cell size.
void sph_constants(float h, float& poly6, float& spiky, float& lap) {
  poly6 = 315.0f / (64.0f * 3.141592f * powf(h, 9.0f));
  spiky = -45.0f / (3.141592f * powf(h, 6.0f));
  lap = 45.0f / (3.141592f * powf(h, 6.0f));
}

void run_solo_air_sph(Ctx& c) {
  const int cur = c.cur, nxt = 1 - c.cur;
  float poly6, spiky, lap;
  sph_constants(c.g.h, poly6, spiky, lap);
  float* dens = c.mass[nxt];  // scratch until the next sort
  float* press = c.temp[nxt];
  launch(c, "sph_density", dim3(nblocks(c.P, 256)), dim3(256), 0, k_sph_density, c.g, c.P, (const f3*)c.pos[cur],
         (const int8_t*)c.flag[cur], (const uint32_t*)c.cstart, (const uint32_t*)c.cend, dens, press,
         c.cfg.airm0 * poly6, c.cfg.rho_liquid);
  launch(c, "sph_force", dim3(nblocks(c.P, 256)), dim3(256), 0, k_sph_force, c.g, c.P, (const f3*)c.pos[cur],
         (const f3*)c.vel[cur], c.vel[nxt], (const int8_t*)c.flag[cur], (const uint32_t*)c.cstart,
         (const uint32_t*)c.cend, (const float*)dens, (const float*)press, lap * c.cfg.sph_viscosity, spiky,
         c.cfg.airm0, c.cfg.dt);
  std::swap(c.vel[cur], c.vel[nxt]);
}

void run_correct_pos(Ctx& c) {
  const int cur = c.cur, nxt = 1 - c.cur;
  launch(c, "correct_pos", dim3(nblocks(c.P, 256)), dim3(256), 0, k_correct_pos, c.g, c.P, (const f3*)c.pos[cur],
         c.pos[nxt], (const float*)c.mass[cur], (const int8_t*)c.flag[cur], (const uint32_t*)c.cstart,
         (const uint32_t*)c.cend, c.cfg.correction_spring, c.cfg.correction_radius, c.cfg.samplespace, c.cfg.dt);
  std::swap(c.pos[cur], c.pos[nxt]);
  c.sorted = c.own_sort = c.mc_sorted = false;
}

void run_correct_bubble_pos(Ctx& c) {
  const int cur = c.cur;
  if (c.cfg.correct_bubble_air_ls) {
    float* ax = c.ftmp[0];  // face scratch (>= G floats each)
    float* ay = c.ftmp[1];
    float* az = c.ftmp[2];
    launch(c, "phigrad_air", dim3(nblocks(c.g.G, 256)), dim3(256), 0, k_phigrad_of, c.g, (const float*)c.lsair, ax,
           ay, az);
    launch(c, "correct_bubble_air", dim3(nblocks(c.P, 256)), dim3(256), 0, k_correct_bubble<true>, c.g, c.P,
           c.pos[cur], (const int8_t*)c.flag[cur], (const float*)c.lsmerge, (const float*)c.phig[0],
           (const float*)c.phig[1], (const float*)c.phig[2], (const float*)c.lsair, (const float*)ax,
           (const float*)ay, (const float*)az);
  } else {
    launch(c, "correct_bubble", dim3(nblocks(c.P, 256)), dim3(256), 0, k_correct_bubble<false>, c.g, c.P,
           c.pos[cur], (const int8_t*)c.flag[cur], (const float*)c.lsmerge, (const float*)c.phig[0],
           (const float*)c.phig[1], (const float*)c.phig[2], (const float*)nullptr, (const float*)nullptr,
           (const float*)nullptr, (const float*)nullptr);
  }
  c.sorted = c.own_sort = c.mc_sorted = false;
}

}  // namespace mpf
