// Particle-side kernels: external force (a5), G2P / FLIP-PIC blend (a14),
// advection with solid collision or RK2 (a15), heat G2P with latent heat
// (a16, particle half) and phase change with nucleation (a17).
#include "mpf_internal.h"
#include "interp.h"

namespace mpf {

namespace {

__device__ __forceinline__ f3 mk3(float x, float y, float z) { return f3{x, y, z}; }
__device__ __forceinline__ f3 add3(f3 a, f3 b) { return mk3(a.x + b.x, a.y + b.y, a.z + b.z); }
__device__ __forceinline__ f3 sub3(f3 a, f3 b) { return mk3(a.x - b.x, a.y - b.y, a.z - b.z); }
__device__ __forceinline__ f3 mul3(f3 a, float s) { return mk3(a.x * s, a.y * s, a.z * s); }
__device__ __forceinline__ float dot3(f3 a, f3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ __forceinline__ float len3(f3 a) { return sqrtf(dot3(a, a)); }

struct Faces3 {
  const float* u[3];
};

// addgravityforce_k spray_k.cu:330-338 ; addbuoyancyforce_vel spray_k.cu:356-371
__global__ void k_forces(int64_t P, f3* __restrict__ vel, const int8_t* __restrict__ flag, float dt, f3 gr, int kick,
                         float vmax, float rate_air, float rate_solo) {
  int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= P) return;
  int8_t fl = flag[p];
  f3 v = vel[p];
  if (fl == FLUID || fl == SOLID) v = add3(v, mul3(gr, dt));
  if (kick) {
    float rate = fmaxf(vmax - v.z, 0.0f) / vmax;
    if (fl == AIR) v.z -= dt * gr.z * rate * rate_air;
    else if (fl == AIRSOLO) v.z -= dt * gr.z * rate * rate_solo;
    else if (fl == SOLID) v.z += dt * gr.z * 0.1f;
  }
  vel[p] = v;
}

// computeDeltaU, spray_k.cu:997-1006 (old <- new - old), one axis
__global__ void k_delta_u(int64_t n, const float* __restrict__ u, float* __restrict__ uo) {
  int64_t f = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (f < n) uo[f] = u[f] - uo[f];
}

// mapvelg2p_flip_bubble spray_k.cu:3525-3540 ; blend of advectparticle_RK2
// spray_k.cu:421-424 when alpha < 1.
__global__ void k_g2p(Grid g, int64_t P, const f3* __restrict__ pos, f3* __restrict__ vel,
                      const int8_t* __restrict__ flag, Faces3 wdu, Faces3 adu, Faces3 wu, Faces3 au, float alpha) {
  int64_t p = xcd_block() * blockDim.x + threadIdx.x;
  if (p >= P) return;
  // the flag, position and velocity are read together, and the phase's face
  // arrays are chosen by value (a pointer into the kernel arguments would be
  // a per-lane load of the pointers after the flag arrives)
  const int8_t fl = flag[p];
  const f3 x = pos[p];
  f3 v = vel[p];
  const bool air = fl == AIR;
  const bool mapped = fl == FLUID || fl == SOLID || air;
  const float* d0 = air ? adu.u[0] : wdu.u[0];
  const float* d1 = air ? adu.u[1] : wdu.u[1];
  const float* d2 = air ? adu.u[2] : wdu.u[2];
  const f3 gv = vel_from_grid(g, x, d0, d1, d2);
  if (alpha == 1.0f) v = add3(v, gv);
  else {
    const float* n0 = air ? au.u[0] : wu.u[0];
    const float* n1 = air ? au.u[1] : wu.u[1];
    const float* n2 = air ? au.u[2] : wu.u[2];
    const f3 u = vel_from_grid(g, x, n0, n1, n2);
    v = add3(mul3(u, 1.0f - alpha), mul3(add3(v, gv), alpha));
  }
  if (mapped) vel[p] = v;
}

// setVelZeroSolid_k spray_k.cu:5213-5218
__global__ void k_zero_solid_vel(int64_t P, f3* __restrict__ vel, const int8_t* __restrict__ flag) {
  int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p < P && flag[p] == SOLID) vel[p] = mk3(0.f, 0.f, 0.f);
}

// CollisionWithSolid_k, spray_k.cu:5276-5345
__global__ void k_collide_move(Grid g, int64_t P, f3* __restrict__ pos, f3* __restrict__ vel,
                               const int8_t* __restrict__ flag, const float* __restrict__ phis, Faces3 su, float dt,
                               float bounce_vel, float bounce_pos) {
  int64_t p = xcd_block() * blockDim.x + threadIdx.x;
  if (p >= P) return;
  // flag, position and velocity are read together; solid particles are
  // computed like the others and not stored
  const int8_t fl = flag[p];
  f3 x = pos[p], v = vel[p];
  const float h = g.h;
  float iphi = scalar_from_grid(g, x, phis);
  if (iphi <= 0.5f && fl != SOLID) {
    f3 sv = vel_from_grid(g, x, su.u[0], su.u[1], su.u[2]);
    f3 rv = sub3(v, sv);
    float d = h * 0.5f;
    f3 gd;
    gd.x = scalar_from_grid(g, add3(x, mk3(d, 0.f, 0.f)), phis) - scalar_from_grid(g, sub3(x, mk3(d, 0.f, 0.f)), phis);
    gd.y = scalar_from_grid(g, add3(x, mk3(0.f, d, 0.f)), phis) - scalar_from_grid(g, sub3(x, mk3(0.f, d, 0.f)), phis);
    gd.z = scalar_from_grid(g, add3(x, mk3(0.f, 0.f, d)), phis) - scalar_from_grid(g, sub3(x, mk3(0.f, 0.f, d)), phis);
    float gl = len3(gd);
    if (gl > 0) {
      gd = mul3(gd, 1.0f / gl);
      float dn = dot3(rv, gd);
      if (dn < 0) v = sub3(v, mul3(gd, bounce_vel * dn));
      x = add3(x, mul3(mul3(mul3(gd, bounce_pos), 0.5f - iphi), h));
    }
  }
  x = add3(x, mul3(v, dt));
  float rate = (fl == AIRSOLO) ? 0.8f : 0.5f;
  float m = h + rate * h;
  f3 lo = mk3(g.gx0 + m, g.gy0 + m, g.gz0 + m), hi = mk3(g.gx1 - m, g.gy1 - m, g.gz1 - m);
  if (x.x <= lo.x) { x.x = lo.x; v.x = 0; }
  if (x.y <= lo.y) { x.y = lo.y; v.y = 0; }
  if (x.z <= lo.z) { x.z = lo.z; v.z = 0; }
  if (x.x >= hi.x) { x.x = hi.x; v.x = 0; }
  if (x.y >= hi.y) { x.y = hi.y; v.y = 0; }
  if (x.z >= hi.z) { x.z = hi.z; v.z = 0; }
  if (fl == SOLID) return;
  vel[p] = v;
  pos[p] = x;
}

// CollisionWithSolid_Freezing, spray_k.cu:5348-5446 (CollisionSolid's
// integrator for the freezing and melting-pour scenes): near a solid
// (phisolid <= 1) a particle pushed by neighbours closer than h/2 leaves
// the solid surface along its normal by the deepest overlap with the normal
// velocity removed, or with neighbours within h only loses 0.1 |v| along
// the normal; then the wall bounce (-0.5 v at gmin + h + 0.3 ss) and
// x += v dt.  Neighbour positions are the pass's input (the reference reads
// them while they are rewritten); the result goes to the other buffer.
__global__ void k_collide_freezing(Grid g, int64_t P, const f3* __restrict__ pin, f3* __restrict__ pout,
                                   f3* __restrict__ vel, const int8_t* __restrict__ flag,
                                   const float* __restrict__ phis, const uint32_t* __restrict__ cstart,
                                   const uint32_t* __restrict__ cend, float dt, float ss) {
  int64_t p = xcd_block() * blockDim.x + threadIdx.x;
  if (p >= P) return;
  f3 x = pin[p];
  if (flag[p] == SOLID) {
    pout[p] = x;
    return;
  }
  f3 v = vel[p];
  const float h = g.h;
  const float qx = (x.x - g.gx0) / h, qy = (x.y - g.gy0) / h, qz = (x.z - g.gz0) / h;
  const int i = (qx >= 0 && qx < (float)g.nx) ? (int)qx : 0;
  const int j = (qy >= 0 && qy < (float)g.ny) ? (int)qy : 0;
  const int k = (qz >= 0 && qz < (float)g.nz) ? (int)qz : 0;
  const float iphi = scalar_from_grid(g, x, phis);
  if (iphi <= 1.0f) {
    const float r = 0.25f * h, hh = 4.0f * r;
    float depth = 0.0f;
    int ncol = 0, nadh = 0;
    for (int di = -1; di <= 1; ++di)
      for (int dj = -1; dj <= 1; ++dj)
        for (int dk = -1; dk <= 1; ++dk) {
          if (!valid(g, i + di, j + dj, k + dk)) continue;
          const int64_t c = cid(g, i + di, j + dj, k + dk);
          const uint32_t s = cstart[c];
          if (s == MPF_CELL_UNDEF) continue;
          const uint32_t e = cend[c];
          for (uint32_t q = s; q < e; ++q) {
            const float dis = len3(sub3(x, pin[q]));
            if (dis > 0 && dis < 2 * r) {
              depth = fmaxf(depth, 2 * r - dis);
              ncol++;
            } else if (dis < hh) {
              nadh++;
            }
          }
        }
    const float d = h * 0.5f;
    f3 n;
    n.x = scalar_from_grid(g, add3(x, mk3(d, 0.f, 0.f)), phis) - scalar_from_grid(g, sub3(x, mk3(d, 0.f, 0.f)), phis);
    n.y = scalar_from_grid(g, add3(x, mk3(0.f, d, 0.f)), phis) - scalar_from_grid(g, sub3(x, mk3(0.f, d, 0.f)), phis);
    n.z = scalar_from_grid(g, add3(x, mk3(0.f, 0.f, d)), phis) - scalar_from_grid(g, sub3(x, mk3(0.f, 0.f, d)), phis);
    const f3 v0 = v;
    const float nl = len3(n);
    if (nl > 0) {
      n = mul3(n, 1.0f / nl);
      if (ncol > 0) {
        v = sub3(v, mul3(n, dot3(v0, n)));
        x = add3(x, mul3(n, depth));
      } else if (nadh > 0) {
        v = sub3(v, mul3(mul3(n, 0.1f), len3(v)));
      }
    }
  }
  const float m = h + 0.3f * ss;
  const f3 lo = mk3(g.gx0 + m, g.gy0 + m, g.gz0 + m), hi = mk3(g.gx1 - m, g.gy1 - m, g.gz1 - m);
  if (x.x > hi.x) { v.x *= -0.5f; x.x = hi.x; }
  if (x.x < lo.x) { v.x *= -0.5f; x.x = lo.x; }
  if (x.y > hi.y) { v.y *= -0.5f; x.y = hi.y; }
  if (x.y < lo.y) { v.y *= -0.5f; x.y = lo.y; }
  if (x.z > hi.z) { v.z *= -0.5f; x.z = hi.z; }
  if (x.z < lo.z) { v.z *= -0.5f; x.z = lo.z; }
  x = add3(x, mul3(v, dt));
  vel[p] = v;
  pout[p] = x;
}

// advectparticle_RK2_bubble spray_k.cu:3463-3523: FLUID particles on the water
// grid, AIR particles on the air grid, every other type untouched; wall
// margin h + 0.5h; clamped components of the velocity zeroed (the FLIP/PIC
// blend of that kernel is done by G2P).  The position is advanced with the
// midpoint step, as advectparticle_RK2 (spray_k.cu:409-459) writes it (the
// bubble kernel's position write is commented out, spray_k.cu:3520-3521).
__global__ void k_advect_rk2(Grid g, int64_t P, f3* __restrict__ pos, f3* __restrict__ vel,
                             const int8_t* __restrict__ flag, Faces3 wu, Faces3 au, float dt) {
  int64_t p = xcd_block() * blockDim.x + threadIdx.x;
  if (p >= P) return;
  int8_t fl = flag[p];
  if (fl != FLUID && fl != AIR) return;
  // the phase's face arrays chosen by value (no per-lane pointer load)
  const bool air = fl == AIR;
  const float* u0 = air ? au.u[0] : wu.u[0];
  const float* u1 = air ? au.u[1] : wu.u[1];
  const float* u2 = air ? au.u[2] : wu.u[2];
  f3 x = pos[p], v = vel[p];
  float m = g.h + 0.5f * g.h;
  f3 lo = mk3(g.gx0 + m, g.gy0 + m, g.gz0 + m), hi = mk3(g.gx1 - m, g.gy1 - m, g.gz1 - m);
  f3 gv = vel_from_grid(g, x, u0, u1, u2);
  f3 mid = add3(x, mul3(mul3(gv, dt), 0.5f));
  f3 gm = vel_from_grid(g, mid, u0, u1, u2);
  x = add3(x, mul3(gm, dt));
  if (x.x <= lo.x) { x.x = lo.x; v.x = 0; }
  if (x.y <= lo.y) { x.y = lo.y; v.y = 0; }
  if (x.z <= lo.z) { x.z = lo.z; v.z = 0; }
  if (x.x >= hi.x) { x.x = hi.x; v.x = 0; }
  if (x.y >= hi.y) { x.y = hi.y; v.y = 0; }
  if (x.z >= hi.z) { x.z = hi.z; v.z = 0; }
  pos[p] = x;
  vel[p] = v;
}

// mapHeatg2p / mapHeatg2p_MeltAndBoil spray_k.cu:4655-4679 + updateLatentHeat_k 5546-5573
__global__ void k_heat_g2p(Grid g, int64_t P, const f3* __restrict__ pos, const int8_t* __restrict__ flag,
                           float* __restrict__ temp, float* __restrict__ lheat, const float* __restrict__ tchange,
                           const float* __restrict__ tgrid, float at, int melt_and_boil, float mp, float bp,
                           float th) {
  int64_t p = xcd_block() * blockDim.x + threadIdx.x;
  if (p >= P) return;
  f3 x = pos[p];
  int8_t fl = flag[p];
  float t = temp[p];
  float nt = at * (t + scalar_from_grid(g, x, tchange)) + (1.0f - at) * scalar_from_grid(g, x, tgrid);
  if (melt_and_boil && fl == SOLID) t = 0.95f * t + 0.05f * nt;
  else t = nt;
  float lh = lheat[p];
  if (fl == SOLID && t > mp) { lh += t - mp; t = mp; }
  if (fl == FLUID) {
    if (t < mp) { lh -= mp - t; t = mp; }
    else if (t > bp) { lh += t - bp; t = bp; }
    else lh = th;
  }
  temp[p] = t;
  lheat[p] = lh;
}

// ---- phase change (a17) ----
// SplitMix64 finalizer (Steele, Lea, Flood, OOPSLA 2014) as a counter hash.
__device__ __forceinline__ uint64_t mix64(uint64_t z) {
  z += 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}
__device__ __forceinline__ float uniform24(uint64_t seed, uint64_t a, uint64_t b, uint64_t c) {
  uint64_t h = mix64(seed ^ mix64(a ^ mix64(b ^ mix64(c))));
  return (float)(h >> 40) * (1.0f / 16777216.0f);
}

// updatesolubility, spray_k.cu:4719-4727
__global__ void k_solubility(int64_t P, const int8_t* __restrict__ flag, const float* __restrict__ temp,
                             float* __restrict__ solub, float rate_m0, float T0) {
  int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p < P && flag[p] == FLUID) solub[p] = rate_m0 * expf(1018.9f * (1.0f / temp[p] - 1.0f / T0));
}

struct GasArgs {
  const f3* pos;
  f3* vel_w;
  f3* pos_w;
  int8_t* flag;
  float* mass;
  float* temp;
  float* lheat;
  float* solub;
  float* gas;
  const uint32_t* cstart;
  const uint32_t* cend;
  const int32_t* seeds;
  int nseeds;
  const float* tgrid;
  float m0, th, vr;
  uint64_t seed;
  int frame;
};

__device__ __forceinline__ bool in_boundary_cell(const Grid& g, int x, int y) {  // spray_k.cu:4214-4221
  const int level = 2;
  return x <= level || x >= g.nx - 1 - level || y <= level || y >= g.ny - 1 - level;
}

// GenerateGasParticle_k spray_k.cu:4730-4817: decision of one cell; returns
// true when it emits and fills the new particle position.
__device__ bool gas_decide(const Grid& g, const GasArgs& a, int64_t idx, f3& npos) {
  int i, j, k;
  ijk_of(g, idx, i, j, k);
  if (k <= 1 || in_boundary_cell(g, i, j)) return false;
  uint32_t s = a.cstart[idx];
  if (s == MPF_CELL_UNDEF) return false;
  float gc = 0, gs = 0, gair = 0, vap = 0;
  int gasc = 0;
  f3 gp = mk3(0.f, 0.f, 0.f);
  for (uint32_t p = s; p < a.cend[idx]; ++p) {
    int8_t fl = a.flag[p];
    if (fl == FLUID) {
      gc += a.gas[p];
      gs += a.solub[p];
      vap += fmaxf(0.0f, a.lheat[p] - a.th) * a.vr * a.m0;
    } else if (fl == AIRSOLO || fl == AIR) {
      gair += a.mass[p];
      gp = add3(gp, a.pos[p]);
      gasc++;
    }
  }
  bool seed = false;
  for (int q = 0; q < a.nseeds; ++q)
    if (a.seeds[q] == (int32_t)idx) seed = true;
  if (!(gc - gs + vap > a.m0 && (seed || gasc > 0))) return false;
  f3 rp = mk3(uniform24(a.seed, (uint64_t)a.frame, (uint64_t)idx, 0), uniform24(a.seed, (uint64_t)a.frame, (uint64_t)idx, 1),
              uniform24(a.seed, (uint64_t)a.frame, (uint64_t)idx, 2));
  float rn = uniform24(a.seed, (uint64_t)a.frame, (uint64_t)idx, 3);
  float r = g.h * 0.25f;
  if (gasc > 0) {
    f3 cen = mk3(gp.x / (float)gasc, gp.y / (float)gasc, gp.z / (float)gasc);
    npos = add3(cen, mul3(mul3(sub3(rp, mk3(0.5f, 0.5f, 0.5f)), fmaxf(0.5f, rn) * r), 2.0f));
  } else {
    npos = add3(add3(mk3((float)i * g.h, (float)j * g.h, (float)k * g.h), mk3(g.gx0, g.gy0, g.gz0)), mul3(rp, g.h));
  }
  return true;
}

__global__ void k_gas_count(Grid g, GasArgs a, uint32_t* __restrict__ add) {
  int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= g.G) return;
  f3 np;
  add[idx] = gas_decide(g, a, idx, np) ? 1u : 0u;
}

__global__ void k_gas_emit(Grid g, GasArgs a, const uint32_t* __restrict__ add, const uint32_t* __restrict__ slot,
                           int64_t P) {
  int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= g.G || !add[idx]) return;
  f3 np;
  gas_decide(g, a, idx, np);
  int64_t q = P + slot[idx];
  a.pos_w[q] = np;
  a.vel_w[q] = mk3(0.f, 0.f, 0.f);
  a.mass[q] = a.m0;
  a.flag[q] = AIRSOLO;
  a.temp[q] = a.tgrid[idx];
  a.lheat[q] = 0;
  a.gas[q] = 0;
  a.solub[q] = 0;
  for (uint32_t p = a.cstart[idx]; p < a.cend[idx]; ++p)
    if (a.flag[p] == FLUID) {
      a.gas[p] = fminf(a.gas[p], a.solub[p]);
      a.lheat[p] = fminf(a.lheat[p], a.th);
    }
}

// MeltingSolidByHeat, spray_k.cu:5059-5072
__global__ void k_melt(int64_t P, int8_t* __restrict__ flag, float* __restrict__ lheat, float th,
                       uint32_t* __restrict__ counter) {
  int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  bool ch = false;
  if (p < P && flag[p] == SOLID && lheat[p] > th) {
    flag[p] = FLUID;
    lheat[p] = th;
    ch = true;
  }
  unsigned long long m = __ballot(ch);
  if (m && __lane_id() == (unsigned)(__ffsll((long long)m) - 1)) atomicAdd(counter, (uint32_t)__popcll(m));
}

// FreezingSolidByHeat, spray_k.cu:5074-5124 (reads the pass's input f0/p0)
__global__ void k_freeze(Grid g, int64_t Pn, const int8_t* __restrict__ f0, const f3* __restrict__ p0,
                         int8_t* __restrict__ flag, f3* __restrict__ pos, float* __restrict__ lheat,
                         const uint32_t* __restrict__ cstart, const uint32_t* __restrict__ cend, float ss,
                         uint32_t* __restrict__ counter, uint32_t* __restrict__ moved) {
  int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  bool ch = false;
  if (p < Pn && f0[p] == FLUID && lheat[p] < 0) {
    f3 x = p0[p];
    float qx = (x.x - g.gx0) / g.h, qy = (x.y - g.gy0) / g.h, qz = (x.z - g.gz0) / g.h;
    int i = (qx >= 0 && qx < (float)g.nx) ? (int)qx : 0;
    int j = (qy >= 0 && qy < (float)g.ny) ? (int)qy : 0;
    int k = (qz >= 0 && qz < (float)g.nz) ? (int)qz : 0;
    float mind = 1000.0f;
    int64_t mi = -1;
    int cs = 0;
    for (int di = -1; di <= 1; ++di)
      for (int dj = -1; dj <= 1; ++dj)
        for (int dk = -1; dk <= 1; ++dk) {
          if (!valid(g, i + di, j + dj, k + dk)) continue;
          int64_t c = cid(g, i + di, j + dj, k + dk);
          uint32_t s = cstart[c];
          if (s == MPF_CELL_UNDEF) continue;
          for (uint32_t q = s; q < cend[c]; ++q) {
            if (f0[q] != SOLID) continue;
            float d = len3(sub3(p0[q], x));
            if (d < g.h) cs++;
            if (d < mind) { mind = d; mi = q; }
          }
        }
    if (mi != -1 && mind < g.h && cs > 2) {
      flag[p] = SOLID;
      lheat[p] = 0;
      ch = true;
      if (mind > ss) {
        f3 y = p0[mi];
        f3 dd = sub3(x, y);
        float l = len3(dd);
        const f3 xn = add3(mul3(mul3(dd, 1.0f / l), ss), y);
        pos[p] = xn;
        // the snapped particle left its cell: the cell lists no longer hold
        // (same cell rule as the hash key)
        const float rx = (xn.x - g.gx0) / g.h, ry = (xn.y - g.gy0) / g.h, rz = (xn.z - g.gz0) / g.h;
        const int ni = (rx >= 0 && rx < (float)g.nx) ? (int)rx : 0;
        const int nj = (ry >= 0 && ry < (float)g.ny) ? (int)ry : 0;
        const int nk = (rz >= 0 && rz < (float)g.nz) ? (int)rz : 0;
        if (ni != i || nj != j || nk != k) atomicOr(moved, 1u);
      }
    }
  }
  unsigned long long m = __ballot(ch);
  if (m && __lane_id() == (unsigned)(__ffsll((long long)m) - 1)) atomicAdd(counter, (uint32_t)__popcll(m));
}


// ---- particle deletion and pouring (f1) ----
// sweepVacuum, spray_k.cu:4117-4132, as a two-buffer pass: an AIR cell with a
// VACUUM cell at one of its eight corner-diagonal neighbours becomes VACUUM.
__global__ void k_sweep_vacuum(Grid g, const int8_t* __restrict__ in, int8_t* __restrict__ out) {
  int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= g.G) return;
  int8_t m = in[idx];
  if (m == AIR) {
    int i, j, k;
    ijk_of(g, idx, i, j, k);
    for (int di = -1; di <= 1; di += 2)
      for (int dj = -1; dj <= 1; dj += 2)
        for (int dk = -1; dk <= 1; dk += 2)
          if (valid(g, i + di, j + dj, k + dk) && in[cid(g, i + di, j + dj, k + dk)] == VACUUM) m = VACUUM;
  }
  out[idx] = m;
}

// markDeleteAirParticle, spray_k.cu:4134-4169: keep[p] = 1 to preserve.
// SOLID stays; FLUID goes when the air level set at it is below -h; AIR and
// AIRSOLO go when a cell of their 3x3x3 neighbourhood is VACUUM or their mass
// is at most 1e-6.
__global__ void k_delete_mark(Grid g, int64_t P, const f3* __restrict__ pos, const int8_t* __restrict__ flag,
                              const float* __restrict__ mass, const int8_t* __restrict__ mark,
                              const float* __restrict__ lsair, uint32_t* __restrict__ keep,
                              uint32_t* __restrict__ fluid_deleted) {
  int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  bool fdel = false;
  if (p < P) {
    const int8_t fl = flag[p];
    uint32_t kp = 1;
    if (fl != SOLID) {
      const f3 x = pos[p];
      if (fl == FLUID) {
        const float lsa = scalar_from_grid(g, x, lsair);
        if (lsa < -1.0f * g.h) {
          kp = 0;
          fdel = true;
        }
      } else {
        const float qx = (x.x - g.gx0) / g.h, qy = (x.y - g.gy0) / g.h, qz = (x.z - g.gz0) / g.h;
        const int i = (qx >= 0 && qx < (float)g.nx) ? (int)qx : 0;
        const int j = (qy >= 0 && qy < (float)g.ny) ? (int)qy : 0;
        const int k = (qz >= 0 && qz < (float)g.nz) ? (int)qz : 0;
        int cnt = 0;
        for (int di = -1; di <= 1; ++di)
          for (int dj = -1; dj <= 1; ++dj)
            for (int dk = -1; dk <= 1; ++dk)
              if (valid(g, i + di, j + dj, k + dk) && mark[cid(g, i + di, j + dj, k + dk)] == VACUUM) cnt++;
        kp = (cnt == 0 && mass[p] > 0.000001f) ? 1u : 0u;
      }
    }
    keep[p] = kp;
  }
  unsigned long long m = __ballot(fdel);
  if (m && __lane_id() == (unsigned)(__ffsll((long long)m) - 1)) atomicAdd(fluid_deleted, (uint32_t)__popcll(m));
}

struct PartPtrs {
  f3 *pos, *vel;
  int8_t* flag;
  float *mass, *temp, *lheat, *solub, *gas;
};

// deleteparticles, spray_k.cu:4172-4194: kept particles to their scan slot
// in the other buffer (order preserved).  off: exclusive scan of keep.
__global__ void k_compact(int64_t P, const uint32_t* __restrict__ off, uint32_t total, PartPtrs in, PartPtrs out) {
  int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= P) return;
  const uint32_t o = off[p];
  const uint32_t nxt = (p + 1 < P) ? off[p + 1] : total;
  if (nxt == o) return;
  out.pos[o] = in.pos[p];
  out.vel[o] = in.vel[p];
  out.flag[o] = in.flag[p];
  out.mass[o] = in.mass[p];
  out.temp[o] = in.temp[p];
  out.lheat[o] = in.lheat[p];
  out.solub[o] = in.solub[p];
  out.gas[o] = in.gas[p];
}

// Cell lists after the compaction: a cell's range [s, e) maps to
// [off[s], off[e]) (order is preserved), empty when nothing of it was kept.
__global__ void k_cells_remap(int64_t G, int64_t P, uint32_t total, const uint32_t* __restrict__ off,
                              uint32_t* __restrict__ cstart, uint32_t* __restrict__ cend) {
  int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= G) return;
  const uint32_t s = cstart[c];
  if (s == MPF_CELL_UNDEF) return;
  const uint32_t e = cend[c];
  const uint32_t ns = off[s], ne = ((int64_t)e < P) ? off[e] : total;
  cstart[c] = (ne > ns) ? ns : MPF_CELL_UNDEF;
  cend[c] = (ne > ns) ? ne : MPF_CELL_UNDEF;
}

// pouringwater, spray_k.cu:5575-5596: n particles appended at the source
// positions with a random offset (pos_rand * samplespace per component, in
// [-1, 1)) and a random velocity perturbation; random numbers from the
// counter hash of (seed, stream tag, source index, component), the same at
// every pour (the reference passes frame 0, spray.cu:3020).
constexpr uint64_t POUR_TAG = 0x706f7572ull << 32;
__global__ void k_pour(int64_t P, int64_t n, const f3* __restrict__ spos, const f3* __restrict__ svel, int8_t fl,
                       PartPtrs out, float pos_rand_ss, float vel_rand, float m0, float T, float th, uint64_t seed) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const f3 rv = mk3(uniform24(seed, POUR_TAG, (uint64_t)i, 0) * 2.0f - 1.0f,
                    uniform24(seed, POUR_TAG, (uint64_t)i, 1) * 2.0f - 1.0f,
                    uniform24(seed, POUR_TAG, (uint64_t)i, 2) * 2.0f - 1.0f);
  const f3 rp = mk3(uniform24(seed, POUR_TAG, (uint64_t)i, 3) * 2.0f - 1.0f,
                    uniform24(seed, POUR_TAG, (uint64_t)i, 4) * 2.0f - 1.0f,
                    uniform24(seed, POUR_TAG, (uint64_t)i, 5) * 2.0f - 1.0f);
  const int64_t q = P + i;
  out.pos[q] = add3(spos[i], mul3(rp, pos_rand_ss));
  out.vel[q] = add3(svel[i], mul3(rv, vel_rand));
  out.mass[q] = m0;
  out.flag[q] = fl;
  out.temp[q] = T;
  out.lheat[q] = th;
  out.gas[q] = 0.0f;
  out.solub[q] = 0.0f;
}

}  // namespace

void run_apply_forces(Ctx& c) {
  const mpf_config& k = c.cfg;
  launch(c, "forces", dim3(nblocks(c.P, 256)), dim3(256), 0, k_forces, c.P, c.vel[c.cur], (const int8_t*)c.flag[c.cur],
         k.dt, f3{k.gravity[0], k.gravity[1], k.gravity[2]}, k.buoyancy_kick, k.bubble_max_vel,
         k.buoyancy_rate_air, k.buoyancy_rate_solo);
}

void run_g2p(Ctx& c) {
  const int64_t fc[3] = {c.g.FX, c.g.FY, c.g.FZ};
  for (int a = 0; a < 3; ++a) {
    launch(c, "delta_u", dim3(nblocks(fc[a], 256)), dim3(256), 0, k_delta_u, fc[a], (const float*)c.wu[a], c.wuo[a]);
    launch(c, "delta_u", dim3(nblocks(fc[a], 256)), dim3(256), 0, k_delta_u, fc[a], (const float*)c.au[a], c.auo[a]);
  }
  Faces3 wdu{{c.wuo[0], c.wuo[1], c.wuo[2]}}, adu{{c.auo[0], c.auo[1], c.auo[2]}};
  Faces3 wu{{c.wu[0], c.wu[1], c.wu[2]}}, au{{c.au[0], c.au[1], c.au[2]}};
  launch(c, "g2p", dim3(nblocks(c.P, 256)), dim3(256), 0, k_g2p, c.g, c.P, (const f3*)c.pos[c.cur], c.vel[c.cur],
         (const int8_t*)c.flag[c.cur], wdu, adu, wu, au, c.cfg.flip_alpha);
}

void run_advect(Ctx& c) {
  const int cur = c.cur;
  if (c.cfg.fix_solids)
    launch(c, "zero_solid_vel", dim3(nblocks(c.P, 256)), dim3(256), 0, k_zero_solid_vel, c.P, c.vel[cur],
           (const int8_t*)c.flag[cur]);
  if (c.cfg.integrator == MPF_ADVECT_RK2) {
    Faces3 wu{{c.wu[0], c.wu[1], c.wu[2]}}, au{{c.au[0], c.au[1], c.au[2]}};
    launch(c, "advect_rk2", dim3(nblocks(c.P, 256)), dim3(256), 0, k_advect_rk2, c.g, c.P, c.pos[cur], c.vel[cur],
           (const int8_t*)c.flag[cur], wu, au, c.cfg.dt);
  } else if (c.cfg.integrator == MPF_ADVECT_FREEZING) {
    run_solid_phi_and_velocity(c);
    const int nxt = 1 - cur;
    launch(c, "collide_freezing", dim3(nblocks(c.P, 256)), dim3(256), 0, k_collide_freezing, c.g, c.P,
           (const f3*)c.pos[cur], c.pos[nxt], c.vel[cur], (const int8_t*)c.flag[cur], (const float*)c.phisolid,
           (const uint32_t*)c.cstart, (const uint32_t*)c.cend, c.cfg.dt, c.cfg.samplespace);
    std::swap(c.pos[cur], c.pos[nxt]);
  } else {
    run_solid_phi_and_velocity(c);
    Faces3 su{{c.su[0], c.su[1], c.su[2]}};
    launch(c, "collide_move", dim3(nblocks(c.P, 256)), dim3(256), 0, k_collide_move, c.g, c.P, c.pos[cur],
           c.vel[cur], (const int8_t*)c.flag[cur], (const float*)c.phisolid, su, c.cfg.dt, c.cfg.bounce_vel,
           c.cfg.bounce_pos);
  }
  c.sorted = false;
  c.own_sort = false;
  c.mc_sorted = false;
}

void run_heat_g2p(Ctx& c) {
  const mpf_config& k = c.cfg;
  launch(c, "heat_g2p", dim3(nblocks(c.P, 256)), dim3(256), 0, k_heat_g2p, c.g, c.P, (const f3*)c.pos[c.cur],
         (const int8_t*)c.flag[c.cur], c.temp[c.cur], c.lheat[c.cur], (const float*)c.tpchange, (const float*)c.tp,
         k.alpha_temp_trans, k.heat_melt_and_boil, k.melting_point, k.boiling_point, k.liquid_heat_th);
}

// GenerateGasParticle, spray.cu:2653-2683: updatesolubility + nucleation with
// slots from a scan over cells.  Returns the particles appended (-1 on error).
static int64_t run_nucleation(Ctx& c, int frame) {
  const mpf_config& k = c.cfg;
  const int cur = c.cur;
  if (!(k.gas_enabled && k.gas_every > 0 && frame % k.gas_every == 0)) return 0;
  launch(c, "solubility", dim3(nblocks(c.P, 256)), dim3(256), 0, k_solubility, c.P, (const int8_t*)c.flag[cur],
         (const float*)c.temp[cur], c.solub[cur], k.dissolve_gas_rate * k.airm0, k.temperature0);
  GasArgs a;
  a.pos = c.pos[cur]; a.pos_w = c.pos[cur]; a.vel_w = c.vel[cur]; a.flag = c.flag[cur]; a.mass = c.mass[cur];
  a.temp = c.temp[cur]; a.lheat = c.lheat[cur]; a.solub = c.solub[cur]; a.gas = c.gas[cur];
  a.cstart = c.cstart; a.cend = c.cend; a.seeds = c.seeds; a.nseeds = c.nseeds; a.tgrid = c.tp;
  a.m0 = k.airm0; a.th = k.liquid_heat_th; a.vr = k.vapor_gen_rate; a.seed = k.seed; a.frame = frame;
  launch(c, "gas_count", dim3(nblocks(c.g.G, 256)), dim3(256), 0, k_gas_count, c.g, a, c.ctmp_u32);
  const int64_t na = run_scan_u32(c, c.ctmp_u32, c.count, c.g.G, true);
  if (na < 0) return -1;
  if (c.P + na > c.cap) {
    c.err = "particle capacity exceeded by nucleation";
    return -2;
  }
  launch(c, "gas_emit", dim3(nblocks(c.g.G, 256)), dim3(256), 0, k_gas_emit, c.g, a, (const uint32_t*)c.ctmp_u32,
         (const uint32_t*)c.count, c.P);
  c.P += na;
  if (na) c.sorted = c.own_sort = c.mc_sorted = false;
  return na;
}

// MeltSolid / Freezing, spray.cu:2816-2875, on the particles [0, Pn) the cell
// lists cover (counters 1 and 2: melted, frozen; 7: frozen and moved).
static void run_melt_freeze(Ctx& c, int64_t Pn) {
  const mpf_config& k = c.cfg;
  const int cur = c.cur;
  if (k.melt_enabled)
    launch(c, "melt", dim3(nblocks(c.P, 256)), dim3(256), 0, k_melt, c.P, c.flag[cur], c.lheat[cur], k.liquid_heat_th,
           c.counters + 1);
  if (k.freeze_enabled && Pn > 0) {
    const int alt = 1 - cur;
    hipMemcpyAsync(c.flag[alt], c.flag[cur], (size_t)Pn, hipMemcpyDeviceToDevice, c.stream);
    hipMemcpyAsync(c.pos[alt], c.pos[cur], (size_t)Pn * sizeof(f3), hipMemcpyDeviceToDevice, c.stream);
    launch(c, "freeze", dim3(nblocks(Pn, 256)), dim3(256), 0, k_freeze, c.g, Pn, (const int8_t*)c.flag[alt],
           (const f3*)c.pos[alt], c.flag[cur], c.pos[cur], c.lheat[cur], (const uint32_t*)c.cstart,
           (const uint32_t*)c.cend, k.samplespace, c.counters + 2, c.counters + 7);
  }
}

static PartPtrs part_ptrs(Ctx& c, int b) {
  return PartPtrs{c.pos[b], c.vel[b], c.flag[b], c.mass[b], c.temp[b], c.lheat[b], c.solub[b], c.gas[b]};
}

// deleteAirFluidParticle, spray.cu:2194-2248, on the cell flags and air level
// set of the last level-set update.  *pPn: particles the cell lists cover
// (their ranges are remapped through the compaction and *pPn updated), or -1
// when the lists are not kept.  Counter 9: FLUID particles deleted.
static int run_delete(Ctx& c, int64_t* pPn, int64_t* deleted, int64_t* fluid_deleted) {
  const mpf_config& k = c.cfg;
  const int64_t P = c.P;
  const int64_t Pn = *pPn;
  *deleted = 0;
  *fluid_deleted = 0;
  // sweepVacuum passes ping-pong between the flags and the solver's per-cell
  // scratch; an odd count ends with a copy back
  int8_t* m0 = c.mark;
  int8_t* m1 = (int8_t*)c.nbm;
  for (int t = 0; t < k.delete_vacuum_sweeps; ++t) {
    launch(c, "sweep_vacuum", dim3(nblocks(c.g.G, 256)), dim3(256), 0, k_sweep_vacuum, c.g, (const int8_t*)m0, m1);
    std::swap(m0, m1);
  }
  if (m0 != c.mark) hipMemcpyAsync(c.mark, m0, (size_t)c.g.G, hipMemcpyDeviceToDevice, c.stream);
  if (P == 0) return MPF_OK;
  const int cur = c.cur, nxt = 1 - c.cur;
  hipMemsetAsync(c.counters + 9, 0, 4, c.stream);
  launch(c, "delete_mark", dim3(nblocks(P, 256)), dim3(256), 0, k_delete_mark, c.g, P, (const f3*)c.pos[cur],
         (const int8_t*)c.flag[cur], (const float*)c.mass[cur], (const int8_t*)c.mark, (const float*)c.lsair, c.key,
         c.counters + 9);
  // in-place exclusive scan: keep[p] = off[p+1] - off[p]
  const int64_t total = run_scan_u32(c, c.key, c.key, P, true);
  if (total < 0) return MPF_HIP_ERROR;
  hipMemcpyAsync(c.counters_host + 9, c.counters + 9, 4, hipMemcpyDeviceToHost, c.stream);
  if (!hip_ok(c, hipStreamSynchronize(c.stream), "delete")) return MPF_HIP_ERROR;
  *fluid_deleted = c.counters_host[9];
  if (total == P) return MPF_OK;
  *deleted = P - total;
  launch(c, "compact", dim3(nblocks(P, 256)), dim3(256), 0, k_compact, P, (const uint32_t*)c.key, (uint32_t)total,
         part_ptrs(c, cur), part_ptrs(c, nxt));
  if (Pn >= 0) {
    launch(c, "cells_remap", dim3(nblocks(c.g.G, 256)), dim3(256), 0, k_cells_remap, c.g.G, P, (uint32_t)total,
           (const uint32_t*)c.key, c.cstart, c.cend);
    if (Pn < P) {
      hipMemcpyAsync(&c.counters_host[10], c.key + Pn, 4, hipMemcpyDeviceToHost, c.stream);
      if (!hip_ok(c, hipStreamSynchronize(c.stream), "delete")) return MPF_HIP_ERROR;
      *pPn = c.counters_host[10];
    } else {
      *pPn = total;
    }
  }
  c.cur = nxt;
  c.P = total;
  c.own_sort = false;
  c.mc_sorted = false;
  if (Pn >= 0) run_cell_mask(c);
  return MPF_OK;
}

int run_phase_change(Ctx& c, int frame, mpf_phase_info* info) {
  const int64_t Pn = c.P;  // particles covered by the cell lists
  hipMemsetAsync(c.counters + 1, 0, 8, c.stream);
  hipMemsetAsync(c.counters + 7, 0, 4, c.stream);
  const int64_t added = run_nucleation(c, frame);
  if (added == -1) return MPF_HIP_ERROR;
  if (added == -2) return MPF_CAPACITY;
  run_melt_freeze(c, Pn);
  hipMemcpyAsync(c.counters_host + 1, c.counters + 1, 7 * 4, hipMemcpyDeviceToHost, c.stream);
  if (!hip_ok(c, hipStreamSynchronize(c.stream), "phase_change")) return MPF_HIP_ERROR;
  if (c.cfg.freeze_enabled && c.counters_host[7]) c.own_sort = c.mc_sorted = false;  // a frozen particle snapped into another cell
  if (info) {
    info->melted = (int32_t)c.counters_host[1];
    info->frozen = (int32_t)c.counters_host[2];
    info->gas_added = (int32_t)added;
    info->deleted = 0;
    info->particles = c.P;
  }
  return MPF_OK;
}

// The step's order (spray.cpp:1155-1180): nucleation, deletion, melting and
// freezing.  Deletion keeps the cell lists valid for the particles they cover.
int run_phase_change_with_delete(Ctx& c, int frame, mpf_phase_info* info, int64_t* deleted) {
  int64_t Pn = c.P;
  hipMemsetAsync(c.counters + 1, 0, 8, c.stream);
  hipMemsetAsync(c.counters + 7, 0, 4, c.stream);
  const int64_t added = run_nucleation(c, frame);
  if (added == -1) return MPF_HIP_ERROR;
  if (added == -2) return MPF_CAPACITY;
  int64_t fl_del = 0;
  if (c.cfg.delete_enabled) {
    const int st = run_delete(c, &Pn, deleted, &fl_del);
    if (st != MPF_OK) return st;
  } else {
    *deleted = 0;
  }
  run_melt_freeze(c, Pn);
  hipMemcpyAsync(c.counters_host + 1, c.counters + 1, 7 * 4, hipMemcpyDeviceToHost, c.stream);
  if (!hip_ok(c, hipStreamSynchronize(c.stream), "phase_change")) return MPF_HIP_ERROR;
  if (c.cfg.freeze_enabled && c.counters_host[7]) c.own_sort = c.mc_sorted = false;  // a frozen particle snapped into another cell
  if (info) {
    info->melted = (int32_t)c.counters_host[1];
    info->frozen = (int32_t)c.counters_host[2];
    info->gas_added = (int32_t)added;
    info->deleted = (int32_t)*deleted;
    info->particles = c.P;
  }
  return MPF_OK;
}

// mpf_delete_particles: the cell lists are remapped when they index every particle.
int run_delete_particles(Ctx& c, int64_t* deleted, int64_t* fluid_deleted) {
  int64_t Pn = c.sorted ? c.P : -1;
  return run_delete(c, &Pn, deleted, fluid_deleted);
}

// pouring / pouringgas, spray.cu:3012-3043.
int run_pour(Ctx& c, int frame, int64_t* added) {
  const mpf_config& k = c.cfg;
  *added = 0;
  if (c.npour <= 0 || k.pour_every <= 0 || frame % k.pour_every != 0) return MPF_OK;
  if (c.P + c.npour > c.cap) return MPF_OK;  // the reference skips the pour when it would not fit
  launch(c, "pour", dim3(nblocks(c.npour, 256)), dim3(256), 0, k_pour, c.P, c.npour, (const f3*)c.pour_pos,
         (const f3*)c.pour_vel, c.pour_flag, part_ptrs(c, c.cur), k.pour_pos_rand * k.samplespace, k.pour_vel_rand,
         k.m0, k.pour_temperature, k.liquid_heat_th, k.seed);
  c.P += c.npour;
  *added = c.npour;
  c.sorted = c.own_sort = c.mc_sorted = false;
  return MPF_OK;
}

}  // namespace mpf
